c     Fortran driver of the drop-in `drv_gv` (or, with the argument fi,
c     `drv_fi`) for the drop-in test.  Reads the coefficient tables, the
c     fields ux .. wz (COMMON /var/), the base flow uxb, uyb, wzb (/blas/)
c     and varg, dvargdx (/vc/) from drv_in.bin (stream access, written by
c     tests/test_gpu_fortran.py), calls drv_gv(dvx, dvy, dvz) or
c     drv_fi(dvx, dvy, dvz) and writes the three results to drv_out.bin.
      program rk_drv

      implicit none
      include 'par.for'
      include 'comm.var'
      include 'comm.coef'
      real*8 uxb(ptsx,jmax), uyb(ptsx,jmax), wzb(ptsx,jmax)
      real*8 varg(ptsx,jmax), dvargdx(ptsx,jmax)
      complex*16 dvx(ptsx,jmax,kfour), dvy(ptsx,jmax,kfour),
     &           dvz(ptsx,jmax,kfour)
      common/blas/ uxb, uyb, wzb
      common/vc/ varg, dvargdx
      character*16 form

      form = 'gv'
      if (command_argument_count() .ge. 1) then
        call get_command_argument(1, form)
      end if

      open(1, file='drv_in.bin', access='stream', form='unformatted',
     &     status='old')
      read(1) fp_fd_coef_e, sp_fd_coef, cp_fd_coef, pp_fd_coef,
     &        lp_fd_coef, fp_sd_coef, sp_sd_coef, cp_sd_coef,
     &        pp_sd_coef, lp_sd_coef, sp_poi_coef, cp_poi_coef,
     &        pp_poi_coef, lp_poi_coef, w_at_w_coef, dwydy_coef
      read(1) ux, uy, uz, wx, wy, wz, uxb, uyb, wzb, varg, dvargdx
      close(1)

      dvx = (0.d0, 0.d0)
      dvy = (0.d0, 0.d0)
      dvz = (0.d0, 0.d0)
      if (trim(form) .eq. 'fi') then
        call drv_fi(dvx, dvy, dvz)
      else
        call drv_gv(dvx, dvy, dvz)
      end if

      open(2, file='drv_out.bin', access='stream', form='unformatted',
     &     status='replace')
      write(2) dvx, dvy, dvz
      close(2)

      end
