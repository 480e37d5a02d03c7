c     Fortran driver of the drop-in `loop` for the drop-in test.  Reads the
c     coefficient tables (with the wall tables), erro, the fields ux, uy,
c     uz, wx, wy, wz, d2uydx2 (COMMON /derw/) and bdfc (COMMON /bd/) from
c     loop_in.bin (stream access, written by tests/test_gpu_fortran.py),
c     calls loop(erro) and writes the six fields to loop_out.bin.
      program rk_loop

      implicit none
      include 'par.for'
      include 'comm.var'
      include 'comm.coef'
      real*8 erro, bdfc(2,imax)
      complex*16 d2uydx2(ptsx,kfour)
      common/bd/ bdfc
      common/derw/ d2uydx2

      open(1, file='loop_in.bin', access='stream', form='unformatted',
     &     status='old')
      read(1) fp_fd_coef_e, sp_fd_coef, cp_fd_coef, pp_fd_coef,
     &        lp_fd_coef, fp_sd_coef, sp_sd_coef, cp_sd_coef,
     &        pp_sd_coef, lp_sd_coef, sp_poi_coef, cp_poi_coef,
     &        pp_poi_coef, lp_poi_coef, w_at_w_coef, dwydy_coef
      read(1) erro
      read(1) ux, uy, uz, wx, wy, wz, d2uydx2, bdfc
      close(1)

      call loop(erro)

      open(2, file='loop_out.bin', access='stream', form='unformatted',
     &     status='replace')
      write(2) ux, uy, uz, wx, wy, wz
      close(2)

      end
