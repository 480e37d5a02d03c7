c     Fortran RK-stage driver for the drop-in test.  Reads the coefficient
c     tables and the stage inputs from stage_in.bin (stream access, written
c     by tests/test_gpu_fortran.py), runs one stage of the hot path through
c     the drop-in subroutines -- four derivatives of uy, then poi_uy -- and
c     writes the results to stage_out.bin.
      program rk_stage

      implicit none
      include 'par.for'
      include 'comm.var'
      include 'comm.coef'
      complex*16 dwzdx(ptsx,jmax,kfour)
      complex*16 d1(ptsx,jmax,kfour), d2(ptsx,jmax,kfour)
      complex*16 d3(ptsx,jmax,kfour), d4(ptsx,jmax,kfour)
      real*8 erromax

      open(1, file='stage_in.bin', access='stream', form='unformatted',
     &     status='old')
      read(1) fp_fd_coef_e, sp_fd_coef, cp_fd_coef, pp_fd_coef,
     &        lp_fd_coef, fp_sd_coef, sp_sd_coef, cp_sd_coef,
     &        pp_sd_coef, lp_sd_coef, sp_poi_coef, cp_poi_coef,
     &        pp_poi_coef, lp_poi_coef
      read(1) erromax
      read(1) wx, dwzdx, uy
      close(1)

      call derparx(d1, uy)
      call derparxx(d2, uy)
      call dery(d3, uy)
      call deryy(d4, uy)
      call poi_uy(dwzdx, erromax)

      open(2, file='stage_out.bin', access='stream', form='unformatted',
     &     status='replace')
      write(2) d1, d2, d3, d4, uy
      close(2)

      end
