cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
c     Driver of the reference's right-hand side and RK4 step (test
c     infrastructure, built by oracle/reference.py into oracle/_ref/).
c
c     Linked with the reference objects compiled from their sources
c     (drv_gv main_par.f:1339-1403, drv_fi 1264-1336, lterms_gv /
c     nlterms_gv / lterms_fi / nlterms_fi with f_to_p, p_to_f, four1
c     nlterms_par.f, gv_pert main_par.f:1577-1610, loop and everything it
c     calls) and a par.for configured for one workload.
c     Set-up as init_gv does it for these routines at one rank (see
c     ref_loop.f); the RK4 update loops are the four stages of the
c     Goertler case, main_par.f:270-341, written out here around the
c     reference's drv_gv, gv_pert and loop.
c
c     usage: ref_rhs COEFS IN OUT WHAT [T]
c       IN   stream file: ux, uy, uz, wx, wy, wz (complex*16
c            (ptsx,jmax,kfour)), uxb, uyb, wzb, varg, dvargdx (real*8
c            (ptsx,jmax)), d2uydx2 (complex*16 (ptsx,kfour)), bdfc
c            (real*8 (2,imax)), fcx2, fcllx2 (real*8 (i2)), dt (real*8)
c       WHAT drv     = drv_gv (dvx, dvy, dvz out)
c            nlterms = nlterms_gv (a, b, c, d out)
c            drvfi   = drv_fi (dvx, dvy, dvz out)
c            nltfi   = nlterms_fi (a, b, c out)
c            ltfi    = lterms_fi (a, b, c out)
c            rk4     = one time step t = T (ux .. wz and d2uydx2 out)
c     stdout: a line 'params Re i1 i2 stpp', then poi_uy's lines.
cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      program ref_rhs

      implicit none
      include 'par.for'
      include 'comm.par'
      include 'comm.var'
      include 'comm.coef'
      include 'mpif.h'
      integer p, i, j, k, t, i_ini
      real*8 dt_run, dt2, dt6, bdfc(2,imax), afil(ptsx), bfil(ptsx),
     &       cfil(ptsx), uxb(ptsx,jmax), uyb(ptsx,jmax),
     &       wzb(ptsx,jmax), varg(ptsx,jmax), dvargdx(ptsx,jmax),
     &       fcx2(i2), fcllx2(i2)
      complex*16 d2uydx2(ptsx,kfour)
      complex*16 dv1x(ptsx,jmax,kfour), dv2x(ptsx,jmax,kfour),
     &           dv1y(ptsx,jmax,kfour), dv2y(ptsx,jmax,kfour),
     &           dv1z(ptsx,jmax,kfour), dv2z(ptsx,jmax,kfour),
     &            wx1(ptsx,jmax,kfour),  wy1(ptsx,jmax,kfour),
     &            wz1(ptsx,jmax,kfour),    e4(ptsx,jmax,kfour)
      character*1024 fcoef, fin, fout, arg, what
      common/bd/ bdfc
      common/derw/ d2uydx2
      common/filt/ afil, bfil, cfil
      common/blas/ uxb, uyb, wzb
      common/vc/ varg, dvargdx
      common/equa2/ fcx2, fcllx2

      call MPI_Init(ierr)
      call MPI_Comm_rank(MPI_COMM_WORLD, my_rank, ierr)
      call MPI_Comm_size(MPI_COMM_WORLD, p, ierr)
      numproc = p - 1
      inter = 2**(msh - 1) * (stencil - 2)
      shift = my_rank * (ptsx - inter - 1)

      call get_command_argument(1, fcoef)
      call get_command_argument(2, fin)
      call get_command_argument(3, fout)
      call get_command_argument(4, what)
      t = 1
      if (command_argument_count() .ge. 5) then
        call get_command_argument(5, arg)
        read(arg, *) t
      end if
      write(*, '(a, es24.16, 3i8)') ' params ', Re, i1, i2, stpp

      open(1, file=trim(fcoef), form='unformatted', status='old')
      read(1) fp_fd_coef_e
      read(1) sp_fd_coef
      read(1) cp_fd_coef
      read(1) pp_fd_coef
      read(1) lp_fd_coef
      read(1) fp_sd_coef
      read(1) sp_sd_coef
      read(1) cp_sd_coef
      read(1) pp_sd_coef
      read(1) lp_sd_coef
      read(1) sp_poi_coef
      read(1) cp_poi_coef
      read(1) pp_poi_coef
      read(1) lp_poi_coef
      read(1) w_at_w_coef
      read(1) dwydy_coef
      close(unit=1)

      call create_ctes
      call derivs_k
      call lhs_tridf(afil, bfil, cfil)

      open(2, file=trim(fin), access='stream', form='unformatted',
     &     status='old')
      read(2) ux, uy, uz, wx, wy, wz
      read(2) uxb, uyb, wzb, varg, dvargdx
      read(2) d2uydx2, bdfc, fcx2, fcllx2, dt_run
      close(unit=2)

      open(3, file=trim(fout), access='stream', form='unformatted',
     &     status='replace')
      if (trim(what) .eq. 'drv') then
        call drv_gv(dv1x, dv1y, dv1z)
        write(3) dv1x, dv1y, dv1z
      else if (trim(what) .eq. 'nlterms') then
        call nlterms_gv(dv1x, dv1y, dv1z, e4)
        write(3) dv1x, dv1y, dv1z, e4
      else if (trim(what) .eq. 'drvfi') then
        call drv_fi(dv1x, dv1y, dv1z)
        write(3) dv1x, dv1y, dv1z
      else if (trim(what) .eq. 'nltfi') then
        call nlterms_fi(dv1x, dv1y, dv1z)
        write(3) dv1x, dv1y, dv1z
      else if (trim(what) .eq. 'ltfi') then
        call lterms_fi(dv1x, dv1y, dv1z)
        write(3) dv1x, dv1y, dv1z
      else if (trim(what) .eq. 'rk4') then
        i_ini = 3
        dt2 = 0.5d0 * dt_run
        dt6 = dt_run / 6.d0
        call drv_gv(dv1x, dv1y, dv1z)
        do k = 1, kfour
          do j = 2, jmax - 1
            do i = i_ini, ptsx
              wx1(i,j,k) = wx(i,j,k)
              wy1(i,j,k) = wy(i,j,k)
              wz1(i,j,k) = wz(i,j,k)
              wx(i,j,k)  = wx1(i,j,k) + dv1x(i,j,k) * dt2
              wy(i,j,k)  = wy1(i,j,k) + dv1y(i,j,k) * dt2
              wz(i,j,k)  = wz1(i,j,k) + dv1z(i,j,k) * dt2
            end do
          end do
        end do
        call gv_pert(t)
        call loop(1d-5)
        call drv_gv(dv2x, dv2y, dv2z)
        do k = 1, kfour
          do j = 2, jmax - 1
            do i = i_ini, ptsx
              wx(i,j,k)   = wx1(i,j,k) + dv2x(i,j,k) * dt2
              wy(i,j,k)   = wy1(i,j,k) + dv2y(i,j,k) * dt2
              wz(i,j,k)   = wz1(i,j,k) + dv2z(i,j,k) * dt2
              dv1x(i,j,k) = dv1x(i,j,k) + 2.d0 * dv2x(i,j,k)
              dv1y(i,j,k) = dv1y(i,j,k) + 2.d0 * dv2y(i,j,k)
              dv1z(i,j,k) = dv1z(i,j,k) + 2.d0 * dv2z(i,j,k)
            end do
          end do
        end do
        call loop(1d-5)
        call drv_gv(dv2x, dv2y, dv2z)
        do k = 1, kfour
          do j = 2, jmax - 1
            do i = i_ini, ptsx
              wx(i,j,k)   = wx1(i,j,k) + dv2x(i,j,k) * dt_run
              wy(i,j,k)   = wy1(i,j,k) + dv2y(i,j,k) * dt_run
              wz(i,j,k)   = wz1(i,j,k) + dv2z(i,j,k) * dt_run
              dv1x(i,j,k) = dv1x(i,j,k) + 2.d0 * dv2x(i,j,k)
              dv1y(i,j,k) = dv1y(i,j,k) + 2.d0 * dv2y(i,j,k)
              dv1z(i,j,k) = dv1z(i,j,k) + 2.d0 * dv2z(i,j,k)
            end do
          end do
        end do
        call gv_pert(t)
        call loop(1d-5)
        call drv_gv(dv2x, dv2y, dv2z)
        do k = 1, kfour
          do j = 2, jmax - 1
            do i = i_ini, ptsx
              wx(i,j,k) = wx1(i,j,k)
     &                  + dt6 * ( dv1x(i,j,k) + dv2x(i,j,k) )
              wy(i,j,k) = wy1(i,j,k)
     &                  + dt6 * ( dv1y(i,j,k) + dv2y(i,j,k) )
              wz(i,j,k) = wz1(i,j,k)
     &                  + dt6 * ( dv1z(i,j,k) + dv2z(i,j,k) )
            end do
          end do
        end do
        call loop(1d-6)
        write(3) ux, uy, uz, wx, wy, wz, d2uydx2
      else
        write(*, *) 'unknown routine ', trim(what)
        call MPI_Abort(MPI_COMM_WORLD, 2, ierr)
      end if
      close(unit=3)

      call MPI_Finalize(ierr)
      end
