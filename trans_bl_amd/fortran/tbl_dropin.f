cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
c     Drop-in replacements of trans_bl's v-Poisson and compact
c     derivative subroutines, running on an MI355X through libtrans_bl_amd.
c
c     Same names and argument lists as the reference (poi_uy,
c     poisson_par.f:6; derparx, derparxx, dery, deryy, deryfv,
c     derivs_par.f:80-177; loop, main_par.f:1762-1785; drv_gv,
c     main_par.f:1339-1403; drv_fi, main_par.f:1264-1336), same hidden
c     inputs: the PARAMETERs of
c     par.for, COMMON /var/, /coefs/, /bd/ (bdfc), /derw/ (d2uydx2),
c     /blas/ (uxb, uyb, wzb) and /vc/ (varg, dvargdx).  Compile this file
c     instead
c     of poisson_par.f's poi_uy and derivs_par.f's derivative routines,
c     with the run's own par.for and comm.* on the include path, and link
c     tbl_iface.o and libtrans_bl_amd.so.  One rank (np = 1).
c
c     The library context is created on the first call, after init_gv has
c     read coefs.bin into COMMON /coefs/ (main_par.f:852-873).
cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      subroutine tbl_setup

      use tbl_iface
      implicit none
      include 'par.for'
      include 'comm.coef'
      type(tbl_grid_t) grid
      type(tbl_coefs_t) coefs
      integer(c_int) rc
      integer l

      if (c_associated(tbl_ctx_handle)) return
      grid%nx = ptsx
      grid%ny = jmax
      grid%nmodes = kfour
      grid%levels = msh
      grid%dx = dx
      grid%dy0 = dy
      grid%stf = stf
      grid%alpha = alpha
      grid%beta = beta
      grid%fac_y = fac_y
      coefs%fp_fd_e = fp_fd_coef_e
      coefs%sp_fd = sp_fd_coef
      coefs%cp_fd = cp_fd_coef
      coefs%pp_fd = pp_fd_coef
      coefs%lp_fd = lp_fd_coef
      coefs%fp_sd = fp_sd_coef
      coefs%sp_sd = sp_sd_coef
      coefs%cp_sd = cp_sd_coef
      coefs%pp_sd = pp_sd_coef
      coefs%lp_sd = lp_sd_coef
      coefs%w_at_w = w_at_w_coef
      coefs%dwydy = dwydy_coef
      coefs%sp_poi = 0.d0
      coefs%cp_poi = 0.d0
      coefs%pp_poi = 0.d0
      coefs%lp_poi = 0.d0
      do l = 1, msh
        coefs%sp_poi(:,l) = sp_poi_coef(:,l)
        coefs%cp_poi(:,l) = cp_poi_coef(:,l)
        coefs%pp_poi(:,l) = pp_poi_coef(:,l)
        coefs%lp_poi(:,l) = lp_poi_coef(:,l)
      end do
      rc = tbl_create(grid, coefs, tbl_ctx_handle)
      if (rc .ne. 0) call tbl_fail('tbl_create', rc)

      return
      end

cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      subroutine tbl_poisson_form(name)

      ! my_form 0 and 4 (immersed, LES) give poi_uy's mean mode a top-row
      ! source from the base flow (duexmdx, uyb; poisson_par.f:145-153)
      ! that this layer does not pass: poi_uy and loop refuse them instead
      ! of solving a different top row (the derivatives and drv_fi serve
      ! every form)
      implicit none
      include 'par.for'
      character*(*) name

      if (my_form .eq. 0 .or. my_form .eq. 4) then
        write(*,*) name, ': my_form ', my_form,
     &             ' not supported (only 1 and 2)'
        stop
      end if

      return
      end

cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      subroutine poi_uy(dwzdx, erromax)

      ! v-Poisson FAS multigrid for every mode (poisson_par.f:6-64):
      ! reads wx and the initial guess uy from COMMON /var/, writes uy on
      ! i = 2..ptsx-1, j = 2..jmax and prints (k, V-cycles, residual)
      use tbl_iface
      implicit none
      include 'par.for'
      include 'comm.var'
      integer k
      real*8 erromax
      complex*16 dwzdx(ptsx,jmax,kfour)
      integer(c_int) rc, cycles(kfour)
      real(c_double) residual(kfour)
      type(tbl_solve_stats_t) stats

      call tbl_poisson_form('poi_uy')
      call tbl_setup
      rc = tbl_poi_uy(tbl_ctx_handle, wx, dwzdx, uy, erromax, tbl_opts,
     &                cycles, residual, stats)
      if (rc .ne. 0) call tbl_fail('poi_uy', rc)
      do k = 1, kfour
        write(*,*) k, cycles(k), residual(k)
      end do

      return
      end

cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      subroutine derparx(ddx, fc)

      ! first derivative in x (derivs_par.f:80-95)
      use tbl_iface
      implicit none
      include 'par.for'
      complex*16 fc(ptsx,jmax,kfour), ddx(ptsx,jmax,kfour)
      integer(c_int) rc

      call tbl_setup
      rc = tbl_derparx(tbl_ctx_handle, fc, ddx)
      if (rc .ne. 0) call tbl_fail('derparx', rc)

      return
      end

cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      subroutine derparxx(d2dx2, fc)

      ! second derivative in x (derivs_par.f:98-113)
      use tbl_iface
      implicit none
      include 'par.for'
      complex*16 fc(ptsx,jmax,kfour), d2dx2(ptsx,jmax,kfour)
      integer(c_int) rc

      call tbl_setup
      rc = tbl_derparxx(tbl_ctx_handle, fc, d2dx2)
      if (rc .ne. 0) call tbl_fail('derparxx', rc)

      return
      end

cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      subroutine dery(ddy, fc)

      ! first derivative in y (derivs_par.f:116-129)
      use tbl_iface
      implicit none
      include 'par.for'
      complex*16 fc(ptsx,jmax,kfour), ddy(ptsx,jmax,kfour)
      integer(c_int) rc

      call tbl_setup
      rc = tbl_dery(tbl_ctx_handle, fc, ddy)
      if (rc .ne. 0) call tbl_fail('dery', rc)

      return
      end

cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      subroutine deryy(d2dy2, fc)

      ! second derivative in y (derivs_par.f:164-177)
      use tbl_iface
      implicit none
      include 'par.for'
      complex*16 fc(ptsx,jmax,kfour), d2dy2(ptsx,jmax,kfour)
      integer(c_int) rc

      call tbl_setup
      rc = tbl_deryy(tbl_ctx_handle, fc, d2dy2)
      if (rc .ne. 0) call tbl_fail('deryy', rc)

      return
      end

cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      subroutine deryfv(ddy, fc)

      ! first derivative in y with dv/dy = 0 at the wall (derivs_par.f:148-161)
      use tbl_iface
      implicit none
      include 'par.for'
      complex*16 fc(ptsx,jmax,kfour), ddy(ptsx,jmax,kfour)
      integer(c_int) rc

      call tbl_setup
      rc = tbl_deryfv(tbl_ctx_handle, fc, ddy)
      if (rc .ne. 0) call tbl_fail('deryfv', rc)

      return
      end

cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      subroutine loop(erro)

      ! the RK-stage update (main_par.f:1762-1785): filter_trid when
      ! erro < 1e-5, bzone, outuy, poi_uy (to erro), poi_ux, poi_uz,
      ! wx_wall, wz_wall, bzonew -- all on the GPU in one call; prints
      ! poi_uy's (k, V-cycles, residual) lines.  th (my_form 2) is
      ! handed over only for that form.
      use tbl_iface
      implicit none
      include 'par.for'
      include 'comm.var'
      real*8 erro, bdfc(2,imax)
      complex*16 d2uydx2(ptsx,kfour)
      common/bd/ bdfc
      common/derw/ d2uydx2
      integer k
      integer(c_int) rc, cycles(kfour)
      real(c_double) residual(kfour)
      type(tbl_solve_stats_t) stats

      call tbl_poisson_form('loop')
      call tbl_setup
      rc = tbl_loop(tbl_ctx_handle, erro, bdfc, d2uydx2, ux, uy, uz,
     &              wx, wy, wz, tbl_th_ptr(my_form .eq. 2, th),
     &              tbl_opts, cycles, residual, stats)
      if (rc .ne. 0) call tbl_fail('loop', rc)
      do k = 1, kfour
        write(*,*) k, cycles(k), residual(k)
      end do

      return
      end

cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      subroutine drv_gv(dvx, dvy, dvz)

      ! the RK right-hand side (main_par.f:1339-1403) with lterms_gv, the
      ! reference's active call: dvx, dvy, dvz on rows 2..jmax from
      ! COMMON /var/, the base flow /blas/ and curvature term /vc/
      use tbl_iface
      implicit none
      include 'par.for'
      include 'comm.var'
      real*8 uxb(ptsx,jmax), uyb(ptsx,jmax), wzb(ptsx,jmax)
      real*8 varg(ptsx,jmax), dvargdx(ptsx,jmax)
      complex*16 dvx(ptsx,jmax,kfour), dvy(ptsx,jmax,kfour),
     &           dvz(ptsx,jmax,kfour)
      common/blas/ uxb, uyb, wzb
      common/vc/ varg, dvargdx
      integer(c_int) rc

      call tbl_setup
      rc = tbl_drv_gv(tbl_ctx_handle, Re, 0, uxb, uyb, wzb, varg,
     &                dvargdx, ux, uy, uz, wx, wy, wz, dvx, dvy, dvz)
      if (rc .ne. 0) call tbl_fail('drv_gv', rc)

      return
      end

cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      subroutine drv_fi(dvx, dvy, dvz)

      ! the RK right-hand side of the flat-plate form (main_par.f:
      ! 1264-1336) with nlterms_fi, the reference's active call (kphys
      ! points in z): dvx, dvy, dvz on rows 2..jmax from COMMON /var/
      use tbl_iface
      implicit none
      include 'par.for'
      include 'comm.var'
      real*8 uxb(ptsx,jmax), uyb(ptsx,jmax), wzb(ptsx,jmax)
      complex*16 dvx(ptsx,jmax,kfour), dvy(ptsx,jmax,kfour),
     &           dvz(ptsx,jmax,kfour)
      common/blas/ uxb, uyb, wzb
      integer(c_int) rc

      call tbl_setup
      rc = tbl_drv_fi(tbl_ctx_handle, Re, kphys, uxb, uyb, wzb,
     &                ux, uy, uz, wx, wy, wz, dvx, dvy, dvz)
      if (rc .ne. 0) call tbl_fail('drv_fi', rc)

      return
      end
