! ISO_C_BINDING interface to libtrans_bl_amd (include/trans_bl_amd.h).
!
! The derived types mirror the C structs field by field; complex arrays are
! passed as complex(c_double_complex) assumed-size dummies, which the C side
! receives as interleaved (re, im) double pointers.  The module also holds
! the one library context that the drop-in subroutines of tbl_dropin.f use.
module tbl_iface
  use, intrinsic :: iso_c_binding
  implicit none

  integer, parameter :: TBL_MAX_LEVELS = 8

  type, bind(C) :: tbl_grid_t
    integer(c_int) :: nx, ny, nmodes, levels
    real(c_double) :: dx, dy0, stf, alpha, beta, fac_y
  end type

  ! COMMON /coefs/ (comm.coef:1-29) in coefs.bin index order; the Poisson
  ! tables are (coef, level) as in sp_poi_coef(7, msh)
  type, bind(C) :: tbl_coefs_t
    real(c_double) :: fp_fd_e(6), sp_fd(9), cp_fd(8), pp_fd(9), lp_fd(7)
    real(c_double) :: fp_sd(8), sp_sd(10), cp_sd(8), pp_sd(10), lp_sd(8)
    real(c_double) :: sp_poi(7, TBL_MAX_LEVELS)
    real(c_double) :: cp_poi(8, TBL_MAX_LEVELS)
    real(c_double) :: pp_poi(4, TBL_MAX_LEVELS)
    real(c_double) :: lp_poi(5, TBL_MAX_LEVELS)
    real(c_double) :: w_at_w(9), dwydy(8)
  end type

  type, bind(C) :: tbl_solve_opts_t
    integer(c_int) :: precision = 0, max_cycles = 0, nu1 = 0, nu2 = 0, ncoarse = 0
    integer(c_int) :: use_graph = 0, guess = 0, fmg = 0
    integer(c_int) :: reserved(4) = 0
  end type

  type, bind(C) :: tbl_solve_stats_t
    integer(c_int) :: cycles_max, cycles_total, nconverged, fallback
    real(c_double) :: residual_max, guess_residual_max, device_ms
  end type

  interface
    function tbl_create(grid, coefs, ctx) bind(C, name="tbl_create") result(rc)
      import :: c_int, c_ptr, tbl_grid_t, tbl_coefs_t
      type(tbl_grid_t), intent(in) :: grid
      type(tbl_coefs_t), intent(in) :: coefs
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function

    function tbl_destroy(ctx) bind(C, name="tbl_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function tbl_poi_uy(ctx, wx, dwzdx, uy, erromax, opts, cycles, residual, stats) &
        bind(C, name="tbl_poi_uy") result(rc)
      import :: c_int, c_ptr, c_double, c_double_complex, tbl_solve_opts_t, tbl_solve_stats_t
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: wx(*), dwzdx(*)
      complex(c_double_complex), intent(inout) :: uy(*)
      real(c_double), value :: erromax
      type(tbl_solve_opts_t), intent(in) :: opts
      integer(c_int), intent(out) :: cycles(*)
      real(c_double), intent(out) :: residual(*)
      type(tbl_solve_stats_t), intent(out) :: stats
      integer(c_int) :: rc
    end function

    function tbl_derparx(ctx, fc, out) bind(C, name="tbl_derparx") result(rc)
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: fc(*)
      complex(c_double_complex), intent(out) :: out(*)
      integer(c_int) :: rc
    end function

    function tbl_derparxx(ctx, fc, out) bind(C, name="tbl_derparxx") result(rc)
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: fc(*)
      complex(c_double_complex), intent(out) :: out(*)
      integer(c_int) :: rc
    end function

    function tbl_dery(ctx, fc, out) bind(C, name="tbl_dery") result(rc)
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: fc(*)
      complex(c_double_complex), intent(out) :: out(*)
      integer(c_int) :: rc
    end function

    function tbl_deryy(ctx, fc, out) bind(C, name="tbl_deryy") result(rc)
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: fc(*)
      complex(c_double_complex), intent(out) :: out(*)
      integer(c_int) :: rc
    end function

    function tbl_deryfv(ctx, fc, out) bind(C, name="tbl_deryfv") result(rc)
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: fc(*)
      complex(c_double_complex), intent(out) :: out(*)
      integer(c_int) :: rc
    end function

    function tbl_last_error() bind(C, name="tbl_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function tbl_loop(ctx, erro, bdfc, d2uydx2, ux, uy, uz, wx, wy, wz, th, opts, cycles, &
        residual, stats) bind(C, name="tbl_loop") result(rc)
      import :: c_int, c_ptr, c_double, c_double_complex, tbl_solve_opts_t, tbl_solve_stats_t
      type(c_ptr), value :: ctx
      real(c_double), value :: erro
      real(c_double), intent(in) :: bdfc(*)
      complex(c_double_complex), intent(in) :: d2uydx2(*)
      complex(c_double_complex), intent(inout) :: ux(*), uy(*), uz(*), wx(*), wy(*), wz(*)
      type(c_ptr), value :: th
      type(tbl_solve_opts_t), intent(in) :: opts
      integer(c_int), intent(out) :: cycles(*)
      real(c_double), intent(out) :: residual(*)
      type(tbl_solve_stats_t), intent(out) :: stats
      integer(c_int) :: rc
    end function

    function tbl_drv_gv(ctx, Re, kphys, uxb, uyb, wzb, varg, dvargdx, ux, uy, uz, wx, wy, wz, &
        dvx, dvy, dvz) bind(C, name="tbl_drv_gv") result(rc)
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), value :: Re
      integer(c_int), value :: kphys
      real(c_double), intent(in) :: uxb(*), uyb(*), wzb(*), varg(*), dvargdx(*)
      complex(c_double_complex), intent(in) :: ux(*), uy(*), uz(*), wx(*), wy(*), wz(*)
      complex(c_double_complex), intent(inout) :: dvx(*), dvy(*), dvz(*)
      integer(c_int) :: rc
    end function
    function tbl_drv_fi(ctx, Re, kphys, uxb, uyb, wzb, ux, uy, uz, wx, wy, wz, dvx, dvy, dvz) &
        bind(C, name="tbl_drv_fi") result(rc)
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), value :: Re
      integer(c_int), value :: kphys
      real(c_double), intent(in) :: uxb(*), uyb(*), wzb(*)
      complex(c_double_complex), intent(in) :: ux(*), uy(*), uz(*), wx(*), wy(*), wz(*)
      complex(c_double_complex), intent(inout) :: dvx(*), dvy(*), dvz(*)
      integer(c_int) :: rc
    end function

  end interface

  type(c_ptr), save :: tbl_ctx_handle = c_null_ptr
  type(tbl_solve_opts_t), save :: tbl_opts

contains

  ! Stop with the library's error text, in the style of the reference's
  ! fatal messages (write to stdout, then stop).
  subroutine tbl_fail(what, rc)
    character(len=*), intent(in) :: what
    integer(c_int), intent(in) :: rc
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: n, i
    type(c_ptr) :: p
    p = tbl_last_error()
    text = ' '
    if (c_associated(p)) then
      call c_f_pointer(p, msg, [512])
      n = 0
      do i = 1, 512
        if (msg(i) == c_null_char) exit
        n = i
        text(i:i) = msg(i)
      end do
    end if
    write(*,*) 'trans_bl_amd: ', what, ' failed, status ', rc, ' ', trim(text)
    stop 1
  end subroutine


  ! address of the th field for tbl_loop: my_form 2 hands it over, any other
  ! form passes a null pointer (th untouched)
  function tbl_th_ptr(with_th, th) result(p)
    logical, intent(in) :: with_th
    complex(c_double_complex), target, intent(in) :: th(*)
    type(c_ptr) :: p
    p = c_null_ptr
    if (with_th) p = c_loc(th(1))
  end function

end module tbl_iface
