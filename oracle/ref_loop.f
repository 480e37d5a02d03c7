cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
c     Driver of the reference's own `loop` routines (test infrastructure,
c     built by oracle/reference.py into oracle/_ref/).
c
c     Linked with the reference objects compiled from their sources (loop
c     main_par.f:1762-1785 with filter_trid, bzone, outuy, poi_uy, poi_ux,
c     poi_uz, wx_wall, wz_wall, bzonew from loop_par.f, filter_par.f,
c     poisson_par.f, penta_par.f, derivs_par.f) and a par.for configured
c     for one workload.  Set-up as init_gv does it for these routines at
c     one rank: /path/ (main_par.f:13-32), the coefficient tables
c     (main_par.f:852-873, with the wall tables of records 15-16),
c     create_ctes, derivs_k, lhs_tridf into /filt/ (main_par.f:590).
c
c     usage: ref_loop COEFS IN OUT ERRO WHAT
c       IN   stream file: ux, uy, uz, wx, wy, wz, th (complex*16
c            (ptsx,jmax,kfour) each), d2uydx2 (complex*16 (ptsx,kfour)),
c            bdfc (real*8 (2,imax))
c       WHAT loop | poi | wall | bzone | bzonew | filter | outuy
c            poi    = poi_ux then poi_uz, extra output duydy
c            wall   = wx_wall then wz_wall, extra output lapv
c            outuy  = extra output dwzdx
c       OUT  stream file: ux, uy, uz, wx, wy, wz, th after the call, then
c            the extra output if any
c     stdout: poi_uy's own line per mode when it runs.
cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      program ref_loop

      implicit none
      include 'par.for'
      include 'comm.par'
      include 'comm.var'
      include 'comm.coef'
      include 'mpif.h'
      integer p
      real*8 erro, bdfc(2,imax), afil(ptsx), bfil(ptsx), cfil(ptsx)
      complex*16 d2uydx2(ptsx,kfour), lapv(ptsx,kfour)
      complex*16 extra(ptsx,jmax,kfour)
      character*1024 fcoef, fin, fout, arg, what
      common/bd/ bdfc
      common/derw/ d2uydx2
      common/filt/ afil, bfil, cfil

      call MPI_Init(ierr)
      call MPI_Comm_rank(MPI_COMM_WORLD, my_rank, ierr)
      call MPI_Comm_size(MPI_COMM_WORLD, p, ierr)
      numproc = p - 1
      inter = 2**(msh - 1) * (stencil - 2)
      shift = my_rank * (ptsx - inter - 1)

      call get_command_argument(1, fcoef)
      call get_command_argument(2, fin)
      call get_command_argument(3, fout)
      call get_command_argument(4, arg)
      read(arg, *) erro
      call get_command_argument(5, what)

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
      read(2) ux, uy, uz, wx, wy, wz, th, d2uydx2, bdfc
      close(unit=2)

      if (trim(what) .eq. 'loop') then
        call loop(erro)
      else if (trim(what) .eq. 'poi') then
        call poi_ux(extra)
        call poi_uz(extra)
      else if (trim(what) .eq. 'wall') then
        call wx_wall(lapv)
        call wz_wall(lapv)
      else if (trim(what) .eq. 'bzone') then
        call bzone
      else if (trim(what) .eq. 'bzonew') then
        call bzonew
      else if (trim(what) .eq. 'filter') then
        call filter_trid
      else if (trim(what) .eq. 'outuy') then
        call outuy(extra)
      else
        write(*, *) 'unknown routine ', trim(what)
        call MPI_Abort(MPI_COMM_WORLD, 2, ierr)
      end if

      open(3, file=trim(fout), access='stream', form='unformatted',
     &     status='replace')
      write(3) ux, uy, uz, wx, wy, wz, th
      if (trim(what) .eq. 'poi' .or. trim(what) .eq. 'outuy') then
        write(3) extra
      else if (trim(what) .eq. 'wall') then
        write(3) lapv
      end if
      close(unit=3)

      call MPI_Finalize(ierr)
      end
