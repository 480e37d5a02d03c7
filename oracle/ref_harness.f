cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
c     Driver of the reference's own hot-path routines (test
c     infrastructure, built by oracle/ref_build.py into oracle/_ref/).
c
c     Linked with the reference objects compiled from their sources
c     under the reference tree (poi_uy poisson_par.f:6-64, the
c     derivatives derivs_par.f:80-177, derivs_k derivs_par.f:6-29,
c     create_ctes main_par.f:1828-1892) and a par.for configured for one
c     workload.  It does at one rank what init_gv and loop do for the
c     path: MPI start-up and /path/ (main_par.f:13-32), the coefficient
c     tables in the record order main_par.f:852-873 reads, create_ctes,
c     derivs_k; then poi_uy on wx, dwzdx and the initial guess uy read
c     from a file, and the five derivatives of the result.
c
c     usage: ref_harness COEFS IN OUT ERROMAX CALLS [T]
c       IN   stream file: wx, dwzdx, uy (complex*16 (ptsx,jmax,kfour))
c       OUT  stream file: uy after the calls, then derparx, derparxx,
c            dery, deryy, deryfv of it
c       CALLS  poi_uy calls in a row (each warm-starts from the last)
c       T    when given: after the calls, ux = wy = dwzdx and uz = wz = wx
c            are set and the reference's checkpoint writer escreve(T)
c            (escreve_par.f:6-33) writes data_00.bin in the working
c            directory
c     stdout: poi_uy's own line per mode (k, V-cycles, residual) and
c     the wall time of the poi_uy calls (mpi_wtime).
cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
      program ref_harness

      implicit none
      include 'par.for'
      include 'comm.par'
      include 'comm.var'
      include 'comm.coef'
      include 'mpif.h'
      integer p, calls, c, tstep
      real*8 erromax, t0, t1
      complex*16 dwzdx(ptsx,jmax,kfour), d(ptsx,jmax,kfour)
      character*1024 fcoef, fin, fout, arg

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
      read(arg, *) erromax
      call get_command_argument(5, arg)
      read(arg, *) calls

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

      open(2, file=trim(fin), access='stream', form='unformatted',
     &     status='old')
      read(2) wx, dwzdx, uy
      close(unit=2)

      t0 = mpi_wtime()
      do c = 1, calls
        call poi_uy(dwzdx, erromax)
      end do
      t1 = mpi_wtime()

      if (command_argument_count() .ge. 6) then
        call get_command_argument(6, arg)
        read(arg, *) tstep
        ux = dwzdx
        wy = dwzdx
        uz = wx
        wz = wx
        call escreve(tstep)
      end if

      open(3, file=trim(fout), access='stream', form='unformatted',
     &     status='replace')
      write(3) uy
      call derparx(d, uy)
      write(3) d
      call derparxx(d, uy)
      write(3) d
      call dery(d, uy)
      write(3) d
      call deryy(d, uy)
      write(3) d
      call deryfv(d, uy)
      write(3) d
      close(unit=3)

      write(*, '(a, f14.6)') ' poi_uy seconds ', t1 - t0
      call MPI_Finalize(ierr)
      end
