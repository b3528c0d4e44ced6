! mlegs_dropin.f90 -- the reference's module names over libmlegs_hip.so, so
! that its drivers (f90/bsnsq_new.f90) compile and link unchanged.
!
! The drivers USE OMP_LIB, MPI and fifteen MOD_* modules (f90/bsnsq_new.f90:
! 5-20) and read module globals (TIM, FILES, MONITORDATA, M, AK, NXCHOP,
! MPI_RANK).  Each module below keeps the reference's name and the public
! names the drivers use; the work goes through the mlegs_hip binding and the
! C ABI.  One GPU, one process:
!   MPI               a serial stand-in (rank 0 of 1; collectives copy)
!   MOD_MISC          P8, ITOA3, PRINT_REAL_TIME, MPRINT, MSAVE of a real matrix
!   MOD_SCALAR3       TYPE(SCALAR), ALLOCATE/DEALLOCATE, CHOPSET/CHOPDO, the
!                     grid globals (NR..NXCHOP, M, AK) and MPI_RANK, IERR
!   MOD_FFT           SETUP_GRID, the transforms, MLOAD/MSAVE of fields
!                     (MPILOADX/MPISAVEX files through the library)
!   MOD_LEGOPS        the spectral operators
!   MOD_BOUSSINESQ    CALC_BOUSSI_FORCE/NONLIN, CALC_BOUSSI_ENERGY with the
!                     especData_K/M.dat writers (src/mod_boussinesq.f90:138-232)
!   MOD_MARCH         SOLVER_T and PT_SOLVER (INITIALIZE, TIME_STEPPING with
!                     its diagnostics, output and INSPECT cadence, FINALIZE),
!                     TIM, FILES, MONITORDATA, DIAGV, DIAGNOST, INSPECT
!   MOD_INIT          SETUP_ENVIRONMENT: read.input through the library's
!                     parser (grid, physics, file names, cadences)
!   MOD_BANDMAT, MOD_EIG, MOD_FD, MOD_LIN_LEGENDRE, MOD_LAYOUT,
!   MOD_DIAGNOSTICS   present for the drivers' USE statements
!   MOD_EVP           SAVE_PERTURB stops: the eigenvalue solvers are not part
!                     of this library
! REMOVE (with the ETD axial sponge for RMV%SW = 2) and the hyperviscosity
! adjustment (NUP_ADJUST: HYP3ADJ + HYP3SET) run at the cadences read.input
! gives them, in TIME_STEPPING's order.

module mpi
  implicit none
  private
  integer, parameter, public :: MPI_COMM_WORLD = 0, MPI_COMM_SELF = 1
  integer, parameter, public :: MPI_CHARACTER = 1, MPI_INTEGER = 2, MPI_DOUBLE_PRECISION = 3, MPI_LOGICAL = 4, &
                                MPI_DOUBLE_COMPLEX = 5, MPI_REAL8 = 3, MPI_COMPLEX16 = 5
  integer, parameter, public :: MPI_SUM = 1, MPI_MAX = 2, MPI_MIN = 3
  integer, parameter, public :: MPI_THREAD_SINGLE = 0, MPI_THREAD_FUNNELED = 1, MPI_THREAD_SERIALIZED = 2, &
                                MPI_THREAD_MULTIPLE = 3
  public :: mpi_init, mpi_init_thread, mpi_comm_rank, mpi_comm_size, mpi_barrier, mpi_abort, mpi_finalize, mpi_wtime
  public :: mpi_bcast
  interface mpi_bcast
    module procedure bcast_c, bcast_i, bcast_iv, bcast_d, bcast_dv, bcast_l
  end interface
contains
  subroutine mpi_init(ierr)
    integer, intent(out) :: ierr
    ierr = 0
  end subroutine mpi_init
  subroutine mpi_init_thread(required, provided, ierr)
    integer, intent(in) :: required
    integer, intent(out) :: provided, ierr
    provided = required
    ierr = 0
  end subroutine mpi_init_thread
  subroutine mpi_comm_rank(comm, rank, ierr)
    integer, intent(in) :: comm
    integer, intent(out) :: rank, ierr
    rank = 0
    ierr = 0
  end subroutine mpi_comm_rank
  subroutine mpi_comm_size(comm, n, ierr)
    integer, intent(in) :: comm
    integer, intent(out) :: n, ierr
    n = 1
    ierr = 0
  end subroutine mpi_comm_size
  subroutine mpi_barrier(comm, ierr)
    integer, intent(in) :: comm
    integer, intent(out) :: ierr
    ierr = 0
  end subroutine mpi_barrier
  subroutine mpi_abort(comm, code, ierr)
    integer, intent(in) :: comm, code
    integer, intent(out) :: ierr
    ierr = 0
    write (*, '(a,i0)') 'MPI_ABORT: error code ', code
    error stop code
  end subroutine mpi_abort
  subroutine mpi_finalize(ierr)
    integer, intent(out) :: ierr
    ierr = 0
  end subroutine mpi_finalize
  double precision function mpi_wtime()
    integer(8) :: cnt, rate
    call system_clock(cnt, rate)
    mpi_wtime = dble(cnt) / dble(rate)
  end function mpi_wtime
  ! one process: a broadcast leaves the buffer as it is
  subroutine bcast_c(buf, n, dtype, root, comm, ierr)
    character(len=*), intent(inout) :: buf
    integer, intent(in) :: n, dtype, root, comm
    integer, intent(out) :: ierr
    ierr = 0
  end subroutine bcast_c
  subroutine bcast_i(buf, n, dtype, root, comm, ierr)
    integer, intent(inout) :: buf
    integer, intent(in) :: n, dtype, root, comm
    integer, intent(out) :: ierr
    ierr = 0
  end subroutine bcast_i
  subroutine bcast_iv(buf, n, dtype, root, comm, ierr)
    integer, intent(inout) :: buf(:)
    integer, intent(in) :: n, dtype, root, comm
    integer, intent(out) :: ierr
    ierr = 0
  end subroutine bcast_iv
  subroutine bcast_d(buf, n, dtype, root, comm, ierr)
    double precision, intent(inout) :: buf
    integer, intent(in) :: n, dtype, root, comm
    integer, intent(out) :: ierr
    ierr = 0
  end subroutine bcast_d
  subroutine bcast_dv(buf, n, dtype, root, comm, ierr)
    double precision, intent(inout) :: buf(:)
    integer, intent(in) :: n, dtype, root, comm
    integer, intent(out) :: ierr
    ierr = 0
  end subroutine bcast_dv
  subroutine bcast_l(buf, n, dtype, root, comm, ierr)
    logical, intent(inout) :: buf
    integer, intent(in) :: n, dtype, root, comm
    integer, intent(out) :: ierr
    ierr = 0
  end subroutine bcast_l
end module mpi

module mod_misc
  use, intrinsic :: iso_c_binding
  implicit none
  private
  integer, parameter, public :: p8 = c_double, p4 = c_int
  public :: itoa3, itoa4, print_real_time, msave
  interface msave
    module procedure msave_real2
  end interface
contains
  function itoa3(i)
    integer, intent(in) :: i
    character(len=3) :: itoa3
    write (itoa3, '(i3.3)') i
  end function itoa3
  function itoa4(i)
    integer, intent(in) :: i
    character(len=4) :: itoa4
    write (itoa4, '(i4.4)') i
  end function itoa4
  subroutine print_real_time()
    character(len=8) :: d
    character(len=10) :: t
    character(len=5) :: z
    integer :: v(8)
    call date_and_time(d, t, z, v)
    write (*, '(" @ ",i4.4,"-",i2.2,"-",i2.2," ",i2.2,":",i2.2,":",i2.2)') v(1), v(2), v(3), v(5), v(6), v(7)
    write (*, *) ''
  end subroutine print_real_time
  ! a real matrix, one row per line (the reference's MSAVEDR writes it with
  ! its own column labels; this writes the values only)
  subroutine msave_real2(a, fn)
    real(p8), intent(in) :: a(:, :)
    character(len=*), intent(in) :: fn
    integer :: u, i
    open (newunit=u, file=trim(fn), status='replace', action='write')
    do i = 1, size(a, 1)
      write (u, '(*(es24.16,1x))') a(i, :)
    end do
    close (u)
  end subroutine msave_real2
end module mod_misc

module mod_bandmat
  implicit none
end module mod_bandmat

module mod_eig
  implicit none
end module mod_eig

module mod_fd
  implicit none
end module mod_fd

module mod_lin_legendre
  implicit none
end module mod_lin_legendre

module mod_scalar3
  use, intrinsic :: iso_c_binding
  use mod_misc, only: p8
  use mlegs_hip, only: scalar, allocate, deallocate, chopdo, to_host, to_device, lib_chopset => chopset, &
                       mlg_params, check, ndimr, ndimth, ndimx, nrchopdim, ntchopdim, nxchopdim, &
                       PPP_SPACE, FFF_SPACE, PFP_SPACE, PFF_SPACE, nrchop_now
  implicit none
  private
  public :: scalar, allocate, deallocate, chopdo, chopset, to_host, to_device, check
  public :: PPP_SPACE, FFF_SPACE, PFP_SPACE, PFF_SPACE
  public :: ndimr, ndimth, ndimx, nrchopdim, ntchopdim, nxchopdim
  ! grid globals of the reference (src/mod_scalar3.f90:15-37), set by SETUP_GRID
  integer, public :: nr = 0, nth = 0, nx = 0, nrchop = 0, ntchop = 0, nxchop = 0, nxchoph = 0
  integer, public :: minc = 1, mklink = 0
  real(p8), public :: ell = 0.0_p8, ell2 = 0.0_p8, zlen = 0.0_p8
  integer, allocatable, public :: m(:)
  real(p8), allocatable, public :: ak(:, :)
  integer, public :: mpi_rank = 0, mpi_comm_ivp = 0, mpi_thread_mode = 0, ierr = 0
  ! grid and physics from read.input (SETUP_ENVIRONMENT), used by SETUP_GRID
  type(mlg_params), public :: run_params
contains
  ! CHOPSET (src/mod_scalar3.f90:165-195): the library's truncation and the
  ! NRCHOP global move together
  subroutine chopset(iof)
    integer, intent(in) :: iof
    call lib_chopset(iof)
    nrchop = nrchop_now()
  end subroutine chopset
end module mod_scalar3

module mod_fft
  use, intrinsic :: iso_c_binding
  use mod_misc, only: p8
  use mod_scalar3
  use mlegs_hip, only: leginit, legfinal, horfft, verfft, rtran, toff, tofp, plnfft, mlegs_ctx, copy_scalar
  implicit none
  private
  public :: setup_grid, leginit, legfinal, horfft, verfft, rtran, toff, tofp, plnfft, mload, msave, mprint, copy_scalar
  interface mload
    module procedure mload_scalar
  end interface
  interface msave
    module procedure msave_scalar
  end interface
  interface
    integer(c_int) function mlg_field_save(f, fn) bind(c)
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: f
      character(kind=c_char), intent(in) :: fn(*)
    end function
    integer(c_int) function mlg_field_load(f, fn, n1, n2) bind(c)
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: f
      character(kind=c_char), intent(in) :: fn(*)
      integer(c_int), value :: n1, n2
    end function
    integer(c_int) function mlg_get_table(c, name, out, n, len) bind(c)
      import :: c_int, c_ptr, c_char, c_double, c_long
      type(c_ptr), value :: c
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: out(*)
      integer(c_long), value :: n
      integer(c_long), intent(out) :: len
    end function
    integer(c_int) function c_mkdir(path, mode) bind(c, name='mkdir')
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: mode
    end function
  end interface
contains
  ! SETUP_GRID (src/mod_fft.f90:165-178): LEGINIT with read.input's
  ! parameters, the grid globals, and the output directory
  subroutine setup_grid(savedir)
    character(len=*), intent(in), optional :: savedir
    integer :: mm, kk, rc
    integer(c_long) :: n
    real(c_double), allocatable :: t(:)
    call leginit(run_params)
    nr = run_params%nr; nth = run_params%nth; nx = run_params%nx
    nrchop = run_params%nrchop; ntchop = run_params%ntchop; nxchop = run_params%nxchop
    nxchoph = nxchop
    minc = max(1, run_params%minc); mklink = run_params%mklink
    ell = run_params%ell; ell2 = ell**2; zlen = run_params%zlen
    if (allocated(m)) deallocate (m)
    if (allocated(ak)) deallocate (ak)
    allocate (m(ntchopdim), ak(ntchopdim, nxchopdim))
    do mm = 1, ntchopdim
      m(mm) = (mm - 1) * minc
    end do
    allocate (t(ntchopdim * nxchopdim))
    call check(mlg_get_table(mlegs_ctx(), 'ak'//c_null_char, t, int(size(t), c_long), n), 'SETUP_GRID')
    do kk = 1, nxchopdim
      do mm = 1, ntchopdim
        ak(mm, kk) = t(mm + ntchopdim * (kk - 1))
      end do
    end do
    if (present(savedir)) then
      rc = c_mkdir(trim(adjustl(savedir))//c_null_char, int(o'755', c_int))
    end if
  end subroutine setup_grid

  ! MLOAD (MLOAD0 -> MPILOADX, src/mod_fft.f90:2390-2642) of a file this
  ! library or a one-process run wrote
  subroutine mload_scalar(fn, a)
    character(len=*), intent(in) :: fn
    type(scalar), intent(inout) :: a
    if (.not. c_associated(a%h)) call allocate(a)
    call check(mlg_field_load(a%h, trim(fn)//c_null_char, 1_c_int, 1_c_int), 'MLOAD '//trim(fn))
    call to_host(a)
  end subroutine mload_scalar

  ! MSAVE (MSAVE0 -> MPISAVEX, src/mod_fft.f90:2224-2377)
  subroutine msave_scalar(a, fn)
    type(scalar), intent(in) :: a
    character(len=*), intent(in) :: fn
    call check(mlg_field_save(a%h, trim(fn)//c_null_char), 'MSAVE '//trim(fn))
  end subroutine msave_scalar

  ! MPRINT (src/mod_fft.f90:2644-2660)
  subroutine mprint(message)
    character(len=*), intent(in) :: message
    if (mpi_rank == 0) write (*, *) trim(message)
  end subroutine mprint
end module mod_fft

module mod_layout
  use mlegs_hip, only: qvortex
  implicit none
  private
  public :: qvortex
end module mod_layout

module mod_legops
  use mlegs_hip, only: delsqh, idelsqh, mulxm, xxdx, del2, idel2, idel2ln, ihelm, ihelmp, grad, pc2vel, pc2vor, project
  implicit none
  private
  public :: delsqh, idelsqh, mulxm, xxdx, del2, idel2, idel2ln, ihelm, ihelmp, grad, pc2vel, pc2vor, project
end module mod_legops

module mod_boussinesq
  use, intrinsic :: iso_c_binding
  use mod_misc, only: p8
  use mod_scalar3
  use mlegs_hip, only: calc_boussi_force, calc_boussi_nonlin, lib_energy => calc_boussi_energy
  implicit none
  private
  public :: calc_boussi_force, calc_boussi_nonlin, calc_boussi_energy
contains
  ! CALC_BOUSSI_ENERGY (src/mod_boussinesq.f90:138-232): KE, PE, DKE per
  ! (m, |k|) and the especData_K.dat / especData_M.dat lines (totals over m
  ! per k, over k per m) appended in SAVEDIR
  subroutine calc_boussi_energy(psi, chi, b, time, savedir)
    type(scalar), intent(in) :: psi, chi, b
    real(p8) :: time
    character(len=72) :: savedir
    real(p8), allocatable :: ke(:, :), pe(:, :), dke(:, :)
    integer :: u, mm, kk, st
    allocate (ke(ntchop, nxchop), pe(ntchop, nxchop), dke(ntchop, nxchop))
    call lib_energy(psi, chi, b, ke, pe, dke)
    if (mpi_rank /= 0) return
    open (newunit=u, file=trim(adjustl(savedir))//'especData_K.dat', status='unknown', iostat=st, position='append')
    if (st == 0) then
      write (u, '(" KE:",F24.16,1X,*(E24.16,1X))') time, (sum(ke(:, kk)), kk=1, nxchop)
      write (u, '(" PE:",F24.16,1X,*(E24.16,1X))') time, (sum(pe(:, kk)), kk=1, nxchop)
      write (u, '("DKE:",F24.16,1X,*(E24.16,1X))') time, (sum(dke(:, kk)), kk=1, nxchop)
      close (u)
    end if
    open (newunit=u, file=trim(adjustl(savedir))//'especData_M.dat', status='unknown', iostat=st, position='append')
    if (st == 0) then
      write (u, '(" KE:",F24.16,1X,*(E24.16,1X))') time, (sum(ke(mm, :)), mm=1, ntchop)
      write (u, '(" PE:",F24.16,1X,*(E24.16,1X))') time, (sum(pe(mm, :)), mm=1, ntchop)
      write (u, '("DKE:",F24.16,1X,*(E24.16,1X))') time, (sum(dke(mm, :)), mm=1, ntchop)
      close (u)
    end if
  end subroutine calc_boussi_energy
end module mod_boussinesq

module mod_diagnostics
  implicit none
end module mod_diagnostics

module mod_march
  use, intrinsic :: iso_c_binding
  use mod_misc, only: p8, print_real_time
  use mod_scalar3
  use mod_fft, only: msave, mprint
  use mod_boussinesq, only: calc_boussi_energy
  use mlegs_hip, only: mlegs_ctx
  implicit none
  private
  ! src/mod_march.f90:17-130
  type, public :: timedata
    real(p8) :: dt = 0.0_p8, t = 0.0_p8, limit = 0.0_p8
    integer :: n = 0
    integer :: scheme = 0
  end type timedata
  type, public :: iofiles
    character(len=72) :: savedir = './', psi0 = 'psi0.dat', chi0 = 'chi0.dat', b0 = 'b0.dat'
    character(len=72) :: psii = ' ', chii = ' ', bi = ' '
    real(p8), pointer :: t(:) => null()
    character(len=72), pointer :: psi(:) => null(), chi(:) => null(), b(:) => null()
    integer :: n = 1, ne = 0
  end type iofiles
  type, public :: diagdata
    logical :: is_initialized = .false.
    integer :: it1 = 40, it2 = 40
    real(p8) :: vori = 0.0_p8, zmomi = 0.0_p8, enei = 0.0_p8, heli = 0.0_p8, angi = 0.0_p8
  end type diagdata
  type, public :: removal
    integer :: sw = 0, int = 10000
  end type removal
  type, public :: viscosity
    integer :: sw = 0, p = 0, adjsw = 0, adjint = 0
    real(p8) :: nu = 0.0_p8, nup = 0.0_p8
  end type viscosity
  type, public :: monitoreigenvalueskit
    integer :: start = 0, interval = 0
    integer, allocatable :: mk(:, :)
  end type monitoreigenvalueskit
  type, public :: solver_t
    type(c_ptr) :: h = c_null_ptr
    logical :: save_variable = .true.
    real(p8) :: time_start = 0.0_p8
  contains
    procedure, pass(this) :: initialize
    procedure, pass(this) :: time_stepping
    procedure, pass(this) :: finalize
  end type solver_t
  type(solver_t), public :: pt_solver
  type(timedata), public :: tim
  type(iofiles), public :: files
  type(diagdata), public :: diagv
  type(removal), public :: rmv
  type(viscosity), public :: visc
  type(monitoreigenvalueskit), public :: monitordata
  public :: diagnost, inspect

  interface
    integer(c_int) function mlg_solver_create(c, s) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      type(c_ptr), intent(out) :: s
    end function
    integer(c_int) function mlg_solver_destroy(s) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function mlg_initialize(s, psi, chi, b) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: s, psi, chi, b
    end function
    integer(c_int) function mlg_time_step(s, psi, chi, b) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: s, psi, chi, b
    end function
    real(c_double) function mlg_solver_time(s) bind(c)
      import :: c_double, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function mlg_solver_nstep(s) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function mlg_diagnost(psi, chi) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: psi, chi
    end function
    integer(c_int) function mlg_diagnost_it1(psi, chi, out) bind(c)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: psi, chi
      real(c_double), intent(out) :: out(12)
    end function
    integer(c_int) function mlg_set_march_controls(s, rmv_sw, rmv_int, adj_sw, adj_int) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: s
      integer(c_int), value :: rmv_sw, rmv_int, adj_sw, adj_int
    end function
    integer(c_int) function mlg_remove(psi, chi, b) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: psi, chi, b
    end function
    integer(c_int) function mlg_hyp3adj(c, psi, chi, b) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: c, psi, chi, b
    end function
    integer(c_int) function mlg_field_m0_columns(f, nk, out, nrad) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: f
      integer(c_int), value :: nk
      type(c_ptr), value :: out
      integer(c_int), intent(out) :: nrad
    end function
  end interface

contains
  ! SOLVER_T%INITIALIZE (src/mod_march.f90:171-240): saves the initial
  ! fields, then the Richardson start of the configured scheme
  subroutine initialize(this, psi, chi, b)
    class(solver_t), intent(inout) :: this
    type(scalar), intent(inout) :: psi, chi, b
    integer(8) :: cnt, rate
    call mprint('SOLVER STARTED')
    if (mpi_rank == 0) call print_real_time()
    call system_clock(cnt, rate)
    this%time_start = dble(cnt) / dble(rate)
    call msave(psi, trim(adjustl(files%savedir))//files%psi0)
    call msave(chi, trim(adjustl(files%savedir))//files%chi0)
    call msave(b, trim(adjustl(files%savedir))//files%b0)
    if (.not. c_associated(this%h)) call check(mlg_solver_create(mlegs_ctx(), this%h), 'SOLVER_T%INITIALIZE')
    ! SPNGINIT for RMV%SW = 2 (:209-215); the REMOVE and HYP3ADJ cadences are
    ! kept here, in TIME_STEPPING's order
    call check(mlg_set_march_controls(this%h, int(rmv%sw, c_int), 0_c_int, 0_c_int, 0_c_int), 'SOLVER_T%INITIALIZE')
    call check(mlg_initialize(this%h, psi%h, chi%h, b%h), 'SOLVER_T%INITIALIZE')
    tim%n = mlg_solver_nstep(this%h)
    tim%t = mlg_solver_time(this%h)
    call mprint('SOLVER: Initialization and Richardson startup complete.')
  end subroutine initialize

  ! SOLVER_T%TIME_STEPPING (src/mod_march.f90:243-307): one step with the
  ! per-step DIAGNOST checks, energies every DIAGV%IT1 steps, snapshots on
  ! FILES%T, INSPECT of the three fields
  subroutine time_stepping(this, psi, chi, b)
    class(solver_t), intent(inout) :: this
    type(scalar), intent(inout) :: psi, chi, b
    call check(mlg_time_step(this%h, psi%h, chi%h, b%h), 'SOLVER_T%TIME_STEPPING')
    tim%n = mlg_solver_nstep(this%h)
    tim%t = mlg_solver_time(this%h)
    call diagnost(psi, chi)
    if (diagv%it1 > 0) then
      if (mod(tim%n, diagv%it1) == 0) call calc_boussi_energy(psi, chi, b, tim%t, files%savedir)
    end if
    if (rmv%sw /= 0 .and. rmv%int > 0) then
      if (mod(tim%n, rmv%int) == 0) then
        ! REMOVE, then the Richardson restart (one step) and its checks
        call check(mlg_remove(psi%h, chi%h, b%h), 'REMOVE')
        call check(mlg_initialize(this%h, psi%h, chi%h, b%h), 'SOLVER_T%TIME_STEPPING')
        tim%n = mlg_solver_nstep(this%h)
        tim%t = mlg_solver_time(this%h)
        call diagnost(psi, chi)
        if (diagv%it1 > 0) then
          if (mod(tim%n, diagv%it1) == 0) call calc_boussi_energy(psi, chi, b, tim%t, files%savedir)
        end if
      end if
    end if
    if (visc%adjsw /= 0 .and. visc%adjint > 0) then
      if (mod(tim%n, visc%adjint) == 0) call check(mlg_hyp3adj(mlegs_ctx(), psi%h, chi%h, b%h), 'HYP3ADJ')
    end if
    if (this%save_variable .and. files%ne > 0 .and. files%n <= files%ne) then
      if (files%t(files%n) <= tim%t) then
        call msave(psi, trim(adjustl(files%savedir))//files%psi(files%n))
        call msave(chi, trim(adjustl(files%savedir))//files%chi(files%n))
        call msave(b, trim(adjustl(files%savedir))//files%b(files%n))
        files%n = files%n + 1
        if (files%n > files%ne) this%save_variable = .false.
      end if
    end if
    call inspect(psi, 1)
    call inspect(chi, 1)
    call inspect(b, 1)
  end subroutine time_stepping

  ! SOLVER_T%FINALIZE (src/mod_march.f90:310-353)
  subroutine finalize(this, psi, chi, b)
    class(solver_t), intent(inout) :: this
    type(scalar), intent(inout), optional :: psi, chi, b
    integer(8) :: cnt, rate
    call mprint('SOLVER: Deallocating ...')
    if (c_associated(this%h)) call check(mlg_solver_destroy(this%h), 'SOLVER_T%FINALIZE')
    this%h = c_null_ptr
    if (present(psi)) call deallocate(psi)
    if (present(chi)) call deallocate(chi)
    if (present(b)) call deallocate(b)
    call system_clock(cnt, rate)
    if (mpi_rank == 0) then
      write (*, *) 'Finalized after ', tim%n, 'steps'
      write (*, *) 'Time-stepping total time: ', dble(cnt) / dble(rate) - this%time_start, 'seconds'
      call print_real_time()
    end if
  end subroutine finalize

  ! DIAGNOST (src/mod_march.f90:1726-1961): the per-step checks (11
  ! dealiasing, 12 divergence end the run as the reference's MPI_ABORT) and,
  ! every DIAGV%IT1 steps, the integrals of vorticity, z-momentum, energy,
  ! helicity and angular momentum with their change since step 0, psi and
  ! chi at r = infinity, and CALCXY's vortex centre and radius (also
  ! appended to SAVEDIR/coresize.dat)
  subroutine diagnost(psi, chi)
    type(scalar), intent(in) :: psi, chi
    real(c_double) :: d(12)
    integer :: u, st
    call check(mlg_diagnost(psi%h, chi%h), 'DIAGNOST')
    if (diagv%it1 <= 0) return
    if (mod(tim%n, diagv%it1) /= 0) return
    call check(mlg_diagnost_it1(psi%h, chi%h, d), 'DIAGNOST')
    if (tim%n == 0) then
      diagv%is_initialized = .true.
      diagv%vori = d(1)
      diagv%zmomi = d(2)
      diagv%enei = d(3)
      diagv%heli = d(4)
      diagv%angi = d(5)
    end if
    if (mpi_rank /= 0) return
    write (*, *) '<<< DIAGNOSTICS >>>'
    write (6, '(A,E13.4,A,I8)') 'TIM%T=', tim%t, ' TIM%N=', tim%n
    write (6, '(A,E13.5,A,E13.5)') 'ADV%X=', 0.0_p8, '  ADV%Y=', 0.0_p8
    if (d(7) > 1.0e-10_p8 .or. d(9) > 1.0e-10_p8) then
      write (6, '(A,E18.9,A,E18.9)') 'PSIAT1(KK=1)=', d(6), ' MAXMIN=', d(7)
      write (6, '(A,E18.9,A,E18.9)') 'CHIAT1(KK=1)=', d(8), ' MAXMIN=', d(9)
    end if
    write (6, *) 'CFL#=', 0.0_p8
    call change_line('VORTICITY', d(1), diagv%vori, .false.)
    call change_line('Z-MOMENTUM', d(2), diagv%zmomi, .false.)
    call change_line('ENERGY', d(3), diagv%enei, .true.)
    call change_line('HELICITY', d(4), diagv%heli, .false.)
    call change_line('INTEG_ANGMOM ', d(5), diagv%angi, .false.)
    write (6, '(A,E10.3,A,E10.3,A,E10.3,A,E10.3)') 'T= ', tim%t, ' <X>= ', d(10), ' <Y>= ', d(11), ' <R>= ', d(12)
    open (newunit=u, file=trim(adjustl(files%savedir))//'coresize.dat', position='append', iostat=st)
    if (st == 0) then
      write (u, '(4E14.6)') tim%t, d(10), d(11), d(12)
      close (u)
    end if
  contains
    ! a value with its fractional change from step 0 (always for the energy)
    subroutine change_line(name, v, v0, always)
      character(len=*), intent(in) :: name
      real(c_double), intent(in) :: v
      real(p8), intent(in) :: v0
      logical, intent(in) :: always
      if (always .or. v0 /= 0.0_p8) then
        write (6, '(A12,A,E23.15,A,E23.15)') name, '=', v, '  FRACTIONAL CHANGE=', -(v0 - v) / v0
      else
        write (6, '(A12,A,E23.15)') name, '=', v
      end if
    end subroutine change_line
  end subroutine diagnost

  ! INSPECT (src/mod_march.f90:3743-3775): appends TIM%T and the m = 0
  ! columns of nk axial modes (the first as real parts, the others as
  ! moduli) to SAVEDIR/inspect.output
  subroutine inspect(field, nk)
    type(scalar), intent(in) :: field
    integer, intent(in), optional :: nk
    integer :: nsave, u, i, nn
    integer(c_int) :: nrad
    complex(c_double_complex), allocatable, target :: col(:, :)
    nsave = 10
    if (present(nk)) nsave = nk
    call check(mlg_field_m0_columns(field%h, 1_c_int, c_null_ptr, nrad), 'INSPECT')
    if (field%space == FFF_SPACE .or. field%space == PFF_SPACE) then
      nsave = min(nsave, nxchopdim)
    else
      nsave = min(nsave, ndimx)
    end if
    allocate (col(nrad, nsave))
    call check(mlg_field_m0_columns(field%h, int(nsave, c_int), c_loc(col), nrad), 'INSPECT')
    if (mpi_rank /= 0) return
    open (newunit=u, file=trim(files%savedir)//'inspect.output', status='unknown', action='write', position='append')
    write (u, *) tim%t
    do i = 1, nsave
      if (i == 1) then
        write (u, '(I4,1x,*(ES14.6,","))') i, (real(col(nn, i)), nn=1, nrad)
      else
        write (u, '(I4,1x,*(ES14.6,","))') i, (abs(col(nn, i)), nn=1, nrad)
      end if
    end do
    close (u)
  end subroutine inspect
end module mod_march

module mod_init
  use, intrinsic :: iso_c_binding
  use mod_misc, only: p8, print_real_time
  use mod_scalar3
  use mlegs_hip, only: mlg_params
  use mod_march, only: tim, files, diagv, rmv, visc, monitordata
  implicit none
  private
  public :: setup_environment
  integer, parameter :: name_len = 72, max_outputs = 512, max_monitor = 32
  type, bind(c) :: run_settings
    character(kind=c_char) :: savedir(name_len), psi0(name_len), chi0(name_len), b0(name_len)
    character(kind=c_char) :: psii(name_len), chii(name_len), bi(name_len)
    integer(c_int) :: it1, it2, rmv_sw, rmv_int, adj_sw, adj_int, noutputs
    real(c_double) :: out_t(max_outputs)
    character(kind=c_char) :: out_psi(name_len, max_outputs), out_chi(name_len, max_outputs), out_b(name_len, max_outputs)
    integer(c_int) :: nmonitor
    integer(c_int) :: monitor_mk(2, max_monitor)
    integer(c_int) :: monitor_start, monitor_interval
  end type run_settings
  interface
    integer(c_int) function mlg_read_input(path, p, t_limit) bind(c)
      import :: c_int, c_char, c_double, mlg_params
      character(kind=c_char), intent(in) :: path(*)
      type(mlg_params), intent(out) :: p
      real(c_double), intent(out) :: t_limit
    end function
    integer(c_int) function mlg_read_run_settings(path, s) bind(c)
      import :: c_int, c_char, run_settings
      character(kind=c_char), intent(in) :: path(*)
      type(run_settings), intent(out) :: s
    end function
  end interface
contains
  function name_of(c) result(s)
    character(kind=c_char), intent(in) :: c(name_len)
    character(len=name_len) :: s
    integer :: i
    do i = 1, name_len
      s(i:i) = c(i)
    end do
  end function name_of

  ! SETUP_ENVIRONMENT (src/mod_init.f90:87-122): read.input in the working
  ! directory through the library's parser (READIN0's blocks)
  subroutine setup_environment(echo_mode)
    character(len=*), optional :: echo_mode
    type(run_settings), allocatable :: s
    real(c_double) :: limit
    integer :: i
    mpi_rank = 0
    write (*, *) 'PARALLEL PROGRAM STARTED'
    call print_real_time()
    if (present(echo_mode)) then
      if (echo_mode == 'ECHO' .or. echo_mode == 'echo') write (*, *) 'read.input'
    end if
    call check(mlg_read_input('read.input'//c_null_char, run_params, limit), 'SETUP_ENVIRONMENT')
    allocate (s)
    call check(mlg_read_run_settings('read.input'//c_null_char, s), 'SETUP_ENVIRONMENT')
    tim%dt = run_params%dt
    tim%limit = limit
    tim%n = 0
    tim%t = 0.0_p8
    files%savedir = name_of(s%savedir)
    files%psi0 = name_of(s%psi0)
    files%chi0 = name_of(s%chi0)
    files%b0 = name_of(s%b0)
    files%psii = name_of(s%psii)
    files%chii = name_of(s%chii)
    files%bi = name_of(s%bi)
    files%ne = s%noutputs
    allocate (files%t(max(1, files%ne)), files%psi(max(1, files%ne)), files%chi(max(1, files%ne)), &
              files%b(max(1, files%ne)))
    files%t = huge(1.0_p8)
    do i = 1, files%ne
      files%t(i) = s%out_t(i)
      files%psi(i) = name_of(s%out_psi(:, i))
      files%chi(i) = name_of(s%out_chi(:, i))
      files%b(i) = name_of(s%out_b(:, i))
    end do
    diagv%it1 = s%it1
    diagv%it2 = s%it2
    rmv%sw = s%rmv_sw
    rmv%int = s%rmv_int
    visc%sw = run_params%visc_sw
    visc%p = run_params%visc_p
    visc%nu = run_params%nu
    visc%nup = run_params%nup
    visc%adjsw = s%adj_sw
    visc%adjint = s%adj_int
    if (allocated(monitordata%mk)) deallocate (monitordata%mk)
    allocate (monitordata%mk(s%nmonitor, 2))
    do i = 1, s%nmonitor
      monitordata%mk(i, 1) = s%monitor_mk(1, i)
      monitordata%mk(i, 2) = s%monitor_mk(2, i)
    end do
    monitordata%start = s%monitor_start
    monitordata%interval = s%monitor_interval
    write (*, *) 'PARAMETERS READ COMPLETED'
  end subroutine setup_environment
end module mod_init

module mod_evp
  implicit none
  private
  public :: save_perturb
contains
  ! the eigenvalue solvers (src/mod_evp.f90) are not part of this library
  subroutine save_perturb(filename)
    character(len=*), intent(in) :: filename
    write (*, *) 'SAVE_PERTURB (', trim(filename), '): the eigenvalue solvers are not provided'
    error stop 30
  end subroutine save_perturb
end module mod_evp
