! mlegs_hip.f90 -- Fortran binding of libmlegs_hip.so through ISO_C_BINDING.
!
! A Fortran driver of MLegS USEs this module where it used MOD_FFT,
! MOD_LEGOPS, MOD_BOUSSINESQ and MOD_MARCH for the time step: the
! procedures keep the reference's names and argument order, and operate on
! TYPE(SCALAR) fields whose coefficients live on the GPU.
!
!   reference                                  this module
!   TYPE(SCALAR) (mod_scalar3.f90:41-47)       TYPE(SCALAR): device handle + host mirror E, LN, SPACE
!   ALLOCATE/DEALLOCATE (mod_fft.f90:136-146)  ALLOCATE(A[,SP]) / DEALLOCATE(A)
!   LEGINIT (mod_fft.f90:180-367)              LEGINIT(P)   (P = grid/physics parameters)
!   HORFFT/VERFFT/RTRAN/TOFF/TOFP/PLNFFT       same names (mod_fft.f90:771-1398)
!   CHOPSET/CHOPDO (mod_scalar3.f90:165-257)   same names
!   DEL2 ... PROJECT (mod_legops.f90)          same names
!   CALC_BOUSSI_FORCE/NONLIN/ENERGY            same names (mod_boussinesq.f90)
!   SOLVER_T%INITIALIZE/TIME_STEPPING          same names (mod_march.f90:171-307)
!
! The device copy is authoritative.  A%E on the host is brought up to date
! by TO_HOST(A); after changing A%E (or A%LN) on the host call TO_DEVICE(A).
! Every library status other than 0 ends the program with ERROR STOP and
! the reference's error code, where the reference calls MPI_ABORT.
module mlegs_hip
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter, public :: PPP_SPACE = 0, FFF_SPACE = 1, PFP_SPACE = 2, PFF_SPACE = 3

  type, bind(c), public :: mlg_params
    integer(c_int) :: nr = 0, nth = 0, nx = 0
    integer(c_int) :: nrchop = 0, ntchop = 0, nxchop = 0
    integer(c_int) :: minc = 1, mklink = 0
    real(c_double) :: ell = 1.0_c_double, zlen = 0.0_c_double
    real(c_double) :: nu = 0.0_c_double, kappa = 0.0_c_double, nup = 0.0_c_double, kappap = 0.0_c_double
    integer(c_int) :: visc_sw = 1, visc_p = 0
    real(c_double) :: omega = 0.0_c_double, bv0 = 0.0_c_double
    integer(c_int) :: adams = 1, hyper_cn = 0
    real(c_double) :: dt = 0.0_c_double
  end type mlg_params

  type, public :: scalar
    type(c_ptr) :: h = c_null_ptr
    complex(c_double_complex), allocatable :: e(:, :, :)
    real(c_double) :: ln = 0.0_c_double
    integer :: space = FFF_SPACE
  end type scalar

  type, public :: solver_t
    type(c_ptr) :: h = c_null_ptr
  contains
    procedure :: initialize => solver_initialize
    procedure :: time_stepping => solver_time_stepping
    procedure :: finalize => solver_finalize
  end type solver_t

  ! grid state of the reference's module globals (one context per program)
  type(c_ptr), save :: ctx = c_null_ptr
  integer, save, public :: ndimr = 0, ndimth = 0, ndimx = 0, nrchopdim = 0, ntchopdim = 0, nxchopdim = 0

  interface allocate
    module procedure salloc
  end interface
  interface deallocate
    module procedure sfree
  end interface

  public :: leginit, legfinal, allocate, deallocate, to_host, to_device, chopset, chopdo, nrchop_now
  public :: horfft, verfft, rtran, toff, tofp, plnfft
  public :: delsqh, idelsqh, mulxm, xxdx, del2, idel2, idel2ln, ihelm, grad, pc2vel, pc2vor, project
  public :: calc_boussi_force, calc_boussi_nonlin, calc_boussi_energy, qvortex, copy_scalar, ihelmp
  public :: check, mlegs_ctx

  interface
    integer(c_int) function c_init(p, c) bind(c, name='mlg_init')
      import :: c_int, c_ptr, mlg_params
      type(mlg_params), intent(in) :: p
      type(c_ptr), intent(out) :: c
    end function
    integer(c_int) function c_finalize(c) bind(c, name='mlg_finalize')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
    end function
    integer(c_int) function c_dims(c, o) bind(c, name='mlg_dims')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), intent(out) :: o(12)
    end function
    integer(c_int) function c_chopset(c, iof) bind(c, name='mlg_chopset')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: iof
    end function
    integer(c_int) function c_nrchop(c) bind(c, name='mlg_nrchop')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
    end function
    integer(c_int) function c_field_create(c, sp, f) bind(c, name='mlg_field_create')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: sp
      type(c_ptr), intent(out) :: f
    end function
    integer(c_int) function c_field_destroy(f) bind(c, name='mlg_field_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: f
    end function
    integer(c_int) function c_field_space(f) bind(c, name='mlg_field_space')
      import :: c_int, c_ptr
      type(c_ptr), value :: f
    end function
    integer(c_int) function c_upload(f, e, sp, ln) bind(c, name='mlg_field_upload')
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: f
      complex(c_double_complex), intent(in) :: e(*)
      integer(c_int), value :: sp
      real(c_double), value :: ln
    end function
    integer(c_int) function c_download(f, e, sp, ln) bind(c, name='mlg_field_download')
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: f
      complex(c_double_complex), intent(out) :: e(*)
      integer(c_int), intent(out) :: sp
      real(c_double), intent(out) :: ln
    end function
  end interface

  interface
    integer(c_int) function mlg_copy(b, a) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: b, a
    end function
    integer(c_int) function mlg_chopdo(a) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: a
    end function
    integer(c_int) function mlg_horfft(a, is) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: a
      integer(c_int), value :: is
    end function
    integer(c_int) function mlg_verfft(a, is) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: a
      integer(c_int), value :: is
    end function
    integer(c_int) function mlg_rtran(a, is) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: a
      integer(c_int), value :: is
    end function
    integer(c_int) function mlg_plnfft(a, is) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: a
      integer(c_int), value :: is
    end function
    integer(c_int) function mlg_toff(a) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: a
    end function
    integer(c_int) function mlg_tofp(a) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: a
    end function
    integer(c_int) function mlg_delsqh(a, b) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: a, b
    end function
    integer(c_int) function mlg_idelsqh(b, a) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: b, a
    end function
    integer(c_int) function mlg_mulxm(a, b) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: a, b
    end function
    integer(c_int) function mlg_xxdx(a, b) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: a, b
    end function
    integer(c_int) function mlg_del2(a, b) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: a, b
    end function
    integer(c_int) function mlg_idel2(b, a, has_ln, ln) bind(c)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b, a
      integer(c_int), value :: has_ln
      real(c_double), value :: ln
    end function
    integer(c_int) function mlg_idel2ln(b, a) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: b, a
    end function
    integer(c_int) function mlg_ihelm(b, a, alp) bind(c)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b, a
      real(c_double), value :: alp
    end function
    integer(c_int) function mlg_ihelmp(b, a, p, alp, nu) bind(c)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b, a
      integer(c_int), value :: p
      real(c_double), value :: alp, nu
    end function
    integer(c_int) function mlg_grad(a, rr, rp, z) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: a, rr, rp, z
    end function
    integer(c_int) function mlg_pc2vel(psi, chi, rur, rup, uz) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: psi, chi, rur, rup, uz
    end function
    integer(c_int) function mlg_pc2vor(psi, chi, ror, rop, oz) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: psi, chi, ror, rop, oz
    end function
    integer(c_int) function mlg_project(rur, rup, uz, psi, chi, has_ln, ln) bind(c)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: rur, rup, uz, psi, chi
      integer(c_int), value :: has_ln
      real(c_double), value :: ln
    end function
    integer(c_int) function mlg_force(psi, chi, b, psin, chin, bn) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: psi, chi, b, psin, chin, bn
    end function
    integer(c_int) function mlg_nonlin(b, rur, rup, uz, bn, ror, rop, oz) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: b, rur, rup, uz, bn, ror, rop, oz
    end function
    integer(c_int) function mlg_energy(psi, chi, b, ke, pe, dke) bind(c)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: psi, chi, b
      real(c_double), intent(out) :: ke(*), pe(*), dke(*)
    end function
    integer(c_int) function mlg_qvortex(c, q, psi, chi) bind(c)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: c
      real(c_double), value :: q
      type(c_ptr), value :: psi, chi
    end function
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
    integer(c_int) function mlg_time_steps(s, psi, chi, b, n) bind(c)
      import :: c_int, c_ptr
      type(c_ptr), value :: s, psi, chi, b
      integer(c_int), value :: n
    end function
  end interface

contains

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc /= 0) then
      write (*, '(a,a,i0)') where, ': MLegS error ', rc
      error stop rc
    end if
  end subroutine check

  ! LEGINIT: grid, tables and truncation state on the current HIP device
  subroutine leginit(p)
    type(mlg_params), intent(in) :: p
    integer(c_int) :: d(12)
    call check(c_init(p, ctx), 'LEGINIT')
    call check(c_dims(ctx, d), 'LEGINIT')
    ndimr = d(1); ndimth = d(2); ndimx = d(3)
    nrchopdim = d(4); ntchopdim = d(5); nxchopdim = d(6)
  end subroutine leginit

  ! the context handle, for modules that call the C ABI directly
  type(c_ptr) function mlegs_ctx()
    mlegs_ctx = ctx
  end function mlegs_ctx

  subroutine legfinal()
    if (c_associated(ctx)) call check(c_finalize(ctx), 'FINALIZE')
    ctx = c_null_ptr
  end subroutine legfinal

  subroutine shape_of(sp, n1, n2, n3)
    integer, intent(in) :: sp
    integer, intent(out) :: n1, n2, n3
    select case (sp)
    case (FFF_SPACE)
      n1 = nrchopdim; n2 = ntchopdim; n3 = nxchopdim
    case (PFF_SPACE)
      n1 = ndimr; n2 = ntchopdim; n3 = nxchopdim
    case default
      n1 = ndimr; n2 = ndimth; n3 = ndimx
    end select
  end subroutine shape_of

  subroutine salloc(a, sp)
    type(scalar), intent(inout) :: a
    integer, intent(in), optional :: sp
    integer :: s, n1, n2, n3
    s = FFF_SPACE
    if (present(sp)) s = sp
    if (.not. c_associated(a%h)) call check(c_field_create(ctx, int(s, c_int), a%h), 'ALLOCATE')
    a%space = s
    a%ln = 0.0_c_double
    call shape_of(s, n1, n2, n3)
    if (allocated(a%e)) deallocate (a%e)
    allocate (a%e(n1, n2, n3))
    a%e = (0.0_c_double, 0.0_c_double)
  end subroutine salloc

  subroutine sfree(a)
    type(scalar), intent(inout) :: a
    if (c_associated(a%h) .and. c_associated(ctx)) call check(c_field_destroy(a%h), 'DEALLOCATE')
    a%h = c_null_ptr
    if (allocated(a%e)) deallocate (a%e)
  end subroutine sfree

  ! host mirror <- device (A%E reshaped to the field's current space)
  subroutine to_host(a)
    type(scalar), intent(inout) :: a
    integer(c_int) :: sp
    integer :: n1, n2, n3
    real(c_double) :: ln
    a%space = c_field_space(a%h)
    call shape_of(a%space, n1, n2, n3)
    if (allocated(a%e)) then
      if (size(a%e, 1) /= n1 .or. size(a%e, 2) /= n2 .or. size(a%e, 3) /= n3) deallocate (a%e)
    end if
    if (.not. allocated(a%e)) allocate (a%e(n1, n2, n3))
    call check(c_download(a%h, a%e, sp, ln), 'TO_HOST')
    a%ln = ln
  end subroutine to_host

  ! device <- host mirror (A%E in space A%SPACE, A%LN)
  subroutine to_device(a)
    type(scalar), intent(inout) :: a
    call check(c_upload(a%h, a%e, int(a%space, c_int), a%ln), 'TO_DEVICE')
  end subroutine to_device

  subroutine copy_scalar(b, a)
    type(scalar), intent(inout) :: b
    type(scalar), intent(in) :: a
    call check(mlg_copy(b%h, a%h), 'COPY')
  end subroutine copy_scalar

  subroutine chopset(iof)
    integer, intent(in) :: iof
    call check(c_chopset(ctx, int(iof, c_int)), 'CHOPSET')
  end subroutine chopset

  integer function nrchop_now()
    nrchop_now = c_nrchop(ctx)
  end function nrchop_now

  subroutine chopdo(a)
    type(scalar), intent(inout) :: a
    call check(mlg_chopdo(a%h), 'CHOPDO')
  end subroutine chopdo

  subroutine horfft(a, is)
    type(scalar), intent(inout) :: a
    integer, intent(in) :: is
    call check(mlg_horfft(a%h, int(is, c_int)), 'HORFFT')
  end subroutine horfft

  subroutine verfft(a, is)
    type(scalar), intent(inout) :: a
    integer, intent(in) :: is
    call check(mlg_verfft(a%h, int(is, c_int)), 'VERFFT')
  end subroutine verfft

  subroutine rtran(a, is)
    type(scalar), intent(inout) :: a
    integer, intent(in) :: is
    call check(mlg_rtran(a%h, int(is, c_int)), 'RTRAN')
  end subroutine rtran

  subroutine plnfft(a, is)
    type(scalar), intent(inout) :: a
    integer, intent(in) :: is
    call check(mlg_plnfft(a%h, int(is, c_int)), 'PLNFFT')
  end subroutine plnfft

  subroutine toff(a)
    type(scalar), intent(inout) :: a
    call check(mlg_toff(a%h), 'TOFF')
  end subroutine toff

  subroutine tofp(a)
    type(scalar), intent(inout) :: a
    call check(mlg_tofp(a%h), 'TOFP')
  end subroutine tofp

  subroutine delsqh(a, b)
    type(scalar), intent(in) :: a
    type(scalar), intent(inout) :: b
    call check(mlg_delsqh(a%h, b%h), 'DELSQH')
  end subroutine delsqh

  subroutine idelsqh(b, a)
    type(scalar), intent(in) :: b
    type(scalar), intent(inout) :: a
    call check(mlg_idelsqh(b%h, a%h), 'IDELSQH')
  end subroutine idelsqh

  subroutine mulxm(a, b)
    type(scalar), intent(in) :: a
    type(scalar), intent(inout) :: b
    call check(mlg_mulxm(a%h, b%h), 'MULXM')
  end subroutine mulxm

  subroutine xxdx(a, b)
    type(scalar), intent(in) :: a
    type(scalar), intent(inout) :: b
    call check(mlg_xxdx(a%h, b%h), 'XXDX')
  end subroutine xxdx

  subroutine del2(a, b)
    type(scalar), intent(in) :: a
    type(scalar), intent(inout) :: b
    call check(mlg_del2(a%h, b%h), 'DEL2')
  end subroutine del2

  subroutine idel2(b, a, ln)
    type(scalar), intent(in) :: b
    type(scalar), intent(inout) :: a
    real(c_double), intent(in), optional :: ln
    if (present(ln)) then
      call check(mlg_idel2(b%h, a%h, 1_c_int, ln), 'IDEL2')
    else
      call check(mlg_idel2(b%h, a%h, 0_c_int, 0.0_c_double), 'IDEL2')
    end if
  end subroutine idel2

  subroutine idel2ln(b, a)
    type(scalar), intent(in) :: b
    type(scalar), intent(inout) :: a
    call check(mlg_idel2ln(b%h, a%h), 'IDEL2LN')
  end subroutine idel2ln

  subroutine ihelm(b, a, alp)
    type(scalar), intent(in) :: b
    type(scalar), intent(inout) :: a
    real(c_double), intent(in) :: alp
    call check(mlg_ihelm(b%h, a%h, alp), 'IHELM')
  end subroutine ihelm

  subroutine ihelmp(p, b, a, alp, nuin)
    integer, intent(in) :: p
    type(scalar), intent(in) :: b
    type(scalar), intent(inout) :: a
    real(c_double), intent(in) :: alp
    real(c_double), intent(in), optional :: nuin
    real(c_double) :: nu
    nu = 0.0_c_double
    if (present(nuin)) nu = nuin
    call check(mlg_ihelmp(b%h, a%h, int(p, c_int), alp, nu), 'IHELMP')
  end subroutine ihelmp

  subroutine grad(a, rr, rp, z)
    type(scalar), intent(in) :: a
    type(scalar), intent(inout) :: rr, rp, z
    call check(mlg_grad(a%h, rr%h, rp%h, z%h), 'GRAD')
  end subroutine grad

  subroutine pc2vel(psi, chi, rur, rup, uz)
    type(scalar), intent(in) :: psi, chi
    type(scalar), intent(inout) :: rur, rup, uz
    call check(mlg_pc2vel(psi%h, chi%h, rur%h, rup%h, uz%h), 'PC2VEL')
  end subroutine pc2vel

  subroutine pc2vor(psi, chi, ror, rop, oz)
    type(scalar), intent(in) :: psi, chi
    type(scalar), intent(inout) :: ror, rop, oz
    call check(mlg_pc2vor(psi%h, chi%h, ror%h, rop%h, oz%h), 'PC2VOR')
  end subroutine pc2vor

  subroutine project(rur, rup, uz, psi, chi, ln)
    type(scalar), intent(inout) :: rur, rup, uz, psi, chi
    real(c_double), intent(in), optional :: ln
    if (present(ln)) then
      call check(mlg_project(rur%h, rup%h, uz%h, psi%h, chi%h, 1_c_int, ln), 'PROJECT')
    else
      call check(mlg_project(rur%h, rup%h, uz%h, psi%h, chi%h, 0_c_int, 0.0_c_double), 'PROJECT')
    end if
  end subroutine project

  subroutine calc_boussi_force(psi, chi, b, psin, chin, bn)
    type(scalar), intent(in) :: psi, chi, b
    type(scalar), intent(inout) :: psin, chin, bn
    call check(mlg_force(psi%h, chi%h, b%h, psin%h, chin%h, bn%h), 'CALC_BOUSSI_FORCE')
  end subroutine calc_boussi_force

  subroutine calc_boussi_nonlin(b, rur, rup, uz, bn, ror, rop, oz)
    type(scalar), intent(in) :: b
    type(scalar), intent(inout) :: rur, rup, uz, bn, ror, rop, oz
    call check(mlg_nonlin(b%h, rur%h, rup%h, uz%h, bn%h, ror%h, rop%h, oz%h), 'CALC_BOUSSI_NONLIN')
  end subroutine calc_boussi_nonlin

  ! KE, PE, DKE over (m = 0..NTCHOP-1, |k| = 0..NXCHOP-1)
  subroutine calc_boussi_energy(psi, chi, b, ke, pe, dke)
    type(scalar), intent(in) :: psi, chi, b
    real(c_double), intent(out) :: ke(:, :), pe(:, :), dke(:, :)
    call check(mlg_energy(psi%h, chi%h, b%h, ke, pe, dke), 'CALC_BOUSSI_ENERGY')
  end subroutine calc_boussi_energy

  subroutine qvortex(q, psi, chi)
    real(c_double), intent(in) :: q
    type(scalar), intent(inout) :: psi, chi
    call check(mlg_qvortex(ctx, q, psi%h, chi%h), 'QVORTEX')
  end subroutine qvortex

  subroutine solver_initialize(this, psi, chi, b)
    class(solver_t), intent(inout) :: this
    type(scalar), intent(inout) :: psi, chi, b
    if (.not. c_associated(this%h)) call check(mlg_solver_create(ctx, this%h), 'SOLVER_T%INITIALIZE')
    call check(mlg_initialize(this%h, psi%h, chi%h, b%h), 'SOLVER_T%INITIALIZE')
  end subroutine solver_initialize

  subroutine solver_time_stepping(this, psi, chi, b, nsteps)
    class(solver_t), intent(inout) :: this
    type(scalar), intent(inout) :: psi, chi, b
    integer, intent(in) :: nsteps
    call check(mlg_time_steps(this%h, psi%h, chi%h, b%h, int(nsteps, c_int)), 'SOLVER_T%TIME_STEPPING')
  end subroutine solver_time_stepping

  subroutine solver_finalize(this)
    class(solver_t), intent(inout) :: this
    if (c_associated(this%h)) call check(mlg_solver_destroy(this%h), 'SOLVER_T%FINALIZE')
    this%h = c_null_ptr
  end subroutine solver_finalize

end module mlegs_hip
