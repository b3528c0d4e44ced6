! Fortran driver of the Boussinesq step through the mlegs_hip module, in the
! shape of the reference's bsnsq_test (f90/bsnsq_test.f90): set up the grid,
! load psi/chi/b, INITIALIZE, TIME_STEPPING, write the state.
!
! usage: drive_bsnsq <state-in> <state-out> <nsteps>
! state files: for psi, chi, b in turn the FFF coefficients
! E(NRCHOPDIM,NTCHOPDIM,NXCHOPDIM) as complex(8) followed by LN as real(8),
! raw stream.  Grid: 32x16x16, NRCHOP=22, NTCHOP=6, NXCHOP=5, ELL=1,
! ZLEN=2 pi, nu=kappa=1e-3, Omega=0.5, Nbar=1, AB2-CN, dt=1e-3.
program drive_bsnsq
  use, intrinsic :: iso_c_binding
  use mlegs_hip
  implicit none
  type(mlg_params) :: p
  type(scalar) :: psi, chi, b
  type(solver_t) :: sol
  character(len=512) :: fin, fout, arg
  integer :: nsteps, u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg)
  read (arg, *) nsteps

  p%nr = 32; p%nth = 16; p%nx = 16
  p%nrchop = 22; p%ntchop = 6; p%nxchop = 5
  p%ell = 1.0d0; p%zlen = 2.0d0 * acos(-1.0d0)
  p%nu = 1.0d-3; p%kappa = 1.0d-3
  p%visc_sw = 1; p%omega = 0.5d0; p%bv0 = 1.0d0
  p%adams = 1; p%dt = 1.0d-3
  call leginit(p)

  call allocate(psi); call allocate(chi); call allocate(b)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) psi%e, psi%ln
  read (u) chi%e, chi%ln
  read (u) b%e, b%ln
  close (u)
  call to_device(psi); call to_device(chi); call to_device(b)

  call sol%initialize(psi, chi, b)
  call sol%time_stepping(psi, chi, b, nsteps)

  call to_host(psi); call to_host(chi); call to_host(b)
  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) psi%e, psi%ln
  write (u) chi%e, chi%ln
  write (u) b%e, b%ln
  close (u)
  call sol%finalize()
  call deallocate(psi); call deallocate(chi); call deallocate(b)
  call legfinal()
  write (*, '(a,i0,a)') 'drive_bsnsq: ', nsteps + 1, ' steps done'
end program drive_bsnsq
