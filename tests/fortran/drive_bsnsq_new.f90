! A driver with the call sequence of the reference's bsnsq_new
! (f90/bsnsq_new.f90:30-99), over the drop-in modules: read.input set-up,
! MLOAD of psi, chi and b from SAVEDIR, the DIAGNOST checks and energies,
! INITIALIZE, TIME_STEPPING up to TIM%LIMIT, FINALIZE.  The GPU tests run it
! in a directory that holds read.input and the initial field files.
program drive_bsnsq_new
  use mpi
  use mod_misc
  use mod_scalar3
  use mod_fft
  use mod_boussinesq
  use mod_march
  use mod_init
  implicit none
  type(scalar) :: psi, chi, b
  integer :: nsteps, it

  call setup_environment('NOECHO')
  call setup_grid(files%savedir)
  call allocate(psi)
  call allocate(chi)
  call allocate(b)
  call mload(trim(adjustl(files%savedir))//files%psi0, psi)
  call mload(trim(adjustl(files%savedir))//files%chi0, chi)
  call mload(trim(adjustl(files%savedir))//files%b0, b)

  call diagnost(psi, chi)
  call calc_boussi_energy(psi, chi, b, tim%t, files%savedir)
  call pt_solver%initialize(psi, chi, b)
  call diagnost(psi, chi)
  call calc_boussi_energy(psi, chi, b, tim%t, files%savedir)

  ! the step count as the reference forms it (real to integer, truncated)
  nsteps = int(tim%limit / tim%dt)
  files%n = 1
  do it = 1, nsteps
    call pt_solver%time_stepping(psi, chi, b)
  end do
  call pt_solver%finalize(psi, chi, b)
end program drive_bsnsq_new
