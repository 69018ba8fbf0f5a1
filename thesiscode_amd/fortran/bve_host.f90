! Standalone Fortran host of the panel RK4 path: reads a problem from a
! binary file, advances it with libbve through the ISO_C_BINDING interface,
! and writes the result.  It is the call pattern of the reference's panel
! drivers (RHWavePanelsMPI.f90:247-320: step, then read totalKE and the
! updated particles and panels) without the reference's grid code.
!
! Usage:  bve_host.exe INPUT OUTPUT
!
! INPUT (stream, native little-endian):
!   int32   nActive, nPassive, nParticles, nSteps, precision, device
!   float64 dt, omega, smooth
!   float64 xActive(3,nActive), relVort(nActive), area(nActive)
!   float64 xPassive(3,nPassive), xParticles(3,nParticles), absVort(nParticles)
! OUTPUT:
!   float64 xActive(3,nActive), relVort(nActive), xPassive(3,nPassive)
!   float64 xParticles(3,nParticles), particle relVort(nParticles)
!   float64 totalKE(nSteps)
program bve_host
	use iso_c_binding
	use BVEInterfaceModule
	implicit none

	integer(c_int) :: nActive, nPassive, nParticles, nSteps, precision, device, step
	real(c_double) :: dt, omega, smooth
	real(c_double), allocatable :: xActive(:,:), relVort(:), area(:), xPassive(:,:)
	real(c_double), allocatable :: xParticles(:,:), absVort(:), relVortP(:), ke(:), totalKE(:)
	character(len=1024) :: inFile, outFile
	type(c_ptr) :: ctx
	integer, parameter :: uin = 31, uout = 32

	call get_command_argument(1, inFile)
	call get_command_argument(2, outFile)
	open(unit=uin, file=trim(inFile), access='stream', form='unformatted', status='old')
	read(uin) nActive, nPassive, nParticles, nSteps, precision, device
	read(uin) dt, omega, smooth
	allocate(xActive(3, nActive), relVort(nActive), area(nActive), xPassive(3, max(nPassive, 1)))
	allocate(xParticles(3, max(nParticles, 1)), absVort(max(nParticles, 1)))
	allocate(relVortP(max(nParticles, 1)), ke(nActive), totalKE(nSteps))
	read(uin) xActive, relVort, area
	if (nPassive > 0) read(uin) xPassive(:, 1:nPassive)
	if (nParticles > 0) read(uin) xParticles(:, 1:nParticles), absVort(1:nParticles)
	close(uin)

	call check(bve_create(ctx, device), 'bve_create')
	call check(bve_set_omega(ctx, omega), 'bve_set_omega')
	call check(bve_set_smooth(ctx, smooth), 'bve_set_smooth')
	call check(bve_set_precision(ctx, precision, 0.02_c_double), 'bve_set_precision')
	call check(bve_load_panels(ctx, nActive, xActive, relVort, area, nPassive, xPassive), &
		'bve_load_panels')
	do step = 1, nSteps
		call check(bve_rk4_step(ctx, dt, nParticles, xParticles, absVort, relVortP, totalKE(step), ke), &
			'bve_rk4_step')
	enddo
	call check(bve_store_panels(ctx, xActive, relVort, xPassive), 'bve_store_panels')
	call check(bve_destroy(ctx), 'bve_destroy')

	open(unit=uout, file=trim(outFile), access='stream', form='unformatted', status='replace')
	write(uout) xActive, relVort
	if (nPassive > 0) write(uout) xPassive(:, 1:nPassive)
	if (nParticles > 0) write(uout) xParticles(:, 1:nParticles), relVortP(1:nParticles)
	write(uout) totalKE
	close(uout)

contains

	subroutine check(rc, what)
		integer(c_int), intent(in) :: rc
		character(len=*), intent(in) :: what
		if (rc /= BVE_OK) then
			write(*, '(A,A,A)') 'bve_host: ', what, ' failed: '//trim(bve_error_message(ctx))
			error stop 1
		endif
	end subroutine

end program
