! Golden-vector harness: drives the reference's own PanelVelocityParallelModule
! (compiled from its unmodified sources by oracle/Makefile.ref) on a uniform
! cubed-sphere grid built by the reference's own grid code, and dumps the
! inputs and outputs of its RK4 step as raw little-endian float64.
!
! Usage:  make_golden LEVEL IC NSTEPS DT SMOOTH OUTFILE [T0]
!   IC = 0  the reference's Rossby-Haurwitz wave 4 (InitRH4Wave)
!   IC = 1  RH4 with omega = K = 7.848e-6*86400, amplitude -30*K (the
!           benchmark's parameterisation), assigned to panels and particles
!   IC = 2  the reference's Gaussian vortex (InitGaussianVortex)
!   IC = 3  the reference's stratosphere model (InitStratosphere), stepped
!           with StratosphereRK4 (Juckes forcing) from time T0
!   IC = 4  the RH4 wave, moved by AdvectionRK4 (NonDivergentWind) from T0
!   IC = 5  the RH4 wave, moved by DivergentAdvectionRK4 (DivergentWind) from T0
!
! Output layout (stream access, native little-endian):
!   int32  nActive, nParticles, nPassive, nSteps
!   float64 dt, smooth, omega
!   state(0), state(1), state(nSteps)       (see write_state)
!   float64 kineticEnergy(nActive)          stage-1 |u|^2 of step 1
!   float64 totalKE(nSteps)                 per step
!   IC >= 3 only: float64 T0, active absVort(nActive)
! where a state is, in the order GatherPanels uses (panel index order):
!   active  x(3,nActive) relVort(nActive) area(nActive)
!   passive x(3,nPassive)
!   particles x(3,nParticles) relVort(nParticles) absVort(nParticles)
program make_golden
	use NumberKindsModule
	use ParticlesEdgesPanelsModule
	use TracerAndVorticityDistributionModule
	use PanelVelocityParallelModule
	implicit none
	include 'mpif.h'

	type(Particles), pointer :: gp
	type(Edges), pointer :: ge
	type(Panels), pointer :: gpan
	integer(kint) :: level, ic, nSteps, step, ierr, rank, nprocs, j, nA, nP, nPP
	real(kreal) :: dt, smoothArg, lambda, z, rhOmega, t0
	real(kreal), allocatable :: keStep1(:), keTotal(:)
	character(len=512) :: arg, outFile
	integer, parameter :: u = 21

	call MPI_INIT(ierr)
	call MPI_COMM_RANK(MPI_COMM_WORLD, rank, ierr)
	call MPI_COMM_SIZE(MPI_COMM_WORLD, nprocs, ierr)

	call get_command_argument(1, arg); read(arg, *) level
	call get_command_argument(2, arg); read(arg, *) ic
	call get_command_argument(3, arg); read(arg, *) nSteps
	call get_command_argument(4, arg); read(arg, *) dt
	call get_command_argument(5, arg); read(arg, *) smoothArg
	call get_command_argument(6, outFile)
	t0 = 0.0_kreal
	if (command_argument_count() >= 7) then
		call get_command_argument(7, arg); read(arg, *) t0
	endif

	call New(gp, ge, gpan, QUAD_PANEL, level, 0, 3, BVE_SOLVER)
	call SetSmooth(smoothArg)

	select case (ic)
	case (0)
		call InitRH4Wave(gp, gpan)
	case (1)
		rhOmega = 7.848e-6_kreal*86400.0_kreal
		do j = 1, gp%N
			z = gp%x(3,j)
			lambda = atan2(gp%x(2,j), gp%x(1,j))
			gp%relVort(j) = 2.0_kreal*rhOmega*z - 30.0_kreal*rhOmega*z*(1.0_kreal - z*z)**2*cos(4.0_kreal*lambda)
			gp%absVort(j) = gp%relVort(j) + 2.0_kreal*Omega*z
		enddo
		do j = 1, gpan%N
			z = gpan%x(3,j)
			lambda = atan2(gpan%x(2,j), gpan%x(1,j))
			gpan%relVort(j) = 2.0_kreal*rhOmega*z - 30.0_kreal*rhOmega*z*(1.0_kreal - z*z)**2*cos(4.0_kreal*lambda)
			gpan%absVort(j) = gpan%relVort(j) + 2.0_kreal*Omega*z
		enddo
	case (2)
		call InitGaussianVortex(gp, gpan)
	case (3)
		call InitStratosphere(gp, gpan)
	case (4, 5)
		call InitRH4Wave(gp, gpan)
	case default
		print *, 'make_golden: unknown IC ', ic
		stop 1
	end select

	nA = gpan%N_Active
	nPP = gpan%N - nA
	nP = gp%N
	allocate(keStep1(nA), keTotal(nSteps))

	open(unit=u, file=trim(outFile), access='stream', form='unformatted', status='replace')
	write(u) int(nA, 4), int(nP, 4), int(nPP, 4), int(nSteps, 4)
	write(u) dt, smoothArg, Omega
	call write_state()

	call InitializeMPIRK4(gp, gpan, rank, nprocs)
	do step = 1, nSteps
		if (ic == 3) then
			call StratosphereRK4(gp, gpan, t0 + real(step - 1, kreal)*dt, dt, rank, nprocs)
		elseif (ic == 4) then
			call AdvectionRK4(gp, gpan, t0 + real(step - 1, kreal)*dt, dt, rank, nprocs)
		elseif (ic == 5) then
			call DivergentAdvectionRK4(gp, gpan, t0 + real(step - 1, kreal)*dt, dt, rank, nprocs)
		else
			call BVERK4(gp, gpan, dt, rank, nprocs)
		endif
		keTotal(step) = totalKE
		if (step == 1) then
			keStep1 = kineticEnergy
			call write_state()
		endif
	enddo
	call write_state()
	write(u) keStep1
	write(u) keTotal
	if (ic >= 3) then
		write(u) t0
		do j = 1, gpan%N
			if (.not. gpan%hasChildren(j)) write(u) gpan%absVort(j)
		enddo
	endif
	close(u)

	call MPI_FINALIZE(ierr)

contains

	subroutine write_state()
		integer(kint) :: k
		do k = 1, gpan%N
			if (.not. gpan%hasChildren(k)) write(u) gpan%x(:,k)
		enddo
		do k = 1, gpan%N
			if (.not. gpan%hasChildren(k)) write(u) gpan%relVort(k)
		enddo
		do k = 1, gpan%N
			if (.not. gpan%hasChildren(k)) write(u) gpan%area(k)
		enddo
		do k = 1, gpan%N
			if (gpan%hasChildren(k)) write(u) gpan%x(:,k)
		enddo
		write(u) gp%x(:,1:gp%N)
		write(u) gp%relVort(1:gp%N)
		write(u) gp%absVort(1:gp%N)
	end subroutine

end program
