! Golden-vector harness for the Voronoi path: drives the reference's own
! VoronoiVelocityParallelModule (compiled from its unmodified sources by
! oracle/Makefile.ref) on an icosahedral Voronoi grid built by the
! reference's own VoronoiPanels / STRIPACK code, and dumps the state around
! each BVERK4 call as raw little-endian float64.
!
! Usage:  make_voronoi_golden INITNEST IC NSTEPS DT SAVEDAREA OUTFILE
!   IC = 0  the reference's Rossby-Haurwitz wave 4 (InitRH4Wave)
!   IC = 1  the reference's Gaussian vortex (InitGaussianVortex)
!   SAVEDAREA = 1 calls UseSavedArea() first, so every step sums with the
!   areas of the initial tessellation (what the RH4 Voronoi driver does
!   without AMR); 0 uses each step's current areas.
!
! Output layout (stream access, native little-endian):
!   int32  n, nSteps, savedArea
!   float64 dt, omega
!   nSteps+1 records s = 0..nSteps, each the state after step s (s = 0: the
!   initial state):  x(n) y(n) z(n) relVort(n) area(n)
!   then per step s = 1..nSteps:  kineticEnergy(n) totalKE
program make_voronoi_golden
	use NumberKindsModule
	use VoronoiPanelsModule
	use TracerAndVorticityDistributionModule
	use VoronoiVelocityParallelModule
	implicit none
	include 'mpif.h'

	type(VorPanels), pointer :: vp
	integer(kint) :: initNest, ic, nSteps, saved, step, ierr, rank, nprocs, n
	real(kreal) :: dt
	real(kreal), allocatable :: ke(:,:), keTotal(:)
	character(len=512) :: arg, outFile
	integer, parameter :: u = 21

	call MPI_INIT(ierr)
	call MPI_COMM_RANK(MPI_COMM_WORLD, rank, ierr)
	call MPI_COMM_SIZE(MPI_COMM_WORLD, nprocs, ierr)

	call get_command_argument(1, arg); read(arg, *) initNest
	call get_command_argument(2, arg); read(arg, *) ic
	call get_command_argument(3, arg); read(arg, *) nSteps
	call get_command_argument(4, arg); read(arg, *) dt
	call get_command_argument(5, arg); read(arg, *) saved
	call get_command_argument(6, outFile)

	call New(vp, initNest, 0, 3)
	select case (ic)
	case (0)
		call InitRH4Wave(vp)
	case (1)
		call InitGaussianVortex(vp)
	case default
		print *, 'make_voronoi_golden: unknown IC ', ic
		stop 1
	end select
	if (saved /= 0) call UseSavedArea()

	n = vp%N
	allocate(ke(n, nSteps), keTotal(nSteps))
	open(unit=u, file=trim(outFile), access='stream', form='unformatted', status='replace')
	write(u) int(n, 4), int(nSteps, 4), int(saved, 4)
	write(u) dt, Omega
	call write_state()

	call InitializeMPIRK4(n, rank, nprocs)
	do step = 1, nSteps
		call BVERK4(vp, dt, rank, nprocs)
		ke(:, step) = kineticEnergy(1:n)
		keTotal(step) = totalKE
		call write_state()
	enddo
	do step = 1, nSteps
		write(u) ke(:, step)
		write(u) keTotal(step)
	enddo
	close(u)

	call MPI_FINALIZE(ierr)

contains

	subroutine write_state()
		write(u) vp%x(1:vp%N)
		write(u) vp%y(1:vp%N)
		write(u) vp%z(1:vp%N)
		write(u) vp%relVort(1:vp%N)
		write(u) vp%area(1:vp%N)
	end subroutine

end program
