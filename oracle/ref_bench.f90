! CPU baseline harness (test infrastructure): times the reference's own
! PanelVelocityParallelModule::BVERK4 (compiled unmodified by
! oracle/Makefile.ref) on a uniform cubed sphere built by the reference's
! grid code, on every MPI rank of the job.  Prints one line on rank 0:
!   ref_bench level nActive nParticles nPassive ranks steps seconds pairs
! where pairs counts every Biot-Savart pair of the steps (active targets
! without self, particles and passive panels against all active sources).
!
! Usage:  mpiexec -n P ref_bench LEVEL NSTEPS
program ref_bench
	use NumberKindsModule
	use ParticlesEdgesPanelsModule
	use TracerAndVorticityDistributionModule
	use PanelVelocityParallelModule
	implicit none
	include 'mpif.h'

	type(Particles), pointer :: gp
	type(Edges), pointer :: ge
	type(Panels), pointer :: gpan
	integer(kint) :: level, nSteps, step, ierr, rank, nprocs, nA, nP, nPP
	real(kreal) :: t0, t1, pairs
	character(len=64) :: arg

	call MPI_INIT(ierr)
	call MPI_COMM_RANK(MPI_COMM_WORLD, rank, ierr)
	call MPI_COMM_SIZE(MPI_COMM_WORLD, nprocs, ierr)
	call get_command_argument(1, arg); read(arg, *) level
	call get_command_argument(2, arg); read(arg, *) nSteps

	call New(gp, ge, gpan, QUAD_PANEL, level, 0, 3, BVE_SOLVER)
	call InitRH4Wave(gp, gpan)
	nA = gpan%N_Active
	nPP = gpan%N - nA
	nP = gp%N
	call InitializeMPIRK4(gp, gpan, rank, nprocs)
	call MPI_BARRIER(MPI_COMM_WORLD, ierr)
	t0 = MPI_WTIME()
	do step = 1, nSteps
		call BVERK4(gp, gpan, 0.005_kreal, rank, nprocs)
	enddo
	call MPI_BARRIER(MPI_COMM_WORLD, ierr)
	t1 = MPI_WTIME()
	pairs = 4.0_kreal*nSteps*(real(nA, kreal)*(nA - 1) + real(nP + nPP, kreal)*nA)
	if (rank == 0) print '(A,4I9,2I5,F12.4,ES16.8)', 'ref_bench', level, nA, nP, nPP, &
		nprocs, nSteps, t1 - t0, pairs
	call FinalizeMPIRK4()
	call MPI_FINALIZE(ierr)
end program
