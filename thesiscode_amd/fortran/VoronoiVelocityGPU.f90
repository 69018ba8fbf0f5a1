! MI355X drop-in for the reference's VoronoiVelocityParallelModule
! (VoronoiVelocityParallel.f90): same module name and public list, so the
! reference's Voronoi drivers (RHWaveVoronoiMPI.f90, GaussVortVoronoi.f90,
! JetVoronoi.f90, RH2Voronoi.f90) compile and link against it unchanged.
! The Biot-Savart sums, the RK4 stage loop and NormalizeGrid run on the GPU
! through libbve (bve_voronoi_rk4_step); the Delaunay triangulation and the
! Voronoi tessellation that follow every step (:480-482) stay the
! reference's own STRIPACK routines on the host.
!
! Semantics kept (VoronoiVelocityParallel.f90:304-483):
!   * the sum weighs each generator by its saved initial area after
!     UseSavedArea (:331-335; saved at the first step, :151-155), else by
!     its current area;
!   * kineticEnergy is |u|^2 at stage 1 and totalKE = 1/2 sum KE * area with
!     the current areas (:358-361), evaluated here with the reference's own
!     expression;
!   * BVERK4 returns without stepping before InitializeMPIRK4 (:311-314).
! procRank and numProcs are accepted; one GPU computes every generator.
! The arithmetic defaults to the reference's own order and rounding
! (BVE_PRECISION_REFERENCE), so the positions handed to STRIPACK are the
! reference's bit for bit: its triangulation is decided by the last bits of
! near-cocircular generators, and on symmetric flows it can loop for ever on
! some last-bit variants of a trajectory.  The environment variable
! BVE_PRECISION selects another mode (fp64, fp64_fast, mixed).
! Voronoi grids are small (the host re-triangulation costs more than the
! sum), so the exact mode costs little here.
! Where the reference's module reads freed memory (a step after ResetRK4,
! which frees the stage arrays and the saved areas but leaves its ready
! flags set, :177-212) this module starts over: arrays and saved areas are
! taken again at the next step.
module VoronoiVelocityParallelModule

use iso_c_binding
use NumberKindsModule
use LoggerModule
use VoronoiPanelsModule
use TracerAndVorticityDistributionModule, only : Omega
use BVEInterfaceModule

implicit none

private
public totalKE
public kineticEnergy
public BVERK4, ResetRK4
public ResetArea
public UseSavedArea
public InitializeMPIRK4, FinalizeMPIRK4, LoadBalance

real(kreal), pointer, save :: kineticEnergy(:) => null()
real(kreal), save :: totalKE = 0.0_kreal
real(kreal), allocatable, save :: savedArea(:)
logical(klog), save :: resetAreas = .False.
logical(klog), save :: saveArea = .False.
logical(klog), save :: rk4IsReady = .False.
logical(klog), save :: mpiIsReady = .False.

type(c_ptr), save :: ctx = c_null_ptr

logical(klog), save :: logInit = .False.
integer(kint), parameter :: logLevel = TRACE_LOGGING_LEVEL
character(len=28), save :: logKey = 'VoronoiVelocityGPU : '
type(Logger) :: log

contains

subroutine UseSavedArea(onOff)
	logical(klog), intent(in), optional :: onOff
	if (present(onOff)) then
		saveArea = onOff
	else
		saveArea = .True.
	endif
end subroutine

subroutine ResetArea(onOff)
	logical(klog), intent(in), optional :: onOff
	if (present(onOff)) then
		resetAreas = onOff
	else
		resetAreas = .True.
	endif
end subroutine

subroutine startLog()
	if (.not. logInit) then
		call New(log, logLevel, 6)
		logInit = .True.
	endif
end subroutine

subroutine check(rc, where)
	integer(c_int), intent(in) :: rc
	character(len=*), intent(in) :: where
	if (rc /= BVE_OK) then
		call LogMessage(log, ERROR_LOGGING_LEVEL, logKey, where//' : '//trim(bve_error_message(ctx)))
		error stop 1
	endif
end subroutine

subroutine InitializeMPIRK4(nPanels, procRank, numProcs)
	integer(kint), intent(in) :: nPanels, procRank, numProcs
	call startLog()
	if (.not. c_associated(ctx)) then
		if (bve_create(ctx, 0_c_int) /= BVE_OK) then
			call LogMessage(log, ERROR_LOGGING_LEVEL, logKey, 'bve_create : '// &
				trim(bve_error_message(c_null_ptr)))
			error stop 1
		endif
		call check(bve_set_precision(ctx, bve_precision_from_env(BVE_PRECISION_REFERENCE), &
			0.02_c_double), 'InitializeMPIRK4 (BVE_PRECISION)')
	endif
	mpiIsReady = .True.
	call LoadBalance(nPanels, numProcs)
end subroutine

subroutine LoadBalance(n, p)
	integer(kint), intent(in) :: n, p
	if (.not. mpiIsReady) then
		call startLog()
		call LogMessage(log, ERROR_LOGGING_LEVEL, logKey, 'LoadBalance ERROR : MPI arrays not allocated.')
	endif
	! one device evaluates every generator: nothing to partition
end subroutine

subroutine FinalizeMPIRK4()
	call ResetRK4()
	if (c_associated(ctx)) call check(bve_destroy(ctx), 'FinalizeMPIRK4')
	ctx = c_null_ptr
	mpiIsReady = .False.
	if (logInit) call Delete(log)
	logInit = .False.
end subroutine

subroutine InitRK4(aPanels)
	type(VorPanels), intent(in) :: aPanels
	allocate(kineticEnergy(aPanels%N_Max))
	kineticEnergy = 0.0_kreal
	if (saveArea .and. .not. allocated(savedArea)) then
		allocate(savedArea(aPanels%N_Max))
		savedArea = aPanels%area
	endif
	rk4IsReady = .True.
end subroutine

subroutine ResetRK4()
	if (.not. rk4IsReady) return
	deallocate(kineticEnergy)
	if (allocated(savedArea)) deallocate(savedArea)
	rk4IsReady = .False.
end subroutine

subroutine BVERK4(aPanels, dt, procRank, numProcs)
	type(VorPanels), intent(inout) :: aPanels
	real(kreal), intent(in) :: dt
	integer(kint) :: procRank, numProcs
	integer(kint) :: n
	real(c_double) :: deviceKE
	if (.not. mpiIsReady) then
		call startLog()
		call LogMessage(log, ERROR_LOGGING_LEVEL, logKey, 'RK4 ERROR : MPI not ready.')
		return
	endif
	if (.not. rk4IsReady) call InitRK4(aPanels)
	n = aPanels%N
	call check(bve_set_omega(ctx, Omega), 'BVERK4')
	if (saveArea) then
		call check(bve_voronoi_rk4_step(ctx, dt, int(n, c_int), aPanels%x, aPanels%y, aPanels%z, &
			aPanels%relVort, savedArea, deviceKE, kineticEnergy), 'BVERK4')
	else
		call check(bve_voronoi_rk4_step(ctx, dt, int(n, c_int), aPanels%x, aPanels%y, aPanels%z, &
			aPanels%relVort, aPanels%area, deviceKE, kineticEnergy), 'BVERK4')
	endif
	! the integral uses the current tessellation's areas, also after UseSavedArea
	totalKE = 0.5_kreal*sum(kineticEnergy(1:aPanels%N)*aPanels%area(1:aPanels%N))
	! connectivity for the new positions, as the reference does after NormalizeGrid
	call DelaunayTri(aPanels)
	call VoronoiGrid(aPanels)
end subroutine

end module VoronoiVelocityParallelModule
