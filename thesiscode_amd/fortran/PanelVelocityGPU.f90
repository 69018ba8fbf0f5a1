! MI355X drop-in for the reference's PanelVelocityParallelModule
! (PanelVelocityParallel.f90): same module name and public list, so the
! reference's drivers (RHWavePanelsMPI.f90, GaussVort.f90, JetPanels.f90,
! ...) compile and link against it unchanged.  The Biot-Savart sums and
! the RK4 stage loop run on the GPU through libbve (include/bve.h); the grid
! types, GatherPanels / ScatterPanels, Omega and the logger stay the
! reference's own.
!
! Semantics kept (PanelVelocityParallel.f90:435-658):
!   * the active / passive panels are gathered on the first BVERK4 after a
!     reset and the device copy is authoritative until ResetRK4; the
!     scatter after every step overwrites the host tree (:656);
!   * particle positions are read from the host at every step (:458) and
!     relVort = absVort - 2 Omega z is set after it (:651);
!   * kineticEnergy / totalKE are those of stage 1 (:495-501), copied to
!     panel tracer 2 when there are at least two tracers.
! StratosphereRK4 (:661-919, Juckes forcing) and AdvectionRK4 /
! DivergentAdvectionRK4 (:922-1279, analytic winds) run on the same cached
! panels.
! procRank and numProcs are accepted; one GPU computes every target.
! The arithmetic is float64 (within 1e-12 of the reference); the environment
! variable BVE_PRECISION selects another mode, and BVE_PRECISION=reference
! reproduces the reference's BVERK4 bit for bit (the sums in the reference's
! order and rounding), so runs through host remeshing follow it exactly.
module PanelVelocityParallelModule

use iso_c_binding
use NumberKindsModule
use LoggerModule
use ParticlesEdgesPanelsModule
use TracerAndVorticityDistributionModule, only : Omega
use BVEInterfaceModule

implicit none

private
public totalKE
public kineticEnergy
public BVERK4, AdvectionRK4, DivergentAdvectionRK4, StratosphereRK4
public ResetRK4
public ResetArea
public SetSmooth, GetSmooth
public InitializeMPIRK4, FinalizeMPIRK4, LoadBalance

real(kreal), allocatable, save :: kineticEnergy(:)
real(kreal), save :: totalKE = 0.0_kreal
real(kreal), save :: smooth = 0.05_kreal
logical(klog), save :: resetAreas = .False.
logical(klog), save :: rk4IsReady = .False.
logical(klog), save :: mpiIsReady = .False.

type(c_ptr), save :: ctx = c_null_ptr
type(Panels), pointer, save :: activePanels => null(), passivePanels => null()
integer(kint), allocatable, save :: activeMap(:), passiveMap(:)

logical(klog), save :: logInit = .False.
integer(kint), parameter :: logLevel = TRACE_LOGGING_LEVEL
character(len=28), save :: logKey = 'PanelVelocityGPU : '
type(Logger) :: log

contains

subroutine SetSmooth(newSmooth)
	real(kreal), intent(in) :: newSmooth
	smooth = newSmooth
	if (c_associated(ctx)) call check(bve_set_smooth(ctx, smooth), 'SetSmooth')
end subroutine

function GetSmooth()
	real(kreal) :: GetSmooth
	GetSmooth = smooth
end function

subroutine ResetArea(onOff)
	logical(klog), intent(in), optional :: onOff
	if (present(onOff)) then
		resetAreas = onOff
	else
		resetAreas = .True.
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

subroutine InitializeMPIRK4(aParticles, aPanels, procRank, numProcs)
	type(Particles), intent(in) :: aParticles
	type(Panels), intent(in) :: aPanels
	integer(kint), intent(in) :: procRank, numProcs
	if (.not. logInit) then
		call New(log, logLevel, 6)
		logInit = .True.
	endif
	if (.not. c_associated(ctx)) then
		if (bve_create(ctx, 0_c_int) /= BVE_OK) then
			call LogMessage(log, ERROR_LOGGING_LEVEL, logKey, 'bve_create : '// &
				trim(bve_error_message(c_null_ptr)))
			error stop 1
		endif
		! float64 unless BVE_PRECISION asks otherwise (reference: the
		! reference's rounding bit for bit)
		call check(bve_set_precision(ctx, bve_precision_from_env(BVE_PRECISION_FP64), &
			0.02_c_double), 'InitializeMPIRK4 (BVE_PRECISION)')
	endif
	mpiIsReady = .True.
	call LoadBalance(aParticles, aPanels, numProcs)
end subroutine

subroutine LoadBalance(aParticles, aPanels, numProcs)
	type(Particles), intent(in) :: aParticles
	type(Panels), intent(in) :: aPanels
	integer(kint), intent(in) :: numProcs
	if (.not. mpiIsReady) then
		call LogMessage(log, ERROR_LOGGING_LEVEL, logKey, 'LoadBalance ERROR : MPI arrays not ready.')
	endif
	! one device evaluates every target: nothing to partition
end subroutine

subroutine FinalizeMPIRK4()
	if (rk4IsReady) call ResetRK4()
	if (c_associated(ctx)) call check(bve_destroy(ctx), 'FinalizeMPIRK4')
	ctx = c_null_ptr
	mpiIsReady = .False.
	if (logInit) call Delete(log)
	logInit = .False.
end subroutine

subroutine InitRK4(aParticles, aPanels)
	type(Particles), intent(in) :: aParticles
	type(Panels), intent(in) :: aPanels
	integer(kint) :: nActive, nPassive
	nActive = aPanels%N_Active
	nPassive = aPanels%N - aPanels%N_Active
	allocate(activePanels, passivePanels)
	allocate(activeMap(nActive), passiveMap(nPassive))
	call New(activePanels, nActive, GetPanelKind(aPanels), GetNTracer(aPanels), GetProblemKind(aPanels))
	call New(passivePanels, nPassive, GetPanelKind(aPanels), GetNTracer(aPanels), GetProblemKind(aPanels))
	activePanels%N = nActive
	activePanels%N_Active = nActive
	passivePanels%N = nPassive
	call GatherPanels(aPanels, activePanels, activeMap, passivePanels, passiveMap)
	call check(bve_set_omega(ctx, Omega), 'InitRK4')
	call check(bve_set_smooth(ctx, smooth), 'InitRK4')
	call check(bve_load_panels(ctx, int(nActive, c_int), activePanels%x, activePanels%relVort, &
		activePanels%area, int(nPassive, c_int), passivePanels%x), 'InitRK4')
	! the forced step (StratosphereRK4) works from the panels' absolute vorticity
	call check(bve_load_panel_absvort(ctx, activePanels%absVort), 'InitRK4')
	allocate(kineticEnergy(nActive))
	kineticEnergy = 0.0_kreal
	rk4IsReady = .True.
end subroutine

subroutine ResetRK4()
	if (.not. rk4IsReady) return
	call check(bve_reset(ctx), 'ResetRK4')
	deallocate(kineticEnergy)
	deallocate(activeMap, passiveMap)
	call Delete(activePanels)
	call Delete(passivePanels)
	deallocate(activePanels, passivePanels)
	rk4IsReady = .False.
end subroutine

subroutine BVERK4(aParticles, aPanels, dt, procRank, numProcs)
	type(Particles), intent(inout) :: aParticles
	type(Panels), intent(inout) :: aPanels
	real(kreal), intent(in) :: dt
	integer(kint), intent(in) :: procRank, numProcs
	integer(kint) :: n
	if (.not. c_associated(ctx)) call InitializeMPIRK4(aParticles, aPanels, procRank, numProcs)
	if (.not. rk4IsReady) call InitRK4(aParticles, aPanels)
	n = aParticles%N
	call check(bve_set_omega(ctx, Omega), 'BVERK4')
	call check(bve_rk4_step(ctx, dt, int(n, c_int), aParticles%x(:,1:n), aParticles%absVort(1:n), &
		aParticles%relVort(1:n), totalKE, kineticEnergy), 'BVERK4')
	call check(bve_store_panels(ctx, activePanels%x, activePanels%relVort, passivePanels%x), 'BVERK4')
	if (GetNTracer(aPanels) >= 2) activePanels%tracer(:,2) = kineticEnergy
	call ScatterPanels(aPanels, activePanels, activeMap, passivePanels, passiveMap)
end subroutine

! StratosphereRK4 (PanelVelocityParallel.f90:661-919): the same cached
! panels and stage loop on the GPU, with every stage's active vorticity
! absVort - 2 Omega z - F(x, t_s) and the Juckes forcing F; afterwards the
! active panels' and particles' relVort = absVort - 2 Omega z - F(x, t + dt).
subroutine StratosphereRK4(aParticles, aPanels, t, dt, procRank, numProcs)
	type(Particles), intent(inout) :: aParticles
	type(Panels), intent(inout) :: aPanels
	real(kreal), intent(in) :: t, dt
	integer(kint), intent(in) :: procRank, numProcs
	integer(kint) :: n
	if (.not. c_associated(ctx)) call InitializeMPIRK4(aParticles, aPanels, procRank, numProcs)
	if (.not. rk4IsReady) call InitRK4(aParticles, aPanels)
	n = aParticles%N
	call check(bve_set_omega(ctx, Omega), 'StratosphereRK4')
	call check(bve_stratosphere_rk4_step(ctx, t, dt, int(n, c_int), aParticles%x(:,1:n), &
		aParticles%absVort(1:n), aParticles%relVort(1:n), totalKE, kineticEnergy), 'StratosphereRK4')
	call check(bve_store_panels(ctx, activePanels%x, activePanels%relVort, passivePanels%x), &
		'StratosphereRK4')
	if (GetNTracer(aPanels) >= 2) activePanels%tracer(:,2) = kineticEnergy
	call ScatterPanels(aPanels, activePanels, activeMap, passivePanels, passiveMap)
end subroutine

! AdvectionRK4 / DivergentAdvectionRK4 (PanelVelocityParallel.f90:922-1279):
! every cached panel and the particles move in the analytic wind; vorticity,
! tracers and the kinetic energy are left as they are.
subroutine AdvectionRK4(aParticles, aPanels, t, dt, procRank, numProcs)
	type(Particles), intent(inout) :: aParticles
	type(Panels), intent(inout) :: aPanels
	real(kreal), intent(in) :: t, dt
	integer(kint), intent(in) :: procRank, numProcs
	call AdvectWithWind(aParticles, aPanels, t, dt, 0_c_int, procRank, numProcs)
end subroutine

subroutine DivergentAdvectionRK4(aParticles, aPanels, t, dt, procRank, numProcs)
	type(Particles), intent(inout) :: aParticles
	type(Panels), intent(inout) :: aPanels
	real(kreal), intent(in) :: t, dt
	integer(kint), intent(in) :: procRank, numProcs
	call AdvectWithWind(aParticles, aPanels, t, dt, 1_c_int, procRank, numProcs)
end subroutine

subroutine AdvectWithWind(aParticles, aPanels, t, dt, divergent, procRank, numProcs)
	type(Particles), intent(inout) :: aParticles
	type(Panels), intent(inout) :: aPanels
	real(kreal), intent(in) :: t, dt
	integer(c_int), intent(in) :: divergent
	integer(kint), intent(in) :: procRank, numProcs
	integer(kint) :: n
	if (.not. c_associated(ctx)) call InitializeMPIRK4(aParticles, aPanels, procRank, numProcs)
	if (.not. rk4IsReady) call InitRK4(aParticles, aPanels)
	n = aParticles%N
	call check(bve_advection_rk4_step(ctx, t, dt, divergent, int(n, c_int), aParticles%x(:,1:n)), &
		'AdvectionRK4')
	call check(bve_store_panels(ctx, activePanels%x, activePanels%relVort, passivePanels%x), &
		'AdvectionRK4')
	call ScatterPanels(aPanels, activePanels, activeMap, passivePanels, passiveMap)
end subroutine

end module PanelVelocityParallelModule
