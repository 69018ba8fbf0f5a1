! MI355X drop-in for the reference's serial panel module
! PanelVelocitySerialModule (PanelVelocitySerial.f90): same module name and
! public list (:12-17), so the serial panel driver (RHWavePanels.f90)
! compiles and links against it unchanged.  (The reference's own module
! does not compile with LLVM flang: it declares InitGaussianVortex,
! GaussianVortexX and HaurwitzStationary4RelVort itself and also takes them
! from TracerAndVorticityDistributionModule, :7.)
!
! BVERK4(aParticles, aPanels, dt) is the one-rank case of the MPI module's
! BVERK4 -- the same gather, stage loop, rounding and scatter
! (PanelVelocitySerial.f90:272-425 against PanelVelocityParallel.f90:435-658)
! -- so it forwards to the GPU PanelVelocityParallelModule with procRank 0
! of 1.  The vorticity profiles the serial module carries (InitRH4,
! InitGaussianVortex and their point functions, :564-655) are the same
! statements as the shared TracerAndVorticityDistributionModule's
! InitRH4WavePanels / InitGaussianVortexPanels
! (TracerAndVorticityDistributions.f90:142-156, :243-276); omega and
! gaussConst are the shared module's variables.  smooth is this module's
! public variable, as in the reference (:24), and is handed to the device
! at every step, so a driver that assigns it directly is followed too.
module PanelVelocitySerialModule

use NumberKindsModule
use ParticlesEdgesPanelsModule
use TracerAndVorticityDistributionModule, only : omega, SetOmega, gaussConst, &
	InitGaussianVortex, GaussianVortexX, InitRH4 => InitRH4Wave, HaurwitzStationary4RelVort
use PanelVelocityParallelModule, only : totalKE, kineticEnergy, ResetArea, SetSmooth, &
	InitializeMPIRK4, ParallelBVERK4 => BVERK4, ParallelResetRK4 => ResetRK4

implicit none

private
public BVERK4, ResetRK4
public omega, smooth
public InitGaussianVortex, GaussianVortexX, gaussConst
public InitRH4, HaurwitzStationary4RelVort
public totalKE, kineticEnergy
public SetOmega, SetSmoothingParameter, ResetArea

real(kreal), save :: smooth = 0.05_kreal
logical(klog), save :: started = .False.

contains

subroutine SetSmoothingParameter(newSmooth)
	real(kreal), intent(in) :: newSmooth
	smooth = newSmooth
end subroutine

subroutine BVERK4(aParticles, aPanels, dt)
	type(Particles), intent(inout) :: aParticles
	type(Panels), intent(inout) :: aPanels
	real(kreal), intent(in) :: dt
	if (.not. started) then
		call InitializeMPIRK4(aParticles, aPanels, 0, 1)
		started = .True.
	endif
	call SetSmooth(smooth)
	call ParallelBVERK4(aParticles, aPanels, dt, 0, 1)
end subroutine

subroutine ResetRK4()
	call ParallelResetRK4()
end subroutine

end module PanelVelocitySerialModule
