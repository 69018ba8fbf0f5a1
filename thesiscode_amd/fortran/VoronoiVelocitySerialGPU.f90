! MI355X drop-in for the reference's serial Voronoi module
! VoronoiVelocitySerialModule (VoronoiVelocitySerial2.f90, and the smaller
! public list of VoronoiVelocitySerial.f90): same module name and public
! list, so the serial Voronoi drivers (RHWaveVoronoi.f90) compile and link
! against it unchanged.
!
! BVERK4(aPanels, dt) is the one-rank case of the MPI module's BVERK4 --
! the same stage loop, rounding and re-triangulation (VoronoiVelocitySerial2
! .f90:242-369 against VoronoiVelocityParallel.f90:304-483) -- so it forwards
! to the GPU VoronoiVelocityParallelModule with procRank 0 of 1.  The
! vorticity profiles the serial module carries (InitRH4Wave,
! InitGaussianVortex and their point functions, :407-462) are the same
! statements as the shared TracerAndVorticityDistributionModule's
! (TracerAndVorticityDistributions.f90:176-183, :306-319, :415-426), and the
! rotation rate is the shared Omega, which SetOmega sets for the profiles
! and the step alike, as the serial module's own Omega does.
module VoronoiVelocitySerialModule

use NumberKindsModule
use VoronoiPanelsModule
use TracerAndVorticityDistributionModule, only : Omega, SetOmega, InitGaussianVortex, &
	GaussianVortexX, InitRH4Wave, HaurwitzStationary4Relvort
use VoronoiVelocityParallelModule, only : totalKE, kineticEnergy, UseSavedArea, ResetArea, &
	InitializeMPIRK4, ParallelBVERK4 => BVERK4, ParallelResetRK4 => ResetRK4

implicit none

private
public totalKE
public kineticEnergy
public BVERK4, ResetRK4
public SetOmega, GetOmega
public ResetArea
public InitGaussianVortex, GaussianVortexX
public InitRH4Wave, HaurwitzStationary4Relvort
public UseSavedArea

logical(klog), save :: started = .False.

contains

function GetOmega()
	real(kreal) :: GetOmega
	GetOmega = Omega
end function

subroutine BVERK4(aPanels, dt)
	type(VorPanels), intent(inout) :: aPanels
	real(kreal), intent(in) :: dt
	if (.not. started) then
		call InitializeMPIRK4(aPanels%N, 0, 1)
		started = .True.
	endif
	call ParallelBVERK4(aPanels, dt, 0, 1)
end subroutine

subroutine ResetRK4()
	call ParallelResetRK4()
end subroutine

end module VoronoiVelocitySerialModule
