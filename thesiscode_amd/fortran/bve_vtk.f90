! vtkOutputNative: the reference's vtkOutput (vtkOutputUnif,
! ParticlesEdgesPanels2.f90:2116-2260) for its own Particles / Panels types,
! written by libbve's native, multi-threaded writer (bve_vtk_write_frame).
! The file is byte for byte the one the reference's routine writes; a driver
! switches by replacing `call vtkOutput(gridParticles, gridPanels, vtkFile)`
! with `call vtkOutputNative(gridParticles, gridPanels, vtkFile)`.
module BVEVTKModule
	use iso_c_binding
	use NumberKindsModule
	use ParticlesEdgesPanelsModule
	use BVEInterfaceModule
	implicit none
	private
	public :: vtkOutputNative

contains

	subroutine vtkOutputNative(aParticles, aPanels, filename)
		type(Particles), intent(in), target :: aParticles
		type(Panels), intent(in), target :: aPanels
		character(len=*), intent(in) :: filename
		type(c_ptr) :: prv, pav, ptr, qtr, qrv, qav, hc, xp, vp
		integer(c_int) :: nTracer, rc, ldP, ldQ

		xp = c_null_ptr
		vp = c_null_ptr
		hc = c_null_ptr
		if (aParticles%N > 0) xp = c_loc(aParticles%x(1,1))
		if (aPanels%N > 0) then
			vp = c_loc(aPanels%vertices(1,1))
			hc = c_loc(aPanels%hasChildren(1))
		endif
		prv = c_null_ptr
		pav = c_null_ptr
		ptr = c_null_ptr
		qtr = c_null_ptr
		qrv = c_null_ptr
		qav = c_null_ptr
		nTracer = 0
		ldP = 0
		ldQ = 0
		if (associated(aParticles%relVort)) prv = c_loc(aParticles%relVort(1))
		if (associated(aParticles%absVort)) pav = c_loc(aParticles%absVort(1))
		if (associated(aParticles%tracer)) then
			nTracer = size(aParticles%tracer, 2)
			ldP = size(aParticles%tracer, 1)
			ptr = c_loc(aParticles%tracer(1,1))
		endif
		if (associated(aPanels%tracer)) then
			ldQ = size(aPanels%tracer, 1)
			qtr = c_loc(aPanels%tracer(1,1))
		endif
		if (associated(aPanels%relVort)) qrv = c_loc(aPanels%relVort(1))
		if (associated(aPanels%absVort)) qav = c_loc(aPanels%absVort(1))
		rc = bve_vtk_write_frame(trim(filename)//c_null_char, int(aParticles%N, c_int), ldP, xp, &
			prv, pav, nTracer, ptr, int(GetPanelKind(aPanels), c_int), int(aPanels%N, c_int), &
			int(aPanels%N_Active, c_int), ldQ, vp, hc, qtr, qrv, qav)
		if (rc /= BVE_OK) then
			print *, 'vtkOutputNative ERROR : ', trim(bve_error_message(c_null_ptr))
			stop 'vtkOutput ERROR: cannot open file.'
		endif
	end subroutine

end module BVEVTKModule
