! Writes the same frame twice, once with the reference's vtkOutput
! (ParticlesEdgesPanels2.f90:2116-2260, compiled from its unmodified
! sources) and once with vtkOutputNative (thesiscode_amd/fortran/bve_vtk.f90
! over libbve), and prints the wall time of each writer.
!
! Usage:  vtk_native_check PANELKIND LEVEL REFFILE NATIVEFILE
!   PANELKIND 3 (icosahedral triangles, advection problem, 2 tracers) or
!   4 (cubed sphere, BVE problem with the RH4 wave, 3 tracers)
program vtk_native_check
	use NumberKindsModule
	use ParticlesEdgesPanelsModule
	use TracerAndVorticityDistributionModule
	use BVEVTKModule
	implicit none

	type(Particles), pointer :: gp
	type(Edges), pointer :: ge
	type(Panels), pointer :: gpan
	integer(kint) :: kind, level, j
	integer(8) :: c0, c1, c2, rate
	character(len=512) :: arg, refFile, nativeFile

	call get_command_argument(1, arg); read(arg, *) kind
	call get_command_argument(2, arg); read(arg, *) level
	call get_command_argument(3, refFile)
	call get_command_argument(4, nativeFile)

	if (kind == TRI_PANEL) then
		call New(gp, ge, gpan, TRI_PANEL, level, 0, 2, ADVECTION_SOLVER)
	else
		call New(gp, ge, gpan, QUAD_PANEL, level, 0, 3, BVE_SOLVER)
		call InitRH4Wave(gp, gpan)
	endif
	do j = 1, gp%N
		gp%tracer(j,1) = atan2(gp%x(2,j), gp%x(1,j))
		gp%tracer(j,2) = gp%x(3,j)**3
		if (kind /= TRI_PANEL) gp%tracer(j,3) = 1.0e-7_kreal*gp%relVort(j)
	enddo
	do j = 1, gpan%N
		gpan%tracer(j,1) = atan2(gpan%x(2,j), gpan%x(1,j))
		gpan%tracer(j,2) = -gpan%x(3,j)**3
		if (kind /= TRI_PANEL) gpan%tracer(j,3) = 1.0e-7_kreal*gpan%relVort(j)
	enddo

	call system_clock(c0, rate)
	call vtkOutput(gp, gpan, trim(refFile))
	call system_clock(c1)
	call vtkOutputNative(gp, gpan, trim(nativeFile))
	call system_clock(c2)
	print '(A,I8,A,I8,A,F10.4,A,F10.4)', 'panels ', gpan%N_Active, ' particles ', gp%N, &
		' reference_s ', real(c1 - c0, 8)/real(rate, 8), ' native_s ', real(c2 - c1, 8)/real(rate, 8)
end program
