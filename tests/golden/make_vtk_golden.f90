! VTK golden-frame harness: builds a grid with the reference's own grid code
! (compiled from its unmodified sources by oracle/Makefile.ref), fills it,
! writes one frame with the reference's vtkOutput (vtkOutputUnif,
! ParticlesEdgesPanels2.f90:2116-2260) and dumps the arrays that frame was
! written from, so a writer fed the same arrays can be compared byte for byte.
!
! Usage:  make_vtk_golden CASE VTKFILE DUMPFILE
!   CASE = 1  cubed sphere level 2 with eight panels divided once more
!             (undivided panels not contiguous in the panel order), RH4
!             wave, one BVERK4 step, three particle and panel tracers
!   CASE = 2  icosahedral triangles level 2, advection problem (no
!             vorticity arrays), two tracers
!   CASE = 3  Voronoi panel kind, synthetic: corner lists of 3 to 10
!             entries with large indices (list-directed records that wrap),
!             and particle values that probe the F24.18 field: signed
!             zeros, ties, sub-normals, values too wide, NaN and infinities
!
! Dump layout (stream access, native little-endian):
!   int32   panelKind, nParticles, nPanels, nActive, nTracer,
!           hasVort (1 when relVort / absVort are present)
!   float64 particles x(3,nParticles)
!   float64 particles relVort(nParticles), absVort(nParticles)    if hasVort
!   float64 particles tracer(nParticles,nTracer)
!   int32   panels vertices(panelKind,nPanels), hasChildren(nPanels) (0/1)
!   float64 panels tracer(nPanels,nTracer)
!   float64 panels relVort(nPanels), absVort(nPanels)              if hasVort
program make_vtk_golden
	use, intrinsic :: ieee_arithmetic
	use NumberKindsModule
	use ParticlesEdgesPanelsModule
	use TracerAndVorticityDistributionModule
	use PanelVelocityParallelModule
	implicit none
	include 'mpif.h'

	type(Particles), pointer :: gp
	type(Edges), pointer :: ge
	type(Panels), pointer :: gpan
	integer(kint) :: icase, ierr, j, k, nT, hasVort, nDiv
	character(len=512) :: arg, vtkFile, dumpFile
	integer(kint), allocatable :: hc(:)
	integer, parameter :: u = 22
	real(kreal) :: special(24)

	call MPI_INIT(ierr)
	call get_command_argument(1, arg); read(arg, *) icase
	call get_command_argument(2, vtkFile)
	call get_command_argument(3, dumpFile)

	select case (icase)
	case (1)
		nT = 3
		call New(gp, ge, gpan, QUAD_PANEL, 2, 1, nT, BVE_SOLVER)
		! divide every 12th undivided panel once more
		nDiv = 0
		do j = 1, gpan%N
			if (.not. gpan%hasChildren(j) .and. mod(j, 12) == 0 .and. nDiv < 8) then
				call DividePanel(gp, ge, gpan, j)
				nDiv = nDiv + 1
			endif
		enddo
		call InitRH4Wave(gp, gpan)
		call InitializeMPIRK4(gp, gpan, 0, 1)
		call BVERK4(gp, gpan, 0.01_kreal, 0, 1)
		do j = 1, gp%N
			gp%tracer(j,1) = gp%x0(3,j)
			gp%tracer(j,2) = 1.0e-9_kreal*gp%relVort(j)
			gp%tracer(j,3) = 12345.678_kreal*gp%x(1,j)
		enddo
		do j = 1, gpan%N
			gpan%tracer(j,1) = gpan%x0(3,j)
			gpan%tracer(j,2) = 1.0e-9_kreal*gpan%relVort(j)
			gpan%tracer(j,3) = -12345.678_kreal*gpan%x(2,j)
		enddo
		hasVort = 1
	case (2)
		nT = 2
		call New(gp, ge, gpan, TRI_PANEL, 2, 0, nT, ADVECTION_SOLVER)
		do j = 1, gp%N
			gp%tracer(j,1) = atan2(gp%x(2,j), gp%x(1,j))
			gp%tracer(j,2) = gp%x(3,j)**3
		enddo
		do j = 1, gpan%N
			gpan%tracer(j,1) = atan2(gpan%x(2,j), gpan%x(1,j))
			gpan%tracer(j,2) = gpan%x(3,j)**3
		enddo
		hasVort = 0
	case (3)
		nT = 1
		allocate(gp, gpan)
		call New(gp, 24, VORONOI_PANEL, nT, BVE_SOLVER)
		! the reference's panel constructor accepts only triangles and
		! quadrilaterals; vtkOutputUnif reads the kind from the first extent
		! of vertices and these six components, so they are allocated here
		allocate(gpan%vertices(VORONOI_PANEL, 12), gpan%hasChildren(12), gpan%tracer(12, nT), &
			gpan%relVort(12), gpan%absVort(12))
		special = [ -0.0_kreal, 1.0e-20_kreal, -1.0e-20_kreal, 2.0_kreal**(-19), -2.0_kreal**(-19), &
		            3.0_kreal*2.0_kreal**(-19), 99999.5_kreal, 100000.0_kreal, -9999.5_kreal, &
		            -10000.0_kreal, 0.1_kreal, 1.0_kreal/3.0_kreal, 4.9e-324_kreal, &
		            99999.99999999999_kreal, -0.5e-18_kreal, 1.5e-18_kreal, 2.5e-18_kreal, &
		            1.0e300_kreal, -123.456_kreal, 0.0_kreal, 0.0_kreal, 0.0_kreal, 7.0_kreal, 1.0_kreal ]
		special(20) = ieee_value(special(20), ieee_quiet_nan)
		special(21) = ieee_value(special(21), ieee_positive_inf)
		special(22) = ieee_value(special(22), ieee_negative_inf)
		gp%N = 24
		do j = 1, gp%N
			gp%x(1,j) = special(j)
			gp%x(2,j) = -special(25 - j)
			gp%x(3,j) = special(mod(j*7, 24) + 1)*0.001_kreal
			gp%relVort(j) = special(mod(j*5, 24) + 1)
			gp%absVort(j) = -special(j)
			gp%tracer(j,1) = real(j, kreal)/7.0_kreal
		enddo
		gpan%N = 12
		gpan%N_Active = 9
		gpan%vertices = 0
		do j = 1, gpan%N
			do k = 1, 3 + mod(j*3, 8)
				gpan%vertices(k,j) = 1 + mod(j*1234567 + k*7654321, 99999999)
			enddo
			gpan%hasChildren(j) = (mod(j, 4) == 0)
			gpan%relVort(j) = special(j + 4)
			gpan%absVort(j) = 1.0_kreal/real(j, kreal)
			gpan%tracer(j,1) = -real(j, kreal)/3.0_kreal
		enddo
		hasVort = 1
	case default
		stop 'make_vtk_golden: CASE must be 1, 2 or 3'
	end select

	call vtkOutput(gp, gpan, trim(vtkFile))

	allocate(hc(gpan%N))
	hc = 0
	where (gpan%hasChildren(1:gpan%N)) hc = 1
	open(unit=u, file=trim(dumpFile), access='stream', form='unformatted', status='replace')
	write(u) int(GetPanelKind(gpan), 4), int(gp%N, 4), int(gpan%N, 4), int(gpan%N_Active, 4), &
		int(nT, 4), int(hasVort, 4)
	write(u) gp%x(:, 1:gp%N)
	if (hasVort == 1) write(u) gp%relVort(1:gp%N), gp%absVort(1:gp%N)
	write(u) gp%tracer(1:gp%N, 1:nT)
	write(u) int(gpan%vertices(:, 1:gpan%N), 4), int(hc, 4)
	write(u) gpan%tracer(1:gpan%N, 1:nT)
	if (hasVort == 1) write(u) gpan%relVort(1:gpan%N), gpan%absVort(1:gpan%N)
	close(u)
	call MPI_FINALIZE(ierr)
end program
