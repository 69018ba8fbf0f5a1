! ISO_C_BINDING interface to libbve (include/bve.h): the MI355X Biot-Savart /
! RK4 library.  Array arguments are the reference's column-major x(3,N)
! position arrays and plain vectors; every function returns BVE_OK (0) or
! an error code whose message bve_error_message() returns.
module BVEInterfaceModule
	use iso_c_binding
	implicit none
	private
	public :: BVE_OK, BVE_PRECISION_FP64, BVE_PRECISION_MIXED, BVE_PRECISION_FP64_FAST
	public :: BVE_PRECISION_MIXED_KAHAN, BVE_PRECISION_REFERENCE, BVE_PRECISION_MIXED_BOUNDED
	public :: bve_precision_from_env, bve_set_error_target
	public :: bve_create, bve_destroy, bve_set_omega, bve_set_smooth, bve_set_precision
	public :: bve_set_renormalize, bve_load_panels, bve_rk4_step, bve_store_panels, bve_reset
	public :: bve_load_panel_absvort, bve_stratosphere_rk4_step, bve_advection_rk4_step
	public :: bve_load_particles, bve_advance, bve_store_particles, bve_synchronize
	public :: bve_voronoi_rk4_step, bve_error_message, bve_vtk_write_frame

	integer(c_int), parameter :: BVE_OK = 0
	integer(c_int), parameter :: BVE_PRECISION_FP64 = 0, BVE_PRECISION_MIXED = 1, &
	                             BVE_PRECISION_FP64_FAST = 2, BVE_PRECISION_MIXED_KAHAN = 3, &
	                             BVE_PRECISION_REFERENCE = 4, BVE_PRECISION_MIXED_BOUNDED = 5

	interface
		integer(c_int) function bve_create(ctx, device) bind(C, name="bve_create")
			import :: c_int, c_ptr
			type(c_ptr), intent(out) :: ctx
			integer(c_int), value :: device
		end function

		integer(c_int) function bve_destroy(ctx) bind(C, name="bve_destroy")
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function

		type(c_ptr) function bve_last_error(ctx) bind(C, name="bve_last_error")
			import :: c_ptr
			type(c_ptr), value :: ctx
		end function

		integer(c_int) function bve_set_omega(ctx, omega) bind(C, name="bve_set_omega")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), value :: omega
		end function

		integer(c_int) function bve_set_smooth(ctx, smooth) bind(C, name="bve_set_smooth")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), value :: smooth
		end function

		integer(c_int) function bve_set_precision(ctx, precision, nearfield) &
				bind(C, name="bve_set_precision")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			integer(c_int), value :: precision
			real(c_double), value :: nearfield
		end function

		integer(c_int) function bve_set_error_target(ctx, rel_err) &
				bind(C, name="bve_set_error_target")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), value :: rel_err
		end function

		integer(c_int) function bve_set_renormalize(ctx, on) bind(C, name="bve_set_renormalize")
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
			integer(c_int), value :: on
		end function

		integer(c_int) function bve_load_panels(ctx, nActive, xActive, relVort, area, nPassive, &
				xPassive) bind(C, name="bve_load_panels")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			integer(c_int), value :: nActive, nPassive
			real(c_double), intent(in) :: xActive(3,*), relVort(*), area(*), xPassive(3,*)
		end function

		integer(c_int) function bve_rk4_step(ctx, dt, nParticles, xParticles, absVort, relVort, &
				totalKE, ke) bind(C, name="bve_rk4_step")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), value :: dt
			integer(c_int), value :: nParticles
			real(c_double), intent(inout) :: xParticles(3,*)
			real(c_double), intent(in) :: absVort(*)
			real(c_double), intent(out) :: relVort(*), totalKE, ke(*)
		end function

		integer(c_int) function bve_load_panel_absvort(ctx, absVort) &
				bind(C, name="bve_load_panel_absvort")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), intent(in) :: absVort(*)
		end function

		integer(c_int) function bve_stratosphere_rk4_step(ctx, t, dt, nParticles, xParticles, &
				absVort, relVort, totalKE, ke) bind(C, name="bve_stratosphere_rk4_step")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), value :: t, dt
			integer(c_int), value :: nParticles
			real(c_double), intent(inout) :: xParticles(3,*)
			real(c_double), intent(in) :: absVort(*)
			real(c_double), intent(out) :: relVort(*), totalKE, ke(*)
		end function

		integer(c_int) function bve_advection_rk4_step(ctx, t, dt, divergent, nParticles, &
				xParticles) bind(C, name="bve_advection_rk4_step")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), value :: t, dt
			integer(c_int), value :: divergent, nParticles
			real(c_double), intent(inout) :: xParticles(3,*)
		end function

		integer(c_int) function bve_store_panels(ctx, xActive, relVort, xPassive) &
				bind(C, name="bve_store_panels")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), intent(out) :: xActive(3,*), relVort(*), xPassive(3,*)
		end function

		integer(c_int) function bve_reset(ctx) bind(C, name="bve_reset")
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function

		integer(c_int) function bve_load_particles(ctx, nParticles, xParticles, absVort) &
				bind(C, name="bve_load_particles")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			integer(c_int), value :: nParticles
			real(c_double), intent(in) :: xParticles(3,*), absVort(*)
		end function

		integer(c_int) function bve_advance(ctx, dt, nSteps) bind(C, name="bve_advance")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), value :: dt
			integer(c_int), value :: nSteps
		end function

		integer(c_int) function bve_store_particles(ctx, xParticles, relVort) &
				bind(C, name="bve_store_particles")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), intent(out) :: xParticles(3,*), relVort(*)
		end function

		integer(c_int) function bve_synchronize(ctx) bind(C, name="bve_synchronize")
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function

		integer(c_int) function bve_voronoi_rk4_step(ctx, dt, n, x, y, z, relVort, area, totalKE, &
				ke) bind(C, name="bve_voronoi_rk4_step")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), value :: dt
			integer(c_int), value :: n
			real(c_double), intent(inout) :: x(*), y(*), z(*), relVort(*)
			real(c_double), intent(in) :: area(*)
			real(c_double), intent(out) :: totalKE, ke(*)
		end function

		! One VTK frame in vtkOutputUnif's format; arrays that may be absent
		! (vorticity, tracers, hasChildren) are passed as c_loc(...) or c_null_ptr.
		integer(c_int) function bve_vtk_write_frame(filename, nParticles, ldParticles, xParticles, &
				relVort, absVort, nTracer, tracer, panelKind, nPanels, nActive, ldPanels, vertices, &
				hasChildren, panelTracer, panelRelVort, panelAbsVort) bind(C, name="bve_vtk_write_frame")
			import :: c_int, c_ptr, c_char
			character(kind=c_char), intent(in) :: filename(*)
			integer(c_int), value :: nParticles, ldParticles, nTracer, panelKind, nPanels, nActive, &
				ldPanels
			type(c_ptr), value :: xParticles, relVort, absVort, tracer, vertices, hasChildren, &
				panelTracer, panelRelVort, panelAbsVort
		end function
	end interface

contains

	! The library's last error message for ctx as a Fortran string.
	function bve_error_message(ctx) result(msg)
		type(c_ptr), intent(in) :: ctx
		character(len=512) :: msg
		character(kind=c_char), pointer :: chars(:)
		type(c_ptr) :: p
		integer :: k
		msg = ' '
		p = bve_last_error(ctx)
		if (.not. c_associated(p)) return
		call c_f_pointer(p, chars, [len(msg)])
		do k = 1, len(msg)
			if (chars(k) == c_null_char) exit
			msg(k:k) = chars(k)
		enddo
	end function

	! Arithmetic the drop-in modules run: the environment variable
	! BVE_PRECISION (fp64, fp64_fast, mixed, mixed_kahan, mixed_bounded or
	! reference) when it is set, else `default`.  An unknown value is an error (-1).
	function bve_precision_from_env(default) result(precision)
		integer(c_int), intent(in) :: default
		integer(c_int) :: precision
		character(len=32) :: v
		integer :: n, stat
		call get_environment_variable('BVE_PRECISION', v, n, stat)
		if (stat /= 0 .or. n == 0) then
			precision = default
			return
		endif
		select case (trim(v))
		case ('fp64')
			precision = BVE_PRECISION_FP64
		case ('fp64_fast')
			precision = BVE_PRECISION_FP64_FAST
		case ('mixed')
			precision = BVE_PRECISION_MIXED
		case ('mixed_kahan')
			precision = BVE_PRECISION_MIXED_KAHAN
		case ('mixed_bounded')
			precision = BVE_PRECISION_MIXED_BOUNDED
		case ('reference')
			precision = BVE_PRECISION_REFERENCE
		case default
			precision = -1
		end select
	end function

end module BVEInterfaceModule
