/*
 * bve.h - C ABI of the MI355X Biot-Savart / RK4 library (libbve.so).
 *
 * Drop-in boundary for the velocity path of the barotropic-vorticity
 * particle-panel solver.  Every entry point takes plain pointers and sizes
 * (host memory unless stated otherwise), returns BVE_OK or an error code,
 * and leaves a message for bve_last_error().  Calls are synchronous at the
 * boundary: host buffers are read or written before the call returns.
 *
 * Array conventions (those of the reference's Fortran types,
 * ParticlesEdgesPanels2.f90:45-89):
 *   positions  double[3*n], point j at [3j, 3j+1, 3j+2]: the column-major x(3,N)
 *   scalars    double[n]
 *   active panels  = leaf panels in GatherPanels order (ParticlesEdgesPanels2.f90:1556-1619)
 *   passive panels = divided (parent) panels in the same order
 *
 * A context is one instance of the reference module's SAVE state: it is not
 * re-entrant; use one per host thread or per ensemble member.
 */
#ifndef BVE_H
#define BVE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bve_ctx bve_ctx;

enum {
    BVE_OK = 0,
    BVE_ERR_ARGUMENT = 1, /* bad size, null pointer or out-of-range value */
    BVE_ERR_STATE = 2,    /* call out of order (e.g. step before panels are loaded) */
    BVE_ERR_DEVICE = 3,   /* a HIP call failed */
    BVE_ERR_MEMORY = 4,   /* device allocation failed */
    BVE_ERR_IO = 5,       /* a file could not be opened or written */
    BVE_ERR_NUMERIC = 6   /* a velocity or position came out NaN or Inf (e.g. a source
                             and a target at the same point, 1 - x.y = 0, where the
                             reference divides 0 by 0); reported by the first
                             synchronising call after the step, with the first
                             offending target in bve_last_error */
};

enum {
    BVE_PRECISION_FP64 = 0,       /* float64 pairs, every 1/d refined to full precision */
    BVE_PRECISION_MIXED = 1,      /* float32 far field, float64 near field and reductions;
                                     float32 sums of each 256-source tile carried in float64 */
    BVE_PRECISION_FP64_FAST = 2,  /* float64 pairs, one reciprocal shared by 2 or 4 pairs */
    BVE_PRECISION_MIXED_KAHAN = 3, /* as MIXED with Kahan-compensated float32 sums */
    BVE_PRECISION_REFERENCE = 4,  /* float64 in the reference's own order and rounding: one
                                     target per lane, sources in index order, each term
                                     ((x_i x x_k) zeta_k) A_k / d_ik with IEEE divisions, no
                                     FMA contraction; BVERK4 results (panel and Voronoi) equal
                                     the reference's bit for bit */
    BVE_PRECISION_MIXED_BOUNDED = 5 /* float64 tile expansions with float32 remainders: each
                                     256-source tile's sum is its second-order float64
                                     expansion about the tile centre plus a float32 sum of
                                     the remainders delta^2 / (E^2 D); a (target group, tile)
                                     pair whose remainder rounding could exceed the error
                                     target (bve_set_error_target) runs in float64 instead */
};

/* ---- lifetime -------------------------------------------------------- */

/* New context on HIP device `device`.  No counterpart in the reference
 * (its state is module SAVE data, PanelVelocityParallel.f90:25-89). */
int bve_create(bve_ctx **ctx, int device);
int bve_destroy(bve_ctx *ctx);
/* Message of the last failed call on ctx; with ctx NULL, of the last failed
 * call that takes no context (bve_create, bve_vtk_write_frame) on this thread. */
const char *bve_last_error(const bve_ctx *ctx);

/* ---- parameters ------------------------------------------------------ */

/* Rotation rate Omega in dzeta/dt = -2 Omega u_z.  Replaces SetOmega,
 * TracerAndVorticityDistributions.f90:136-139 (default 2 PI per day, :34-35). */
int bve_set_omega(bve_ctx *ctx, double omega);
/* Regularisation of the passive-panel sum, denominator (1 - x.y) + smooth^2.
 * Replaces SetSmooth / GetSmooth, PanelVelocityParallel.f90:96-105 (default 0.05, :27). */
int bve_set_smooth(bve_ctx *ctx, double smooth);
int bve_get_smooth(const bve_ctx *ctx, double *smooth);
/* Arithmetic of the pair sum: BVE_PRECISION_FP64 (default),
 * BVE_PRECISION_FP64_FAST (float64, relative error per pair ~1e-15),
 * BVE_PRECISION_MIXED (float32 far field, float64 for pairs closer than
 * nearfield_radius, float64 target reduction) or BVE_PRECISION_REFERENCE
 * (the reference's rounding, bit for bit; about 6x the cost of FP64, for
 * runs that must reproduce a reference run exactly, e.g. through host
 * re-triangulation, which is sensitive to the last bits of the positions).
 * No reference counterpart. */
int bve_set_precision(bve_ctx *ctx, int precision, double nearfield_radius);
/* Normwise error target of BVE_PRECISION_MIXED_BOUNDED, relative to the
 * largest velocity of the pass (default 1e-10): each float32 remainder tile
 * is held to 0.3 of it times that velocity by a rounding bound, the rest
 * runs in float64.  No reference counterpart. */
int bve_set_error_target(bve_ctx *ctx, double rel_err);
/* 1 = project every position back onto the unit sphere after each step
 * (RenormalizeGrid, ParticlesEdgesPanels2.f90:1690-1702, applied on the
 * device); 0 (default) = the reference's BVERK4, which never does. */
int bve_set_renormalize(bve_ctx *ctx, int on);
/* Launch geometry of the pair kernel: targets held per lane (1, 2, 4),
 * sources per LDS tile (256, 512), the work split and, for
 * BVE_PRECISION_FP64_FAST, the pairs sharing one reciprocal (2, 4); 0
 * keeps the built-in choice.  The float64 and mixed modes run a
 * persistent launch over `target_workgroups` segments of the (target tile,
 * source tile) sequence (the mixed modes visit each target tile's source
 * tiles in a stride order).  Used by the autotuner. */
int bve_set_plan(bve_ctx *ctx, int targets_per_lane, int tile, int target_workgroups,
                 int recip);
/* chunk > 0: segments of `chunk` sources' worth of work (chunk / tile
 * consecutive (target tile, source tile) items of the tile-major sequence,
 * chunk rounded up to whole 256-source tiles), in every arithmetic mode --
 * the split ensemble members use, so a problem run alone with the same
 * chunk and targets per lane reproduces a member bit for bit.  0 (the
 * default) restores the built-in segment count. */
int bve_set_chunk(bve_ctx *ctx, int chunk);

/* The pair kernel's current launch geometry for the loaded problem:
 * targets per lane, sources per LDS tile, sources' worth of work per segment
 * (segment length x tile), segments (for the reference-order pass: one
 * chunk of all sources), pairs per shared reciprocal.  Any pointer may be
 * NULL. */
int bve_get_plan(const bve_ctx *ctx, int *targets_per_lane, int *tile, int *chunk,
                 int *n_chunks, int *recip);
/* Launch shape of the pair kernel: segments (0 for a chunked plan),
 * launched workgroups, target tiles, and the bytes of partial sums one pass
 * writes (each is read once by the stage update).  Any pointer may be NULL. */
int bve_get_launch(const bve_ctx *ctx, int *segments, int *grid, int *n_tiles,
                   long long *partial_bytes);

/* ---- panel RK4 path (drop-in for PanelVelocityParallelModule) --------- */

/* InitRK4 + GatherPanels payload (PanelVelocityParallel.f90:236-318): the
 * active panels' positions, relative vorticity and areas and the passive
 * panels' positions are copied to the device and cached there until
 * bve_reset, exactly as the module caches its private panel copies.  The
 * device holds the passive panels in a spatial order of its own (Hilbert
 * order per cube face); bve_store_panels returns them in the caller's. */
int bve_load_panels(bve_ctx *ctx, int n_active, const double *x_active,
                    const double *relvort_active, const double *area_active,
                    int n_passive, const double *x_passive);

/* BVERK4(aParticles, aPanels, dt, procRank, numProcs),
 * PanelVelocityParallel.f90:435-658: one classical RK4 step of the cached
 * panels and of the n_particles particles (read from x_particles every
 * call and overwritten with the new positions).  If absvort is given, the
 * particles' relative vorticity absVort - 2 Omega z is written to relvort
 * (:651).  total_ke (scalar) and ke (n_active) receive totalKE and
 * kineticEnergy of stage 1 (:495-501); any of them may be NULL. */
int bve_rk4_step(bve_ctx *ctx, double dt, int n_particles, double *x_particles,
                 const double *absvort, double *relvort, double *total_ke, double *ke);

/* The active panels' absolute vorticity (n_active values, gathered order),
 * which the forced step needs: activePanels%absVort of
 * PanelVelocityParallel.f90:686.  Cached until the next bve_load_panels. */
int bve_load_panel_absvort(bve_ctx *ctx, const double *absvort_active);

/* StratosphereRK4(aParticles, aPanels, t, dt, procRank, numProcs),
 * PanelVelocityParallel.f90:661-919: the RK4 step with the Juckes forcing
 * (TracerAndVorticityDistributions.f90:1223-1231).  Each stage's active
 * vorticity is absVort - 2 Omega z - F(x, t_s) at the stage positions,
 * t_s = t, t + dt/2, t + dt/2, t + dt; afterwards the active panels' and,
 * if absvort is given, the particles' relative vorticity is
 * absVort - 2 Omega z - F(x, t + dt) (:906-915).  Arguments otherwise as
 * bve_rk4_step; needs bve_load_panel_absvort. */
int bve_stratosphere_rk4_step(bve_ctx *ctx, double t, double dt, int n_particles,
                              double *x_particles, const double *absvort, double *relvort,
                              double *total_ke, double *ke);

/* AdvectionRK4 (divergent = 0) and DivergentAdvectionRK4 (divergent = 1)
 * (aParticles, aPanels, t, dt, procRank, numProcs),
 * PanelVelocityParallel.f90:922-1279: one RK4 step of every cached panel
 * and of the n_particles particles (read from x_particles and overwritten)
 * in the analytic NonDivergentWind / DivergentWind (:1415-1459), stages at
 * t, t + dt/2, t + dt/2, t + dt.  Vorticity and KE are not touched. */
int bve_advection_rk4_step(bve_ctx *ctx, double t, double dt, int divergent, int n_particles,
                           double *x_particles);

/* ScatterPanels payload (ParticlesEdgesPanels2.f90:1622-1637): current
 * active positions and relative vorticity and passive positions.  Any
 * pointer may be NULL. */
int bve_store_panels(bve_ctx *ctx, double *x_active, double *relvort_active, double *x_passive);

/* ResetRK4, PanelVelocityParallel.f90:321-380: drop the cached panels and
 * particles (call after a remesh). */
int bve_reset(bve_ctx *ctx);

/* ---- device-resident stepping (no host transfer per step) ------------ */

/* Upload particles once (and their absolute vorticity, may be NULL). */
int bve_load_particles(bve_ctx *ctx, int n_particles, const double *x_particles,
                       const double *absvort);
/* n_steps RK4 steps on the resident state; asynchronous on the context's
 * stream until bve_synchronize or a store call. */
int bve_advance(bve_ctx *ctx, double dt, int n_steps);
int bve_store_particles(bve_ctx *ctx, double *x_particles, double *relvort);
int bve_get_kinetic_energy(bve_ctx *ctx, double *total_ke, double *ke);
int bve_synchronize(bve_ctx *ctx);

/* ---- single evaluation (BVEActiveRHS / BVEPassiveRHS / BVESmoothVelocity) */

/* u (3*n_targets) at the targets from the n_sources sources with weight
 * vort*area:  targets [0, self_end) skip the source of the same index
 * (BVEActiveRHS, PanelVelocityParallel.f90:1287-1334, self_end <= n_sources),
 * targets [smooth_begin, smooth_end) use the smoothed denominator
 * (BVESmoothVelocity, :1376-1412), all others the singular one
 * (BVEPassiveRHS, :1337-1373).  dvort (self_end entries, may be NULL)
 * receives -2 Omega u_z (:1333). */
int bve_velocity(bve_ctx *ctx, int n_targets, const double *x_targets, int self_end,
                 int smooth_begin, int smooth_end, int n_sources, const double *x_sources,
                 const double *vort, const double *area, double *u, double *dvort);

/* No counterpart in the reference.  bve_velocity for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops): the same
 * sums, but every pointer is device memory (x_targets 3*n_targets and
 * x_sources 3*n_sources AoS, vort and area n_sources, u 3*n_targets,
 * dvort self_end or NULL), the work is enqueued on the context's stream
 * without waiting (bve_synchronize waits and reports numeric status), and
 * nothing is allocated: the kernels' scratch (target rows, source records,
 * queue, partial sums) is a static device array of ~46 MB per GPU, held
 * until the process ends and shared by every call on that GPU, so calls
 * that overlap on one GPU race.  At most 262144 targets and 262144
 * sources; float64, float64-fast and reference-order modes (the others
 * return BVE_ERR_ARGUMENT).  Segmented plans are capped at 1024 segments.
 * Its pass is not timed by bve_set_profiling (timing events would be
 * created on demand). */
int bve_velocity_noalloc(bve_ctx *ctx, int n_targets, const double *x_targets, int self_end,
                         int smooth_begin, int smooth_end, int n_sources, const double *x_sources,
                         const double *vort, const double *area, double *u, double *dvort);

/* Velocities of every resident target at the current state (the step-start
 * positions and vorticity), evaluated by the same launch plan and
 * arithmetic as a step's pair passes: BVEActiveRHS / BVEPassiveRHS /
 * BVESmoothVelocity over the cached panels, as BVERK4's first stage
 * evaluates them (PanelVelocityParallel.f90:465-473, :1287-1412).
 * u_active (3*n_active), u_passive (3*n_passive, the caller's passive
 * order) and u_particles (3*n_particles) may each be NULL. */
int bve_resident_velocity(bve_ctx *ctx, double *u_active, double *u_passive, double *u_particles);

/* ---- Voronoi path (drop-in for VoronoiVelocity{Serial,Parallel}Module) */

/* BVERK4(aPanels, dt[, procRank, numProcs]) of VoronoiVelocityParallel.f90:304-483
 * on n generators in SoA form (x, y, z, relvort in/out, area in), including
 * NormalizeGrid (:480); the host re-triangulates afterwards.  total_ke, ke
 * as for the panel step (may be NULL).  The generators replace the
 * context's loaded problem; its device arrays are kept for the next call
 * with the same n. */
int bve_voronoi_rk4_step(bve_ctx *ctx, double dt, int n, double *x, double *y, double *z,
                         double *relvort, const double *area, double *total_ke, double *ke);

/* ---- adjoint (gradient of a final-time functional of the active panels) */

/* No reference counterpart.  Runs n_steps RK4 steps of the active panels
 * only (passive panels and particles do not act on them) from the resident
 * state, keeping float64 checkpoints every checkpoint_every steps, and
 * returns the final state x_T (3*n_active) and zeta_T (n_active, may be
 * NULL).  The resident state is left at its starting value.  Float64
 * precision modes only, renormalisation off. */
int bve_adjoint_forward(bve_ctx *ctx, double dt, int n_steps, int checkpoint_every,
                        double *x_T, double *zeta_T);
/* Reverse sweep for the cotangent (dJ/dx_T, dJ/dzeta_T) of a functional J of
 * the final active state: recomputes each checkpoint segment's stage inputs
 * and applies the transposed RK4 stages with the hand-written vector-
 * Jacobian product of the pair sum.  Returns dJ/dx_0 (3*n_active, may be
 * NULL) and dJ/dzeta_0 (n_active). */
int bve_adjoint_backward(bve_ctx *ctx, const double *lambda_x, const double *lambda_zeta,
                         double *grad_x0, double *grad_zeta0);
/* No reference counterpart.  Arithmetic of the reverse sweep's pair pass
 * (the forward sweep and its recomputation keep the context's float64
 * mode): BVE_PRECISION_FP64 (the default) or BVE_PRECISION_MIXED, float32
 * pair terms wherever the bounding spheres of a 64-panel target group and a
 * 256-panel source tile are far apart (radii sum < ratio x centre
 * distance, 0 < ratio < 1; 0.7 is the tested value), float64 elsewhere and
 * in every carry; the recomputation of checkpoint segments then shares one
 * float64 reciprocal among four pairs.  The float32 terms bound the
 * cotangents and sums to float32's range (~3e38).  A repeated reverse sweep
 * after one bve_adjoint_forward recomputes the last segment, which the
 * first took from the forward sweep, so in this mode the two agree to
 * rounding (~1e-13), not bit for bit.  Takes effect at the next
 * bve_adjoint_forward or bve_adjoint_backward call. */
int bve_set_adjoint_precision(bve_ctx *ctx, int precision, double ratio);

/* No reference counterpart.  Exact finite differences of the same active-
 * panel RK4 map: from the resident state (x0, zeta0) and a vorticity
 * perturbation dzeta0 (n_active), advances both the trajectory of
 * (x0, zeta0) and the difference delta = trajectory(x0, zeta0 + dzeta0) -
 * trajectory(x0, zeta0) for n_steps steps, the difference in difference form
 * (every term built from differences, so it carries rounding relative to
 * itself, not to the state: no cancellation at small perturbations).
 * Returns x_T (3*n_active, may be NULL), zeta_T (may be NULL), dx_T and
 * dzeta_T.  Float64; the resident state is left as it is.  Used to check the
 * adjoint against central differences at small epsilon. */
int bve_difference_forward(bve_ctx *ctx, double dt, int n_steps, const double *dzeta0,
                           double *x_T, double *zeta_T, double *dx_T, double *dzeta_T);

/* ---- ensembles (many independent problems, shared launches) ----------- */

/* No reference counterpart (the reference runs one problem per MPI job).
 * An ensemble holds n_members panel problems of any sizes and advances all
 * of them with one persistent pair-sum launch per RK4 stage (work items of
 * every member, largest first, taken from a queue; one launch per distinct
 * targets-per-lane value among the members, so one with the autotuner's
 * plans) and one stage-update launch; the steps can be captured once into a
 * HIP graph and replayed. */
typedef struct bve_ensemble bve_ensemble;
int bve_ensemble_create(bve_ensemble **ens, int device, int n_members);
int bve_ensemble_destroy(bve_ensemble *ens);
const char *bve_ensemble_last_error(const bve_ensemble *ens);
/* Omega, smooth and arithmetic of every member (see bve_set_*). */
int bve_ensemble_set_params(bve_ensemble *ens, double omega, double smooth, int precision,
                            double nearfield_radius);
/* The member's panels (as bve_load_panels) and particles (as bve_load_particles). */
int bve_ensemble_load_member(bve_ensemble *ens, int member, int n_active, const double *x_active,
                             const double *relvort_active, const double *area_active,
                             int n_passive, const double *x_passive, int n_particles,
                             const double *x_particles, const double *absvort);
/* Launch geometry of one member: targets per lane (1, 2, 4; 0 = built-in
 * choice) and sources per work item (a multiple of 256).  chunk = 0 (the
 * default) gives every member of a targets-per-lane value one segment
 * length from the whole ensemble's work (see
 * bve_ensemble_set_items_per_workgroup).  Members with the same targets per
 * lane share one persistent pair launch (the autotuner gives all members
 * one value). */
int bve_ensemble_set_member_plan(bve_ensemble *ens, int member, int targets_per_lane, int chunk);
/* The member's launch geometry as the next advance runs it: targets per
 * lane, sources per work item and work items (any pointer may be NULL). */
int bve_ensemble_get_member_plan(bve_ensemble *ens, int member, int *targets_per_lane, int *chunk,
                                 int *work_items);
/* Members with chunk 0: work items per resident workgroup of their shared
 * pair launch (0 = 16); more items balance the launch's end better, fewer
 * write fewer partial rows. */
int bve_ensemble_set_items_per_workgroup(bve_ensemble *ens, int items);
/* Workgroups of each persistent pair launch (0 = as many as the kernel's
 * occupancy keeps resident). */
int bve_ensemble_set_grid(bve_ensemble *ens, int workgroups);
/* n_steps RK4 steps of every member; use_graph = 1 captures the steps into
 * a HIP graph (re-captured when dt or n_steps change) and launches it.
 * Asynchronous until bve_ensemble_synchronize or a store. */
int bve_ensemble_advance(bve_ensemble *ens, double dt, int n_steps, int use_graph);
int bve_ensemble_synchronize(bve_ensemble *ens);
/* The member's state (any pointer may be NULL); total_ke of the last step. */
int bve_ensemble_store_member(bve_ensemble *ens, int member, double *x_active,
                              double *relvort_active, double *x_passive, double *x_particles,
                              double *relvort_particles, double *total_ke);
/* HIP events around every pair launch (or every graph launch). */
int bve_ensemble_set_profiling(bve_ensemble *ens, int on);
int bve_ensemble_get_profile(bve_ensemble *ens, long long *launches, double *total_ms);
int bve_ensemble_pairs_per_step(const bve_ensemble *ens, double *pairs);

/* ---- output ----------------------------------------------------------- */

/* One legacy-VTK ASCII POLYDATA frame, byte for byte the file the
 * reference's vtkOutputUnif writes (ParticlesEdgesPanels2.f90:2116-2260,
 * reached through its vtkOutput interface, :133-137): particle positions as
 * POINTS (F24.18), the undivided panels' corner lists as POLYGONS (I8; for
 * Voronoi panels every panel, list-directed), then per-particle tracers,
 * relVort and absVort (POINT_DATA) and per-undivided-panel tracers,
 * relVort and absVort (CELL_DATA).  Host memory only, no device work.
 *
 * The arrays are the reference types' own storage (column-major):
 *   particles_x        x(3, N_Max) of Particles; the first n_particles columns
 *   particles_relvort, particles_absvort   (N_Max) or NULL when absent
 *                      (advection problems carry no vorticity)
 *   particles_tracer   tracer(ld_particles, n_tracer) or NULL (no tracers);
 *                      n_tracer is the reference's nTracers and also sets
 *                      how many panel tracers are written
 *   panel_kind         3 (TRI_PANEL), 4 (QUAD_PANEL) or 10 (VORONOI_PANEL)
 *   n_panels, n_active aPanels%N, aPanels%N_Active
 *   vertices           vertices(panel_kind, N_Max), 1-based particle indices
 *   has_children       (N_Max) Fortran default logicals, nonzero = divided;
 *                      NULL = no panel is divided
 *   panels_tracer      tracer(ld_panels, n_tracer) or NULL
 *   panels_relvort, panels_absvort   (N_Max) or NULL
 * Returns BVE_ERR_IO when the file cannot be written (the reference stops). */
int bve_vtk_write_frame(const char *filename, int n_particles, int ld_particles,
                        const double *particles_x, const double *particles_relvort,
                        const double *particles_absvort, int n_tracer,
                        const double *particles_tracer, int panel_kind, int n_panels,
                        int n_active, int ld_panels, const int *vertices, const int *has_children,
                        const double *panels_tracer, const double *panels_relvort,
                        const double *panels_absvort);

/* ---- measurement ------------------------------------------------------ */

/* 1 = bracket every pair-kernel launch with HIP events on the context's
 * stream; bve_get_profile returns the launches and their summed duration
 * since profiling was switched on. */
int bve_set_profiling(bve_ctx *ctx, int on);
int bve_get_profile(bve_ctx *ctx, long long *launches, double *total_ms);
/* Pairs evaluated by one RK4 step of the loaded problem: 4 (N_a (N_a - 1) + (N_p + N_pp) N_a). */
int bve_pairs_per_step(const bve_ctx *ctx, double *pairs);

#ifdef __cplusplus
}
#endif
#endif /* BVE_H */
