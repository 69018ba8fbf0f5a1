// Host-side launchers of the Biot-Savart / RK4 kernels.
#pragma once
#include "bve_common.h"

namespace bve {

// One pair-sum pass: targets SoA (rows tx, ty, tz of length t_stride),
// sources AoS records (n_src_pad, zero-weight padding), partial sums
// partial[(chunk * 3 + c) * t_stride + i] (chunked plans) or
// partial[((segment + tile) * 3 + c) * tile_targets + i - tile * tile_targets]
// (segmented plans, tile_targets = kBlock * targets_per_lane).  queue: two
// counters, zero between launches (segmented plans only).
struct PairArgs {
    const double *tx, *ty, *tz;
    int t_stride;
    TargetClasses cls;
    const Source *src;
    int n_src_pad;
    PairPlan plan;
    double *partial;
    unsigned *queue;
    // reference-order pass: the sources' areas (their records then carry the
    // vorticity itself in w) and the counts of real targets and sources
    const double *area;
    int n_targets, n_sources;
};

// One mixed-precision pair-sum pass: the float64 pass's arguments plus the
// per-stage float32 source records, bounding spheres and near-field radius.
struct MixedArgs {
    PairArgs base;
    int n_targets;        // real targets (padding is left out of the spheres)
    int n_sources;        // real sources
    float4 *src32;        // per source: o = x - tile centre (float32), |o|^2 / 2 + c
                          // (bounded mode: o' = x - sub-tile centre, (w x).x)
    float4 *srcw;         // per source: w o, w (float32; bounded mode: float2
                          // records (w x).y, (w x).z in the first half)
    float *srcc;          // per source: c = (1 - |x|^2) / 2
    double4 *tile_box;    // per 256-source tile: centre, radius
    double4 *group_box;   // per 64 targets: centre, radius
    double rnf;           // near-field radius (float64 below it)
    int kahan;            // 1: Kahan-compensated float32 sums
    // Bounded mode (BVE_PRECISION_MIXED_BOUNDED): per 256-source tile the
    // float64 moments {W0, m (3), P (6), W, pad, sub-tile offsets (8)} of
    // kTileMoments doubles
    // (W0 = sum w, m = sum w o, P = sum w o o^T as xx yy zz xy xz yz,
    // W = sum |w|, o = x - tile centre), the velocity scale max |x x S| the
    // probe pass finds, and the error target relative to it.
    double *tile_mom;
    double *scale;
    double err_target;
    // Bounded mode, per 64-source sub-tile (kSubMoments doubles): its
    // centre c' (3), radius, the float64 moments {W0, m (3), P (6), W} of
    // o' = x - c', and a pad.  The bounded pass's float32 records src32 are
    // o', relative to the sub-tile; each tile's entry in tile_mom ends with
    // the float32 offsets {fl32(c' - c), 0} of its four sub-tiles from its
    // own centre c (four float4 in the last eight doubles).
    double *sub_mom;
    // Bounded mode: the routing of every (target group, source tile) pair
    // of the pass, n_groups per source tile: bit 0 the whole tile in
    // float64, bit 1 + s sub-tile s in float64 (when the tile is retried)
    unsigned short *route;
    int n_groups;
};

constexpr int kGroup = 64;   // targets per bounding sphere (one wave's lanes)
constexpr int kTileMoments = 20;
constexpr int kSubTile = 64;
constexpr int kSubMoments = 16;

// Record of one active panel in the adjoint's pair pass: position x, the
// weighted position w x (w = zeta A) and n = m x x (m its velocity
// cotangent); 80 B, the position in the first 32.
struct __align__(16) AdjRec {
    double x, y, z, pad, wx, wy, wz, nx, ny, nz;
};

// Partial rows per (segment, target tile) of the adjoint's pair pass:
// V (3), N (3).
constexpr int kAdjRows = 6;

// Float32 record of one active panel for the mixed adjoint pair pass, taken
// relative to the centre c of its 256-record tile: o = {fl32(x - c), B} with
// B = |o|^2 / 2 + (1 - |x|^2) / 2 (float64 from the rounded o, rounded
// once), x = {fl32(x), 0}, w = {fl32(w x), 0}, n = {fl32(n), 0}; 64 B.
struct __align__(16) AdjRec32 {
    float4 o, x, w, n;
};

// One reverse RK4 stage of the adjoint on n active panels.  Row arrays have
// `stride` entries per row (rows x, y, z, zeta); source and record arrays
// have n_pad (a multiple of 1024) entries, zero beyond n.
struct AdjStage {
    int n, n_pad, stride;
    double h, omega;
    double b;              // RK4 weight of this stage's k in the update
    double c;              // coefficient of this stage's k in the next stage's input
    const double *lam;     // lambda' (cotangent of the step's output), 4 rows
    double *ybar;          // in: Ybar of the next stage; out: Ybar of this stage, 4 rows
    double *lam_acc;       // lambda' + sum of the Ybar so far, 4 rows
    double *m;             // 3 rows
    AdjRec *rec;
    const Source *ys;      // this stage's input as source records {x, zeta A}
    const double *sums;    // forward sums S of this stage's input, 3 rows
    const double *area;
    PairPlan plan;         // segmented plan of the pair pass over n_pad records
    double *partial;       // pair pass: kAdjRows rows per (segment + tile)
    unsigned *queue;       // segment queue (two counters, zero between launches)
    // Mixed pair pass when rec32 is set: its float32 records and the
    // spheres of the n_pad / 256 record tiles.  A (target tile, source
    // tile) pair runs in float32 when, for each 256-record part of the
    // target tile, radii sum < ratio * centre distance.
    AdjRec32 *rec32;
    double4 *tile_box;
    double ratio;
};

hipError_t launch_adjoint_stage(const AdjStage &a, hipStream_t s);
// Workgroups of the adjoint pair kernel one CU holds at once (mixed: the
// float32/float64 kernel).
int vjp_blocks_per_cu(int targets_per_lane, int recip, bool mixed = false);

// Device state of a panel RK4 step.  Targets are [active | passive | particles]
// in SoA rows of length t_stride.
struct StepState {
    int n_targets, t_stride, n_active;
    int particle_begin, particle_end;
    double *x0;      // step-start positions (3 rows)
    double *xs;      // current stage positions (3 rows)
    double *acc;     // RK4 accumulator x + sum b_s k_s (3 rows)
    double *zeta0;   // active relative vorticity at step start
    double *zacc;    // its RK4 accumulator
    const double *area;
    Source *src;     // active source records of the current stage
    const double *partial;
    double *ke;      // stage-1 |u|^2 of active panels
    double *ke_part; // per-block partial sums of ke * area
    double *ke_total;
    const double *particle_absvort;   // may be null
    double *particle_relvort;
    // Forced step (StratosphereRK4) when panel_absvort is set: every stage's
    // active vorticity is absVort - 2 Omega z - F(x, t) at the stage input.
    const double *panel_absvort;
    double t_step;    // time at the start of the step
    double omega;
    // 1: source records carry the vorticity itself (w = zeta; the reference-
    // order pass multiplies by the area term by term), 0: w = zeta * A
    int unweighted;
    // numeric status word (may be null) and the member index its keys carry
    unsigned long long *status;
    unsigned member;
    // if set (bounded mode), the stage update folds max 4 PI |u| = max
    // |x_i x S_i| over its targets into this non-negative double (atomicMax
    // on its bits): the velocity scale of the next pair pass
    double *vel_max;
};

struct StageParams {
    int stage;        // 0..3
    PairPlan plan;    // plan of the pass whose partial sums the stage reduces
    double dt, omega;
    double a_next;    // 0.5, 0.5, 1.0 (next stage input x + a k); unused at stage 3
    double divisor;   // 6, 3, 3, 6
    int renormalize;
    double t_next;    // forced step: time of the next stage input (of the result at stage 3)
    // if set, the factored sums S of active targets go here (3 rows of
    // sums_stride; the plan's rows must hold S, not crossed sums)
    double *sums_out;
    int sums_stride;
};

bool plan_supported(int targets_per_lane, int tile);
// Workgroups of the segmented float64 pair kernel one CU holds at once (0 if
// the occupancy query fails).
int pair_f64_blocks_per_cu(int targets_per_lane, int tile, int recip);
hipError_t launch_pair_sum_f64(const PairArgs &a, hipStream_t s);
// Workgroups of the segmented mixed pair kernel one CU holds at once (0 if
// the occupancy query fails).
int pair_mixed_blocks_per_cu(int targets_per_lane, int kahan);
hipError_t launch_mixed_prep(const MixedArgs &a, hipStream_t s);
hipError_t launch_pair_sum_mixed(const MixedArgs &a, hipStream_t s);
// Bounded mixed mode: preparation (records, tile moments, group spheres,
// velocity-scale probe) and the pair pass.
hipError_t launch_bounded_prep(const MixedArgs &a, hipStream_t s, bool probe = true);
// Clears the velocity-scale word (before a stage update folds max |u| into it).
hipError_t launch_scale_reset(double *scale, hipStream_t s);
hipError_t launch_pair_sum_bounded(const MixedArgs &a, hipStream_t s);
// Workgroups per CU of the bounded pass over n_src_pad sources (its kernel
// depends on the source count).
int pair_bounded_blocks_per_cu(int targets_per_lane, int n_src_pad);
// Reference-order pass (a.plan.exact): bitwise the reference's sums.
hipError_t launch_pair_sum_exact(const PairArgs &a, hipStream_t s);
hipError_t launch_stage_update(const StepState &st, const StageParams &sp, hipStream_t s);
// One AdvectionRK4 (divergent = 0) / DivergentAdvectionRK4 (1) step of every target.
hipError_t launch_advection_step(const StepState &st, double t, double dt, int divergent,
                                 hipStream_t s);
hipError_t launch_prepare_step(const StepState &st, hipStream_t s);
int ke_partial_count(int n_targets);

hipError_t launch_aos_to_soa(const double *aos, int n, double *dst, int stride, int off,
                             hipStream_t s);
hipError_t launch_soa_to_aos(const double *src, int stride, int off, int n, double *aos,
                             hipStream_t s);
hipError_t launch_make_sources(const double *xyz, const double *vort, const double *area, int n,
                               int unweighted, Source *src, hipStream_t s);
hipError_t launch_finalize_velocity(const double *partial, const PairPlan &plan, const double *t,
                                    int stride, int n, double omega, double *u, double *dvort,
                                    int n_dvort, unsigned long long *status, hipStream_t s);
// Allocation-free single evaluation: AoS targets into SoA rows of `stride`
// (zeros beyond n), source records for n_pad slots (zero weight beyond n),
// and zeroed queue counters, all by kernels (no memset nodes).
hipError_t launch_aos_to_soa_padded(const double *aos, int n, double *dst, int stride,
                                    hipStream_t s);
hipError_t launch_make_sources_padded(const double *xyz, const double *vort, const double *area,
                                      int n, int n_pad, int unweighted, Source *src,
                                      hipStream_t s);
hipError_t launch_zero_words(unsigned *p, int n, hipStream_t s);

// Exact-difference run of the active panels (bve_difference_forward): the
// trajectory of (x, zeta) and the difference to the trajectory of
// (x + dx, zeta + dzeta), both advanced by the same RK4 map.
struct __align__(16) DiffRec {
    double x, y, z;      // stage input of the base trajectory
    double dx, dy, dz;   // its difference to the perturbed trajectory's
    double w, dw;        // zeta A and delta zeta A of the stage input
};
struct DiffState {
    int n, stride;                     // panels; row stride (multiple of 1024)
    double *x0, *z0, *acc, *zacc;      // base: step start (3 + 1 rows), RK4 accumulators
    double *dx0, *dz0, *dacc, *dzacc;  // the same for the difference
    const double *area;
    DiffRec *rec;                      // stage inputs, zero beyond n
    double *partial;                   // per source chunk: S (3 rows), delta S (3 rows)
};
hipError_t launch_diff_prepare(const DiffState &st, hipStream_t s);
hipError_t launch_diff_pair(const DiffState &st, const PairPlan &plan, hipStream_t s);
hipError_t launch_diff_stage(const DiffState &st, const PairPlan &plan, int stage, double dt,
                             double omega, hipStream_t s);

// Ensemble: one work item of the persistent pair launch, a segment of a
// member's segmented plan.
struct EnsItem {
    int member, segment;
};

// Device tables of an ensemble launch (all pointers are device memory).
// Item groups of the ensemble's pair launches: members with 1, 2 and 4
// targets per lane.
constexpr int kEnsGroups = 3;

struct EnsLaunch {
    int n_members;
    const EnsItem *items;            // grouped by targets per lane, largest first in a group
    int group_begin[kEnsGroups], group_count[kEnsGroups];
    int group_grid[kEnsGroups];      // persistent workgroups of each group's launch
    unsigned *queue;                 // two counters per group, zero between launches
    const PairArgs *pair_members;    // float64 modes
    const MixedArgs *mixed_members;  // mixed modes
    int mixed, kahan, recip;
    const StepState *states;
    const PairPlan *plans;           // per member (segmented)
    const int2 *stage_map;           // (member, block) of every stage-update block
    int n_stage_blocks;
    const int2 *src_map;             // (member, source tile) of the mixed preparation
    int n_src_blocks;
    const int2 *tgt_map;             // (member, 256-target block) of the mixed preparation
    int n_tgt_blocks;
};

// Workgroups of the ensemble pair kernel for T targets per lane that one CU
// holds at once (0 if the occupancy query fails).
int ensemble_pair_blocks_per_cu(int T, int mixed, int kahan, int recip);
hipError_t launch_ensemble_pairs(const EnsLaunch &e, hipStream_t s);
hipError_t launch_ensemble_mixed_prep(const EnsLaunch &e, hipStream_t s);
hipError_t launch_ensemble_stage(const EnsLaunch &e, const StageParams &sp, hipStream_t s);
hipError_t launch_ensemble_prepare(const EnsLaunch &e, hipStream_t s);

}  // namespace bve
