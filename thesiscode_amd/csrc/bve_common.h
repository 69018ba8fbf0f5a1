// Shared definitions of the Biot-Savart / RK4 HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstddef>
#include <cstdint>

namespace bve {

// PI of the reference's NumberKinds.f90:12, rounded to float64
constexpr double kPi = 3.141592653589793238462643383279502797479068098137295573004504331874296718662975536062731407582759857177734375;

constexpr int kBlock = 256;        // threads per workgroup of the pair kernels (4 wave64)
constexpr int kTargetAlign = 1024; // target arrays are padded to this many entries
constexpr int kSourceAlign = 512;  // source arrays are padded to this many entries

// One source panel as the pair kernels read it: position and weight zeta*A.
struct __align__(32) Source {
    double x, y, z, w;
};

// Which targets a pair launch treats specially.  Targets are one array of
// n_targets points; targets [0, self_end) exclude the source of the same
// index (active panels, BVEActiveRHS), targets [smooth_begin, smooth_end)
// use the denominator (1 - x.y) + eps2 (divided parent panels,
// BVESmoothVelocity); all others use 1 - x.y (BVEPassiveRHS).
struct TargetClasses {
    int self_end;
    int smooth_begin;
    int smooth_end;
    double eps2;
    // reference-order pass only: 1 = the Voronoi module's denominator
    // ((1 - x_j x_k) - y_j y_k) - z_j z_k (VoronoiVelocityParallel.f90:502),
    // 0 = the panel module's 1 - ((x_k x_j + y_k y_j) + z_k z_j) (:1320)
    int voronoi;
};

// Launch geometry of one pair-sum pass (fixed per context and size).
struct PairPlan {
    int targets_per_lane;   // 1, 2 or 4 targets held in registers per lane
    int tile;               // sources per LDS tile
    int recip;              // sources sharing one float64 reciprocal: 1, 2 or 4
    int n_tiles;            // target tiles of kBlock * targets_per_lane
    int chunk;              // sources per chunk (multiple of tile)
    int n_chunks;           // source chunks; partial sums per chunk
    int n_work;             // n_tiles * n_chunks
    int grid;               // launched workgroups (n_work rounded up to 8 XCDs)
    // Segmented launch (segments > 0; chunk / n_chunks unused): the
    // tile-major sequence of (target tile, source tile) items is cut into
    // `segments` runs of seg_len items, which a persistent grid of `grid`
    // workgroups takes from a queue.  A segment writes one partial row per
    // target tile it touches, row (segment + target tile).
    int segments;
    int n_src_tiles;        // n_src_pad / tile
    int seg_len;
    // Mixed segmented pass: item k of a target tile's row is source tile
    // (k * src_step) mod n_src_tiles (src_step coprime with n_src_tiles), so
    // the careful tiles of a target group, adjacent in the source order, fall
    // into different segments.  Float64 passes: 1 (source order).
    int src_step;
    // Reference-order pass (exact = 1): one target per lane, every source in
    // index order, the reference's own expression and rounding per pair; one
    // chunk, whose rows hold the sums of the cross-product terms themselves
    // (u = S / (-4 PI)) rather than the factored S = sum (w/d) x.
    int exact;
    // Float64 pass (crossed = 1): the rows hold sums of x_i x S_t, each
    // source tile's factored sum crossed with the target before it is added,
    // so u = rows / (-4 PI) as in the reference-order pass.
    int crossed;
};

__host__ __device__ inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

// Numeric status word of a context (device memory, 0 = clean).  A kernel
// that produces a non-finite velocity or position records
// key = (member << 32) | target with atomicMax(status, ~key), so the word
// ends up holding the complement of the smallest key: the first offending
// target of the lowest member.  The host reads and clears it at the next
// synchronisation.
__device__ __forceinline__ void flag_nonfinite(unsigned long long *status, unsigned member,
                                               int target)
{
    if (status)
        atomicMax(status, ~(((unsigned long long)member << 32) | (unsigned)target));
}

__host__ __device__ inline bool finite3(double a, double b, double c)
{
    // false for NaN and +-Inf (NaN compares false)
    return fabs(a) <= 1.7976931348623157e308 && fabs(b) <= 1.7976931348623157e308 &&
           fabs(c) <= 1.7976931348623157e308;
}
__host__ __device__ inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Doubles of partial sums one pass with plan p writes: chunked, 3 rows of
// t_stride per chunk; segmented, 3 rows of one target tile per (segment,
// tile) piece, whose index segment + tile is below segments + n_tiles.
__host__ __device__ inline size_t partial_doubles(const PairPlan &p, int t_stride)
{
    if (p.segments > 0)
        return (size_t)(p.segments + p.n_tiles) * 3 * kBlock * p.targets_per_lane;
    return (size_t)p.n_chunks * 3 * t_stride;
}

}  // namespace bve
