// Device-side bodies of the Biot-Savart pair sums and the RK4 stage update,
// shared by the per-context launches and the persistent ensemble launch.
//
// The velocity at a target x_i is (PanelVelocityParallel.f90:1287-1412)
//
//     u_i = -1/(4 PI) sum_j w_j (x_i x x_j) / d_ij,   d_ij = (1 - x_i.x_j) + eps2,
//
// with w_j = zeta_j A_j.  The float64 body evaluates the factored form
// S_i = sum_j (w_j / d_ij) x_j and the stage body forms u_i = (x_i x S_i) / (-4 PI),
// which saves the per-pair cross product.  The weight is folded into the
// source when a tile is staged in LDS ({x_j, w_j x_j}), so a pair costs 3 FMA
// for the denominator, the reciprocal (below) and 3 FMA into S.
//
// Work decomposition: a work item is (target tile, source chunk).  Targets
// are held in registers (T per lane), sources stream through LDS in tiles
// (HBM records {x, y, z, w}), and each item writes its own partial sums,
// which the stage body adds in a fixed order: results do not depend on
// scheduling and reruns are bitwise reproducible.
#pragma once
#include "bve_common.h"
#include "bve_kernels.h"

namespace bve {

// ---------------------------------------------------------------------------
// Reciprocals
// ---------------------------------------------------------------------------

// 1/d: v_rcp_f64 seed (accurate to about 2^-24) refined by two Newton
// steps r <- r + r (1 - d r): correctly rounded to within an ulp or two.
__device__ __forceinline__ double recip_f64(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

// 1/p with one Newton step: relative error about 2^-48.
__device__ __forceinline__ double recip1_f64(double p)
{
    const double r = __builtin_amdgcn_rcp(p);
    const double e = __builtin_fma(-p, r, 1.0);
    return __builtin_fma(r, e, r);
}

// Maps workgroup g of a grid of `grid` workgroups (a multiple of 8) to a
// work item so that XCD (g % 8) runs a contiguous run of the chunk-major work
// list: the source chunks an XCD reads stay in its L2.
__device__ __forceinline__ int xcd_work_item(int g, int grid)
{
    return (g % 8) * (grid / 8) + g / 8;
}

// Wave-uniform: does source tile [jt, jt + tile) hold the own index of one
// of the 64 lanes whose targets start at lo (targets [0, self_end) are
// self-excluded)?
__device__ __forceinline__ bool diag_tile(int lo, int self_end, int jt, int tile)
{
    const int hi = min(lo + 64, self_end);
    return (lo < hi) && (lo < jt + tile) && (jt < hi);
}

// ---------------------------------------------------------------------------
// Reference-order pair body
// ---------------------------------------------------------------------------

// The sums of BVEActiveRHS / BVEPassiveRHS / BVESmoothVelocity
// (PanelVelocityParallel.f90:1287-1412) and of the Voronoi BVEVelocity
// (VoronoiVelocityParallel.f90:487-515) in the reference's own order and
// rounding: target i (one per lane) adds, for k = 0, 1, ... in index order
// (skipping k = i for the self-excluded targets), the terms
//     ((x_i,b x_k,c - x_i,c x_k,b) zeta_k) A_k / denom
// each divided separately (IEEE division), into sums that start at 0, with
// no contraction into FMA: the float64 results equal the reference's
// (compiled without FMA) bit for bit.  The records carry w = zeta_k and the
// areas come from a.area; sources stream through LDS 256 at a time, and all
// lanes of a wave read the same source (an LDS broadcast).
__device__ __forceinline__ void pair_exact_body(const PairArgs &a, double4 *lds_src,
                                                double *lds_area)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < a.n_targets;
    const double X0 = live ? a.tx[i] : 0.0, X1 = live ? a.ty[i] : 0.0, X2 = live ? a.tz[i] : 0.0;
    const bool self = i < a.cls.self_end;
    const bool smoothed = i >= a.cls.smooth_begin && i < a.cls.smooth_end;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int base = 0; base < a.n_sources; base += kBlock) {
        __syncthreads();
        const int k = base + threadIdx.x;
        if (k < a.n_sources) {
            const Source q = a.src[k];
            lds_src[threadIdx.x] = make_double4(q.x, q.y, q.z, q.w);
            lds_area[threadIdx.x] = a.area[k];
        }
        __syncthreads();
        const int m = min(kBlock, a.n_sources - base);
        if (!live) continue;
        for (int kk = 0; kk < m; ++kk) {
            if (self && base + kk == i) continue;
            const double4 q = lds_src[kk];
            const double ar = lds_area[kk];
            double denom;
            if (a.cls.voronoi) {
                denom = 1.0 - X0 * q.x - X1 * q.y - X2 * q.z;
            } else {
                denom = 1.0 - ((q.x * X0 + q.y * X1) + q.z * X2);
                if (smoothed) denom = denom + a.cls.eps2;
            }
            a0 = a0 + ((X1 * q.z - X2 * q.y) * q.w) * ar / denom;
            a1 = a1 + ((X2 * q.x - X0 * q.z) * q.w) * ar / denom;
            a2 = a2 + ((X0 * q.y - X1 * q.x) * q.w) * ar / denom;
        }
    }
    if (live) {
        a.partial[i] = a0;
        a.partial[a.t_stride + i] = a1;
        a.partial[2 * (size_t)a.t_stride + i] = a2;
    }
}

// Velocity of target (x, y, z) from its reduced sums: u = (x x S) / (-4 PI)
// for the factored sums of the mixed pass, u = S / (-4 PI) when the rows hold
// cross-product sums already (the reference-order pass, whose terms are the
// reference's own, PanelVelocityParallel.f90:1332, and the float64 pass,
// which crosses each source tile's sum).
__device__ __forceinline__ void velocity_from_sums(int exact, double x, double y, double z,
                                                   double Sx, double Sy, double Sz, double &ux,
                                                   double &uy, double &uz)
{
#pragma clang fp contract(off)
    const double c4 = -4.0 * kPi;
    if (exact) {
        ux = Sx / c4;
        uy = Sy / c4;
        uz = Sz / c4;
    } else {
        ux = (y * Sz - z * Sy) / c4;
        uy = (z * Sx - x * Sz) / c4;
        uz = (x * Sy - y * Sx) / c4;
    }
}

// Weight of an active panel's source record: zeta A, or zeta itself when the
// reference-order pass multiplies by the area term by term.
__device__ __forceinline__ double source_weight(const StepState &st, double zeta, int i)
{
    return st.unweighted ? zeta : zeta * st.area[i];
}

// ---------------------------------------------------------------------------
// float64 pair body
// ---------------------------------------------------------------------------

// A source as staged in LDS: position and weighted position w x.
struct alignas(16) StagedSource {
    double x, y, z, wx, wy, wz;
};

// A source record as the float64 pair loop uses it: staged (as is) or
// straight from its HBM record.
__device__ __forceinline__ StagedSource staged_of(const StagedSource &s) { return s; }
__device__ __forceinline__ StagedSource staged_of(const Source &s)
{
    return StagedSource{s.x, s.y, s.z, s.w * s.x, s.w * s.y, s.w * s.z};
}

// Stages `n` HBM records starting at src into LDS (all threads of the block).
__device__ __forceinline__ void stage_sources(StagedSource *lds, const Source *__restrict__ src,
                                              int n)
{
    for (int k = threadIdx.x; k < n; k += kBlock) {
        const Source s = src[k];
        lds[k] = StagedSource{s.x, s.y, s.z, s.w * s.x, s.w * s.y, s.w * s.z};
    }
}

// One LDS tile of sources against the T targets of this lane.
//
// R = 1: every pair gets its own reciprocal (recip_f64).
// R = 2, 4: R consecutive sources share one reciprocal of the product of
// their denominators (1/d_a = d_b / (d_a d_b)), refined by one Newton step,
// which trades R-1 v_rcp_f64 and Newton steps for 2(R-1) multiplies.
// Denominators lie in [~1e-6, 2 + eps2], so products of four stay far from
// float64 underflow and overflow.
//
// MASK: the tile holds the own-index source of some target of this wave;
// that pair's denominator (~0) is replaced by 1 and its 1/d by 0.
template <int T, int TILE, int R, bool MASK, class Rec>
__device__ __forceinline__ void tile_f64(const Rec *__restrict__ lds, int jt,
                                         const double (&nx)[T], const double (&ny)[T],
                                         const double (&nz)[T], const double (&c0)[T],
                                         const int (&self)[T], double (&sx)[T],
                                         double (&sy)[T], double (&sz)[T])
{
    constexpr int kUnroll = R == 1 ? 2 : 1;
#pragma unroll kUnroll
    for (int k = 0; k < TILE; k += R) {
        StagedSource s[R];
#pragma unroll
        for (int r = 0; r < R; ++r) s[r] = staged_of(lds[k + r]);
#pragma unroll
        for (int t = 0; t < T; ++t) {
            double d[R], q[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                d[r] = __builtin_fma(nx[t], s[r].x, c0[t]);
                d[r] = __builtin_fma(ny[t], s[r].y, d[r]);
                d[r] = __builtin_fma(nz[t], s[r].z, d[r]);
                if (MASK) d[r] = (jt + k + r == self[t]) ? 1.0 : d[r];
            }
            if constexpr (R == 1) {
                q[0] = recip_f64(d[0]);
            } else if constexpr (R == 2) {
                const double rr = recip1_f64(d[0] * d[1]);
                q[0] = d[1] * rr;
                q[1] = d[0] * rr;
            } else {
                const double p01 = d[0] * d[1], p23 = d[2] * d[3];
                const double rr = recip1_f64(p01 * p23);
                const double r01 = rr * p23, r23 = rr * p01;
                q[0] = d[1] * r01;
                q[1] = d[0] * r01;
                q[2] = d[3] * r23;
                q[3] = d[2] * r23;
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (MASK) q[r] = (jt + k + r == self[t]) ? 0.0 : q[r];
                sx[t] = __builtin_fma(q[r], s[r].wx, sx[t]);
                sy[t] = __builtin_fma(q[r], s[r].wy, sy[t]);
                sz[t] = __builtin_fma(q[r], s[r].wz, sz[t]);
            }
        }
    }
}

// Target tile ttile against sources [j_begin, j_end) (multiples of TILE) of
// a float64 pass; the sums of target i go to out[i - out_shift] and the rows
// out_stride and 2 out_stride after it.  lds holds TILE records.
//
// With plan.crossed, each source tile's factored sum S_t = sum_j (w_j /
// d_ij) x_j is crossed with x_i before it joins the target's running sum,
// which is therefore sum_t x_i x S_t (the rows hold cross-product sums).
// S_t has a large component along x_i for close pairs (w / d with d ~ r^2 /
// 2), which x_i x S_t removes; carried across tiles it would round every
// later addition at eps |S| (2e-12 of max |u| at level 8), crossed per tile
// it rounds only the additions of its own tile.  Without it the rows hold S
// itself (the adjoint's forward sweep keeps S for its reverse sweep).
template <int T, int TILE, int R>
__device__ __forceinline__ void pair_f64_range(const PairArgs &a, int ttile, int j_begin,
                                               int j_end, double *out, int out_stride,
                                               int out_shift, StagedSource *lds)
{
    const int tile_base = ttile * (kBlock * T);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const TargetClasses cls = a.cls;

    double nx[T], ny[T], nz[T], c0[T], ux[T], uy[T], uz[T];
    int self[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x;
        nx[t] = -a.tx[i];
        ny[t] = -a.ty[i];
        nz[t] = -a.tz[i];
        c0[t] = (i >= cls.smooth_begin && i < cls.smooth_end) ? 1.0 + cls.eps2 : 1.0;
        self[t] = (i < cls.self_end) ? i : -1;
        ux[t] = uy[t] = uz[t] = 0.0;
    }

    for (int jt = j_begin; jt < j_end; jt += TILE) {
        __syncthreads();
        stage_sources(lds, a.src + jt, TILE);
        __syncthreads();
        bool diag = false;
#pragma unroll
        for (int t = 0; t < T; ++t)
            diag |= diag_tile(tile_base + t * kBlock + wave * 64, cls.self_end, jt, TILE);
        double sx[T], sy[T], sz[T];
#pragma unroll
        for (int t = 0; t < T; ++t) sx[t] = sy[t] = sz[t] = 0.0;
        if (diag)
            tile_f64<T, TILE, R, true>(lds, jt, nx, ny, nz, c0, self, sx, sy, sz);
        else
            tile_f64<T, TILE, R, false>(lds, jt, nx, ny, nz, c0, self, sx, sy, sz);
        if (a.plan.crossed) {
            // u += x_i x S_t with x_i = -n
#pragma unroll
            for (int t = 0; t < T; ++t) {
                ux[t] = __builtin_fma(nz[t], sy[t], __builtin_fma(-ny[t], sz[t], ux[t]));
                uy[t] = __builtin_fma(nx[t], sz[t], __builtin_fma(-nz[t], sx[t], uy[t]));
                uz[t] = __builtin_fma(ny[t], sx[t], __builtin_fma(-nx[t], sy[t], uz[t]));
            }
        } else {
#pragma unroll
            for (int t = 0; t < T; ++t) {
                ux[t] += sx[t];
                uy[t] += sy[t];
                uz[t] += sz[t];
            }
        }
    }

#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x - out_shift;
        out[i] = ux[t];
        out[out_stride + i] = uy[t];
        out[2 * (size_t)out_stride + i] = uz[t];
    }
}

// Work queue of the persistent launches: the workgroup takes item indices
// from queue[0] until n are handed out, running f(k) for each; the last
// workgroup to finish clears both counters for the next launch (so graph
// replays need no memset).
template <class F>
__device__ __forceinline__ void run_queue(unsigned *queue, int n, int *slot, F &&f)
{
    for (;;) {
        if (threadIdx.x == 0) *slot = (int)atomicAdd(&queue[0], 1u);
        __syncthreads();
        const int k = *slot;
        __syncthreads();
        if (k >= n) break;
        f(k);
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

// Segment s of a segmented pass: items [s seg_len, (s + 1) seg_len) of the
// tile-major sequence (item = target tile * n_src_tiles + source tile);
// range(t, j_begin, j_end, s) runs every maximal run of it inside one
// target tile t, and writes row s + t (unique: s + t increases along the
// sequence).
template <int TILE, class Range>
__device__ __forceinline__ void run_segment(const PairPlan &p, int seg, Range &&range)
{
    const int nst = p.n_src_tiles;
    const int n_items = p.n_tiles * nst;
    int b = seg * p.seg_len;
    const int e = min(b + p.seg_len, n_items);
    while (b < e) {
        const int t = b / nst;
        const int k0 = b - t * nst;
        const int k1 = min(nst, k0 + (e - b));
        range(t, k0 * TILE, k1 * TILE, seg);
        b += k1 - k0;
    }
}

// The persistent passes take their segments from a queue.
template <int TILE, class Range>
__device__ __forceinline__ void run_segments(const PairPlan &p, unsigned *queue, int *seg_slot,
                                             Range &&range)
{
    run_queue(queue, p.segments, seg_slot, [&](int seg) { run_segment<TILE>(p, seg, range); });
}

// Segment `seg` of a float64 pass (one work item of the persistent and the
// ensemble launches).
template <int T, int TILE, int R>
__device__ __forceinline__ void pair_f64_segment(const PairArgs &a, int seg, StagedSource *lds)
{
    constexpr int TT = kBlock * T;
    run_segment<TILE>(a.plan, seg, [&](int t, int j0, int j1, int s) {
        pair_f64_range<T, TILE, R>(a, t, j0, j1, a.partial + (size_t)(s + t) * 3 * TT, TT, t * TT,
                                   lds);
    });
}

// Persistent body of a segmented float64 pass.
template <int T, int TILE, int R>
__device__ __forceinline__ void pair_f64_segments(const PairArgs &a, StagedSource *lds,
                                                  int *seg_slot)
{
    run_queue(a.queue, a.plan.segments, seg_slot,
              [&](int seg) { pair_f64_segment<T, TILE, R>(a, seg, lds); });
}

// ---------------------------------------------------------------------------
// Mixed-precision pair body (float32 far field, float64 near field)
// ---------------------------------------------------------------------------
//
// The sum is written in difference form,
//
//     S_i = sum_j q_ij (x_j - x_i),  q_ij = w_j / D_ij,
//     D_ij = |x_i - x_j|^2 / 2 + c_i + c_j (+ eps2),  c = (1 - |x|^2) / 2,
//
// which is 1 - x_i.x_j (+ eps2) exactly, without its cancellation for close
// pairs, and x_i x S_i is the same cross product as in the float64 body
// (x_i x x_i = 0).  Float32 differences are formed from per-tile float64
// origins: every source tile (256 sources) and every group of 64 targets
// has a float64 centre c; records hold o = fl32(x - c), and a target's
// offset from a source tile's centre is a = fl32(c_g - c_s) + o_i, so its
// rounding error scales with the tile sizes and the distance of the tiles,
// not with |x| = 1.
//
// Far-field tiles run in moment form.  With A_i = |a|^2 / 2 + c_i and
// B_j = |o_j|^2 / 2 + c_j (float64, rounded once, from the rounded a and
// o_j),
//
//     D_ij = (A_i + B_j) - a . o_j,
//     sum_j q_ij (o_j - a) = M_i - a Q_i,
//     M_i = sum_j (w_j o_j) / D_ij,  Q_i = sum_j w_j / D_ij,
//
// so a pair costs an add and three FMAs for D, one v_rcp_f32 and four FMAs,
// against three subtractions more and a weight product in the difference
// form.  M - a Q is formed per tile in float64 with a = x_i - c_s in
// float64.  A non-careful tile lies at least r_nf + r_s from the target
// (r_s its radius), so |a| / |a - o_j| stays below (r_nf + r_s) / r_nf and
// the cancellation in D stays within that factor squared of one rounding.
//
// Pairs closer than the near-field radius are evaluated in float64 from
// the float64 source records.  A tile pair is sent down the careful path
// (per-pair distance test, difference form) only when the bounding spheres
// of the target group and the source tile come within the radius, or when
// the tile holds a target's own index; all other tiles run the float32
// moment loop.  Float32 partial sums are carried into float64 accumulators
// after every tile, or after every pair with Kahan compensation when
// requested.

constexpr int kMixedTile = 256;   // sources per LDS tile and per source bounding sphere

// One source tile of the mixed pass in LDS.
struct MixedLds {
    float4 pos[kMixedTile];   // o_j, B_j
    float4 wgt[kMixedTile];   // w_j o_j, w_j
    float c[kMixedTile];      // c_j
};

template <bool KAHAN>
__device__ __forceinline__ void add_f32(float &acc, float &comp, float term)
{
    if (KAHAN) {
        const float y = term - comp;
        const float t = acc + y;
        comp = (t - acc) - y;
        acc = t;
    } else {
        acc += term;
    }
}

// Moment-form sum of the source tile in LDS (centre sb) for N targets with
// float32 offsets a from the tile centre; adds M - a Q (a = x_i - c_s in
// float64) to the float64 accumulators.
template <int N, bool KAHAN>
__device__ __forceinline__ void moment_tile(const MixedLds &lds, const double4 &sb,
                                            const float *ax, const float *ay, const float *az,
                                            const double *ci, const double *xi,
                                            const double *yi, const double *zi, double *S64x,
                                            double *S64y, double *S64z)
{
    float nax[N], nay[N], naz[N], A[N];
    float Mx[N], My[N], Mz[N], Q[N], Cx[N], Cy[N], Cz[N], CQ[N];
#pragma unroll
    for (int t = 0; t < N; ++t) {
        nax[t] = -ax[t];
        nay[t] = -ay[t];
        naz[t] = -az[t];
        A[t] = (float)(0.5 * ((double)ax[t] * ax[t] + (double)ay[t] * ay[t] +
                              (double)az[t] * az[t]) + ci[t]);
        Mx[t] = My[t] = Mz[t] = Q[t] = Cx[t] = Cy[t] = Cz[t] = CQ[t] = 0.f;
    }
    // One v_rcp_f32 per pair: sharing it between two or four pairs (products
    // of denominators) measured no faster on gfx950 and adds rounding.
#pragma unroll 4
    for (int k = 0; k < kMixedTile; ++k) {
        const float4 s = lds.pos[k];
        const float4 v = lds.wgt[k];
#pragma unroll
        for (int t = 0; t < N; ++t) {
            const float D = __builtin_fmaf(
                nax[t], s.x, __builtin_fmaf(nay[t], s.y, __builtin_fmaf(naz[t], s.z, A[t] + s.w)));
            const float r = __builtin_amdgcn_rcpf(D);
            add_f32<KAHAN>(Mx[t], Cx[t], r * v.x);
            add_f32<KAHAN>(My[t], Cy[t], r * v.y);
            add_f32<KAHAN>(Mz[t], Cz[t], r * v.z);
            add_f32<KAHAN>(Q[t], CQ[t], r * v.w);
        }
    }
#pragma unroll
    for (int t = 0; t < N; ++t) {
        const double q = (double)Q[t] - (double)CQ[t];
        S64x[t] += ((double)Mx[t] - (double)Cx[t]) - (xi[t] - sb.x) * q;
        S64y[t] += ((double)My[t] - (double)Cy[t]) - (yi[t] - sb.y) * q;
        S64z[t] += ((double)Mz[t] - (double)Cz[t]) - (zi[t] - sb.z) * q;
    }
}

// Careful sum of the source tile in LDS (first source jt) for one target:
// float32 difference form with the pairs below the near-field radius and
// the target's own index masked out (selects, no branch), and those near
// pairs in float64 from the float64 records (a branch the wave skips when
// none of its lanes has one).
template <bool KAHAN>
__device__ __forceinline__ void careful_tile(const MixedArgs &a, const MixedLds &lds, int jt,
                                             float ax, float ay, float az, float cif, int self,
                                             double xi, double yi, double zi, double ci,
                                             float rnf2, double &S64x, double &S64y,
                                             double &S64z)
{
    float Sx = 0.f, Sy = 0.f, Sz = 0.f, Cx = 0.f, Cy = 0.f, Cz = 0.f;
    for (int k = 0; k < kMixedTile; ++k) {
        const float4 s = lds.pos[k];
        const float dx = ax - s.x, dy = ay - s.y, dz = az - s.z;
        const float h = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
        const int j = jt + k;
        const bool near = h < rnf2;
        const float D = __builtin_fmaf(0.5f, h, cif + lds.c[k]);
        const float q = (near || j == self) ? 0.f : lds.wgt[k].w * __builtin_amdgcn_rcpf(D);
        add_f32<KAHAN>(Sx, Cx, -q * dx);
        add_f32<KAHAN>(Sy, Cy, -q * dy);
        add_f32<KAHAN>(Sz, Cz, -q * dz);
        if (near && j != self) {
            const Source sj = a.base.src[j];
            const double ex = sj.x - xi, ey = sj.y - yi, ez = sj.z - zi;
            const double cj64 = 0.5 * (1.0 - (sj.x * sj.x + sj.y * sj.y + sj.z * sj.z));
            const double D64 = 0.5 * (ex * ex + ey * ey + ez * ez) + (ci + cj64);
            const double q64 = sj.w / D64;
            S64x = __builtin_fma(q64, ex, S64x);
            S64y = __builtin_fma(q64, ey, S64y);
            S64z = __builtin_fma(q64, ez, S64z);
        }
    }
    S64x += (double)Sx - (double)Cx;
    S64y += (double)Sy - (double)Cy;
    S64z += (double)Sz - (double)Cz;
}

// Source tiles (first + k step) mod n_src_tiles for k in [k0, k1), for target
// tile `ttile`: the sums of target i go to out[i - out_off],
// out[out_stride + i - out_off] and out[2 out_stride + i - out_off].
template <int T, bool KAHAN>
__device__ __forceinline__ void pair_mixed_range(const MixedArgs &a, int ttile, int first,
                                                 int step, int k0, int k1, double *out,
                                                 int out_stride, int out_off, MixedLds &lds)
{
    constexpr int TILE = kMixedTile;
    const int tile_base = ttile * (kBlock * T);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const TargetClasses cls = a.base.cls;

    double xi[T], yi[T], zi[T], ci[T], gx[T], gy[T], gz[T], gr[T];
    double S64x[T], S64y[T], S64z[T];
    float oix[T], oiy[T], oiz[T], cif[T];
    int self[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x;
        xi[t] = a.base.tx[i];
        yi[t] = a.base.ty[i];
        zi[t] = a.base.tz[i];
        const bool smooth = i >= cls.smooth_begin && i < cls.smooth_end;
        ci[t] = 0.5 * (1.0 - (xi[t] * xi[t] + yi[t] * yi[t] + zi[t] * zi[t])) +
                (smooth ? cls.eps2 : 0.0);
        cif[t] = (float)ci[t];
        self[t] = (i < cls.self_end) ? i : -1;
        const double4 gb = a.group_box[(tile_base + t * kBlock) / kGroup + wave];
        gx[t] = gb.x;
        gy[t] = gb.y;
        gz[t] = gb.z;
        gr[t] = gb.w;
        oix[t] = (float)(xi[t] - gx[t]);
        oiy[t] = (float)(yi[t] - gy[t]);
        oiz[t] = (float)(zi[t] - gz[t]);
        S64x[t] = S64y[t] = S64z[t] = 0.0;
    }
    const float rnf2 = (float)(a.rnf * a.rnf);

    const int nst = a.base.n_src_pad / TILE;
    for (int k = k0; k < k1; ++k) {
        const int jt = (int)(((long long)k * step + first) % nst) * TILE;
        __syncthreads();
        lds.pos[threadIdx.x] = a.src32[jt + threadIdx.x];
        lds.wgt[threadIdx.x] = a.srcw[jt + threadIdx.x];
        lds.c[threadIdx.x] = a.srcc[jt + threadIdx.x];
        __syncthreads();
        const double4 sb = a.tile_box[jt / TILE];
        float ax[T], ay[T], az[T];
        bool care[T];
        bool any = false;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const double dx = gx[t] - sb.x, dy = gy[t] - sb.y, dz = gz[t] - sb.z;
            ax[t] = (float)dx + oix[t];
            ay[t] = (float)dy + oiy[t];
            az[t] = (float)dz + oiz[t];
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            // The decision is per target group (one wave's lanes hold one
            // group per t), hence uniform over the wave.  1e-6: slack for the
            // float32 rounding of the per-pair test.
            care[t] = (sb.w >= 0.0 && dist < gr[t] + sb.w + a.rnf + 1e-6) ||
                      diag_tile(tile_base + t * kBlock + wave * 64, cls.self_end, jt, TILE);
            any |= care[t];
        }
        if (!any) {
            moment_tile<T, KAHAN>(lds, sb, ax, ay, az, ci, xi, yi, zi, S64x, S64y, S64z);
        } else {
#pragma unroll
            for (int t = 0; t < T; ++t) {
                if (care[t])
                    careful_tile<KAHAN>(a, lds, jt, ax[t], ay[t], az[t], cif[t], self[t], xi[t],
                                        yi[t], zi[t], ci[t], rnf2, S64x[t], S64y[t], S64z[t]);
                else
                    moment_tile<1, KAHAN>(lds, sb, &ax[t], &ay[t], &az[t], &ci[t], &xi[t],
                                          &yi[t], &zi[t], &S64x[t], &S64y[t], &S64z[t]);
            }
        }
    }

#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x - out_off;
        out[i] = S64x[t];
        out[out_stride + i] = S64y[t];
        out[2 * (size_t)out_stride + i] = S64z[t];
    }
}

// Segment `seg` of a mixed pass, each target tile's source tiles visited in
// the plan's stride order (tile (k src_step) mod n_src_tiles for item k):
// the careful tiles of a target group, its neighbours in the source order,
// fall into different segments, so segments cost alike.
template <int T, bool KAHAN>
__device__ __forceinline__ void pair_mixed_segment(const MixedArgs &a, int seg, MixedLds &lds)
{
    constexpr int TT = kBlock * T;
    run_segment<kMixedTile>(a.base.plan, seg, [&](int t, int j0, int j1, int s) {
        pair_mixed_range<T, KAHAN>(a, t, 0, a.base.plan.src_step, j0 / kMixedTile,
                                   j1 / kMixedTile, a.base.partial + (size_t)(s + t) * 3 * TT, TT,
                                   t * TT, lds);
    });
}

// Persistent body of a segmented mixed pass.
template <int T, bool KAHAN>
__device__ __forceinline__ void pair_mixed_segments(const MixedArgs &a, MixedLds &lds,
                                                    int *seg_slot)
{
    run_queue(a.base.queue, a.base.plan.segments, seg_slot,
              [&](int seg) { pair_mixed_segment<T, KAHAN>(a, seg, lds); });
}

// ---------------------------------------------------------------------------
// Bounded mixed-precision pair body: float32 remainders of a float64 tile
// expansion, float64 where a rounding bound exceeds the error target
// ---------------------------------------------------------------------------
//
// For target x_i and a source tile with float64 centre c write
// D_ij = c0_i - x_i.x_j = E - delta_j, with E = c0_i - x_i.c and
// delta_j = x_i.o_j, o_j = x_j - c (the reference's own denominator,
// PanelVelocityParallel.f90:1320, :1405, split at the tile centre).  Exactly
//
//     1/D = 1/E + delta/E^2 + delta^2 / (E^2 D),
//
// so the tile's factored sum (x_j = c + o_j; only x_i x S is used) is
//
//     S = sum_j w_j x_j / D_ij
//       = (W0 c + m) / E + (c (m.x_i) + P x_i + R) / E^2
//
// with the float64 tile moments W0 = sum w, m = sum w o, P = sum w o o^T
// and the float32 remainder sum R = sum_j (delta_j^2/D_ij) w_j x_j: per pair
// three products for delta, the difference D, a quarter of a shared
// reciprocal, delta^2 / D and three FMAs into R.
// The remainder is t^2 of the plain term (t = |delta| / E, at most
// 2R/rho + R^2/rho^2 for a tile of radius R whose centre lies rho from
// the target: with unit vectors, E >= |x_i - c|^2 / 2 and
// |x_i.o_j| <= rho R + |c.o_j|, |c.o_j| <= R^2 / 2 as
// 1 - |c|^2 = mean_j |o_j|^2 <= R^2), so its float32 rounding is t^2 of a
// float32 pair's: about
// 8 eps32 t^2 |w| / d for the rounding of delta^2 / D (x_i x x_j has length
// about d = |x_i - x_j|) and 8 eps32 t^2 |w| / d^2 for that of the records
// w x_j and of the sums (the whole vector, radial part included).  A
// (target group, tile) pair runs the float64 pass instead when the tile
// holds a target's own index, the group's sphere reaches the tile's, t may
// exceed kBoundedTmax, or the rounding bound 8 eps32 t^2 sum|w| /
// (d_min min(d_min, 1)) of the remainder exceeds tau = 0.3 err_target S,
// d_min the least distance between the two spheres and S the velocity scale
// max |x_i x S_i| (4 PI |u|): over the previous RK4 stage's targets inside
// an advance (the field moves by O(dt) between stages), else over the probe
// pass's targets (a lower bound of the true maximum, so tau errs on the
// safe side): vortex-core tiles and near tiles run in float64, the
// background far field in float32.
//
// A (group, tile) pair that fails these tests is tried again per 64-source
// sub-tile, each with its own centre c', radius and float64 moments: about
// half the radius, so about a quarter of the rounding bound.  A CPU
// emulation of this routing at level 7 (Gaussian vortex, target 1e-10)
// puts 4.7% of the pairs in float64, against 15.7% at whole tiles.  A
// sub-tile runs in float64 for every target of a lane when it does for one
// of them.  The float32 records hold o' = x_j - c' (relative to the
// sub-tile), so that the sub-tiles need no records of their own (a second
// set would crowd the source records out of the L2); a whole-tile
// expansion about c takes delta = x_i.o' + x_i.(c' - c), the second term
// once per target and sub-tile, and its rounding then follows
// |o'| + |c' - c| <= 2R in place of |o| <= R.  Four pairs share one
// v_rcp_f32 of the product of their denominators (all >= (1 - kBoundedTmax) E,
// so the product stays far from float32 underflow).

constexpr float kBoundedTmax = 0.9f;
constexpr double kBoundedEps32 = 5.9604644775390625e-08;   // 2^-24

struct BoundedLds {
    float4 pos[kMixedTile];         // o'_j, (w_j x_j).x; SCALAR passes: staged with f64
    float2 wgt[kMixedTile];         // (w_j x_j).y, .z; SCALAR passes: even tiles
    StagedSource f64[kMixedTile];   // float64 records, staged when a target needs them
    int4 need[2];                   // per wave: float64 records needed (alternate tiles)
    float2 wgt_alt[kMixedTile];     // SCALAR passes: wgt of odd tiles
};

// Sources up to which a bounded pass reads its position records through
// scalar loads: each wave then fetches every tile's positions itself (16 B
// per source), which pays while the records stay in an XCD's 4 MB L2.
// Level 7 (98,304 sources): 5.46 against 5.66 ms per launch with every
// record staged in LDS (32 B records, whole-tile routing); level 8
// (393,216): 80.4 against 77.1 ms with the present records.
constexpr int kBoundedScalarMaxSources = 131072;

// Remainder sums R over NS float32 records (pos = {o', (w x).x},
// wgt = {(w x).y, (w x).z}) for N targets, from LDS or (SCALAR passes) straight from global memory at
// wave-uniform addresses, where they arrive through scalar loads into
// SGPRs.  delta = x_i.o' + kf: kf = x_i.(c' - c) when the expansion is about
// the tile centre c, 0 when it is about the sub-tile's own centre c'.
template <int N, int NS>
__device__ __forceinline__ void remainder_tile(const float4 *__restrict__ pos,
                                               const float2 *__restrict__ wgt, const float *xf,
                                               const float *yf, const float *zf, const float *Ef,
                                               const float *kf, float *Rx, float *Ry, float *Rz)
{
#pragma unroll 2
    for (int k = 0; k < NS; k += 4) {
        float4 s[4];
        float2 v[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            s[r] = pos[k + r];
            v[r] = wgt[k + r];
        }
#pragma unroll
        for (int t = 0; t < N; ++t) {
            float dl[4], D[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                dl[r] = __builtin_fmaf(xf[t], s[r].x,
                                       __builtin_fmaf(yf[t], s[r].y,
                                                      __builtin_fmaf(zf[t], s[r].z, kf[t])));
                D[r] = Ef[t] - dl[r];
            }
            const float p01 = D[0] * D[1], p23 = D[2] * D[3];
            const float rr = __builtin_amdgcn_rcpf(p01 * p23);
            const float r01 = rr * p23, r23 = rr * p01;
            float f[4];
            f[0] = (dl[0] * dl[0]) * (D[1] * r01);
            f[1] = (dl[1] * dl[1]) * (D[0] * r01);
            f[2] = (dl[2] * dl[2]) * (D[3] * r23);
            f[3] = (dl[3] * dl[3]) * (D[2] * r23);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                Rx[t] = __builtin_fmaf(f[r], s[r].w, Rx[t]);
                Ry[t] = __builtin_fmaf(f[r], v[r].x, Ry[t]);
                Rz[t] = __builtin_fmaf(f[r], v[r].y, Rz[t]);
            }
        }
    }
}

// u += x_i x S of a remainder tile, S = k c + e m + e^2 (P x_i + R) with
// e = 1/E and k = (W0 + e (m.x_i)) e: the tile's float64 expansion plus its
// float32 remainder sum.  mom = {W0, m (3), P (xx yy zz xy xz yz), W}.
__device__ __forceinline__ void bounded_combine(const double *__restrict__ mom, const double4 &sb,
                                                double xi, double yi, double zi, double E,
                                                float Rx, float Ry, float Rz,
                                                double &ux, double &uy, double &uz)
{
    const double W0 = mom[0], m0 = mom[1], m1 = mom[2], m2 = mom[3];
    const double Pxx = mom[4], Pyy = mom[5], Pzz = mom[6], Pxy = mom[7], Pxz = mom[8],
                 Pyz = mom[9];
    const double e = recip_f64(E);
    const double k = (W0 + (m0 * xi + m1 * yi + m2 * zi) * e) * e;
    const double e2 = e * e;
    const double Bx = (Pxx * xi + Pxy * yi + Pxz * zi) + (double)Rx;
    const double By = (Pxy * xi + Pyy * yi + Pyz * zi) + (double)Ry;
    const double Bz = (Pxz * xi + Pyz * yi + Pzz * zi) + (double)Rz;
    const double Sx = k * sb.x + (e * m0 + e2 * Bx);
    const double Sy = k * sb.y + (e * m1 + e2 * By);
    const double Sz = k * sb.z + (e * m2 + e2 * Bz);
    ux += yi * Sz - zi * Sy;
    uy += zi * Sx - xi * Sz;
    uz += xi * Sy - yi * Sx;
}

// Whether a (target group, source block) pair of a bounded pass runs in
// float64: the block holds a target's own index (diag), the group sphere gb
// reaches the block sphere sb, t may exceed kBoundedTmax, or the remainder's
// rounding bound 8 eps32 t^2 W / (d_min min(d_min, 1)) exceeds tau
// (W = sum |w| over the block).
__device__ __forceinline__ bool bounded_route64(const double4 &gb, const double4 &sb, double W,
                                                bool diag, double tau)
{
    const double dx = gb.x - sb.x, dy = gb.y - sb.y, dz = gb.z - sb.z;
    const double rho_c = sqrt(dx * dx + dy * dy + dz * dz) - gb.w;
    const double R = sb.w;
    if (diag || !(rho_c > R)) return true;
    const double q = R / rho_c;
    const double tb = q * (2.0 + q);
    const double dmin = rho_c - R;
    return tb >= (double)kBoundedTmax ||
           8.0 * kBoundedEps32 * tb * tb * W > tau * dmin * fmin(dmin, 1.0);
}

// Error threshold tau of a bounded pass from its velocity scale.
__device__ __forceinline__ double bounded_tau(const MixedArgs &a)
{
    return 0.3 * a.err_target * *a.scale;
}

// Routing of target group g against source tile tl (bounded_route64 of the
// group sphere against the tile and against each of its sub-tiles) ->
// a.route[tl * n_groups + g]; one pass over every pair before the pair
// kernel, which then reads the wave's bits instead of testing.
__device__ __forceinline__ void bounded_route_body(const MixedArgs &a, int item, double tau)
{
    const int tl = item / a.n_groups, g = item - tl * a.n_groups;
    const int jt = tl * kMixedTile, lo = g * kGroup, self_end = a.base.cls.self_end;
    const double4 gb = a.group_box[g];
    unsigned bits = bounded_route64(gb, a.tile_box[tl], a.tile_mom[(size_t)tl * kTileMoments + 10],
                                    diag_tile(lo, self_end, jt, kMixedTile), tau)
                        ? 1u
                        : 0u;
#pragma unroll
    for (int s = 0; s < kMixedTile / kSubTile; ++s) {
        const double *sm = a.sub_mom + (size_t)(jt / kSubTile + s) * kSubMoments;
        if (bounded_route64(gb, make_double4(sm[0], sm[1], sm[2], sm[3]), sm[14],
                            diag_tile(lo, self_end, jt + s * kSubTile, kSubTile), tau))
            bits |= 2u << s;
    }
    a.route[item] = (unsigned short)bits;
}

// u += x_i x S over NS float32 records (pos, wgt) and their float64
// expansion (sphere sb, moments mom), for N targets.  TILE: the expansion
// is about the tile centre and the records span the tile's sub-tiles
// (offsets c' - c in mom); else the records are one sub-tile's and the
// expansion is about its own centre.
template <int N, int NS, bool TILE>
__device__ __forceinline__ void remainder_block(const float4 *__restrict__ pos,
                                                const float2 *__restrict__ wgt, const double4 &sb,
                                                const double *__restrict__ mom, const double *c0,
                                                const double *xi, const double *yi,
                                                const double *zi, const float *xf,
                                                const float *yf, const float *zf, double *ux,
                                                double *uy, double *uz)
{
    float Rx[N], Ry[N], Rz[N], Ef[N], kf[N];
    double E[N];
#pragma unroll
    for (int t = 0; t < N; ++t) {
        E[t] = c0[t] - ((xi[t] * sb.x + yi[t] * sb.y) + zi[t] * sb.z);
        Ef[t] = (float)E[t];
        Rx[t] = Ry[t] = Rz[t] = 0.f;
        kf[t] = 0.f;
    }
    if constexpr (TILE) {
        // unrolled, so that the offsets arrive with the tile's moments
#pragma unroll
        for (int s = 0; s < NS / kSubTile; ++s) {
            const float4 dc = ((const float4 *)(mom + 12))[s];
#pragma unroll
            for (int t = 0; t < N; ++t)
                kf[t] = __builtin_fmaf(xf[t], dc.x, __builtin_fmaf(yf[t], dc.y, zf[t] * dc.z));
            remainder_tile<N, kSubTile>(pos + s * kSubTile, wgt + s * kSubTile, xf, yf, zf, Ef,
                                        kf, Rx, Ry, Rz);
        }
    } else {
        remainder_tile<N, NS>(pos, wgt, xf, yf, zf, Ef, kf, Rx, Ry, Rz);
    }
#pragma unroll
    for (int t = 0; t < N; ++t)
        bounded_combine(mom, sb, xi[t], yi[t], zi[t], E[t], Rx[t], Ry[t], Rz[t], ux[t], uy[t],
                        uz[t]);
}

// u += x_i x S over NS float64 records (staged in LDS, or HBM records read
// through scalar loads; the first is source jt), for N targets; diag: the
// block holds the own index of some target.
template <int N, int NS, class Rec>
__device__ __forceinline__ void f64_block(const Rec *__restrict__ lds, int jt, bool diag,
                                          const double *c0, const int *self, const double *xi,
                                          const double *yi, const double *zi, double *ux,
                                          double *uy, double *uz)
{
    double nx[N], ny[N], nz[N], cc[N], sx[N], sy[N], sz[N];
    int sf[N];
#pragma unroll
    for (int t = 0; t < N; ++t) {
        nx[t] = -xi[t];
        ny[t] = -yi[t];
        nz[t] = -zi[t];
        cc[t] = c0[t];
        sf[t] = self[t];
        sx[t] = sy[t] = sz[t] = 0.0;
    }
    if (diag)
        tile_f64<N, NS, 4, true>(lds, jt, nx, ny, nz, cc, sf, sx, sy, sz);
    else
        tile_f64<N, NS, 4, false>(lds, jt, nx, ny, nz, cc, sf, sx, sy, sz);
#pragma unroll
    for (int t = 0; t < N; ++t) {
        ux[t] += yi[t] * sz[t] - zi[t] * sy[t];
        uy[t] += zi[t] * sx[t] - xi[t] * sz[t];
        uz[t] += xi[t] * sy[t] - yi[t] * sx[t];
    }
}

// Source tiles (first + k step) mod n_src_tiles, k in [k0, k1), for target
// tile `ttile` of a bounded pass: the sums x_i x S_i of target i go to
// out[i - out_off] and the rows out_stride, 2 out_stride after it.
template <int T, bool SCALAR>
__device__ __forceinline__ void pair_bounded_range(const MixedArgs &a, int ttile, int first,
                                                   int step, int k0, int k1, double *out,
                                                   int out_stride, int out_off, BoundedLds &lds,
                                                   const float4 *__restrict__ g32,
                                                   const float2 *__restrict__ gw,
                                                   const double *__restrict__ gmom,
                                                   const double *__restrict__ gsub,
                                                   const unsigned short *__restrict__ groute)
{
    // g32, gw: the float32 records a.src32, a.srcw, and gmom, gsub: the
    // moments a.tile_mom, a.sub_mom, as the kernel's no-alias arguments
    // (read-only here, so the compiler may use scalar loads for them); the
    // float32 records are read from LDS instead when not SCALAR
    const int tile_base = ttile * (kBlock * T);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int lane = threadIdx.x & 63;
    const TargetClasses cls = a.base.cls;

    // per-target state kept small (4 waves per SIMD): the group spheres are
    // re-read per tile (cached), c0 and the self index recomputed where
    // needed
    double xi[T], yi[T], zi[T], ux[T], uy[T], uz[T];
    float xf[T], yf[T], zf[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + wave * (64 * T) + t * 64 + lane;
        // targets and partial rows stream (non-temporal): the L2 keeps the
        // source records, which every workgroup re-reads (level 7: FETCH_SIZE
        // 273 against 483 MiB per launch, 5.33 against 5.46 ms, whole-tile
        // routing)
        xi[t] = __builtin_nontemporal_load(&a.base.tx[i]);
        yi[t] = __builtin_nontemporal_load(&a.base.ty[i]);
        zi[t] = __builtin_nontemporal_load(&a.base.tz[i]);
        xf[t] = (float)xi[t];
        yf[t] = (float)yi[t];
        zf[t] = (float)zi[t];
        ux[t] = uy[t] = uz[t] = 0.0;
    }
    auto c0_of = [&](int t) {
        const int i = tile_base + wave * (64 * T) + t * 64 + lane;
        return (i >= cls.smooth_begin && i < cls.smooth_end) ? 1.0 + cls.eps2 : 1.0;
    };
    auto self_of = [&](int t) {
        const int i = tile_base + wave * (64 * T) + t * 64 + lane;
        return i < cls.self_end ? i : -1;
    };
    // first target of this wave's group for t: wave w holds the T
    // neighbouring groups from tile_base + 64 T w (lane l target
    // tile_base + 64 (T w + t) + l), whose routing agrees more often than
    // that of groups kBlock apart (level 7: 5.08 against 5.13 ms per launch)
    auto group_of = [&](int t) { return tile_base + wave * (64 * T) + t * 64; };
    // every wave has read the need slots of the previous range
    __syncthreads();

    const int nst = a.base.n_src_pad / kMixedTile;
    for (int k = k0; k < k1; ++k) {
        const int tl = (int)(((long long)k * step + first) % nst);
        const int jt = tl * kMixedTile;
        const int par = (k - k0) & 1;
        // SCALAR passes read o' through scalar loads and w x from LDS (the
        // scalar cache alone would carry both for 16 waves per CU), staged
        // for every tile in two alternating buffers that the per-tile
        // barrier below publishes
        float2 *wgt_lds = (SCALAR && par) ? lds.wgt_alt : lds.wgt;
        if (!SCALAR) {
            __syncthreads();
            lds.pos[threadIdx.x] = g32[jt + threadIdx.x];
        }
        wgt_lds[threadIdx.x] = gw[jt + threadIdx.x];
        const double4 sb = a.tile_box[tl];
        const double *mom = gmom + (size_t)tl * kTileMoments;
        const double *sub = gsub + (size_t)(jt / kSubTile) * kSubMoments;
        // the routing of the wave's T groups (16 bits each, adjacent: one
        // aligned scalar load; the first group's index is a multiple of T)
        const unsigned short *rp = groute + (size_t)tl * a.n_groups + group_of(0) / kGroup;
        unsigned long long rw;
        if constexpr (T == 1)
            rw = rp[0];
        else if constexpr (T == 2)
            rw = *(const unsigned *)rp;
        else
            rw = *(const unsigned long long *)rp;
        bool any = false;
        unsigned sub4 = 0;   // sub-tiles in float64 for some group of the wave
#pragma unroll
        for (int t = 0; t < T; ++t) {
            any |= (rw >> (16 * t)) & 1;
            sub4 |= (unsigned)(rw >> (16 * t + 1)) & 15u;
        }
        // one barrier: each wave posts its need in this tile's slot, which
        // is next written two tiles on, after another barrier (and ranges
        // start with one)
        if (lane == 0) (&lds.need[par].x)[wave] = any ? 1 : 0;
        __syncthreads();
        const int4 need = lds.need[par];
        if (need.x | need.y | need.z | need.w) {
            stage_sources(lds.f64, a.base.src + jt, kMixedTile);
            if (SCALAR) {
                // the sub-tiles' float32 loops read LDS too: short runs of
                // scalar loads would each wait on the L2
                lds.pos[threadIdx.x] = g32[jt + threadIdx.x];
            }
            __syncthreads();
        }
        if (!any) {
            // every target of the lane: the whole tile in float32
            double c0[T];
#pragma unroll
            for (int t = 0; t < T; ++t) c0[t] = c0_of(t);
            remainder_block<T, kMixedTile, true>(SCALAR ? g32 + jt : lds.pos, wgt_lds, sb, mom, c0,
                                                 xi, yi, zi, xf, yf, zf, ux, uy, uz);
            continue;
        }
        // the tile again per sub-tile, each in float32 or float64 (for every
        // target of the lane when for one of them: the lane's targets share
        // its loops)
#pragma unroll 1
        for (int s = 0; s < kMixedTile / kSubTile; ++s) {
            const int js = jt + s * kSubTile;
            const double *sm = sub + s * kSubMoments;
            double c0[T];
#pragma unroll
            for (int t = 0; t < T; ++t) c0[t] = c0_of(t);
            if ((sub4 >> s) & 1) {
                int self[T];
                bool diag = false;
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    self[t] = self_of(t);
                    diag |= diag_tile(group_of(t), cls.self_end, js, kSubTile);
                }
                f64_block<T, kSubTile>(lds.f64 + s * kSubTile, js, diag, c0, self, xi, yi, zi, ux,
                                       uy, uz);
            } else {
                remainder_block<T, kSubTile, false>(
                    lds.pos + s * kSubTile, wgt_lds + s * kSubTile,
                    make_double4(sm[0], sm[1], sm[2], sm[3]), sm + 4, c0, xi, yi, zi, xf, yf, zf,
                    ux, uy, uz);
            }
        }
    }

#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + wave * (64 * T) + t * 64 + lane - out_off;
        __builtin_nontemporal_store(ux[t], &out[i]);
        __builtin_nontemporal_store(uy[t], &out[out_stride + i]);
        __builtin_nontemporal_store(uz[t], &out[2 * (size_t)out_stride + i]);
    }
}

// Persistent body of a segmented bounded pass (source tiles in stride order).
template <int T, bool SCALAR>
__device__ __forceinline__ void pair_bounded_segments(const MixedArgs &a, BoundedLds &lds,
                                                      int *seg_slot,
                                                      const float4 *__restrict__ g32,
                                                      const float2 *__restrict__ gw,
                                                      const double *__restrict__ gmom,
                                                      const double *__restrict__ gsub,
                                                      const unsigned short *__restrict__ groute)
{
    constexpr int TT = kBlock * T;
    run_segments<kMixedTile>(a.base.plan, a.base.queue, seg_slot,
                             [&](int t, int j0, int j1, int seg) {
        pair_bounded_range<T, SCALAR>(a, t, 0, a.base.plan.src_step, j0 / kMixedTile, j1 / kMixedTile,
                              a.base.partial + (size_t)(seg + t) * 3 * TT, TT, t * TT, lds, g32, gw,
                              gmom, gsub, groute);
    });
}

// ---------------------------------------------------------------------------
// Block reductions (fixed order: deterministic)
// ---------------------------------------------------------------------------

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_max(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// Mixed-precision per-stage preparation bodies
// ---------------------------------------------------------------------------

// Source tile `tile` (kMixedTile threads): centre c (mean of the valid
// sources), radius, and the float32 records pos = {o, B}, wgt = {w o, w}
// and c_j, with o = fl32(x - c), c_j = (1 - |x|^2) / 2 and
// B = |o|^2 / 2 + c_j, w o formed in float64 from the rounded o.
__device__ __forceinline__ void mixed_source_prep_body(const Source *__restrict__ src,
                                                       int n_sources, float4 *__restrict__ src32,
                                                       float4 *__restrict__ srcw,
                                                       float *__restrict__ srcc,
                                                       double4 *__restrict__ tile_box, int tile,
                                                       double *sh)
{
    const int j = tile * kMixedTile + threadIdx.x;
    const Source s = src[j];
    const bool valid = j < n_sources;
    const double cnt = block_sum(valid ? 1.0 : 0.0, sh);
    const double mx = block_sum(valid ? s.x : 0.0, sh);
    const double my = block_sum(valid ? s.y : 0.0, sh);
    const double mz = block_sum(valid ? s.z : 0.0, sh);
    const double inv = cnt > 0 ? 1.0 / cnt : 0.0;
    const double cx = mx * inv, cy = my * inv, cz = mz * inv;
    const double dx = s.x - cx, dy = s.y - cy, dz = s.z - cz;
    const double r = block_max(valid ? sqrt(dx * dx + dy * dy + dz * dz) : -1.0, sh);
    if (threadIdx.x == 0) tile_box[tile] = make_double4(cx, cy, cz, r);
    const float ox = (float)dx, oy = (float)dy, oz = (float)dz;
    const double cj = 0.5 * (1.0 - (s.x * s.x + s.y * s.y + s.z * s.z));
    const double o2 = (double)ox * ox + (double)oy * oy + (double)oz * oz;
    src32[j] = make_float4(ox, oy, oz, (float)(0.5 * o2 + cj));
    srcw[j] = make_float4((float)(s.w * ox), (float)(s.w * oy), (float)(s.w * oz), (float)s.w);
    srcc[j] = (float)cj;
}

// Targets [blk * 256, blk * 256 + 256), one wave per group of kGroup:
// centre and radius of each group.
__device__ __forceinline__ void mixed_target_prep_body(const PairArgs &b, int n_targets,
                                                       double4 *__restrict__ group_box, int blk,
                                                       double *sh)
{
    const int i = blk * 256 + threadIdx.x;
    const double x = b.tx[i], y = b.ty[i], z = b.tz[i];
    // a non-finite target (already reported) is left out of its group's
    // sphere, so that it does not spoil its neighbours' sums
    const bool valid = i < n_targets && finite3(x, y, z);
    const int lane = threadIdx.x & 63, base = threadIdx.x & ~63;
    double v[4] = {valid ? 1.0 : 0.0, valid ? x : 0.0, valid ? y : 0.0, valid ? z : 0.0};
    double m[4];
    for (int q = 0; q < 4; ++q) {
        sh[threadIdx.x] = v[q];
        __syncthreads();
        for (int s = 32; s > 0; s >>= 1) {
            if (lane < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        m[q] = sh[base];
        __syncthreads();
    }
    const double inv = m[0] > 0 ? 1.0 / m[0] : 0.0;
    const double cx = m[1] * inv, cy = m[2] * inv, cz = m[3] * inv;
    const double dx = x - cx, dy = y - cy, dz = z - cz;
    sh[threadIdx.x] = valid ? sqrt(dx * dx + dy * dy + dz * dz) : -1.0;
    __syncthreads();
    for (int s = 32; s > 0; s >>= 1) {
        if (lane < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (lane == 0) group_box[i / kGroup] = make_double4(cx, cy, cz, sh[base]);
    __syncthreads();
}

// Sums over a wave of N values per lane by a butterfly: both lanes of a
// step add the same two values, so every lane ends with the same sums.
template <int N>
__device__ __forceinline__ void wave_sums(double (&v)[N])
{
#pragma unroll
    for (int q = 0; q < N; ++q)
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) v[q] += __shfl_xor(v[q], m, 64);
}

// Sums over a workgroup of 256 threads (4 waves) of N values per thread:
// the wave sums (wave_sums), then added in wave order.  Every thread gets
// the N totals; fixed order, deterministic.  sh: at least 4 N doubles.
template <int N>
__device__ __forceinline__ void block_sums(double (&v)[N], double *sh)
{
    wave_sums<N>(v);
    const int wave = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int q = 0; q < N; ++q) sh[wave * N + q] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = ((sh[q] + sh[N + q]) + sh[2 * N + q]) + sh[3 * N + q];
    __syncthreads();
}

// Source tile `tile` of a bounded pass (kMixedTile threads): the centre c
// (mean of the valid sources) and radius -> tile_box, the float64 moments
// {W0 = sum w, m = sum w o, P = sum w o o^T (xx yy zz xy xz yz),
// W = sum |w|, 0} of o = x - c -> mom[tile * kTileMoments ...], and
// wgt = {fl32((w x).y), fl32((w x).z)} (float2 records in srcw's first
// half).  Each wave does the same for
// its 64-source sub-tile: centre c' and radius and the moments of
// o' = x - c' -> sub_mom[sub-tile * kSubMoments ...], the offset
// {fl32(c' - c), 0} -> the tile's mom entry (doubles 12 + 2 wave,
// 13 + 2 wave), and the float32 records pos = {fl32(o'), fl32((w x).x)}.
__device__ __forceinline__ void bounded_source_prep_body(const Source *__restrict__ src,
                                                         int n_sources, float4 *__restrict__ src32,
                                                         float4 *__restrict__ srcw,
                                                         double4 *__restrict__ tile_box,
                                                         double *__restrict__ mom,
                                                         double *__restrict__ sub_mom, int tile,
                                                         double *sh)
{
    const int j = tile * kMixedTile + threadIdx.x;
    const Source s = src[j];
    const bool valid = j < n_sources;
    const double w = valid ? s.w : 0.0;
    double cs[4] = {valid ? 1.0 : 0.0, valid ? s.x : 0.0, valid ? s.y : 0.0, valid ? s.z : 0.0};
    double ws[4] = {cs[0], cs[1], cs[2], cs[3]};
    block_sums<4>(cs, sh);
    const double inv = cs[0] > 0 ? 1.0 / cs[0] : 0.0;
    const double cx = cs[1] * inv, cy = cs[2] * inv, cz = cs[3] * inv;
    const double ox = s.x - cx, oy = s.y - cy, oz = s.z - cz;
    const double r = block_max(valid ? sqrt(ox * ox + oy * oy + oz * oz) : -1.0, sh);
    ((float2 *)srcw)[j] = make_float2((float)(w * s.y), (float)(w * s.z));
    double v[11] = {w, w * ox, w * oy, w * oz, w * ox * ox, w * oy * oy, w * oz * oz,
                    w * ox * oy, w * ox * oz, w * oy * oz, fabs(w)};
    block_sums<11>(v, sh);
    if (threadIdx.x == 0) {
        tile_box[tile] = make_double4(cx, cy, cz, r);
        double *m = mom + (size_t)tile * kTileMoments;
#pragma unroll
        for (int q = 0; q < 11; ++q) m[q] = v[q];
        m[11] = 0.0;
    }

    // the wave's sub-tile (all lanes hold the same wave sums)
    wave_sums<4>(ws);
    const double winv = ws[0] > 0 ? 1.0 / ws[0] : 0.0;
    const double sx = ws[1] * winv, sy = ws[2] * winv, sz = ws[3] * winv;
    const double px = s.x - sx, py = s.y - sy, pz = s.z - sz;
    double sr = valid ? sqrt(px * px + py * py + pz * pz) : -1.0;
#pragma unroll
    for (int k = 32; k > 0; k >>= 1) sr = fmax(sr, __shfl_xor(sr, k, 64));
    src32[j] = make_float4((float)px, (float)py, (float)pz, (float)(w * s.x));
    double u[11] = {w, w * px, w * py, w * pz, w * px * px, w * py * py, w * pz * pz,
                    w * px * py, w * px * pz, w * py * pz, fabs(w)};
    wave_sums<11>(u);
    if ((threadIdx.x & 63) == 0) {
        double *m = sub_mom + (size_t)(j / kSubTile) * kSubMoments;
        m[0] = sx;
        m[1] = sy;
        m[2] = sz;
        m[3] = sr;
#pragma unroll
        for (int q = 0; q < 11; ++q) m[4 + q] = u[q];
        m[15] = 0.0;
        ((float4 *)(mom + (size_t)tile * kTileMoments + 12))[threadIdx.x / 64] =
            make_float4((float)(sx - cx), (float)(sy - cy), (float)(sz - cz), 0.f);
    }
}

// Velocity-scale probe of the bounded pass: for target row `row` (the
// first of a target group), |x_i x S_i| over every source, in float32
// (a scale, not a result: ~1e-6 relative is plenty), the self pair and
// coincident points left out; combined into *scale with atomicMax on the
// bits of a non-negative double.  One workgroup of 256 threads per probe.
__device__ __forceinline__ void bounded_probe_body(const PairArgs &b, int n_targets,
                                                   int n_sources, int row,
                                                   unsigned long long *scale_bits, double *sh)
{
    if (row >= n_targets) return;
    const double x = b.tx[row], y = b.ty[row], z = b.tz[row];
    if (!finite3(x, y, z)) return;
    const TargetClasses cls = b.cls;
    const float e2 = (row >= cls.smooth_begin && row < cls.smooth_end) ? (float)cls.eps2 : 0.f;
    const int self = row < cls.self_end ? row : -1;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int j = threadIdx.x; j < n_sources; j += 256) {
        const Source s = b.src[j];
        const float dx = (float)(s.x - x), dy = (float)(s.y - y), dz = (float)(s.z - z);
        const float h = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
        const float q = (j == self || !(h > 0.f)) ? 0.f
                                                  : (float)s.w * __builtin_amdgcn_rcpf(0.5f * h + e2);
        sx = __builtin_fmaf(q, dx, sx);
        sy = __builtin_fmaf(q, dy, sy);
        sz = __builtin_fmaf(q, dz, sz);
    }
    const double Sx = block_sum(sx, sh), Sy = block_sum(sy, sh), Sz = block_sum(sz, sh);
    if (threadIdx.x == 0) {
        const double ux = y * Sz - z * Sy, uy = z * Sx - x * Sz, uz = x * Sy - y * Sx;
        const double m = sqrt(ux * ux + uy * uy + uz * uz);
        if (m <= 1.7976931348623157e308)
            atomicMax(scale_bits, (unsigned long long)__double_as_longlong(m));
    }
}

// ---------------------------------------------------------------------------
// RK4 stage update body: chunk reduction, velocity, k_s, accumulator,
// next-stage inputs and source records, stage-1 kinetic energy.
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Juckes forcing of the stratosphere model (TracerAndVorticityDistributions.f90)
// ---------------------------------------------------------------------------

// Longitude in [0, 2 PI) by the case split of atan4 (SphereGeom.f90:213-248).
__device__ __forceinline__ double longitude_0_2pi(double y, double x)
{
    if (x == 0.0) return y > 0.0 ? kPi / 2.0 : (y < 0.0 ? 3.0 * kPi / 2.0 : 0.0);
    if (y == 0.0) return x > 0.0 ? 0.0 : kPi;
    const double theta = atan2(fabs(y), fabs(x));
    if (x > 0.0) return y > 0.0 ? theta : 2.0 * kPi - theta;
    return y > 0.0 ? kPi - theta : kPi + theta;
}

// F0 A(t) B(lat) cos(lon) (Juckes_Forcing, :1223-1231): F0 = 0.3 * 4 PI; the
// ramp A (:550-568) rises over [0, 4), holds on [4, 11), falls over
// [11, 15); B (:606-617) = tan^2(PI/3)/tan^2(lat) exp(1 - tan^2(PI/3)/tan^2(lat))
// north of the equator, 0 south of it.
__device__ __forceinline__ double juckes_forcing(double x, double y, double z, double t)
{
    const double tfull = 4.0, tend = 15.0, b = kPi / tfull;
    double a = 0.0;
    if (0.0 <= t && t < tfull)
        a = 0.5 * (1.0 - cos(b * t));
    else if (tfull <= t && t < tend - tfull)
        a = 1.0;
    else if (tend - tfull <= t && t < tend)
        a = 0.5 * (1.0 - cos(b * (t - (tend - tfull)) + kPi));
    const double lat = atan2(z, sqrt(x * x + y * y));
    double bl = 0.0;
    if (lat > 0.0) {
        const double th1 = kPi / 3.0;
        const double tan1sq = tan(th1) * tan(th1);
        const double tl = tan(lat);
        bl = tan1sq / (tl * tl) * exp(1.0 - tan1sq / (tl * tl));
    }
    return 0.3 * 4.0 * kPi * a * bl * cos(longitude_0_2pi(y, x));
}

// Relative vorticity of a forced step at position (x, y, z) and time t:
// absVort - 2 Omega z - F(x, t) (PanelVelocityParallel.f90:686-689, :906-915).
__device__ __forceinline__ double forced_relvort(double absvort, double omega, double x, double y,
                                                 double z, double t)
{
    return absvort - 2.0 * omega * z - juckes_forcing(x, y, z, t);
}

// ---------------------------------------------------------------------------
// Analytic winds of the advection tests (PanelVelocityParallel.f90:1415-1459)
// ---------------------------------------------------------------------------

// NonDivergentWind (divergent = 0) or DivergentWind (divergent = 1) at
// (x, y, z), time t: R = 1, T = 5, zero within 1e-14 of the polar axis.
__device__ __forceinline__ void analytic_wind(int divergent, double x, double y, double z,
                                              double t, double &wx, double &wy, double &wz)
{
    const double RR = 1.0, TT = 5.0;
    const double raxis = sqrt(x * x + y * y);
    if (!(raxis > 1e-14)) {
        wx = wy = wz = 0.0;
        return;
    }
    const double lat = atan2(z, sqrt(x * x + y * y));
    const double lon = longitude_0_2pi(y, x);
    const double phase = lon - 2.0 * kPi * t / TT;
    double u, v;
    if (!divergent) {
        u = 10.0 * RR / TT * sin(phase) * sin(phase) * sin(2.0 * lat) * cos(kPi * t / TT) +
            2.0 * kPi * RR / TT * cos(lat);
        v = 10.0 * RR / TT * sin(2.0 * phase) * cos(lat) * cos(kPi * t / TT);
    } else {
        u = -5.0 * RR / TT * sin(phase / 2.0) * sin(phase / 2.0) * sin(2.0 * lat) * cos(lat) *
                cos(lat) * cos(kPi * t / TT) +
            2.0 * kPi * RR * cos(lat) / TT;
        v = 5.0 * RR / (2.0 * TT) * sin(phase) * cos(lat) * cos(lat) * cos(lat) * cos(kPi * t / TT);
    }
    wx = -u * y / raxis - v * x * z / raxis;
    wy = u * x / raxis - v * y * z / raxis;
    wz = v * raxis;
}

// One AdvectionRK4 / DivergentAdvectionRK4 step of target i
// (PanelVelocityParallel.f90:922-1279): the pointwise wind needs no other
// target, so the four stages run in registers; an active panel's source
// record follows its position (its vorticity is unchanged).
__device__ __forceinline__ void advection_step_body(const StepState &st, int i, double t,
                                                    double dt, int divergent)
{
    const int stride = st.t_stride;
    const double p[3] = {st.x0[i], st.x0[stride + i], st.x0[2 * stride + i]};
    double k[4][3], in[3];
    analytic_wind(divergent, p[0], p[1], p[2], t, k[0][0], k[0][1], k[0][2]);
    static constexpr double a[3] = {0.5, 0.5, 1.0};
    for (int s = 0; s < 3; ++s) {
        for (int c = 0; c < 3; ++c) k[s][c] = dt * k[s][c];
        for (int c = 0; c < 3; ++c) in[c] = a[s] == 1.0 ? p[c] + k[s][c] : p[c] + a[s] * k[s][c];
        const double ts = s < 2 ? t + 0.5 * dt : t + dt;
        analytic_wind(divergent, in[0], in[1], in[2], ts, k[s + 1][0], k[s + 1][1], k[s + 1][2]);
    }
    for (int c = 0; c < 3; ++c) k[3][c] = dt * k[3][c];
    double q[3];
    for (int c = 0; c < 3; ++c)
        q[c] = p[c] + k[0][c] / 6.0 + k[1][c] / 3.0 + k[2][c] / 3.0 + k[3][c] / 6.0;
    if (!finite3(q[0], q[1], q[2])) flag_nonfinite(st.status, st.member, i);
    for (int c = 0; c < 3; ++c) {
        st.x0[c * stride + i] = q[c];
        st.xs[c * stride + i] = q[c];
    }
    if (i < st.n_active) st.src[i] = Source{q[0], q[1], q[2], source_weight(st, st.zeta0[i], i)};
}

// Sum of the partials of target i of a pass with plan p, in a fixed order
// (chunk 0 first, or the first segment touching i's tile first).
__device__ __forceinline__ void reduce_partials(const double *__restrict__ partial,
                                                const PairPlan &p, int stride, int i,
                                                double &Sx, double &Sy, double &Sz)
{
    Sx = Sy = Sz = 0.0;
    if (p.segments > 0) {
        const int tt = kBlock * p.targets_per_lane;
        const int t = i / tt;
        const int loc = i - t * tt;
        const int first = t * p.n_src_tiles;
        const int s_hi = (first + p.n_src_tiles - 1) / p.seg_len;
        for (int s = first / p.seg_len; s <= s_hi; ++s) {
            const double *q = partial + (size_t)(s + t) * 3 * tt;
            Sx += q[loc];
            Sy += q[tt + loc];
            Sz += q[2 * tt + loc];
        }
        return;
    }
    for (int c = 0; c < p.n_chunks; ++c) {
        const double *q = partial + (size_t)(3 * c) * stride;
        Sx += q[i];
        Sy += q[stride + i];
        Sz += q[2 * (size_t)stride + i];
    }
}

// Targets [blk * 256, blk * 256 + 256) of one step state; at stage 0 the
// block's sum of ke * area goes to st.ke_part[blk].
__device__ __forceinline__ void stage_update_body(const StepState &st, const StageParams &sp,
                                                  int blk, double *sh)
{
    // every operation rounded on its own, as the reference's BVERK4 is
#pragma clang fp contract(off)
    const int i = blk * 256 + threadIdx.x;
    const int stride = st.t_stride;
    double ke_local = 0.0, u_abs = 0.0;
    if (i < st.n_targets) {
        double Sx, Sy, Sz;
        reduce_partials(st.partial, sp.plan, stride, i, Sx, Sy, Sz);
        if (sp.sums_out && i < st.n_active) {   // kept for the adjoint's reverse sweep
            sp.sums_out[i] = Sx;
            sp.sums_out[sp.sums_stride + i] = Sy;
            sp.sums_out[2 * (size_t)sp.sums_stride + i] = Sz;
        }
        const double x = st.xs[i], y = st.xs[stride + i], z = st.xs[2 * stride + i];
        // u = sum_j w_j (x_i x x_j)/d_ij / (-4 PI)    (PanelVelocityParallel.f90:1332)
        double ux, uy, uz;
        velocity_from_sums(sp.plan.exact | sp.plan.crossed, x, y, z, Sx, Sy, Sz, ux, uy, uz);
        // a coincident source and target (1 - x.y = 0) makes the sums NaN or
        // Inf, as the reference's 0/0 does: reported, not silently carried
        if (!finite3(ux, uy, uz))
            flag_nonfinite(st.status, st.member, i);
        else if (st.vel_max)   // in the pair sums' units: |x_i x S_i| = 4 PI |u|
            u_abs = 4.0 * kPi * sqrt((ux * ux + uy * uy) + uz * uz);
        const double kx = sp.dt * ux, ky = sp.dt * uy, kz = sp.dt * uz;
        const double *base = (sp.stage == 0) ? st.x0 : st.acc;
        // x + k1/6 + k2/3 + k3/3 + k4/6, left to right (:637-644)
        const double ax = base[i] + kx / sp.divisor;
        const double ay = base[stride + i] + ky / sp.divisor;
        const double az = base[2 * stride + i] + kz / sp.divisor;
        const bool active = i < st.n_active;
        const bool forced = st.panel_absvort != nullptr;
        double zeta_next = 0.0, zacc = 0.0;
        if (active) {
            if (!forced) {
                const double dvort = -(2.0 * sp.omega * uz);   // :1333
                const double kv = sp.dt * dvort;
                zacc = ((sp.stage == 0) ? st.zeta0[i] : st.zacc[i]) + kv / sp.divisor;
                st.zacc[i] = zacc;
                zeta_next = st.zeta0[i] + sp.a_next * kv;
            }
            if (sp.stage == 0) {
                const double e = (ux * ux + uy * uy) + uz * uz;   // :495-497
                st.ke[i] = e;
                ke_local = e * st.area[i];
            }
        }
        if (sp.stage < 3) {
            st.acc[i] = ax;
            st.acc[stride + i] = ay;
            st.acc[2 * stride + i] = az;
            // next stage input x + a k  (:513-516, :555-558, :597-600)
            const double nx = st.x0[i] + sp.a_next * kx;
            const double ny = st.x0[stride + i] + sp.a_next * ky;
            const double nz = st.x0[2 * stride + i] + sp.a_next * kz;
            st.xs[i] = nx;
            st.xs[stride + i] = ny;
            st.xs[2 * stride + i] = nz;
            if (active) {
                if (forced)   // stage inputs' vorticity (:744-747, :792-795, :837-840)
                    zeta_next = forced_relvort(st.panel_absvort[i], sp.omega, nx, ny, nz, sp.t_next);
                st.src[i] = Source{nx, ny, nz, source_weight(st, zeta_next, i)};
            }
        } else {
            double nx = ax, ny = ay, nz = az;
            if (sp.renormalize) {   // RenormalizeGrid, ParticlesEdgesPanels2.f90:1690-1702
                const double r = sqrt((nx * nx + ny * ny) + nz * nz);
                nx /= r;
                ny /= r;
                nz /= r;
            }
            st.x0[i] = nx;
            st.x0[stride + i] = ny;
            st.x0[2 * stride + i] = nz;
            st.xs[i] = nx;
            st.xs[stride + i] = ny;
            st.xs[2 * stride + i] = nz;
            if (active) {
                if (forced)   // :906-911
                    zacc = forced_relvort(st.panel_absvort[i], sp.omega, nx, ny, nz, sp.t_next);
                st.zeta0[i] = zacc;
                st.src[i] = Source{nx, ny, nz, source_weight(st, zacc, i)};
            }
            // particles carry absolute vorticity: relVort = absVort - 2 Omega z (:651),
            // minus the forcing in a forced step (:912-915)
            if (st.particle_absvort && i >= st.particle_begin && i < st.particle_end) {
                const int p = i - st.particle_begin;
                st.particle_relvort[p] =
                    forced ? forced_relvort(st.particle_absvort[p], sp.omega, nx, ny, nz, sp.t_next)
                           : st.particle_absvort[p] - 2.0 * sp.omega * nz;
            }
        }
    }
    if (sp.stage == 0) {
        const double b = block_sum(ke_local, sh);
        if (threadIdx.x == 0) st.ke_part[blk] = b;
    }
    if (st.vel_max) {
        const double m = block_max(u_abs, sh);
        if (threadIdx.x == 0 && m > 0.0 && m <= 1.7976931348623157e308)
            atomicMax((unsigned long long *)st.vel_max, (unsigned long long)__double_as_longlong(m));
    }
}

// totalKE of the reference-order mode: 1/2 sum_j ke_j A_j added one panel
// after the other, as the reference's sum intrinsic does (:498).  The
// workgroup forms 256 products ke_j A_j at a time (each rounded on its own,
// as in the reference) into sh; lane 0 adds them in index order.
__device__ __forceinline__ void ke_total_sequential_body(const double *__restrict__ ke,
                                                         const double *__restrict__ area, int n,
                                                         double *__restrict__ out, double *sh)
{
#pragma clang fp contract(off)
    double t = 0.0;
    for (int base = 0; base < n; base += 256) {
        const int j = base + threadIdx.x;
        if (j < n) sh[threadIdx.x] = ke[j] * area[j];
        __syncthreads();
        if (threadIdx.x == 0) {
            const int m = min(256, n - base);
            for (int k = 0; k < m; ++k) t = t + sh[k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = 0.5 * t;
}

// totalKE = 1/2 sum_j ke_j A_j (:498), the n block partials added in order.
__device__ __forceinline__ void ke_total_body(const double *__restrict__ part, int n,
                                              double *__restrict__ out, double *sh)
{
    double v = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) v += part[k];
    const double b = block_sum(v, sh);
    if (threadIdx.x == 0) *out = 0.5 * b;
}

// Stage-1 inputs of a step from the step-start state: xs = x0 and the
// source records {x, zeta A} of the active panels.
__device__ __forceinline__ void prepare_step_body(const StepState &st, int blk)
{
#pragma clang fp contract(off)
    const int i = blk * 256 + threadIdx.x;
    if (i >= st.n_targets) return;
    const int stride = st.t_stride;
    const double x = st.x0[i], y = st.x0[stride + i], z = st.x0[2 * stride + i];
    st.xs[i] = x;
    st.xs[stride + i] = y;
    st.xs[2 * stride + i] = z;
    if (i < st.n_active) {
        // a forced step's first stage uses absVort - 2 Omega z - F(x, t) (:686-689)
        const double w = st.panel_absvort
                             ? forced_relvort(st.panel_absvort[i], st.omega, x, y, z, st.t_step)
                             : st.zeta0[i];
        st.src[i] = Source{x, y, z, source_weight(st, w, i)};
    }
}

}  // namespace bve
