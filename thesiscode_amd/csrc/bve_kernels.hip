// Biot-Savart pair sums and the fused RK4 stage update, hand-written for
// gfx950 (CDNA4, wave64).
//
// The velocity at a target x_i is (PanelVelocityParallel.f90:1287-1412)
//
//     u_i = -1/(4 PI) sum_j w_j (x_i x x_j) / d_ij,   d_ij = (1 - x_i.x_j) + eps2,
//
// with w_j = zeta_j A_j.  The pair kernel evaluates the factored form
// S_i = sum_j (w_j / d_ij) x_j and the stage kernel forms u_i = (x_i x S_i) / (-4 PI),
// which saves the per-pair cross product.  Per pair that is 3 FMA for the
// denominator, one v_rcp_f64 plus two Newton steps (4 FMA) for 1/d, one
// multiply by w and 3 FMA into S.
//
// Work decomposition: targets are held in registers (T per lane), sources
// stream through LDS in tiles of AoS {x, y, z, w} records, and the source
// range is cut into chunks so that a launch has enough workgroups for 256
// CUs.  Each (target tile, source chunk) work item writes its own partial
// sums; the stage kernel adds the chunks in a fixed order, so results do
// not depend on scheduling and reruns are bitwise reproducible.
#include "bve_common.h"
#include "bve_kernels.h"

namespace bve {

// ---------------------------------------------------------------------------
// Pair kernel, float64
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
// that pair's denominator (~0) is replaced by 1 and its weight by 0.
template <int T, int TILE, int R, bool MASK>
__device__ __forceinline__ void tile_f64(const Source *__restrict__ lds, int jt,
                                         const double (&nx)[T], const double (&ny)[T],
                                         const double (&nz)[T], const double (&c0)[T],
                                         const int (&self)[T], double (&sx)[T],
                                         double (&sy)[T], double (&sz)[T])
{
#pragma unroll(R == 1 ? 2 : 1)
    for (int k = 0; k < TILE; k += R) {
        Source s[R];
#pragma unroll
        for (int r = 0; r < R; ++r) s[r] = lds[k + r];
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
                q[0] = s[0].w * recip_f64(d[0]);
            } else if constexpr (R == 2) {
                const double rr = recip1_f64(d[0] * d[1]);
                q[0] = s[0].w * d[1] * rr;
                q[1] = s[1].w * d[0] * rr;
            } else {
                const double p01 = d[0] * d[1], p23 = d[2] * d[3];
                const double rr = recip1_f64(p01 * p23);
                const double r01 = rr * p23, r23 = rr * p01;
                q[0] = s[0].w * d[1] * r01;
                q[1] = s[1].w * d[0] * r01;
                q[2] = s[2].w * d[3] * r23;
                q[3] = s[3].w * d[2] * r23;
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (MASK) q[r] = (jt + k + r == self[t]) ? 0.0 : q[r];
                sx[t] = __builtin_fma(q[r], s[r].x, sx[t]);
                sy[t] = __builtin_fma(q[r], s[r].y, sy[t]);
                sz[t] = __builtin_fma(q[r], s[r].z, sz[t]);
            }
        }
    }
}

// Maps workgroup g to a work item so that XCD (g % 8) runs a contiguous run
// of the chunk-major work list: the source chunks an XCD reads stay in its L2.
__device__ __forceinline__ int xcd_work_item(int g, int grid)
{
    return (g % 8) * (grid / 8) + g / 8;
}

template <int T, int TILE, int R>
__global__ __launch_bounds__(kBlock) void pair_sum_f64_kernel(
    const double *__restrict__ tx, const double *__restrict__ ty, const double *__restrict__ tz,
    int t_stride, TargetClasses cls, const Source *__restrict__ src, int n_src_pad,
    PairPlan plan, double *__restrict__ partial)
{
    const int work = xcd_work_item(blockIdx.x, plan.grid);
    if (work >= plan.n_work) return;
    const int chunk = work / plan.n_tiles;
    const int ttile = work - chunk * plan.n_tiles;
    const int tile_base = ttile * (kBlock * T);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);

    double nx[T], ny[T], nz[T], c0[T], sx[T], sy[T], sz[T];
    int self[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x;
        nx[t] = -tx[i];
        ny[t] = -ty[i];
        nz[t] = -tz[i];
        c0[t] = (i >= cls.smooth_begin && i < cls.smooth_end) ? 1.0 + cls.eps2 : 1.0;
        self[t] = (i < cls.self_end) ? i : -1;
        sx[t] = sy[t] = sz[t] = 0.0;
    }

    __shared__ Source lds[TILE];
    const int j_begin = chunk * plan.chunk;
    const int j_end = min(j_begin + plan.chunk, n_src_pad);
    for (int jt = j_begin; jt < j_end; jt += TILE) {
        __syncthreads();
        for (int k = threadIdx.x; k < TILE; k += kBlock) lds[k] = src[jt + k];
        __syncthreads();
        // Does this tile hold the own-index source of a self-excluded target
        // of this wave?  (wave-uniform: depends on the wave index only)
        bool diag = false;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int lo = tile_base + t * kBlock + wave * 64;
            const int hi = min(lo + 64, cls.self_end);
            diag |= (lo < hi) && (lo < jt + TILE) && (jt < hi);
        }
        if (diag)
            tile_f64<T, TILE, R, true>(lds, jt, nx, ny, nz, c0, self, sx, sy, sz);
        else
            tile_f64<T, TILE, R, false>(lds, jt, nx, ny, nz, c0, self, sx, sy, sz);
    }

    double *px = partial + (size_t)(chunk * 3) * t_stride;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x;
        px[i] = sx[t];
        px[t_stride + i] = sy[t];
        px[2 * (size_t)t_stride + i] = sz[t];
    }
}

template <int T, int TILE, int R>
static hipError_t launch_f64(const PairArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((pair_sum_f64_kernel<T, TILE, R>), dim3(a.plan.grid), dim3(kBlock), 0, s,
                       a.tx, a.ty, a.tz, a.t_stride, a.cls, a.src, a.n_src_pad, a.plan, a.partial);
    return hipGetLastError();
}

template <int T, int TILE>
static hipError_t launch_f64_r(const PairArgs &a, hipStream_t s)
{
    switch (a.plan.recip) {
    case 1: return launch_f64<T, TILE, 1>(a, s);
    case 2: return launch_f64<T, TILE, 2>(a, s);
    case 4: return launch_f64<T, TILE, 4>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_pair_sum_f64(const PairArgs &a, hipStream_t s)
{
    const int T = a.plan.targets_per_lane, TL = a.plan.tile;
    if (T == 1 && TL == 256) return launch_f64_r<1, 256>(a, s);
    if (T == 2 && TL == 256) return launch_f64_r<2, 256>(a, s);
    if (T == 4 && TL == 256) return launch_f64_r<4, 256>(a, s);
    if (T == 1 && TL == 512) return launch_f64_r<1, 512>(a, s);
    if (T == 2 && TL == 512) return launch_f64_r<2, 512>(a, s);
    if (T == 4 && TL == 512) return launch_f64_r<4, 512>(a, s);
    return hipErrorInvalidValue;
}

bool plan_supported(int targets_per_lane, int tile)
{
    return (targets_per_lane == 1 || targets_per_lane == 2 || targets_per_lane == 4) &&
           (tile == 256 || tile == 512);
}

// ---------------------------------------------------------------------------
// RK4 stage update (fused): chunk reduction, velocity, k_s, accumulator,
// next-stage inputs and source records, stage-1 kinetic energy.
// ---------------------------------------------------------------------------

// Block sum in a fixed order (deterministic), result valid in thread 0.
__device__ double block_sum_256(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

// Sum of the chunk partials of target i, chunk 0 first.
__device__ __forceinline__ void reduce_chunks(const double *__restrict__ partial, int n_chunks,
                                              int stride, int i, double &Sx, double &Sy,
                                              double &Sz)
{
    Sx = Sy = Sz = 0.0;
    for (int c = 0; c < n_chunks; ++c) {
        const double *p = partial + (size_t)(3 * c) * stride;
        Sx += p[i];
        Sy += p[stride + i];
        Sz += p[2 * (size_t)stride + i];
    }
}

__global__ __launch_bounds__(256) void stage_update_kernel(StepState st, StageParams sp)
{
    __shared__ double sh[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int stride = st.t_stride;
    double ke_local = 0.0;
    if (i < st.n_targets) {
        double Sx, Sy, Sz;
        reduce_chunks(st.partial, sp.n_chunks, stride, i, Sx, Sy, Sz);
        const double x = st.xs[i], y = st.xs[stride + i], z = st.xs[2 * stride + i];
        const double c4 = -4.0 * kPi;
        // u = sum_j w_j (x_i x x_j)/d_ij / (-4 PI)    (PanelVelocityParallel.f90:1332)
        const double ux = (y * Sz - z * Sy) / c4;
        const double uy = (z * Sx - x * Sz) / c4;
        const double uz = (x * Sy - y * Sx) / c4;
        const double kx = sp.dt * ux, ky = sp.dt * uy, kz = sp.dt * uz;
        const double *base = (sp.stage == 0) ? st.x0 : st.acc;
        // x + k1/6 + k2/3 + k3/3 + k4/6, left to right (:637-644)
        const double ax = base[i] + kx / sp.divisor;
        const double ay = base[stride + i] + ky / sp.divisor;
        const double az = base[2 * stride + i] + kz / sp.divisor;
        const bool active = i < st.n_active;
        double zeta_next = 0.0, zacc = 0.0;
        if (active) {
            const double dvort = -(2.0 * sp.omega * uz);   // :1333
            const double kv = sp.dt * dvort;
            zacc = ((sp.stage == 0) ? st.zeta0[i] : st.zacc[i]) + kv / sp.divisor;
            st.zacc[i] = zacc;
            zeta_next = st.zeta0[i] + sp.a_next * kv;
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
            if (active) st.src[i] = Source{nx, ny, nz, zeta_next * st.area[i]};
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
                st.zeta0[i] = zacc;
                st.src[i] = Source{nx, ny, nz, zacc * st.area[i]};
            }
            // particles carry absolute vorticity: relVort = absVort - 2 Omega z (:651)
            if (st.particle_absvort && i >= st.particle_begin && i < st.particle_end) {
                const int p = i - st.particle_begin;
                st.particle_relvort[p] = st.particle_absvort[p] - 2.0 * sp.omega * nz;
            }
        }
    }
    if (sp.stage == 0) {
        const double b = block_sum_256(ke_local, sh);
        if (threadIdx.x == 0) st.ke_part[blockIdx.x] = b;
    }
}

// totalKE = 1/2 sum_j ke_j A_j (:498), block partials added in order.
__global__ __launch_bounds__(256) void ke_total_kernel(const double *__restrict__ part, int n,
                                                       double *__restrict__ out)
{
    __shared__ double sh[256];
    double v = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) v += part[k];
    const double b = block_sum_256(v, sh);
    if (threadIdx.x == 0) *out = 0.5 * b;
}

hipError_t launch_stage_update(const StepState &st, const StageParams &sp, hipStream_t s)
{
    const int blocks = ceil_div(st.n_targets, 256);
    hipLaunchKernelGGL(stage_update_kernel, dim3(blocks), dim3(256), 0, s, st, sp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || sp.stage != 0) return e;
    hipLaunchKernelGGL(ke_total_kernel, dim3(1), dim3(256), 0, s, st.ke_part, blocks, st.ke_total);
    return hipGetLastError();
}

int ke_partial_count(int n_targets) { return ceil_div(n_targets, 256); }

// Writes the stage-1 inputs of a step from the step-start state: Xs = X0
// and the source records {x, zeta A} of the active panels.
__global__ __launch_bounds__(256) void prepare_step_kernel(StepState st)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= st.n_targets) return;
    const int stride = st.t_stride;
    const double x = st.x0[i], y = st.x0[stride + i], z = st.x0[2 * stride + i];
    st.xs[i] = x;
    st.xs[stride + i] = y;
    st.xs[2 * stride + i] = z;
    if (i < st.n_active) st.src[i] = Source{x, y, z, st.zeta0[i] * st.area[i]};
}

hipError_t launch_prepare_step(const StepState &st, hipStream_t s)
{
    hipLaunchKernelGGL(prepare_step_kernel, dim3(ceil_div(st.n_targets, 256)), dim3(256), 0, s, st);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Layout helpers
// ---------------------------------------------------------------------------

// AoS x(3, n) -> SoA rows of dst (stride), rows 0..2, entries [off, off + n).
__global__ void aos_to_soa_kernel(const double *__restrict__ aos, int n, double *__restrict__ dst,
                                  int stride, int off)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    dst[off + i] = aos[3 * i];
    dst[stride + off + i] = aos[3 * i + 1];
    dst[2 * (size_t)stride + off + i] = aos[3 * i + 2];
}

__global__ void soa_to_aos_kernel(const double *__restrict__ src, int stride, int off, int n,
                                  double *__restrict__ aos)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    aos[3 * i] = src[off + i];
    aos[3 * i + 1] = src[stride + off + i];
    aos[3 * i + 2] = src[2 * (size_t)stride + off + i];
}

hipError_t launch_aos_to_soa(const double *aos, int n, double *dst, int stride, int off,
                             hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(aos_to_soa_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, aos, n, dst,
                       stride, off);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(const double *src, int stride, int off, int n, double *aos,
                             hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(soa_to_aos_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, src, stride,
                       off, n, aos);
    return hipGetLastError();
}

// Source records from AoS positions, vorticity and area: w = vort * area.
__global__ void make_sources_kernel(const double *__restrict__ xyz, const double *__restrict__ vort,
                                    const double *__restrict__ area, int n,
                                    Source *__restrict__ src)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    src[i] = Source{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], vort[i] * area[i]};
}

hipError_t launch_make_sources(const double *xyz, const double *vort, const double *area, int n,
                               Source *src, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(make_sources_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, xyz, vort,
                       area, n, src);
    return hipGetLastError();
}

// Single evaluation: u (AoS) from the chunk partials and the SoA targets,
// and dVort = -2 Omega u_z for the first n_dvort targets.
__global__ void finalize_velocity_kernel(const double *__restrict__ partial, int n_chunks,
                                         const double *__restrict__ t, int stride, int n,
                                         double omega, double *__restrict__ u,
                                         double *__restrict__ dvort, int n_dvort)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double Sx, Sy, Sz;
    reduce_chunks(partial, n_chunks, stride, i, Sx, Sy, Sz);
    const double x = t[i], y = t[stride + i], z = t[2 * stride + i];
    const double c4 = -4.0 * kPi;
    const double uz = (x * Sy - y * Sx) / c4;
    u[3 * i] = (y * Sz - z * Sy) / c4;
    u[3 * i + 1] = (z * Sx - x * Sz) / c4;
    u[3 * i + 2] = uz;
    if (dvort && i < n_dvort) dvort[i] = -(2.0 * omega * uz);
}

hipError_t launch_finalize_velocity(const double *partial, int n_chunks, const double *t,
                                    int stride, int n, double omega, double *u, double *dvort,
                                    int n_dvort, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(finalize_velocity_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, partial,
                       n_chunks, t, stride, n, omega, u, dvort, n_dvort);
    return hipGetLastError();
}

// x / |x| on AoS points (NormalizeGrid of the Voronoi path, VoronoiPanels.f90:395-406).
__global__ void normalize_soa_kernel(double *__restrict__ t, int stride, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double x = t[i], y = t[stride + i], z = t[2 * stride + i];
    const double r = sqrt((x * x + y * y) + z * z);
    t[i] = x / r;
    t[stride + i] = y / r;
    t[2 * stride + i] = z / r;
}

hipError_t launch_normalize_soa(double *t, int stride, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(normalize_soa_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, t, stride, n);
    return hipGetLastError();
}

}  // namespace bve
