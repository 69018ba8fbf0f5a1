// Per-context launches of the float64 Biot-Savart pair sum and the fused RK4
// stage update (gfx950, wave64), plus layout helpers.  The kernel bodies
// live in bve_device.h, shared with the persistent ensemble launch.
#include "bve_common.h"
#include "bve_device.h"
#include "bve_kernels.h"

namespace bve {

// ---------------------------------------------------------------------------
// float64 pair pass
// ---------------------------------------------------------------------------

// Segmented pass: a persistent grid (about one workgroup per resident slot)
// takes segments from the queue.
template <int T, int TILE, int R>
__global__ __launch_bounds__(kBlock) void pair_sum_f64_seg_kernel(PairArgs a)
{
    __shared__ StagedSource lds[TILE];
    __shared__ int seg;
    pair_f64_segments<T, TILE, R>(a, lds, &seg);
}

template <int T, int TILE, int R>
static hipError_t launch_f64(const PairArgs &a, hipStream_t s)
{
    if (a.plan.segments <= 0 || !a.queue) return hipErrorInvalidValue;
    hipLaunchKernelGGL((pair_sum_f64_seg_kernel<T, TILE, R>), dim3(a.plan.grid), dim3(kBlock), 0,
                       s, a);
    return hipGetLastError();
}

template <int T, int TILE, int R>
static int seg_blocks_per_cu()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pair_sum_f64_seg_kernel<T, TILE, R>,
                                                     kBlock, 0) != hipSuccess)
        return 0;
    return n;
}

template <int T, int TILE>
static int seg_blocks_per_cu_r(int recip)
{
    switch (recip) {
    case 1: return seg_blocks_per_cu<T, TILE, 1>();
    case 2: return seg_blocks_per_cu<T, TILE, 2>();
    case 4: return seg_blocks_per_cu<T, TILE, 4>();
    default: return 0;
    }
}

int pair_f64_blocks_per_cu(int T, int tile, int recip)
{
    if (T == 1 && tile == 256) return seg_blocks_per_cu_r<1, 256>(recip);
    if (T == 2 && tile == 256) return seg_blocks_per_cu_r<2, 256>(recip);
    if (T == 4 && tile == 256) return seg_blocks_per_cu_r<4, 256>(recip);
    if (T == 1 && tile == 512) return seg_blocks_per_cu_r<1, 512>(recip);
    if (T == 2 && tile == 512) return seg_blocks_per_cu_r<2, 512>(recip);
    if (T == 4 && tile == 512) return seg_blocks_per_cu_r<4, 512>(recip);
    return 0;
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

// ---------------------------------------------------------------------------
// Reference-order pair pass: one target per lane, one workgroup per 256
// targets, every source in index order
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void pair_sum_exact_kernel(PairArgs a)
{
    __shared__ double4 lds_src[kBlock];
    __shared__ double lds_area[kBlock];
    pair_exact_body(a, lds_src, lds_area);
}

hipError_t launch_pair_sum_exact(const PairArgs &a, hipStream_t s)
{
    if (!a.plan.exact || (!a.area && a.n_sources > 0) || a.plan.n_chunks != 1 ||
        a.n_targets > a.t_stride)
        return hipErrorInvalidValue;
    if (a.n_targets <= 0) return hipSuccess;
    hipLaunchKernelGGL(pair_sum_exact_kernel, dim3(ceil_div(a.n_targets, kBlock)), dim3(kBlock),
                       0, s, a);
    return hipGetLastError();
}

bool plan_supported(int targets_per_lane, int tile)
{
    return (targets_per_lane == 1 || targets_per_lane == 2 || targets_per_lane == 4) &&
           (tile == 256 || tile == 512);
}

// ---------------------------------------------------------------------------
// RK4 stage update
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void stage_update_kernel(StepState st, StageParams sp)
{
    __shared__ double sh[256];
    stage_update_body(st, sp, blockIdx.x, sh);
}

__global__ __launch_bounds__(256) void advection_step_kernel(StepState st, double t, double dt,
                                                             int divergent)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < st.n_targets) advection_step_body(st, i, t, dt, divergent);
}

hipError_t launch_advection_step(const StepState &st, double t, double dt, int divergent,
                                 hipStream_t s)
{
    if (st.n_targets <= 0) return hipSuccess;
    hipLaunchKernelGGL(advection_step_kernel, dim3(ceil_div(st.n_targets, 256)), dim3(256), 0, s,
                       st, t, dt, divergent);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void ke_total_kernel(const double *__restrict__ part, int n,
                                                       double *__restrict__ out)
{
    __shared__ double sh[256];
    ke_total_body(part, n, out, sh);
}

__global__ __launch_bounds__(256) void ke_total_sequential_kernel(const double *__restrict__ ke,
                                                                  const double *__restrict__ area,
                                                                  int n, double *__restrict__ out)
{
    __shared__ double sh[256];
    ke_total_sequential_body(ke, area, n, out, sh);
}

hipError_t launch_stage_update(const StepState &st, const StageParams &sp, hipStream_t s)
{
    const int blocks = ceil_div(st.n_targets, 256);
    hipLaunchKernelGGL(stage_update_kernel, dim3(blocks), dim3(256), 0, s, st, sp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || sp.stage != 0) return e;
    if (sp.plan.exact)
        hipLaunchKernelGGL(ke_total_sequential_kernel, dim3(1), dim3(256), 0, s, st.ke, st.area,
                           st.n_active, st.ke_total);
    else
        hipLaunchKernelGGL(ke_total_kernel, dim3(1), dim3(256), 0, s, st.ke_part, blocks,
                           st.ke_total);
    return hipGetLastError();
}

int ke_partial_count(int n_targets) { return ceil_div(n_targets, 256); }

__global__ __launch_bounds__(256) void prepare_step_kernel(StepState st)
{
    prepare_step_body(st, blockIdx.x);
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

// Source records from AoS positions, vorticity and area: w = vort * area
// (unweighted: w = vort, for the reference-order pass).
__global__ void make_sources_kernel(const double *__restrict__ xyz, const double *__restrict__ vort,
                                    const double *__restrict__ area, int n, int unweighted,
                                    Source *__restrict__ src)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    src[i] = Source{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2],
                    unweighted ? vort[i] : vort[i] * area[i]};
}

hipError_t launch_make_sources(const double *xyz, const double *vort, const double *area, int n,
                               int unweighted, Source *src, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(make_sources_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, xyz, vort,
                       area, n, unweighted, src);
    return hipGetLastError();
}

// Single evaluation: u (AoS) from the pass's partials and the SoA targets,
// and dVort = -2 Omega u_z for the first n_dvort targets.
__global__ void finalize_velocity_kernel(const double *__restrict__ partial, PairPlan plan,
                                         const double *__restrict__ t, int stride, int n,
                                         double omega, double *__restrict__ u,
                                         double *__restrict__ dvort, int n_dvort,
                                         unsigned long long *status)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double Sx, Sy, Sz;
    reduce_partials(partial, plan, stride, i, Sx, Sy, Sz);
    const double x = t[i], y = t[stride + i], z = t[2 * stride + i];
    double ux, uy, uz;
    velocity_from_sums(plan.exact | plan.crossed, x, y, z, Sx, Sy, Sz, ux, uy, uz);
    if (!finite3(ux, uy, uz)) flag_nonfinite(status, 0, i);
    u[3 * i] = ux;
    u[3 * i + 1] = uy;
    u[3 * i + 2] = uz;
    if (dvort && i < n_dvort) dvort[i] = -(2.0 * omega * uz);
}

__global__ void aos_to_soa_padded_kernel(const double *__restrict__ aos, int n,
                                         double *__restrict__ dst, int stride)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= stride) return;
    const bool live = i < n;
    dst[i] = live ? aos[3 * i] : 0.0;
    dst[stride + i] = live ? aos[3 * i + 1] : 0.0;
    dst[2 * (size_t)stride + i] = live ? aos[3 * i + 2] : 0.0;
}

hipError_t launch_aos_to_soa_padded(const double *aos, int n, double *dst, int stride,
                                    hipStream_t s)
{
    if (stride <= 0) return hipSuccess;
    hipLaunchKernelGGL(aos_to_soa_padded_kernel, dim3(ceil_div(stride, 256)), dim3(256), 0, s, aos,
                       n, dst, stride);
    return hipGetLastError();
}

__global__ void make_sources_padded_kernel(const double *__restrict__ xyz,
                                           const double *__restrict__ vort,
                                           const double *__restrict__ area, int n, int n_pad,
                                           int unweighted, Source *__restrict__ src)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n_pad) return;
    src[i] = i < n ? Source{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2],
                            unweighted ? vort[i] : vort[i] * area[i]}
                   : Source{0.0, 0.0, 0.0, 0.0};
}

hipError_t launch_make_sources_padded(const double *xyz, const double *vort, const double *area,
                                      int n, int n_pad, int unweighted, Source *src,
                                      hipStream_t s)
{
    if (n_pad <= 0) return hipSuccess;
    hipLaunchKernelGGL(make_sources_padded_kernel, dim3(ceil_div(n_pad, 256)), dim3(256), 0, s,
                       xyz, vort, area, n, n_pad, unweighted, src);
    return hipGetLastError();
}

__global__ void zero_words_kernel(unsigned *p, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_zero_words(unsigned *p, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_words_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_finalize_velocity(const double *partial, const PairPlan &plan, const double *t,
                                    int stride, int n, double omega, double *u, double *dvort,
                                    int n_dvort, unsigned long long *status, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(finalize_velocity_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, partial,
                       plan, t, stride, n, omega, u, dvort, n_dvort, status);
    return hipGetLastError();
}

}  // namespace bve
