// Per-context launches of the mixed-precision Biot-Savart pair sum (float32
// far field, float64 near field and reductions) and of its per-stage
// preparation (bounding spheres, float32 source records).  The bodies and
// the precision design are in bve_device.h.
#include "bve_common.h"
#include "bve_device.h"
#include "bve_kernels.h"

namespace bve {

__global__ __launch_bounds__(kMixedTile) void mixed_source_prep_kernel(
    const Source *__restrict__ src, int n_sources, float4 *__restrict__ src32,
    float4 *__restrict__ srcw, float *__restrict__ srcc, double4 *__restrict__ tile_box)
{
    __shared__ double sh[kMixedTile];
    mixed_source_prep_body(src, n_sources, src32, srcw, srcc, tile_box, blockIdx.x, sh);
}

__global__ __launch_bounds__(256) void mixed_target_prep_kernel(const PairArgs b, int n_targets,
                                                                double4 *__restrict__ group_box)
{
    __shared__ double sh[256];
    mixed_target_prep_body(b, n_targets, group_box, blockIdx.x, sh);
}

template <int T, bool KAHAN>
__global__ __launch_bounds__(kBlock) void pair_sum_mixed_seg_kernel(MixedArgs a)
{
    __shared__ MixedLds lds;
    __shared__ int seg;
    pair_mixed_segments<T, KAHAN>(a, lds, &seg);
}

template <int T, bool KAHAN>
static hipError_t launch_mixed_tk(const MixedArgs &a, hipStream_t s)
{
    if (a.base.plan.segments <= 0 || !a.base.queue || a.base.plan.src_step < 1)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL((pair_sum_mixed_seg_kernel<T, KAHAN>), dim3(a.base.plan.grid), dim3(kBlock),
                       0, s, a);
    return hipGetLastError();
}

template <int T>
static hipError_t launch_mixed_t(const MixedArgs &a, hipStream_t s)
{
    return a.kahan ? launch_mixed_tk<T, true>(a, s) : launch_mixed_tk<T, false>(a, s);
}

template <int T, bool KAHAN>
static int mixed_seg_blocks()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pair_sum_mixed_seg_kernel<T, KAHAN>,
                                                     kBlock, 0) != hipSuccess)
        return 0;
    return n;
}

int pair_mixed_blocks_per_cu(int T, int kahan)
{
    switch (T) {
    case 1: return kahan ? mixed_seg_blocks<1, true>() : mixed_seg_blocks<1, false>();
    case 2: return kahan ? mixed_seg_blocks<2, true>() : mixed_seg_blocks<2, false>();
    case 4: return kahan ? mixed_seg_blocks<4, true>() : mixed_seg_blocks<4, false>();
    default: return 0;
    }
}

hipError_t launch_mixed_prep(const MixedArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(mixed_source_prep_kernel, dim3(a.base.n_src_pad / kMixedTile),
                       dim3(kMixedTile), 0, s, a.base.src, a.n_sources, a.src32, a.srcw,
                       a.srcc, a.tile_box);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mixed_target_prep_kernel, dim3(a.base.t_stride / 256), dim3(256), 0, s,
                       a.base, a.n_targets, a.group_box);
    return hipGetLastError();
}

hipError_t launch_pair_sum_mixed(const MixedArgs &a, hipStream_t s)
{
    if (a.base.plan.tile != kMixedTile || a.base.n_src_pad % kMixedTile != 0 ||
        a.base.t_stride % 256 != 0)
        return hipErrorInvalidValue;
    switch (a.base.plan.targets_per_lane) {
    case 1: return launch_mixed_t<1>(a, s);
    case 2: return launch_mixed_t<2>(a, s);
    case 4: return launch_mixed_t<4>(a, s);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// Bounded mixed mode
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kMixedTile) void bounded_source_prep_kernel(
    const Source *__restrict__ src, int n_sources, float4 *__restrict__ src32,
    float4 *__restrict__ srcw, double4 *__restrict__ tile_box, double *__restrict__ mom,
    double *__restrict__ sub_mom)
{
    __shared__ double sh[kMixedTile];
    bounded_source_prep_body(src, n_sources, src32, srcw, tile_box, mom, sub_mom, blockIdx.x, sh);
}

// The probe's maximum starts from zero every pass (a kernel, not a memset,
// so that captured graphs replay it).
__global__ void bounded_scale_reset_kernel(double *scale)
{
    if (threadIdx.x == 0) *scale = 0.0;
}

__global__ __launch_bounds__(256) void bounded_probe_kernel(const PairArgs b, int n_targets,
                                                            int n_sources, int group_stride,
                                                            double *scale)
{
    __shared__ double sh[256];
    bounded_probe_body(b, n_targets, n_sources, blockIdx.x * group_stride * kGroup,
                       (unsigned long long *)scale, sh);
}

// 4 waves per SIMD (at most 128 VGPRs): the remainder loop needs the
// latency hiding more than the registers
template <int T, bool SCALAR>
__global__ __launch_bounds__(kBlock, 4) void pair_sum_bounded_seg_kernel(
    MixedArgs a, const float4 *__restrict__ g32, const float2 *__restrict__ gw,
    const double *__restrict__ gmom, const double *__restrict__ gsub,
    const unsigned short *__restrict__ groute)
{
    __shared__ BoundedLds lds;
    __shared__ int seg;
    pair_bounded_segments<T, SCALAR>(a, lds, &seg, g32, gw, gmom, gsub, groute);
}

__global__ __launch_bounds__(256) void bounded_route_kernel(const MixedArgs a, int n_items)
{
    const double tau = bounded_tau(a);
    for (int item = blockIdx.x * 256 + threadIdx.x; item < n_items; item += gridDim.x * 256)
        bounded_route_body(a, item, tau);
}

template <int T, bool SCALAR>
static hipError_t launch_bounded_ts(const MixedArgs &a, hipStream_t s)
{
    if (a.base.plan.segments <= 0 || !a.base.queue || a.base.plan.src_step < 1)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL((pair_sum_bounded_seg_kernel<T, SCALAR>), dim3(a.base.plan.grid),
                       dim3(kBlock), 0, s, a, (const float4 *)a.src32, (const float2 *)a.srcw,
                       (const double *)a.tile_mom, (const double *)a.sub_mom,
                       (const unsigned short *)a.route);
    return hipGetLastError();
}

template <int T>
static hipError_t launch_bounded_t(const MixedArgs &a, hipStream_t s)
{
    return a.base.n_src_pad <= kBoundedScalarMaxSources ? launch_bounded_ts<T, true>(a, s)
                                                        : launch_bounded_ts<T, false>(a, s);
}

template <int T, bool SCALAR>
static int bounded_seg_blocks()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pair_sum_bounded_seg_kernel<T, SCALAR>,
                                                     kBlock, 0) != hipSuccess)
        return 0;
    return n;
}

int pair_bounded_blocks_per_cu(int T, int n_src_pad)
{
    const bool scalar = n_src_pad <= kBoundedScalarMaxSources;
    switch (T) {
    case 1: return scalar ? bounded_seg_blocks<1, true>() : bounded_seg_blocks<1, false>();
    case 2: return scalar ? bounded_seg_blocks<2, true>() : bounded_seg_blocks<2, false>();
    case 4: return scalar ? bounded_seg_blocks<4, true>() : bounded_seg_blocks<4, false>();
    default: return 0;
    }
}

hipError_t launch_scale_reset(double *scale, hipStream_t s)
{
    if (!scale) return hipErrorInvalidValue;
    hipLaunchKernelGGL(bounded_scale_reset_kernel, dim3(1), dim3(64), 0, s, scale);
    return hipGetLastError();
}

// The routing bits of every (target group, source tile) pair, from the
// spheres, moments and velocity scale of this pass.
static hipError_t launch_bounded_route(const MixedArgs &a, hipStream_t s)
{
    if (!a.route || a.n_groups != a.base.t_stride / kGroup) return hipErrorInvalidValue;
    const int n_items = a.n_groups * (a.base.n_src_pad / kMixedTile);
    const int blocks = ceil_div(n_items, 256);
    hipLaunchKernelGGL(bounded_route_kernel, dim3(blocks < 4096 ? blocks : 4096), dim3(256), 0,
                       s, a, n_items);
    return hipGetLastError();
}

// probe = false: the scale word already holds the velocity scale (max |u|
// of the previous RK4 stage, folded in by its stage update).  Ends with the
// routing pass, which needs the scale.
hipError_t launch_bounded_prep(const MixedArgs &a, hipStream_t s, bool probe)
{
    if (!a.tile_mom || !a.scale || !a.sub_mom) return hipErrorInvalidValue;
    hipLaunchKernelGGL(bounded_source_prep_kernel, dim3(a.base.n_src_pad / kMixedTile),
                       dim3(kMixedTile), 0, s, a.base.src, a.n_sources, a.src32, a.srcw,
                       a.tile_box, a.tile_mom, a.sub_mom);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mixed_target_prep_kernel, dim3(a.base.t_stride / 256), dim3(256), 0, s,
                       a.base, a.n_targets, a.group_box);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (!probe) return launch_bounded_route(a, s);
    e = launch_scale_reset(a.scale, s);
    if (e != hipSuccess) return e;
    if (a.n_targets <= 0) return launch_bounded_route(a, s);
    // the first target of every group_stride-th group of 64: at most ~1024
    // probes, a target every ~220 at level 7 (the groups are spatially
    // compact runs, so the probes cover the sphere); ~0.05 ms at level 7
    const int groups = ceil_div(a.n_targets, kGroup);
    const int stride = ceil_div(groups, 1024);
    hipLaunchKernelGGL(bounded_probe_kernel, dim3(ceil_div(groups, stride)), dim3(256), 0, s,
                       a.base, a.n_targets, a.n_sources, stride, a.scale);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_bounded_route(a, s);
}

hipError_t launch_pair_sum_bounded(const MixedArgs &a, hipStream_t s)
{
    if (a.base.plan.tile != kMixedTile || a.base.n_src_pad % kMixedTile != 0 ||
        a.base.t_stride % 1024 != 0 || !a.tile_mom || !a.scale || !a.sub_mom || !a.route ||
        a.n_groups != a.base.t_stride / kGroup ||
        a.base.plan.n_tiles * kBlock * a.base.plan.targets_per_lane > a.base.t_stride)
        return hipErrorInvalidValue;
    switch (a.base.plan.targets_per_lane) {
    case 1: return launch_bounded_t<1>(a, s);
    case 2: return launch_bounded_t<2>(a, s);
    case 4: return launch_bounded_t<4>(a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace bve
