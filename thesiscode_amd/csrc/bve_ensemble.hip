// Batched ensemble launches: one persistent pair-sum launch over every
// member's (target tile, source chunk) work items, and block-mapped launches
// of the RK4 stage update and the per-stage preparation for all members.
//
// Members have uneven sizes (e.g. cubed spheres at several levels).  The
// host lists the work items in decreasing cost, and the persistent grid
// walks the list round-robin, so each workgroup gets a similar share of
// pairs; every member keeps its own launch geometry (targets per lane and
// chunk length), chosen per size by the autotuner.
#include "bve_common.h"
#include "bve_device.h"
#include "bve_kernels.h"

namespace bve {

template <int R>
__global__ __launch_bounds__(kBlock) void ens_pair_f64_kernel(const EnsItem *__restrict__ items,
                                                              int n_items,
                                                              const PairArgs *__restrict__ members)
{
    __shared__ StagedSource lds[256];
    for (int k = blockIdx.x; k < n_items; k += gridDim.x) {
        const EnsItem it = items[k];
        const PairArgs a = members[it.member];
        switch (a.plan.targets_per_lane) {
        case 1: pair_f64_item<1, 256, R>(a, it.chunk, it.ttile, lds); break;
        case 2: pair_f64_item<2, 256, R>(a, it.chunk, it.ttile, lds); break;
        default: pair_f64_item<4, 256, R>(a, it.chunk, it.ttile, lds); break;
        }
    }
}

template <bool KAHAN>
__global__ __launch_bounds__(kBlock) void ens_pair_mixed_kernel(const EnsItem *__restrict__ items,
                                                                int n_items,
                                                                const MixedArgs *__restrict__ members)
{
    __shared__ MixedLds lds;
    for (int k = blockIdx.x; k < n_items; k += gridDim.x) {
        const EnsItem it = items[k];
        const MixedArgs a = members[it.member];
        switch (a.base.plan.targets_per_lane) {
        case 1: pair_mixed_item<1, KAHAN>(a, it.chunk, it.ttile, lds); break;
        case 2: pair_mixed_item<2, KAHAN>(a, it.chunk, it.ttile, lds); break;
        default: pair_mixed_item<4, KAHAN>(a, it.chunk, it.ttile, lds); break;
        }
    }
}

__global__ __launch_bounds__(256) void ens_stage_kernel(const int2 *__restrict__ map,
                                                        const StepState *__restrict__ states,
                                                        const int *__restrict__ n_chunks,
                                                        StageParams sp)
{
    __shared__ double sh[256];
    const int2 mb = map[blockIdx.x];
    // ensemble members run chunked plans
    sp.plan.segments = 0;
    sp.plan.n_chunks = n_chunks[mb.x];
    stage_update_body(states[mb.x], sp, mb.y, sh);
}

__global__ __launch_bounds__(256) void ens_ke_kernel(const StepState *__restrict__ states)
{
    __shared__ double sh[256];
    const StepState &st = states[blockIdx.x];
    ke_total_body(st.ke_part, ceil_div(st.n_targets, 256), st.ke_total, sh);
}

__global__ __launch_bounds__(256) void ens_prepare_kernel(const int2 *__restrict__ map,
                                                          const StepState *__restrict__ states)
{
    const int2 mb = map[blockIdx.x];
    prepare_step_body(states[mb.x], mb.y);
}

__global__ __launch_bounds__(kMixedTile) void ens_mixed_src_prep_kernel(
    const int2 *__restrict__ map, const MixedArgs *__restrict__ members)
{
    __shared__ double sh[kMixedTile];
    const int2 mb = map[blockIdx.x];
    const MixedArgs &a = members[mb.x];
    mixed_source_prep_body(a.base.src, a.n_sources, a.src32, a.srcw, a.srcc, a.tile_box, mb.y,
                           sh);
}

__global__ __launch_bounds__(256) void ens_mixed_tgt_prep_kernel(
    const int2 *__restrict__ map, const MixedArgs *__restrict__ members)
{
    __shared__ double sh[256];
    const int2 mb = map[blockIdx.x];
    const MixedArgs &a = members[mb.x];
    mixed_target_prep_body(a.base, a.n_targets, a.group_box, mb.y, sh);
}

hipError_t launch_ensemble_pairs(const EnsLaunch &e, hipStream_t s)
{
    if (e.n_items <= 0) return hipSuccess;
    if (e.mixed) {
        if (e.kahan)
            hipLaunchKernelGGL(ens_pair_mixed_kernel<true>, dim3(e.grid), dim3(kBlock), 0, s,
                               e.items, e.n_items, e.mixed_members);
        else
            hipLaunchKernelGGL(ens_pair_mixed_kernel<false>, dim3(e.grid), dim3(kBlock), 0, s,
                               e.items, e.n_items, e.mixed_members);
    } else if (e.recip == 4) {
        hipLaunchKernelGGL(ens_pair_f64_kernel<4>, dim3(e.grid), dim3(kBlock), 0, s, e.items,
                           e.n_items, e.pair_members);
    } else if (e.recip == 2) {
        hipLaunchKernelGGL(ens_pair_f64_kernel<2>, dim3(e.grid), dim3(kBlock), 0, s, e.items,
                           e.n_items, e.pair_members);
    } else {
        hipLaunchKernelGGL(ens_pair_f64_kernel<1>, dim3(e.grid), dim3(kBlock), 0, s, e.items,
                           e.n_items, e.pair_members);
    }
    return hipGetLastError();
}

hipError_t launch_ensemble_mixed_prep(const EnsLaunch &e, hipStream_t s)
{
    hipLaunchKernelGGL(ens_mixed_src_prep_kernel, dim3(e.n_src_blocks), dim3(kMixedTile), 0, s,
                       e.src_map, e.mixed_members);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(ens_mixed_tgt_prep_kernel, dim3(e.n_tgt_blocks), dim3(256), 0, s,
                       e.tgt_map, e.mixed_members);
    return hipGetLastError();
}

hipError_t launch_ensemble_stage(const EnsLaunch &e, const StageParams &sp, hipStream_t s)
{
    hipLaunchKernelGGL(ens_stage_kernel, dim3(e.n_stage_blocks), dim3(256), 0, s, e.stage_map,
                       e.states, e.n_chunks, sp);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || sp.stage != 0) return err;
    hipLaunchKernelGGL(ens_ke_kernel, dim3(e.n_members), dim3(256), 0, s, e.states);
    return hipGetLastError();
}

hipError_t launch_ensemble_prepare(const EnsLaunch &e, hipStream_t s)
{
    hipLaunchKernelGGL(ens_prepare_kernel, dim3(e.n_stage_blocks), dim3(256), 0, s, e.stage_map,
                       e.states);
    return hipGetLastError();
}

}  // namespace bve
