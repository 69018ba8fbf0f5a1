// Batched ensemble launches: one persistent pair-sum launch over every
// member's segments (runs of its tile-major (target tile, source tile)
// sequence, the single-context pass's work items), and block-mapped launches
// of the RK4 stage update and the per-stage preparation for all members.
//
// Members have uneven sizes (e.g. cubed spheres at several levels).  The
// host lists the work items in decreasing cost, and the persistent grid
// takes them from a queue, so each workgroup gets a similar share of the
// work whatever the items cost (the mixed kernel's careful tiles cost more
// than its far-field ones).  By default the members of one launch share a
// segment length derived from the launch's whole work (a tuned number of
// segments per resident workgroup); an explicit chunk keeps a member's
// own.  The kernels are compiled per targets-per-lane
// count T, so each holds only its own registers (the mixed loop needs four
// waves per SIMD): one launch per T present among the members, one in all
// when the autotuner's common T is used.
#include "bve_common.h"
#include "bve_device.h"
#include "bve_kernels.h"

namespace bve {

template <int T, int R>
__global__ __launch_bounds__(kBlock) void ens_pair_f64_kernel(const EnsItem *__restrict__ items,
                                                              int n_items,
                                                              const PairArgs *__restrict__ members,
                                                              unsigned *queue)
{
    __shared__ StagedSource lds[256];
    __shared__ int slot;
    run_queue(queue, n_items, &slot, [&](int k) {
        const EnsItem it = items[k];
        const PairArgs a = members[it.member];
        pair_f64_segment<T, 256, R>(a, it.segment, lds);
    });
}

template <int T, bool KAHAN>
__global__ __launch_bounds__(kBlock) void ens_pair_mixed_kernel(const EnsItem *__restrict__ items,
                                                                int n_items,
                                                                const MixedArgs *__restrict__ members,
                                                                unsigned *queue)
{
    __shared__ MixedLds lds;
    __shared__ int slot;
    run_queue(queue, n_items, &slot, [&](int k) {
        const EnsItem it = items[k];
        const MixedArgs a = members[it.member];
        pair_mixed_segment<T, KAHAN>(a, it.segment, lds);
    });
}

__global__ __launch_bounds__(256) void ens_stage_kernel(const int2 *__restrict__ map,
                                                        const StepState *__restrict__ states,
                                                        const PairPlan *__restrict__ plans,
                                                        StageParams sp)
{
    __shared__ double sh[256];
    const int2 mb = map[blockIdx.x];
    // each member's own segmented plan (its rows and their order)
    sp.plan = plans[mb.x];
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

// Pair kernel of the ensemble for targets-per-lane T in the ensemble's
// arithmetic.
template <int T>
static const void *ens_pair_kernel(int mixed, int kahan, int recip)
{
    if (mixed)
        return kahan ? (const void *)ens_pair_mixed_kernel<T, true>
                     : (const void *)ens_pair_mixed_kernel<T, false>;
    if (recip == 4) return (const void *)ens_pair_f64_kernel<T, 4>;
    if (recip == 2) return (const void *)ens_pair_f64_kernel<T, 2>;
    return (const void *)ens_pair_f64_kernel<T, 1>;
}

static const void *ens_pair_kernel_for(int T, int mixed, int kahan, int recip)
{
    switch (T) {
    case 1: return ens_pair_kernel<1>(mixed, kahan, recip);
    case 2: return ens_pair_kernel<2>(mixed, kahan, recip);
    case 4: return ens_pair_kernel<4>(mixed, kahan, recip);
    default: return nullptr;
    }
}

int ensemble_pair_blocks_per_cu(int T, int mixed, int kahan, int recip)
{
    const void *k = ens_pair_kernel_for(T, mixed, kahan, recip);
    int n = 0;
    if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, kBlock, 0) != hipSuccess)
        return 0;
    return n;
}

hipError_t launch_ensemble_pairs(const EnsLaunch &e, hipStream_t s)
{
    static const int kT[kEnsGroups] = {1, 2, 4};
    for (int g = 0; g < kEnsGroups; ++g) {
        if (e.group_count[g] <= 0) continue;
        if (!e.queue || e.group_grid[g] <= 0) return hipErrorInvalidValue;
        const void *k = ens_pair_kernel_for(kT[g], e.mixed, e.kahan, e.recip);
        const EnsItem *items = e.items + e.group_begin[g];
        int n = e.group_count[g];
        unsigned *queue = e.queue + 2 * g;
        const void *members =
            e.mixed ? (const void *)e.mixed_members : (const void *)e.pair_members;
        void *args[] = {(void *)&items, (void *)&n, (void *)&members, (void *)&queue};
        hipError_t err = hipLaunchKernel(k, dim3(e.group_grid[g]), dim3(kBlock), args, 0, s);
        if (err != hipSuccess) return err;
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
                       e.states, e.plans, sp);
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
