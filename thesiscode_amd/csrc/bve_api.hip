// C ABI of libbve: contexts, device state of the panel and Voronoi RK4
// paths, single velocity evaluations and launch-level profiling.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "bve_common.h"
#include "bve_host_error.h"
#include "bve_kernels.h"
#include "../../include/bve.h"

using namespace bve;

namespace {


template <class P>
hipError_t dev_alloc(P **p, size_t count)
{
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void **)p, count * sizeof(P));
}

template <class P>
void dev_free(P *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of the allocation-free single evaluation (bve_velocity_noalloc):
// one file-scope device array per GPU, held until the process ends.  Host
// code never uses the array's name as an address: each device's address
// comes from hipGetSymbolAddress, resolved once per device with that device
// current (bve_create) and kept in scratch_addr.
//
// Layout for at most kNoallocMaxTargets targets and kNoallocMaxSources
// sources: target rows (3 x stride doubles, stride = targets rounded up to
// 1024), source records (sources rounded up to 512), the segment queue
// (2 words, zeroed by a kernel at every call) and the partial rows of a
// segmented pass of at most kNoallocMaxSegments segments (for the
// reference-order pass, 3 x stride doubles).
// ---------------------------------------------------------------------------
namespace {

constexpr int kNoallocMaxTargets = 262144;
constexpr int kNoallocMaxSources = 262144;
constexpr int kNoallocMaxSegments = 1024;
constexpr size_t kNoallocTargetBytes = 3 * sizeof(double) * (size_t)kNoallocMaxTargets;
constexpr size_t kNoallocSourceBytes = sizeof(Source) * (size_t)kNoallocMaxSources;
constexpr size_t kNoallocQueueBytes = 256;
// (segments + target tiles) rows of 3 x kBlock T doubles, largest at T = 4
constexpr size_t kNoallocPartialBytes =
    (size_t)(kNoallocMaxSegments + kNoallocMaxTargets / (4 * kBlock)) * 3 * 4 * kBlock *
    sizeof(double);
constexpr size_t kNoallocScratchBytes =
    kNoallocTargetBytes + kNoallocSourceBytes + kNoallocQueueBytes + kNoallocPartialBytes;
constexpr int kMaxDevices = 64;

std::mutex scratch_mutex;
void *scratch_addr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

// The scratch array's address on `device` (which must be current); resolved
// once per device.  Null with an error string if the lookup fails.
void *scratch_for_device(int device, std::string &err)
{
    if (device < 0 || device >= kMaxDevices) {
        err = "device index beyond the scratch table";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_addr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        scratch_addr[device] = ptr;
    }
    return scratch_addr[device];
}

}  // namespace

struct bve_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;

    // parameters
    double omega = 2.0 * kPi;
    double smooth = 0.05;
    int precision = BVE_PRECISION_FP64;
    double nearfield = 0.02;
    double err_target = 1e-10;   // BVE_PRECISION_MIXED_BOUNDED
    int renormalize = 0;
    int voronoi = 0;   // 1 while a Voronoi step runs (its module's denominator form)
    int plan_t = 0, plan_tile = 0, plan_wg = 0, plan_recip = 0;
    int plan_chunk = 0;   // explicit source chunk length (ensemble members), 0 = derived
    bool chunked_only = false;   // ensemble members: the batched launch runs chunked plans
    int n_cu = 256;              // compute units of the device
    unsigned *queue = nullptr;   // segment queue of the segmented pair pass (2 counters)
    // this device's address of the allocation-free scratch (null if the
    // symbol lookup failed; scratch_error says why)
    unsigned char *scratch = nullptr;
    std::string scratch_error;
    // numeric status word (device; see flag_nonfinite) and its host copy
    unsigned long long *status = nullptr;
    unsigned long long status_host = 0;

    // Pinned host arena for the per-step host transfers: uploads are copied
    // in and sent asynchronously; downloads land here and are copied to the
    // caller's arrays after the one synchronisation of the call (host_sync).
    char *pinned = nullptr;
    size_t pinned_cap = 0, pinned_used = 0;
    struct Download {
        void *host;
        const char *pinned;
        size_t bytes;
        const int *order;   // non-null: xyz record k goes to host record order[k]
    };
    std::vector<Download> downloads;

    // problem sizes
    bool panels_loaded = false;
    int n_active = 0, n_passive = 0, n_particles = 0;
    int n_targets = 0, t_stride = 0, n_src_pad = 0;

    // device state (targets [active | passive | particles], SoA rows)
    double *x0 = nullptr, *xs = nullptr, *acc = nullptr;
    double *zeta0 = nullptr, *zacc = nullptr, *area = nullptr;
    double *absvort = nullptr, *relvort = nullptr;
    double *panel_absvort = nullptr;   // active panels' absolute vorticity (forced steps)
    Source *src = nullptr;
    std::vector<int> passive_order;   // passive row k holds the caller's panel passive_order[k]
    double *partial = nullptr;
    size_t partial_cap = 0;   // doubles
    double *ke = nullptr, *ke_part = nullptr, *ke_total = nullptr;
    int ke_part_cap = 0;
    double *staging = nullptr;   // AoS staging for uploads/downloads
    float4 *src32 = nullptr;     // mixed precision: float32 source records
    float4 *srcw = nullptr;
    float *srcc = nullptr;
    double4 *tile_box = nullptr, *group_box = nullptr;
    double *tile_mom = nullptr;   // bounded mode: kTileMoments per 256-source tile
    double *sub_mom = nullptr;    // bounded mode: kSubMoments per 64-source sub-tile
    unsigned short *route = nullptr;   // bounded mode: routing per (target group, source tile)
    size_t route_cap = 0;
    double *scale = nullptr;      // bounded mode: velocity scale of the current pass
    // set between a bounded stage update, which folded max |u| into scale,
    // and the next pair pass of the same run, which then needs no probe
    bool scale_from_stage = false;
    int mixed_src_cap = 0, mixed_tgt_cap = 0, mixed_src_cap_mom = 0;
    size_t staging_cap = 0;      // doubles
    bool particles_have_absvort = false;
    PairPlan plan{};

    // adjoint (active panels only)
    int adj_n = 0, adj_pad = 0, adj_steps = 0, adj_every = 0, adj_ckpts = 0;
    double adj_dt = 0.0;
    bool adj_ready = false;
    bool adj_last_stored = false;   // adj_stages holds the last segment's stages
    double *adj_ckpt = nullptr;     // adj_ckpts x 4 rows (x, y, z, zeta) x adj_pad
    Source *adj_stages = nullptr;   // adj_every steps x 4 stages x adj_pad records
    double *adj_sums = nullptr;     // their forward sums S: adj_every x 4 x 3 rows of adj_pad
    double *adj_lam = nullptr, *adj_lam_acc = nullptr, *adj_ybar = nullptr, *adj_m = nullptr;
    AdjRec *adj_rec = nullptr;
    double *adj_pt = nullptr;   // partial rows of the adjoint's pair pass
    size_t adj_pt_cap = 0, adj_ckpt_cap = 0, adj_stage_cap = 0, adj_sums_cap = 0;
    // mixed pair pass of the adjoint (bve_set_adjoint_precision)
    bool adj_mixed = false;
    double adj_ratio = 0.7;
    AdjRec32 *adj_rec32 = nullptr;
    double4 *adj_tbox = nullptr;
    size_t adj_rec32_cap = 0, adj_tbox_cap = 0;

    // profiling
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, event_pool;
    long long prof_launches = 0;
    double prof_ms = 0.0;

    int fail(int code, const char *fmt, ...)
    {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        error = buf;
        return code;
    }
    int hip_fail(hipError_t e, const char *what)
    {
        return fail(e == hipErrorOutOfMemory ? BVE_ERR_MEMORY : BVE_ERR_DEVICE, "%s: %s", what,
                    hipGetErrorString(e));
    }
};

#define HIP_TRY(ctx, call)                                       \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return (ctx)->hip_fail(e_, #call); \
    } while (0)

namespace {

int choose_targets_per_lane(int n_targets)
{
    if (n_targets >= 65536) return 4;
    if (n_targets >= 16384) return 2;
    return 1;
}

// Launch geometry: target tiles of kBlock*T, and a source chunk length that
// gives about target_wg work items (or exactly `chunk` sources when > 0).
PairPlan make_plan(int n_targets, int n_src_pad, int T, int tile, int target_wg, int recip,
                   int chunk = 0)
{
    PairPlan p{};
    p.targets_per_lane = T;
    p.tile = tile;
    p.recip = recip;
    p.n_tiles = std::max(1, ceil_div(n_targets, kBlock * T));
    const int want_chunks = std::max(1, ceil_div(target_wg, p.n_tiles));
    p.chunk = chunk > 0 ? round_up(chunk, tile)
                        : round_up(std::max(tile, ceil_div(n_src_pad, want_chunks)), tile);
    p.n_chunks = std::max(1, ceil_div(n_src_pad, p.chunk));
    p.n_work = p.n_tiles * p.n_chunks;
    p.grid = round_up(p.n_work, 8);
    return p;
}

// Segmented geometry: about `segments` equal runs of the tile-major item
// sequence, run by a persistent grid of one workgroup per resident slot
// (slots = 0: one workgroup per segment).
PairPlan make_segmented_plan(int n_targets, int n_src_pad, int T, int tile, int segments,
                             int recip, int slots)
{
    PairPlan p{};
    p.targets_per_lane = T;
    p.tile = tile;
    p.recip = recip;
    p.n_tiles = std::max(1, ceil_div(n_targets, kBlock * T));
    p.n_src_tiles = n_src_pad / tile;
    const int n_items = p.n_tiles * p.n_src_tiles;
    p.seg_len = std::max(1, ceil_div(n_items, std::max(1, segments)));
    p.segments = ceil_div(n_items, p.seg_len);
    p.grid = slots > 0 ? std::min(slots, p.segments) : p.segments;
    p.chunk = p.seg_len * tile;
    p.n_chunks = 0;
    p.n_work = p.segments;
    p.src_step = 1;
    return p;
}

// Segmented geometry with segments of an explicit length: seg_len items of
// the tile-major sequence each (an explicit chunk of seg_len * tile
// sources), run by a persistent grid of at most `slots` workgroups.
PairPlan make_len_plan(int n_targets, int n_src_pad, int T, int tile, int seg_len, int recip,
                       int slots)
{
    PairPlan p{};
    p.targets_per_lane = T;
    p.tile = tile;
    p.recip = recip;
    p.n_tiles = std::max(1, ceil_div(n_targets, kBlock * T));
    p.n_src_tiles = n_src_pad / tile;
    const int n_items = p.n_tiles * p.n_src_tiles;
    p.seg_len = std::max(1, std::min(seg_len, n_items));
    p.segments = ceil_div(n_items, p.seg_len);
    p.grid = slots > 0 ? std::min(slots, p.segments) : p.segments;
    p.chunk = p.seg_len * tile;
    p.n_chunks = 0;
    p.n_work = p.segments;
    p.src_step = 1;
    return p;
}

// A stride near 0.618 n that is coprime with n (1 for n <= 2): k -> k * step
// mod n visits every index once, consecutive k far apart.
int coprime_step(int n)
{
    if (n <= 2) return 1;
    int p = std::max(1, (int)std::lround(0.618 * n));
    while (std::gcd(p, n) != 1) ++p;
    return p;
}

int ensure_staging(bve_ctx *c, size_t doubles)
{
    if (doubles <= c->staging_cap) return BVE_OK;
    dev_free(c->staging);
    size_t cap = std::max(doubles, (size_t)4096);
    HIP_TRY(c, dev_alloc(&c->staging, cap));
    c->staging_cap = cap;
    return BVE_OK;
}

int ensure_partial(bve_ctx *c, size_t doubles)
{
    if (doubles <= c->partial_cap) return BVE_OK;
    dev_free(c->partial);
    HIP_TRY(c, dev_alloc(&c->partial, doubles));
    c->partial_cap = doubles;
    return BVE_OK;
}

void free_state(bve_ctx *c)
{
    dev_free(c->x0);
    dev_free(c->xs);
    dev_free(c->acc);
    dev_free(c->zeta0);
    dev_free(c->zacc);
    dev_free(c->area);
    dev_free(c->absvort);
    dev_free(c->relvort);
    dev_free(c->panel_absvort);
    dev_free(c->src);
    dev_free(c->ke);
    dev_free(c->ke_part);
    dev_free(c->ke_total);
    c->ke_part_cap = 0;
    c->panels_loaded = false;
    c->n_active = c->n_passive = c->n_particles = 0;
    c->passive_order.clear();
    c->n_targets = c->t_stride = c->n_src_pad = 0;
    c->particles_have_absvort = false;
}

int recip_for(const bve_ctx *c)
{
    if (c->precision != BVE_PRECISION_FP64_FAST) return 1;
    return c->plan_recip ? c->plan_recip : 4;
}

// Reference-order pass: one chunk of all sources, 256 targets per workgroup.
PairPlan make_exact_plan(int n_targets, int n_src_pad)
{
    PairPlan p{};
    p.targets_per_lane = 1;
    p.tile = kBlock;
    p.recip = 1;
    p.n_tiles = std::max(1, ceil_div(n_targets, kBlock));
    p.chunk = n_src_pad;
    p.n_chunks = 1;
    p.n_work = p.n_tiles;
    p.grid = p.n_tiles;
    p.exact = 1;
    return p;
}

PairPlan plan_for(const bve_ctx *c, int n_targets, int n_src_pad)
{
    if (c->precision == BVE_PRECISION_REFERENCE) return make_exact_plan(n_targets, n_src_pad);
    if (c->precision == BVE_PRECISION_MIXED_BOUNDED) {
        // the mixed pass's geometry: 256-source tiles (the expansion and
        // error-bound granularity), stride-ordered segment queue, or
        // interleaved chunks when a chunk is set; rows hold x_i x S sums
        const int T = c->plan_t ? c->plan_t : (n_targets >= 131072 ? 2 : 1);
        const int slots = pair_bounded_blocks_per_cu(T, n_src_pad) * c->n_cu;
        PairPlan p;
        if (c->plan_chunk > 0) {
            p = make_len_plan(n_targets, n_src_pad, T, 256, round_up(c->plan_chunk, 256) / 256, 1,
                              slots);
        } else {
            const int segments = c->plan_wg ? c->plan_wg : 28 * (slots > 0 ? slots : 4 * c->n_cu);
            p = make_segmented_plan(n_targets, n_src_pad, T, 256, segments, 1, slots);
        }
        p.src_step = coprime_step(p.n_src_tiles);
        p.crossed = 1;
        return p;
    }
    const bool mixed = c->precision == BVE_PRECISION_MIXED ||
                       c->precision == BVE_PRECISION_MIXED_KAHAN;
    // mixed: 2 targets per lane from ~130k targets (level 7: 5.3 against
    // 5.8 ms per launch), 1 below (level 6: 0.40 against 0.45 ms); never 4
    // (3 waves per SIMD instead of 4)
    const int T = c->plan_t ? c->plan_t
                            : (mixed ? (n_targets >= 131072 ? 2 : 1)
                                     : choose_targets_per_lane(n_targets));
    // the mixed kernel's tile is its bounding-sphere granularity: fixed at 256
    const int tile = mixed ? 256 : (c->plan_tile ? c->plan_tile : 256);
    if (mixed && !c->chunked_only && c->plan_chunk == 0) {
        // Mixed: the float64 pass's segment queue with the source tiles of a
        // target tile's row visited in stride order, so that each segment
        // holds a spread of them and the careful ones (neighbours of the
        // target group) do not pile up in a few slow segments.  28 segments
        // per resident workgroup: level 7, 4.90 ms per launch and 19.95 ms
        // per step, against 4.87 / 20.2 at 56 (twice the partial rows) and
        // 5.12 / 20.7 at 16; chunked with interleaved chunks 5.02 / 20.75.
        const int slots =
            pair_mixed_blocks_per_cu(T, c->precision == BVE_PRECISION_MIXED_KAHAN) * c->n_cu;
        const int segments = c->plan_wg ? c->plan_wg : 28 * (slots > 0 ? slots : 4 * c->n_cu);
        PairPlan p = make_segmented_plan(n_targets, n_src_pad, T, tile, segments, 1, slots);
        p.src_step = coprime_step(p.n_src_tiles);
        return p;
    }
    if (mixed || c->chunked_only || c->plan_chunk > 0) {
        // Segments of an explicit length (an explicit chunk, or the
        // ensemble's members, whose segments the batched launch deals out):
        // chunk / tile items each, the length the former chunked split gave
        // its work items by default (~64k items for the mixed pass, ~16k for
        // float64), so every member's rows number segments + tiles.  The
        // mixed pass visits each target tile's source tiles in stride order.
        const int wg = c->plan_wg ? c->plan_wg : (mixed ? 65536 : 16384);
        const int chunk = c->plan_chunk > 0
                              ? round_up(c->plan_chunk, tile)
                              : make_plan(n_targets, n_src_pad, T, tile, wg, 1).chunk;
        const int recip = recip_for(c);
        const int slots = (mixed ? pair_mixed_blocks_per_cu(T, c->precision == BVE_PRECISION_MIXED_KAHAN)
                                 : pair_f64_blocks_per_cu(T, tile, recip)) * c->n_cu;
        PairPlan p = make_len_plan(n_targets, n_src_pad, T, tile, chunk / tile, mixed ? 1 : recip,
                                   slots);
        if (mixed) p.src_step = coprime_step(p.n_src_tiles);
        p.crossed = mixed ? 0 : 1;
        return p;
    }
    // float64: items cost the same, so a persistent grid of one workgroup
    // per resident slot works through ~8 segments each (the queue evens out
    // the self-exclusion tiles); partial rows number segments + tiles
    // instead of tiles x chunks.  512-source tiles halve the LDS barriers.
    const int recip = recip_for(c);
    const int tile64 = c->plan_tile ? c->plan_tile : 512;
    const int slots = pair_f64_blocks_per_cu(T, tile64, recip) * c->n_cu;
    const int segments = c->plan_wg ? c->plan_wg : 8 * (slots > 0 ? slots : 2 * c->n_cu);
    PairPlan p = make_segmented_plan(n_targets, n_src_pad, T, tile64, segments, recip, slots);
    p.crossed = 1;
    return p;
}

int update_plan(bve_ctx *c)
{
    c->plan = plan_for(c, c->n_targets, c->n_src_pad);
    const int T = c->plan.targets_per_lane;
    if (c->plan.n_tiles * kBlock * T > c->t_stride)
        return c->fail(BVE_ERR_STATE, "internal: target tiles exceed the padded target arrays");
    return ensure_partial(c, partial_doubles(c->plan, c->t_stride));
}

// Resizes the target arrays for n_particles particles, keeping the panel part.
int pinned_reserve(bve_ctx *c, size_t bytes);

int ensure_particles(bve_ctx *c, int n_particles)
{
    if (n_particles == c->n_particles && c->x0) return BVE_OK;
    // one step's host traffic through the pinned arena: particles and their
    // absolute vorticity up; particles, relative vorticity, the panels and
    // the kinetic energy down (plus per-transfer alignment)
    int rc = pinned_reserve(c, 64 * (size_t)n_particles + 48 * (size_t)c->n_active +
                                   24 * (size_t)c->n_passive + 4096);
    if (rc) return rc;
    const int n_panel = c->n_active + c->n_passive;
    const int nt = n_panel + n_particles;
    const int stride = round_up(std::max(nt, 1), kTargetAlign);
    if (stride != c->t_stride) {
        double *x0 = nullptr, *xs = nullptr, *acc = nullptr;
        HIP_TRY(c, dev_alloc(&x0, 3 * (size_t)stride));
        HIP_TRY(c, dev_alloc(&xs, 3 * (size_t)stride));
        HIP_TRY(c, dev_alloc(&acc, 3 * (size_t)stride));
        HIP_TRY(c, hipMemsetAsync(x0, 0, 3 * sizeof(double) * stride, c->stream));
        HIP_TRY(c, hipMemsetAsync(xs, 0, 3 * sizeof(double) * stride, c->stream));
        HIP_TRY(c, hipMemsetAsync(acc, 0, 3 * sizeof(double) * stride, c->stream));
        if (c->x0 && n_panel > 0)
            for (int r = 0; r < 3; ++r)
                HIP_TRY(c, hipMemcpyAsync(x0 + (size_t)r * stride, c->x0 + (size_t)r * c->t_stride,
                                          sizeof(double) * n_panel, hipMemcpyDeviceToDevice,
                                          c->stream));
        dev_free(c->x0);
        dev_free(c->xs);
        dev_free(c->acc);
        c->x0 = x0;
        c->xs = xs;
        c->acc = acc;
        c->t_stride = stride;
    }
    if (n_particles != c->n_particles || !c->absvort) {
        dev_free(c->absvort);
        dev_free(c->relvort);
        HIP_TRY(c, dev_alloc(&c->absvort, std::max(n_particles, 1)));
        HIP_TRY(c, dev_alloc(&c->relvort, std::max(n_particles, 1)));
    }
    c->n_particles = n_particles;
    c->n_targets = nt;
    const int need = ke_partial_count(nt);
    if (need > c->ke_part_cap) {
        dev_free(c->ke_part);
        HIP_TRY(c, dev_alloc(&c->ke_part, need));
        c->ke_part_cap = need;
    }
    return update_plan(c);
}

StepState step_state(bve_ctx *c)
{
    StepState st{};
    st.n_targets = c->n_targets;
    st.t_stride = c->t_stride;
    st.n_active = c->n_active;
    st.particle_begin = c->n_active + c->n_passive;
    st.particle_end = c->n_targets;
    st.x0 = c->x0;
    st.xs = c->xs;
    st.acc = c->acc;
    st.zeta0 = c->zeta0;
    st.zacc = c->zacc;
    st.area = c->area;
    st.src = c->src;
    st.partial = c->partial;
    st.ke = c->ke;
    st.ke_part = c->ke_part;
    st.ke_total = c->ke_total;
    st.particle_absvort = c->particles_have_absvort ? c->absvort : nullptr;
    st.particle_relvort = c->relvort;
    st.unweighted = c->precision == BVE_PRECISION_REFERENCE;
    st.status = c->status;
    st.member = 0;
    return st;
}

std::pair<hipEvent_t, hipEvent_t> take_events(bve_ctx *c)
{
    if (!c->event_pool.empty()) {
        auto ev = c->event_pool.back();
        c->event_pool.pop_back();
        return ev;
    }
    hipEvent_t a = nullptr, b = nullptr;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    return {a, b};
}

// Adds the durations of completed profiled launches (call after a sync).
void collect_profile(bve_ctx *c)
{
    for (auto &ev : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) {
            c->prof_ms += ms;
            c->prof_launches += 1;
        }
        c->event_pool.push_back(ev);
    }
    c->pending.clear();
}

bool is_mixed(const bve_ctx *c)
{
    return c->precision == BVE_PRECISION_MIXED || c->precision == BVE_PRECISION_MIXED_KAHAN;
}

bool is_bounded(const bve_ctx *c) { return c->precision == BVE_PRECISION_MIXED_BOUNDED; }

// Buffers of the mixed-precision pass for n_src_pad sources and t_stride targets.
int ensure_mixed(bve_ctx *c, int n_src_pad, int t_stride)
{
    if (n_src_pad > c->mixed_src_cap) {
        dev_free(c->src32);
        dev_free(c->srcw);
        dev_free(c->srcc);
        dev_free(c->tile_box);
        HIP_TRY(c, dev_alloc(&c->src32, n_src_pad));
        HIP_TRY(c, dev_alloc(&c->srcw, n_src_pad));
        HIP_TRY(c, dev_alloc(&c->srcc, n_src_pad));
        HIP_TRY(c, dev_alloc(&c->tile_box, n_src_pad / 256));
        c->mixed_src_cap = n_src_pad;
    }
    if (t_stride > c->mixed_tgt_cap) {
        dev_free(c->group_box);
        HIP_TRY(c, dev_alloc(&c->group_box, t_stride / 64));
        c->mixed_tgt_cap = t_stride;
    }
    if (is_bounded(c)) {
        if (!c->tile_mom || n_src_pad > c->mixed_src_cap_mom) {
            dev_free(c->tile_mom);
            dev_free(c->sub_mom);
            HIP_TRY(c, dev_alloc(&c->tile_mom, (size_t)(n_src_pad / 256) * kTileMoments));
            HIP_TRY(c, dev_alloc(&c->sub_mom, (size_t)(n_src_pad / kSubTile) * kSubMoments));
            c->mixed_src_cap_mom = n_src_pad;
        }
        const size_t route_n = (size_t)(t_stride / kGroup) * (size_t)(n_src_pad / 256);
        if (!c->route || route_n > c->route_cap) {
            dev_free(c->route);
            c->route_cap = 0;
            HIP_TRY(c, dev_alloc(&c->route, route_n));
            c->route_cap = route_n;
        }
        if (!c->scale) HIP_TRY(c, dev_alloc(&c->scale, 1));
    }
    return BVE_OK;
}

// Mixed-pass arguments for the float64 pass `a` over n_targets / n_sources.
MixedArgs mixed_args_of(const bve_ctx *c, const PairArgs &a, int n_targets, int n_sources)
{
    MixedArgs ma{};
    ma.base = a;
    ma.n_targets = n_targets;
    ma.n_sources = n_sources;
    ma.src32 = c->src32;
    ma.srcw = c->srcw;
    ma.srcc = c->srcc;
    ma.tile_box = c->tile_box;
    ma.group_box = c->group_box;
    ma.rnf = c->nearfield;
    ma.kahan = c->precision == BVE_PRECISION_MIXED_KAHAN;
    ma.tile_mom = c->tile_mom;
    ma.sub_mom = c->sub_mom;
    ma.route = c->route;
    ma.n_groups = a.t_stride / kGroup;
    ma.scale = c->scale;
    ma.err_target = c->err_target;
    return ma;
}

// Pair-pass arguments of the context's resident state with launch plan `plan`.
PairArgs pair_args_of(const bve_ctx *c, const PairPlan &plan)
{
    PairArgs pa{};
    pa.tx = c->xs;
    pa.ty = c->xs + c->t_stride;
    pa.tz = c->xs + 2 * (size_t)c->t_stride;
    pa.t_stride = c->t_stride;
    pa.cls.self_end = c->n_active;
    pa.cls.smooth_begin = c->n_active;
    pa.cls.smooth_end = c->n_active + c->n_passive;
    pa.cls.eps2 = c->smooth * c->smooth;
    pa.cls.voronoi = c->voronoi;
    pa.src = c->src;
    pa.n_src_pad = c->n_src_pad;
    pa.plan = plan;
    pa.partial = c->partial;
    pa.queue = c->queue;
    pa.area = c->area;
    return pa;
}

// One pair-sum pass over n_targets targets and n_sources sources in the
// context's arithmetic (the mixed pass first builds its per-stage records).
// profile = false: no timing events even when profiling is on (they may be
// created on demand, which the allocation-free entry point must not do).
int pair_pass(bve_ctx *c, const PairArgs &a_in, int n_targets, int n_sources,
              bool profile = true)
{
    PairArgs a = a_in;
    a.n_targets = n_targets;
    a.n_sources = n_sources;
    MixedArgs ma{};
    if (is_mixed(c) || is_bounded(c)) {
        int rc = ensure_mixed(c, a.n_src_pad, a.t_stride);
        if (rc) return rc;
        ma = mixed_args_of(c, a, n_targets, n_sources);
        HIP_TRY(c, is_bounded(c) ? launch_bounded_prep(ma, c->stream, !c->scale_from_stage)
                                 : launch_mixed_prep(ma, c->stream));
        c->scale_from_stage = false;
    }
    std::pair<hipEvent_t, hipEvent_t> ev{};
    const bool timed = profile && c->profiling;
    if (timed) {
        ev = take_events(c);
        HIP_TRY(c, hipEventRecord(ev.first, c->stream));
    }
    if (is_mixed(c))
        HIP_TRY(c, launch_pair_sum_mixed(ma, c->stream));
    else if (is_bounded(c))
        HIP_TRY(c, launch_pair_sum_bounded(ma, c->stream));
    else if (a.plan.exact)
        HIP_TRY(c, launch_pair_sum_exact(a, c->stream));
    else
        HIP_TRY(c, launch_pair_sum_f64(a, c->stream));
    if (timed) {
        HIP_TRY(c, hipEventRecord(ev.second, c->stream));
        c->pending.push_back(ev);
    }
    return BVE_OK;
}

// n_steps RK4 steps of the first n_targets targets ([active | passive |
// particles] prefix) with launch plan `plan`.  If stage_store is given, the
// source records of every stage input are copied there, stage_stride records
// apart, in step-major order, and the active targets' factored sums S of
// every stage go to sum_store, 3 rows of stage_stride per stage (the
// adjoint's forward sweep and recomputation; the plan must not be crossed).
// forced: the StratosphereRK4 variant, step n starting at time t + n dt.
int run_steps_on(bve_ctx *c, double dt, int n_steps, int n_targets, const PairPlan &plan,
                 Source *stage_store, size_t stage_stride, bool forced = false, double t = 0.0,
                 double *sum_store = nullptr)
{
    if (sum_store && (plan.crossed || plan.exact)) return c->fail(BVE_ERR_STATE, "run_steps_on: sums of a crossed plan");
    StepState st = step_state(c);
    st.n_targets = n_targets;
    // bounded mode: the run's first pass probes the velocity scale; each
    // stage update then folds max |u| of its stage into the scale word for
    // the next pass (the field moves by O(dt) between stages, well inside
    // the routing's margin)
    const bool fold_scale = is_bounded(c) && c->scale;
    c->scale_from_stage = false;
    struct ScaleFlagClear {   // no later pass outside this run reuses the scale
        bve_ctx *c;
        ~ScaleFlagClear() { c->scale_from_stage = false; }
    } clear_on_exit{c};
    if (forced) {
        st.panel_absvort = c->panel_absvort;
        st.t_step = t;
        st.omega = c->omega;
    }
    const PairArgs pa = pair_args_of(c, plan);
    static const double a_next[4] = {0.5, 0.5, 1.0, 0.0};
    static const double divisor[4] = {6.0, 3.0, 3.0, 6.0};
    static const double t_frac[4] = {0.5, 0.5, 1.0, 1.0};
    HIP_TRY(c, launch_prepare_step(st, c->stream));
    for (int n = 0; n < n_steps; ++n) {
        const double tn = t + n * dt;
        for (int s = 0; s < 4; ++s) {
            if (stage_store)
                HIP_TRY(c, hipMemcpyAsync(stage_store + (size_t)(4 * n + s) * stage_stride, c->src,
                                          sizeof(Source) * c->n_src_pad, hipMemcpyDeviceToDevice,
                                          c->stream));
            int rc = pair_pass(c, pa, n_targets, c->n_active);
            if (rc) return rc;
            StageParams sp{};
            sp.stage = s;
            sp.plan = plan;
            sp.dt = dt;
            sp.omega = c->omega;
            sp.a_next = a_next[s];
            sp.divisor = divisor[s];
            sp.renormalize = c->renormalize;
            // t + dt/2, t + dt/2, t + dt, t + dt (PanelVelocityParallel.f90:746, :794, :839, :908)
            sp.t_next = t_frac[s] == 1.0 ? tn + dt : tn + t_frac[s] * dt;
            if (sum_store) {
                sp.sums_out = sum_store + (size_t)(4 * n + s) * 3 * stage_stride;
                sp.sums_stride = (int)stage_stride;
            }
            const bool fold = fold_scale && !(n == n_steps - 1 && s == 3);
            st.vel_max = fold ? c->scale : nullptr;
            if (fold) HIP_TRY(c, launch_scale_reset(c->scale, c->stream));
            HIP_TRY(c, launch_stage_update(st, sp, c->stream));
            c->scale_from_stage = fold;
        }
    }
    return BVE_OK;
}

int run_steps(bve_ctx *c, double dt, int n_steps)
{
    return run_steps_on(c, dt, n_steps, c->n_targets, c->plan, nullptr, 0);
}

// ---- adjoint helpers -------------------------------------------------------

void free_adjoint(bve_ctx *c)
{
    dev_free(c->adj_ckpt);
    dev_free(c->adj_stages);
    dev_free(c->adj_sums);
    dev_free(c->adj_lam);
    dev_free(c->adj_lam_acc);
    dev_free(c->adj_ybar);
    dev_free(c->adj_m);
    dev_free(c->adj_rec);
    dev_free(c->adj_pt);
    dev_free(c->adj_rec32);
    dev_free(c->adj_tbox);
    c->adj_pt_cap = c->adj_ckpt_cap = c->adj_stage_cap = c->adj_sums_cap = 0;
    c->adj_rec32_cap = c->adj_tbox_cap = 0;
    c->adj_ready = false;
    c->adj_last_stored = false;
    c->adj_n = c->adj_pad = 0;
}

template <class P>
int grow(bve_ctx *c, P *&p, size_t &cap, size_t count, bool zero)
{
    if (count <= cap && p) return BVE_OK;
    dev_free(p);
    HIP_TRY(c, dev_alloc(&p, count));
    if (zero) HIP_TRY(c, hipMemsetAsync(p, 0, sizeof(P) * count, c->stream));
    cap = count;
    return BVE_OK;
}

// Copies the active panels' positions and vorticity (x0 rows, zeta0) to or
// from 4 rows of adj_pad entries.
int copy_active_state(bve_ctx *c, double *rows, bool save)
{
    const size_t n = (size_t)c->n_active, P = (size_t)c->adj_pad;
    for (int r = 0; r < 3; ++r) {
        double *dev = c->x0 + (size_t)r * c->t_stride;
        if (save)
            HIP_TRY(c, hipMemcpyAsync(rows + r * P, dev, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        else
            HIP_TRY(c, hipMemcpyAsync(dev, rows + r * P, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    }
    if (save)
        HIP_TRY(c, hipMemcpyAsync(rows + 3 * P, c->zeta0, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    else
        HIP_TRY(c, hipMemcpyAsync(c->zeta0, rows + 3 * P, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    return BVE_OK;
}

// Plans of the adjoint: the active-only forward pass, the reverse sweep's
// recomputation pass and the VJP pass.  With the mixed adjoint the
// recomputation shares one float64 reciprocal among four pairs (the
// fp64_fast pass, ~1e-13 relative): it only rebuilds the stage inputs and
// sums the reverse sweep linearises about, while the forward sweep (the
// functional and the checkpoints) keeps the context's arithmetic.
int adjoint_plans(bve_ctx *c, PairPlan &fwd, PairPlan &vjp, PairPlan *recompute = nullptr)
{
    fwd = plan_for(c, c->n_active, c->n_src_pad);
    fwd.crossed = 0;   // the rows hold S, which the reverse sweep reuses
    int rc = ensure_partial(c, partial_doubles(fwd, c->t_stride));
    if (rc) return rc;
    if (recompute) {
        *recompute = fwd;
        if (c->adj_mixed && c->precision == BVE_PRECISION_FP64) {
            c->precision = BVE_PRECISION_FP64_FAST;
            *recompute = plan_for(c, c->n_active, c->n_src_pad);
            c->precision = BVE_PRECISION_FP64;
            recompute->crossed = 0;
            rc = ensure_partial(c, partial_doubles(*recompute, c->t_stride));
            if (rc) return rc;
        }
    }
    const int T = c->plan_t ? c->plan_t : choose_targets_per_lane(c->n_active);
    // The adjoint's pair pass shares one float64 reciprocal (refined once)
    // among four pairs in both float64 modes: ~1e-13 relative, far inside
    // what a gradient needs, and 4-6% off the gradient at level 6.  It runs
    // the forward pass's segment queue with 2 segments per resident
    // workgroup: its items are few and alike (level 6: 4,608 of them for
    // 768 slots), and every segment writes 6 partial rows per target.  The
    // mixed pass takes 4: its float64 tiles (near the diagonal) cost about
    // twice a float32 one, and shorter segments spread them (level 6
    // gradient: 71.5 ms at 2, 73.2 at 1, 69.0-69.2 at 4, 6 or 8).
    const int slots = vjp_blocks_per_cu(T, 4, c->adj_mixed) * c->n_cu;
    const int segments = c->plan_wg ? c->plan_wg
                                    : (c->adj_mixed ? 4 : 2) * (slots > 0 ? slots : 2 * c->n_cu);
    vjp = make_segmented_plan(c->n_active, c->adj_pad, T, 256, segments, 4, slots);
    rc = grow(c, c->adj_pt, c->adj_pt_cap,
              (size_t)(vjp.segments + vjp.n_tiles) * kAdjRows * kBlock * T, false);
    if (!rc && c->adj_mixed) {
        const size_t P = (size_t)c->adj_pad;
        rc = grow(c, c->adj_rec32, c->adj_rec32_cap, P, true);
        if (!rc) rc = grow(c, c->adj_tbox, c->adj_tbox_cap, P / 256, true);
    }
    return rc;
}

// One reverse RK4 step: lam <- lam + sum_s Ybar_s with the stage inputs in
// `stages` (4 x adj_pad source records) and their forward sums in `sums`
// (4 x 3 rows of adj_pad).
int adjoint_reverse_step(bve_ctx *c, const Source *stages, const double *sums, const PairPlan &vjp)
{
    static const double B[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    static const double C[4] = {0.5, 0.5, 1.0, 0.0};   // a of the next stage's input
    const size_t P = (size_t)c->adj_pad;
    HIP_TRY(c, hipMemcpyAsync(c->adj_lam_acc, c->adj_lam, sizeof(double) * 4 * P,
                              hipMemcpyDeviceToDevice, c->stream));
    for (int s = 3; s >= 0; --s) {
        AdjStage a{};
        a.n = c->n_active;
        a.n_pad = c->adj_pad;
        a.stride = c->adj_pad;
        a.h = c->adj_dt;
        a.omega = c->omega;
        a.b = B[s];
        a.c = C[s];
        a.lam = c->adj_lam;
        a.ybar = c->adj_ybar;
        a.lam_acc = c->adj_lam_acc;
        a.m = c->adj_m;
        a.rec = c->adj_rec;
        a.ys = stages + (size_t)s * P;
        a.sums = sums + (size_t)s * 3 * P;
        a.area = c->area;
        a.plan = vjp;
        a.partial = c->adj_pt;
        a.queue = c->queue;
        if (c->adj_mixed) {
            a.rec32 = c->adj_rec32;
            a.tile_box = c->adj_tbox;
            a.ratio = c->adj_ratio;
        }
        HIP_TRY(c, launch_adjoint_stage(a, c->stream));
    }
    std::swap(c->adj_lam, c->adj_lam_acc);
    return BVE_OK;
}

// Waits for the context's stream, then completes the pending downloads and
// frees the pinned arena for reuse.
int host_sync(bve_ctx *c)
{
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (const auto &d : c->downloads) {
        if (!d.order) {
            std::memcpy(d.host, d.pinned, d.bytes);
            continue;
        }
        const size_t rec = 3 * sizeof(double);
        for (size_t k = 0; k < d.bytes / rec; ++k)
            std::memcpy((char *)d.host + rec * (size_t)d.order[k], d.pinned + rec * k, rec);
    }
    c->downloads.clear();
    c->pinned_used = 0;
    return BVE_OK;
}

// `bytes` of the pinned arena; when it is full, the transfers in flight
// are completed first and the arena grown to hold everything that was
// asked for since the last wait, so a repeated call pattern (one step's
// uploads and downloads) soon fits without waiting midway.
int pinned_take(bve_ctx *c, size_t bytes, char **out)
{
    const size_t need = (bytes + 255) / 256 * 256;
    if (c->pinned_used + need > c->pinned_cap) {
        const size_t demand = c->pinned_used + need;
        int rc = host_sync(c);
        if (rc) return rc;
        if (c->pinned_cap < demand) {
            if (c->pinned) (void)hipHostFree(c->pinned);
            c->pinned = nullptr;
            const size_t cap = std::max({demand, 2 * c->pinned_cap, (size_t)4 << 20});
            c->pinned_cap = 0;
            HIP_TRY(c, hipHostMalloc((void **)&c->pinned, cap, hipHostMallocDefault));
            c->pinned_cap = cap;
        }
    }
    *out = c->pinned + c->pinned_used;
    c->pinned_used += need;
    return BVE_OK;
}

// Makes the pinned arena hold `bytes` without growing later (pinning host
// memory takes milliseconds: done when the problem size is set, not per step).
int pinned_reserve(bve_ctx *c, size_t bytes)
{
    if (c->pinned_cap >= bytes) return BVE_OK;
    int rc = host_sync(c);
    if (rc) return rc;
    if (c->pinned) (void)hipHostFree(c->pinned);
    c->pinned = nullptr;
    c->pinned_cap = 0;
    HIP_TRY(c, hipHostMalloc((void **)&c->pinned, bytes, hipHostMallocDefault));
    c->pinned_cap = bytes;
    return BVE_OK;
}

// Host -> device through the pinned arena (the host array is free for the
// caller as soon as this returns).
int host_up(bve_ctx *c, void *dev, const void *host, size_t bytes)
{
    if (bytes == 0) return BVE_OK;
    char *p = nullptr;
    int rc = pinned_take(c, bytes, &p);
    if (rc) return rc;
    std::memcpy(p, host, bytes);
    HIP_TRY(c, hipMemcpyAsync(dev, p, bytes, hipMemcpyHostToDevice, c->stream));
    return BVE_OK;
}

// Device -> host through the pinned arena; the host array is written by the
// next host_sync.
int host_down(bve_ctx *c, void *host, const void *dev, size_t bytes, const int *order = nullptr)
{
    if (bytes == 0 || !host) return BVE_OK;
    char *p = nullptr;
    int rc = pinned_take(c, bytes, &p);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(p, dev, bytes, hipMemcpyDeviceToHost, c->stream));
    c->downloads.push_back({host, p, bytes, order});
    return BVE_OK;
}

// Enqueues the download of the numeric status word; the next host_sync
// completes it and check_status reads it.
int enqueue_status(bve_ctx *c)
{
    c->status_host = 0;
    return host_down(c, &c->status_host, c->status, sizeof(unsigned long long));
}

// After host_sync: BVE_ERR_NUMERIC if a kernel flagged a non-finite result
// since the last check (the word is cleared).  resident: the flagged index
// is a row of the resident targets [active | passive | particles], named
// by class and the caller's index; otherwise a target of the call.
int check_status(bve_ctx *c, const char *what, bool resident)
{
    const unsigned long long w = c->status_host;
    c->status_host = 0;
    if (w == 0) return BVE_OK;
    HIP_TRY(c, hipMemsetAsync(c->status, 0, sizeof(unsigned long long), c->stream));
    const long long i = (long long)(~w & 0xffffffffull);
    if (!resident)
        return c->fail(BVE_ERR_NUMERIC, "%s: non-finite velocity at target %lld", what, i);
    const long long na = c->n_active, np = c->n_passive;
    if (i < na)
        return c->fail(BVE_ERR_NUMERIC, "%s: non-finite velocity or position at active panel %lld "
                       "(a coincident source?)", what, i);
    if (i < na + np) {
        const long long k = i - na;
        const long long panel = k < (long long)c->passive_order.size() ? c->passive_order[k] : k;
        return c->fail(BVE_ERR_NUMERIC, "%s: non-finite velocity or position at passive panel "
                       "%lld", what, panel);
    }
    return c->fail(BVE_ERR_NUMERIC, "%s: non-finite velocity or position at particle %lld "
                   "(a coincident source?)", what, i - na - np);
}

// host_sync plus the numeric status of the work it waited for.
int host_sync_checked(bve_ctx *c, const char *what)
{
    int rc = enqueue_status(c);
    if (!rc) rc = host_sync(c);
    if (!rc) rc = check_status(c, what, true);
    return rc;
}

// Spatial order of n points (AoS xyz): indices sorted by
// cube face (axis of the largest |component| and its sign), then by the
// Hilbert index of the point's gnomonic face coordinates on a 1024 x 1024
// grid (the curve has no jumps inside a face), so that runs of 64 points
// are spatially compact.
std::vector<int> hilbert_order(const double *x, int n)
{
    constexpr uint32_t N = 1024;
    auto cell = [](double w) {
        const double c = (w + 1.0) * 0.5 * N;
        return (uint32_t)std::min((double)(N - 1), std::max(0.0, c));   // NaN -> 0
    };
    std::vector<std::pair<uint64_t, int>> key((size_t)n);
    for (int i = 0; i < n; ++i) {
        const double *p = x + 3 * (size_t)i;
        int ax = 0;
        for (int d = 1; d < 3; ++d)
            if (std::fabs(p[d]) > std::fabs(p[ax])) ax = d;
        const double m = std::fabs(p[ax]);
        const uint64_t face = 2 * (uint64_t)ax + (p[ax] < 0.0 ? 1 : 0);
        uint32_t hx = cell(m > 0.0 ? p[ax == 0 ? 1 : 0] / m : 0.0);
        uint32_t hy = cell(m > 0.0 ? p[ax == 2 ? 1 : 2] / m : 0.0);
        uint64_t d = 0;
        for (uint32_t s = N / 2; s > 0; s /= 2) {
            const uint32_t rx = (hx & s) ? 1u : 0u, ry = (hy & s) ? 1u : 0u;
            d += (uint64_t)s * s * ((3u * rx) ^ ry);
            if (ry == 0) {   // rotate the quadrant
                if (rx == 1) {
                    hx = N - 1 - hx;
                    hy = N - 1 - hy;
                }
                std::swap(hx, hy);
            }
        }
        key[i] = {face * N * N + d, i};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> perm((size_t)n);
    for (int i = 0; i < n; ++i) perm[i] = key[i].second;
    return perm;
}

int upload_aos(bve_ctx *c, const double *host, int n, double *dst_rows, int stride, int off)
{
    if (n <= 0) return BVE_OK;
    int rc = ensure_staging(c, 3 * (size_t)n);
    if (rc) return rc;
    rc = host_up(c, c->staging, host, 3 * sizeof(double) * n);
    if (rc) return rc;
    HIP_TRY(c, launch_aos_to_soa(c->staging, n, dst_rows, stride, off, c->stream));
    return BVE_OK;
}

// Enqueues the download; the host array is written by the next host_sync.
// order non-null: row off + k goes to host record order[k].
int download_aos(bve_ctx *c, const double *src_rows, int stride, int off, int n, double *host,
                 const int *order = nullptr)
{
    if (n <= 0 || !host) return BVE_OK;
    int rc = ensure_staging(c, 3 * (size_t)n);
    if (rc) return rc;
    HIP_TRY(c, launch_soa_to_aos(src_rows, stride, off, n, c->staging, c->stream));
    return host_down(c, host, c->staging, 3 * sizeof(double) * n, order);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

int bve_create(bve_ctx **out, int device)
{
    if (!out) {
        context_free_error() = "bve_create: null output pointer";
        return BVE_ERR_ARGUMENT;
    }
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        context_free_error() = std::string("bve_create: no HIP device: ") + hipGetErrorString(e);
        return BVE_ERR_DEVICE;
    }
    if (device < 0 || device >= n) {
        context_free_error() = "bve_create: device index out of range";
        return BVE_ERR_ARGUMENT;
    }
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        context_free_error() = std::string("bve_create: hipSetDevice: ") + hipGetErrorString(e);
        return BVE_ERR_DEVICE;
    }
    bve_ctx *c = new bve_ctx;
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        context_free_error() = std::string("bve_create: hipStreamCreate: ") + hipGetErrorString(e);
        delete c;
        return BVE_ERR_DEVICE;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus > 0)
        c->n_cu = cus;
    // set-up of the allocation-free entry point on this device (current)
    c->scratch = (unsigned char *)scratch_for_device(device, c->scratch_error);
    // the segment queue starts at zero and every segmented launch leaves it so
    e = hipMalloc(&c->queue, 2 * sizeof(unsigned));
    if (e == hipSuccess) e = hipMemset(c->queue, 0, 2 * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc(&c->status, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->status, 0, sizeof(unsigned long long));
    if (e != hipSuccess) {
        context_free_error() = std::string("bve_create: queue: ") + hipGetErrorString(e);
        if (c->queue) (void)hipFree(c->queue);
        if (c->status) (void)hipFree(c->status);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return BVE_ERR_MEMORY;
    }
    *out = c;
    return BVE_OK;
}

int bve_destroy(bve_ctx *c)
{
    if (!c) return BVE_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_state(c);
    free_adjoint(c);
    dev_free(c->partial);
    dev_free(c->staging);
    dev_free(c->src32);
    dev_free(c->srcw);
    dev_free(c->srcc);
    dev_free(c->tile_box);
    dev_free(c->group_box);
    dev_free(c->tile_mom);
    dev_free(c->sub_mom);
    dev_free(c->route);
    dev_free(c->scale);
    if (c->queue) (void)hipFree(c->queue);
    if (c->status) (void)hipFree(c->status);
    if (c->pinned) (void)hipHostFree(c->pinned);
    for (auto &ev : c->pending) c->event_pool.push_back(ev);
    for (auto &ev : c->event_pool) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    (void)hipStreamDestroy(c->stream);
    delete c;
    return BVE_OK;
}

const char *bve_last_error(const bve_ctx *c)
{
    return c ? c->error.c_str() : context_free_error().c_str();
}

int bve_set_omega(bve_ctx *c, double omega)
{
    if (!c) return BVE_ERR_ARGUMENT;
    c->omega = omega;
    return BVE_OK;
}

int bve_set_smooth(bve_ctx *c, double smooth)
{
    if (!c) return BVE_ERR_ARGUMENT;
    c->smooth = smooth;
    return BVE_OK;
}

int bve_get_smooth(const bve_ctx *c, double *smooth)
{
    if (!c || !smooth) return BVE_ERR_ARGUMENT;
    *smooth = c->smooth;
    return BVE_OK;
}

int bve_set_precision(bve_ctx *c, int precision, double nearfield_radius)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (precision != BVE_PRECISION_FP64 && precision != BVE_PRECISION_MIXED &&
        precision != BVE_PRECISION_FP64_FAST && precision != BVE_PRECISION_MIXED_KAHAN &&
        precision != BVE_PRECISION_REFERENCE && precision != BVE_PRECISION_MIXED_BOUNDED)
        return c->fail(BVE_ERR_ARGUMENT, "bve_set_precision: unknown precision %d", precision);
    if (!(nearfield_radius >= 0.0))
        return c->fail(BVE_ERR_ARGUMENT, "bve_set_precision: nearfield_radius must be >= 0");
    c->precision = precision;
    c->nearfield = nearfield_radius;
    if (c->panels_loaded && c->x0) {
        HIP_TRY(c, hipSetDevice(c->device));
        return update_plan(c);
    }
    return BVE_OK;
}

int bve_set_error_target(bve_ctx *c, double rel_err)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!(rel_err >= 0.0) || !(rel_err < 1.0))
        return c->fail(BVE_ERR_ARGUMENT, "bve_set_error_target: target must be in [0, 1)");
    c->err_target = rel_err;
    return BVE_OK;
}

int bve_set_adjoint_precision(bve_ctx *c, int precision, double ratio)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (precision != BVE_PRECISION_FP64 && precision != BVE_PRECISION_MIXED)
        return c->fail(BVE_ERR_ARGUMENT,
                       "bve_set_adjoint_precision: precision must be BVE_PRECISION_FP64 or BVE_PRECISION_MIXED");
    if (precision == BVE_PRECISION_MIXED && !(ratio > 0.0 && ratio < 1.0))
        return c->fail(BVE_ERR_ARGUMENT, "bve_set_adjoint_precision: ratio must be in (0, 1)");
    c->adj_mixed = precision == BVE_PRECISION_MIXED;
    if (c->adj_mixed) c->adj_ratio = ratio;
    return BVE_OK;
}

int bve_set_renormalize(bve_ctx *c, int on)
{
    if (!c) return BVE_ERR_ARGUMENT;
    c->renormalize = on ? 1 : 0;
    return BVE_OK;
}

int bve_set_plan(bve_ctx *c, int targets_per_lane, int tile, int target_workgroups, int recip)
{
    if (!c) return BVE_ERR_ARGUMENT;
    const int T = targets_per_lane ? targets_per_lane : 1;
    const int TL = tile ? tile : 256;
    if (!plan_supported(T, TL) || target_workgroups < 0 ||
        !(recip == 0 || recip == 2 || recip == 4))
        return c->fail(BVE_ERR_ARGUMENT, "bve_set_plan: unsupported plan (T=%d, tile=%d, wg=%d, "
                       "recip=%d)", targets_per_lane, tile, target_workgroups, recip);
    c->plan_t = targets_per_lane;
    c->plan_tile = tile;
    c->plan_wg = target_workgroups;
    c->plan_recip = recip;
    if (c->panels_loaded && c->x0) {
        HIP_TRY(c, hipSetDevice(c->device));
        return update_plan(c);
    }
    return BVE_OK;
}

int bve_set_chunk(bve_ctx *c, int chunk)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (chunk < 0) return c->fail(BVE_ERR_ARGUMENT, "bve_set_chunk: negative chunk %d", chunk);
    c->plan_chunk = chunk;
    if (c->panels_loaded && c->x0) {
        HIP_TRY(c, hipSetDevice(c->device));
        return update_plan(c);
    }
    return BVE_OK;
}

int bve_load_panels(bve_ctx *c, int n_active, const double *x_active, const double *relvort,
                    const double *area, int n_passive, const double *x_passive)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (n_active < 1 || n_passive < 0 || !x_active || !relvort || !area ||
        (n_passive > 0 && !x_passive))
        return c->fail(BVE_ERR_ARGUMENT, "bve_load_panels: bad sizes or null arrays "
                       "(n_active=%d, n_passive=%d)", n_active, n_passive);
    HIP_TRY(c, hipSetDevice(c->device));
    c->downloads.clear();
    int rc = host_sync(c);   // uploads in flight into the old arrays
    if (rc) return rc;
    free_adjoint(c);
    // Same sizes as the loaded problem (a Voronoi step, a reload after a
    // remesh that kept the counts): the device arrays are reused, only the
    // per-problem state is dropped; otherwise they are allocated afresh.
    const bool reuse = c->panels_loaded && n_active == c->n_active && n_passive == c->n_passive;
    if (reuse) {
        dev_free(c->panel_absvort);
        c->particles_have_absvort = false;
        c->panels_loaded = false;
    } else {
        free_state(c);
        c->n_active = n_active;
        c->n_passive = n_passive;
        c->n_src_pad = round_up(n_active, kSourceAlign);
        HIP_TRY(c, dev_alloc(&c->zeta0, n_active));
        HIP_TRY(c, dev_alloc(&c->zacc, n_active));
        HIP_TRY(c, dev_alloc(&c->area, n_active));
        HIP_TRY(c, dev_alloc(&c->ke, n_active));
        HIP_TRY(c, dev_alloc(&c->ke_total, 1));
        HIP_TRY(c, dev_alloc(&c->src, c->n_src_pad));
    }
    HIP_TRY(c, hipMemsetAsync(c->src, 0, sizeof(Source) * c->n_src_pad, c->stream));
    rc = host_up(c, c->zeta0, relvort, sizeof(double) * n_active);
    if (!rc) rc = host_up(c, c->area, area, sizeof(double) * n_active);
    if (rc) return rc;
    c->n_particles = -1;   // forces the target arrays to be built
    rc = ensure_particles(c, 0);
    if (rc) return rc;
    rc = upload_aos(c, x_active, n_active, c->x0, c->t_stride, 0);
    if (rc) return rc;
    // Passive panels come in tree order, parents of all levels mixed; the
    // device holds them in Hilbert order (passive row k = the caller's
    // panel passive_order[k]), so that runs of targets are spatially
    // compact -- the mixed pass's 64-target groups need that.  Every
    // target's sum is independent of its row, and bve_store_panels puts
    // them back in the caller's order.
    c->passive_order = hilbert_order(x_passive, n_passive);
    {
        std::vector<double> sorted(3 * (size_t)n_passive);
        for (int k = 0; k < n_passive; ++k)
            std::memcpy(&sorted[3 * (size_t)k], x_passive + 3 * (size_t)c->passive_order[k],
                        3 * sizeof(double));
        rc = upload_aos(c, sorted.data(), n_passive, c->x0, c->t_stride, n_active);
    }
    if (rc) return rc;
    rc = host_sync(c);
    if (rc) return rc;
    c->panels_loaded = true;
    return BVE_OK;
}

int bve_load_particles(bve_ctx *c, int n_particles, const double *x, const double *absvort)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded) return c->fail(BVE_ERR_STATE, "bve_load_particles: no panels loaded");
    if (n_particles < 0 || (n_particles > 0 && !x))
        return c->fail(BVE_ERR_ARGUMENT, "bve_load_particles: bad arguments (n=%d)", n_particles);
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = ensure_particles(c, n_particles);
    if (rc) return rc;
    const int off = c->n_active + c->n_passive;
    rc = upload_aos(c, x, n_particles, c->x0, c->t_stride, off);
    if (rc) return rc;
    c->particles_have_absvort = absvort != nullptr;
    if (absvort && n_particles > 0) {
        rc = host_up(c, c->absvort, absvort, sizeof(double) * n_particles);
        if (rc) return rc;
    }
    // no wait: the host arrays are already copied into the pinned arena
    return BVE_OK;
}

int bve_advance(bve_ctx *c, double dt, int n_steps)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded) return c->fail(BVE_ERR_STATE, "bve_advance: no panels loaded");
    if (n_steps < 0) return c->fail(BVE_ERR_ARGUMENT, "bve_advance: n_steps < 0");
    HIP_TRY(c, hipSetDevice(c->device));
    return run_steps(c, dt, n_steps);
}

}  // extern "C"

namespace {

// Downloads of the particles (positions, relative vorticity) and of the
// stage-1 kinetic energy, completed by the caller's host_sync.
int enqueue_store_particles(bve_ctx *c, double *x, double *relvort)
{
    const int off = c->n_active + c->n_passive;
    int rc = download_aos(c, c->x0, c->t_stride, off, c->n_particles, x);
    if (rc) return rc;
    if (relvort && c->n_particles > 0) {
        if (!c->particles_have_absvort)
            return c->fail(BVE_ERR_STATE, "bve_store_particles: particles were loaded without absvort");
        rc = host_down(c, relvort, c->relvort, sizeof(double) * c->n_particles);
    }
    return rc;
}

int enqueue_kinetic_energy(bve_ctx *c, double *total_ke, double *ke)
{
    int rc = BVE_OK;
    if (total_ke) rc = host_down(c, total_ke, c->ke_total, sizeof(double));
    if (!rc && ke) rc = host_down(c, ke, c->ke, sizeof(double) * c->n_active);
    return rc;
}

// End of a host-driven step: particles, their relative vorticity and the
// kinetic energy back to the caller with one wait.
int finish_step(bve_ctx *c, double *x_particles, double *relvort, double *total_ke, double *ke)
{
    c->downloads.clear();
    int rc = enqueue_store_particles(c, x_particles, relvort);
    if (!rc) rc = enqueue_kinetic_energy(c, total_ke, ke);
    if (!rc) rc = host_sync_checked(c, "RK4 step");
    if (rc) return rc;
    collect_profile(c);
    return BVE_OK;
}

}  // namespace

extern "C" {

int bve_store_particles(bve_ctx *c, double *x, double *relvort)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded) return c->fail(BVE_ERR_STATE, "bve_store_particles: no panels loaded");
    HIP_TRY(c, hipSetDevice(c->device));
    c->downloads.clear();
    int rc = enqueue_store_particles(c, x, relvort);
    if (!rc) rc = host_sync_checked(c, "bve_store_particles");
    if (rc) return rc;
    collect_profile(c);
    return BVE_OK;
}

int bve_get_kinetic_energy(bve_ctx *c, double *total_ke, double *ke)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded) return c->fail(BVE_ERR_STATE, "bve_get_kinetic_energy: no panels loaded");
    HIP_TRY(c, hipSetDevice(c->device));
    c->downloads.clear();
    int rc = enqueue_kinetic_energy(c, total_ke, ke);
    if (!rc) rc = host_sync_checked(c, "bve_get_kinetic_energy");
    if (rc) return rc;
    collect_profile(c);
    return BVE_OK;
}

int bve_synchronize(bve_ctx *c)
{
    if (!c) return BVE_ERR_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device));
    const int rc = host_sync_checked(c, "bve_synchronize");
    if (rc) return rc;
    collect_profile(c);
    return BVE_OK;
}

int bve_rk4_step(bve_ctx *c, double dt, int n_particles, double *x_particles,
                 const double *absvort, double *relvort, double *total_ke, double *ke)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded) return c->fail(BVE_ERR_STATE, "bve_rk4_step: no panels loaded");
    if (n_particles < 0 || (n_particles > 0 && !x_particles))
        return c->fail(BVE_ERR_ARGUMENT, "bve_rk4_step: bad particle arguments (n=%d)", n_particles);
    if (relvort && !absvort)
        return c->fail(BVE_ERR_ARGUMENT, "bve_rk4_step: relvort requested without absvort");
    int rc = bve_load_particles(c, n_particles, x_particles, absvort);
    if (rc) return rc;
    rc = run_steps(c, dt, 1);
    if (rc) return rc;
    return finish_step(c, x_particles, relvort, total_ke, ke);
}

int bve_load_panel_absvort(bve_ctx *c, const double *absvort_active)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded)
        return c->fail(BVE_ERR_STATE, "bve_load_panel_absvort: no panels loaded");
    if (!absvort_active) return c->fail(BVE_ERR_ARGUMENT, "bve_load_panel_absvort: null array");
    HIP_TRY(c, hipSetDevice(c->device));
    if (!c->panel_absvort) HIP_TRY(c, dev_alloc(&c->panel_absvort, (size_t)c->n_active));
    return host_up(c, c->panel_absvort, absvort_active, sizeof(double) * c->n_active);
}

int bve_advection_rk4_step(bve_ctx *c, double t, double dt, int divergent, int n_particles,
                           double *x_particles)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded)
        return c->fail(BVE_ERR_STATE, "bve_advection_rk4_step: no panels loaded");
    if (n_particles < 0 || (n_particles > 0 && !x_particles) || (divergent != 0 && divergent != 1))
        return c->fail(BVE_ERR_ARGUMENT, "bve_advection_rk4_step: bad arguments (n=%d, "
                       "divergent=%d)", n_particles, divergent);
    int rc = bve_load_particles(c, n_particles, x_particles, nullptr);
    if (rc) return rc;
    HIP_TRY(c, launch_advection_step(step_state(c), t, dt, divergent, c->stream));
    c->downloads.clear();
    rc = enqueue_store_particles(c, x_particles, nullptr);
    if (!rc) rc = host_sync_checked(c, "advection RK4 step");
    return rc;
}

int bve_stratosphere_rk4_step(bve_ctx *c, double t, double dt, int n_particles,
                              double *x_particles, const double *absvort, double *relvort,
                              double *total_ke, double *ke)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded)
        return c->fail(BVE_ERR_STATE, "bve_stratosphere_rk4_step: no panels loaded");
    if (!c->panel_absvort)
        return c->fail(BVE_ERR_STATE, "bve_stratosphere_rk4_step: no panel absolute vorticity "
                       "(bve_load_panel_absvort)");
    if (n_particles < 0 || (n_particles > 0 && !x_particles))
        return c->fail(BVE_ERR_ARGUMENT, "bve_stratosphere_rk4_step: bad particle arguments "
                       "(n=%d)", n_particles);
    if (relvort && !absvort)
        return c->fail(BVE_ERR_ARGUMENT, "bve_stratosphere_rk4_step: relvort requested without "
                       "absvort");
    int rc = bve_load_particles(c, n_particles, x_particles, absvort);
    if (rc) return rc;
    rc = run_steps_on(c, dt, 1, c->n_targets, c->plan, nullptr, 0, true, t);
    if (rc) return rc;
    return finish_step(c, x_particles, relvort, total_ke, ke);
}

int bve_store_panels(bve_ctx *c, double *x_active, double *relvort, double *x_passive)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded) return c->fail(BVE_ERR_STATE, "bve_store_panels: no panels loaded");
    HIP_TRY(c, hipSetDevice(c->device));
    c->downloads.clear();
    int rc = download_aos(c, c->x0, c->t_stride, 0, c->n_active, x_active);
    if (!rc)
        rc = download_aos(c, c->x0, c->t_stride, c->n_active, c->n_passive, x_passive,
                          c->passive_order.data());
    if (!rc && relvort) rc = host_down(c, relvort, c->zeta0, sizeof(double) * c->n_active);
    if (!rc) rc = host_sync_checked(c, "bve_store_panels");
    return rc;
}

int bve_reset(bve_ctx *c)
{
    if (!c) return BVE_ERR_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device));
    c->downloads.clear();
    const int rc = host_sync(c);
    if (rc) return rc;
    collect_profile(c);
    free_state(c);
    free_adjoint(c);
    return BVE_OK;
}

int bve_velocity(bve_ctx *c, int n_targets, const double *x_targets, int self_end,
                 int smooth_begin, int smooth_end, int n_sources, const double *x_sources,
                 const double *vort, const double *area, double *u, double *dvort)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (n_targets < 0 || n_sources < 0 || (n_targets > 0 && (!x_targets || !u)) ||
        (n_sources > 0 && (!x_sources || !vort || !area)))
        return c->fail(BVE_ERR_ARGUMENT, "bve_velocity: bad sizes or null arrays");
    if (self_end < 0 || self_end > n_targets || self_end > n_sources || smooth_begin < 0 ||
        smooth_end < smooth_begin || smooth_end > n_targets)
        return c->fail(BVE_ERR_ARGUMENT, "bve_velocity: bad target classes (self_end=%d, "
                       "smooth=[%d,%d), n_targets=%d, n_sources=%d)", self_end, smooth_begin,
                       smooth_end, n_targets, n_sources);
    if (n_targets == 0) return BVE_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    const int stride = round_up(n_targets, kTargetAlign);
    const int ns_pad = round_up(std::max(n_sources, 1), kSourceAlign);
    double *t = nullptr, *xs = nullptr, *vv = nullptr, *aa = nullptr, *uu = nullptr, *dv = nullptr;
    Source *src = nullptr;
    int rc = BVE_OK;
    hipError_t e = hipSuccess;
    auto cleanup = [&]() {
        dev_free(t); dev_free(xs); dev_free(vv); dev_free(aa); dev_free(uu); dev_free(dv);
        dev_free(src);
    };
#define VTRY(call)                                       \
    do {                                                 \
        e = (call);                                      \
        if (e != hipSuccess) {                           \
            rc = c->hip_fail(e, #call);                  \
            (void)hipStreamSynchronize(c->stream);       \
            cleanup();                                   \
            return rc;                                   \
        }                                                \
    } while (0)
    VTRY(dev_alloc(&t, 3 * (size_t)stride));
    VTRY(dev_alloc(&src, ns_pad));
    VTRY(dev_alloc(&uu, 3 * (size_t)n_targets));
    VTRY(hipMemsetAsync(t, 0, 3 * sizeof(double) * stride, c->stream));
    VTRY(hipMemsetAsync(src, 0, sizeof(Source) * ns_pad, c->stream));
    if (n_sources > 0) {
        VTRY(dev_alloc(&xs, 3 * (size_t)n_sources));
        VTRY(dev_alloc(&vv, n_sources));
        VTRY(dev_alloc(&aa, n_sources));
        VTRY(hipMemcpyAsync(xs, x_sources, 3 * sizeof(double) * n_sources, hipMemcpyHostToDevice, c->stream));
        VTRY(hipMemcpyAsync(vv, vort, sizeof(double) * n_sources, hipMemcpyHostToDevice, c->stream));
        VTRY(hipMemcpyAsync(aa, area, sizeof(double) * n_sources, hipMemcpyHostToDevice, c->stream));
        VTRY(launch_make_sources(xs, vv, aa, n_sources, c->precision == BVE_PRECISION_REFERENCE,
                                 src, c->stream));
    }
    rc = ensure_staging(c, 3 * (size_t)n_targets);
    if (rc) { cleanup(); return rc; }
    VTRY(hipMemcpyAsync(c->staging, x_targets, 3 * sizeof(double) * n_targets, hipMemcpyHostToDevice, c->stream));
    VTRY(launch_aos_to_soa(c->staging, n_targets, t, stride, 0, c->stream));
    const PairPlan plan = plan_for(c, n_targets, ns_pad);
    rc = ensure_partial(c, partial_doubles(plan, stride));
    if (rc) { cleanup(); return rc; }
    PairArgs pa{};
    pa.tx = t;
    pa.ty = t + stride;
    pa.tz = t + 2 * (size_t)stride;
    pa.t_stride = stride;
    pa.cls.self_end = self_end;
    pa.cls.smooth_begin = smooth_begin;
    pa.cls.smooth_end = smooth_end;
    pa.cls.eps2 = c->smooth * c->smooth;
    pa.src = src;
    pa.n_src_pad = ns_pad;
    pa.plan = plan;
    pa.partial = c->partial;
    pa.queue = c->queue;
    pa.area = aa;
    rc = pair_pass(c, pa, n_targets, n_sources);
    if (rc) { (void)hipStreamSynchronize(c->stream); cleanup(); return rc; }
    if (dvort && self_end > 0) VTRY(dev_alloc(&dv, self_end));
    VTRY(launch_finalize_velocity(c->partial, plan, t, stride, n_targets, c->omega, uu,
                                  dv, self_end, c->status, c->stream));
    VTRY(hipMemcpyAsync(u, uu, 3 * sizeof(double) * n_targets, hipMemcpyDeviceToHost, c->stream));
    if (dv) VTRY(hipMemcpyAsync(dvort, dv, sizeof(double) * self_end, hipMemcpyDeviceToHost, c->stream));
    VTRY(hipStreamSynchronize(c->stream));
    VTRY(hipMemcpy(&c->status_host, c->status, sizeof(unsigned long long), hipMemcpyDeviceToHost));
#undef VTRY
    collect_profile(c);
    cleanup();
    return check_status(c, "bve_velocity", false);
}

int bve_velocity_noalloc(bve_ctx *c, int n_targets, const double *x_targets, int self_end,
                         int smooth_begin, int smooth_end, int n_sources, const double *x_sources,
                         const double *vort, const double *area, double *u, double *dvort)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (n_targets < 0 || n_sources < 0 || (n_targets > 0 && (!x_targets || !u)) ||
        (n_sources > 0 && (!x_sources || !vort || !area)))
        return c->fail(BVE_ERR_ARGUMENT, "bve_velocity_noalloc: bad sizes or null arrays");
    if (n_targets > kNoallocMaxTargets || n_sources > kNoallocMaxSources)
        return c->fail(BVE_ERR_ARGUMENT, "bve_velocity_noalloc: %d targets / %d sources exceed the "
                       "scratch's maximum of %d / %d", n_targets, n_sources, kNoallocMaxTargets,
                       kNoallocMaxSources);
    if (self_end < 0 || self_end > n_targets || self_end > n_sources || smooth_begin < 0 ||
        smooth_end < smooth_begin || smooth_end > n_targets)
        return c->fail(BVE_ERR_ARGUMENT, "bve_velocity_noalloc: bad target classes (self_end=%d, "
                       "smooth=[%d,%d), n_targets=%d, n_sources=%d)", self_end, smooth_begin,
                       smooth_end, n_targets, n_sources);
    if (c->precision != BVE_PRECISION_FP64 && c->precision != BVE_PRECISION_FP64_FAST &&
        c->precision != BVE_PRECISION_REFERENCE)
        return c->fail(BVE_ERR_ARGUMENT, "bve_velocity_noalloc: runs the float64 and "
                       "reference-order modes (precision %d)", c->precision);
    if (!c->scratch)
        return c->fail(BVE_ERR_DEVICE, "bve_velocity_noalloc: no scratch on device %d: %s",
                       c->device, c->scratch_error.c_str());
    if (n_targets == 0) return BVE_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    const int stride = round_up(n_targets, kTargetAlign);
    const int ns_pad = round_up(std::max(n_sources, 1), kSourceAlign);
    double *t = (double *)c->scratch;
    Source *src = (Source *)(c->scratch + kNoallocTargetBytes);
    unsigned *queue = (unsigned *)(c->scratch + kNoallocTargetBytes + kNoallocSourceBytes);
    double *partial = (double *)(c->scratch + kNoallocTargetBytes + kNoallocSourceBytes +
                                 kNoallocQueueBytes);
    PairPlan plan = plan_for(c, n_targets, ns_pad);
    if (!plan.exact && plan.segments > kNoallocMaxSegments) {
        // the same geometry with fewer, longer segments
        plan = make_segmented_plan(n_targets, ns_pad, plan.targets_per_lane, plan.tile,
                                   kNoallocMaxSegments, plan.recip,
                                   std::min(plan.grid, kNoallocMaxSegments));
        plan.crossed = 1;
    }
    if (partial_doubles(plan, stride) * sizeof(double) > kNoallocPartialBytes)
        return c->fail(BVE_ERR_ARGUMENT, "bve_velocity_noalloc: the launch plan's partial sums "
                       "exceed the scratch (use the default plan)");
    HIP_TRY(c, launch_aos_to_soa_padded(x_targets, n_targets, t, stride, c->stream));
    HIP_TRY(c, launch_make_sources_padded(x_sources, vort, area, n_sources, ns_pad,
                                          c->precision == BVE_PRECISION_REFERENCE, src,
                                          c->stream));
    HIP_TRY(c, launch_zero_words(queue, 2, c->stream));
    PairArgs pa{};
    pa.tx = t;
    pa.ty = t + stride;
    pa.tz = t + 2 * (size_t)stride;
    pa.t_stride = stride;
    pa.cls.self_end = self_end;
    pa.cls.smooth_begin = smooth_begin;
    pa.cls.smooth_end = smooth_end;
    pa.cls.eps2 = c->smooth * c->smooth;
    pa.src = src;
    pa.n_src_pad = ns_pad;
    pa.plan = plan;
    pa.partial = partial;
    pa.queue = queue;
    pa.area = area;
    const int rc = pair_pass(c, pa, n_targets, n_sources, false);
    if (rc) return rc;
    HIP_TRY(c, launch_finalize_velocity(partial, plan, t, stride, n_targets, c->omega, u,
                                        dvort, dvort ? self_end : 0, c->status, c->stream));
    return BVE_OK;
}

int bve_resident_velocity(bve_ctx *c, double *u_active, double *u_passive, double *u_particles)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded)
        return c->fail(BVE_ERR_STATE, "bve_resident_velocity: no panels loaded");
    HIP_TRY(c, hipSetDevice(c->device));
    c->downloads.clear();
    int rc = host_sync(c);   // earlier uploads and downloads complete
    if (rc) return rc;
    const int n = c->n_targets, na = c->n_active, np = c->n_passive;
    rc = ensure_staging(c, 3 * (size_t)std::max(n, 1));
    if (rc) return rc;
    // the step-start state as the first stage's input, then one pass of
    // the step's own plan
    const StepState st = step_state(c);
    HIP_TRY(c, launch_prepare_step(st, c->stream));
    rc = pair_pass(c, pair_args_of(c, c->plan), n, na);
    if (rc) return rc;
    HIP_TRY(c, launch_finalize_velocity(c->partial, c->plan, c->xs, c->t_stride, n, c->omega,
                                        c->staging, nullptr, 0, c->status, c->stream));
    rc = host_down(c, u_active, c->staging, 3 * sizeof(double) * na);
    if (!rc)
        rc = host_down(c, u_passive, c->staging + 3 * (size_t)na, 3 * sizeof(double) * np,
                       c->passive_order.data());
    if (!rc)
        rc = host_down(c, u_particles, c->staging + 3 * (size_t)(na + np),
                       3 * sizeof(double) * (size_t)std::max(c->n_particles, 0));
    if (!rc) rc = host_sync_checked(c, "bve_resident_velocity");
    if (rc) return rc;
    collect_profile(c);
    return BVE_OK;
}

int bve_voronoi_rk4_step(bve_ctx *c, double dt, int n, double *x, double *y, double *z,
                         double *relvort, const double *area, double *total_ke, double *ke)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (n < 1 || !x || !y || !z || !relvort || !area)
        return c->fail(BVE_ERR_ARGUMENT, "bve_voronoi_rk4_step: bad sizes or null arrays (n=%d)", n);
    std::vector<double> aos(3 * (size_t)n);
    for (int j = 0; j < n; ++j) {
        aos[3 * (size_t)j] = x[j];
        aos[3 * (size_t)j + 1] = y[j];
        aos[3 * (size_t)j + 2] = z[j];
    }
    const int saved_renorm = c->renormalize;
    int rc = bve_load_panels(c, n, aos.data(), relvort, area, 0, nullptr);
    if (rc) return rc;
    rc = ensure_particles(c, 0);
    if (rc) return rc;
    c->renormalize = 1;   // NormalizeGrid after every Voronoi step (VoronoiVelocityParallel.f90:480)
    c->voronoi = 1;
    rc = run_steps(c, dt, 1);
    c->renormalize = saved_renorm;
    c->voronoi = 0;
    if (rc) return rc;
    rc = bve_store_panels(c, aos.data(), relvort, nullptr);
    if (rc) return rc;
    rc = bve_get_kinetic_energy(c, total_ke, ke);
    if (rc) return rc;
    for (int j = 0; j < n; ++j) {
        x[j] = aos[3 * (size_t)j];
        y[j] = aos[3 * (size_t)j + 1];
        z[j] = aos[3 * (size_t)j + 2];
    }
    // the device arrays stay allocated for the next step of the same size
    return BVE_OK;
}

int bve_adjoint_forward(bve_ctx *c, double dt, int n_steps, int checkpoint_every, double *x_T,
                        double *zeta_T)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded) return c->fail(BVE_ERR_STATE, "bve_adjoint_forward: no panels loaded");
    if (n_steps < 1 || checkpoint_every < 1)
        return c->fail(BVE_ERR_ARGUMENT, "bve_adjoint_forward: n_steps and checkpoint_every must be >= 1");
    if (c->precision != BVE_PRECISION_FP64 && c->precision != BVE_PRECISION_FP64_FAST)
        return c->fail(BVE_ERR_STATE, "bve_adjoint_forward: the adjoint runs in the float64 modes");
    if (c->renormalize)
        return c->fail(BVE_ERR_STATE, "bve_adjoint_forward: renormalisation is not differentiated; switch it off");
    HIP_TRY(c, hipSetDevice(c->device));
    const int every = std::min(checkpoint_every, n_steps);
    const int ckpts = ceil_div(n_steps, every);
    if (c->adj_n != c->n_active) free_adjoint(c);
    c->adj_n = c->n_active;
    c->adj_pad = round_up(c->n_active, 1024);
    const size_t P = (size_t)c->adj_pad;
    int rc = grow(c, c->adj_ckpt, c->adj_ckpt_cap, (size_t)ckpts * 4 * P, true);
    if (!rc) rc = grow(c, c->adj_stages, c->adj_stage_cap, (size_t)every * 4 * P, true);
    if (!rc) rc = grow(c, c->adj_sums, c->adj_sums_cap, (size_t)every * 4 * 3 * P, true);
    size_t cap4 = 0;
    if (!rc && !c->adj_lam) {
        rc = grow(c, c->adj_lam, cap4, 4 * P, true);
        cap4 = 0;
        if (!rc) rc = grow(c, c->adj_lam_acc, cap4, 4 * P, true);
        cap4 = 0;
        if (!rc) rc = grow(c, c->adj_ybar, cap4, 4 * P, true);
        cap4 = 0;
        if (!rc) rc = grow(c, c->adj_m, cap4, 3 * P, true);
        cap4 = 0;
        if (!rc) rc = grow(c, c->adj_rec, cap4, P, true);
    }
    if (rc) return rc;
    PairPlan fwd, vjp;
    rc = adjoint_plans(c, fwd, vjp);
    if (rc) return rc;
    // the last segment's stage inputs are kept: the reverse sweep starts
    // there and need not recompute them
    c->adj_last_stored = false;
    for (int k = 0; k < ckpts; ++k) {
        const bool last = k == ckpts - 1;
        rc = copy_active_state(c, c->adj_ckpt + (size_t)k * 4 * P, true);
        if (!rc) rc = run_steps_on(c, dt, std::min(every, n_steps - k * every), c->n_active, fwd,
                                   last ? c->adj_stages : nullptr, last ? P : 0, false, 0.0,
                                   last ? c->adj_sums : nullptr);
        if (rc) return rc;
    }
    c->adj_last_stored = true;
    c->downloads.clear();
    rc = download_aos(c, c->x0, c->t_stride, 0, c->n_active, x_T);
    if (!rc && zeta_T) rc = host_down(c, zeta_T, c->zeta0, sizeof(double) * c->n_active);
    if (rc) return rc;
    rc = copy_active_state(c, c->adj_ckpt, false);   // the resident state is left at t = 0
    if (!rc) rc = host_sync_checked(c, "bve_adjoint_forward");
    if (rc) return rc;
    collect_profile(c);
    c->adj_dt = dt;
    c->adj_steps = n_steps;
    c->adj_every = every;
    c->adj_ckpts = ckpts;
    c->adj_ready = true;
    return BVE_OK;
}

int bve_difference_forward(bve_ctx *c, double dt, int n_steps, const double *dzeta0,
                           double *x_T, double *zeta_T, double *dx_T, double *dzeta_T)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded)
        return c->fail(BVE_ERR_STATE, "bve_difference_forward: no panels loaded");
    if (n_steps < 0 || !dzeta0 || !dx_T || !dzeta_T)
        return c->fail(BVE_ERR_ARGUMENT, "bve_difference_forward: bad arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    const int n = c->n_active;
    const size_t R = (size_t)round_up(n, kTargetAlign);
    PairPlan plan = make_plan(n, n, 1, kBlock, 4096, 1);
    DiffState st{};
    st.n = n;
    st.stride = (int)R;
    st.area = c->area;
    double *rows = nullptr;
    DiffRec *rec = nullptr;
    double *partial = nullptr;
    int rc = BVE_OK;
    hipError_t e = hipSuccess;
    auto cleanup = [&]() {
        dev_free(rows);
        dev_free(rec);
        dev_free(partial);
    };
#define DTRY(call)                                       \
    do {                                                 \
        e = (call);                                      \
        if (e != hipSuccess) {                           \
            rc = c->hip_fail(e, #call);                  \
            (void)hipStreamSynchronize(c->stream);       \
            cleanup();                                   \
            return rc;                                   \
        }                                                \
    } while (0)
    DTRY(dev_alloc(&rows, 16 * R));
    DTRY(dev_alloc(&rec, R));
    DTRY(dev_alloc(&partial, (size_t)plan.n_chunks * 6 * R));
    DTRY(hipMemsetAsync(rows, 0, sizeof(double) * 16 * R, c->stream));
    DTRY(hipMemsetAsync(rec, 0, sizeof(DiffRec) * R, c->stream));
    st.x0 = rows;
    st.z0 = rows + 3 * R;
    st.acc = rows + 4 * R;
    st.zacc = rows + 7 * R;
    st.dx0 = rows + 8 * R;
    st.dz0 = rows + 11 * R;
    st.dacc = rows + 12 * R;
    st.dzacc = rows + 15 * R;
    st.rec = rec;
    st.partial = partial;
    // the resident active state is the base; it is left as it is
    for (int r = 0; r < 3; ++r)
        DTRY(hipMemcpyAsync(st.x0 + r * R, c->x0 + (size_t)r * c->t_stride, sizeof(double) * n,
                            hipMemcpyDeviceToDevice, c->stream));
    DTRY(hipMemcpyAsync(st.z0, c->zeta0, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    DTRY(hipMemcpyAsync(st.dz0, dzeta0, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    DTRY(launch_diff_prepare(st, c->stream));
    for (int k = 0; k < n_steps; ++k)
        for (int s = 0; s < 4; ++s) {
            DTRY(launch_diff_pair(st, plan, c->stream));
            DTRY(launch_diff_stage(st, plan, s, dt, c->omega, c->stream));
        }
    std::vector<double> h(16 * R);
    DTRY(hipMemcpyAsync(h.data(), rows, sizeof(double) * 16 * R, hipMemcpyDeviceToHost, c->stream));
    DTRY(hipStreamSynchronize(c->stream));
#undef DTRY
    cleanup();
    for (int i = 0; i < n; ++i)
        for (int r = 0; r < 3; ++r) {
            if (x_T) x_T[3 * (size_t)i + r] = h[r * R + i];
            dx_T[3 * (size_t)i + r] = h[(8 + r) * R + i];
        }
    for (int i = 0; i < n; ++i) {
        if (zeta_T) zeta_T[i] = h[3 * R + i];
        dzeta_T[i] = h[11 * R + i];
    }
    return BVE_OK;
}

int bve_adjoint_backward(bve_ctx *c, const double *lambda_x, const double *lambda_zeta,
                         double *grad_x0, double *grad_zeta0)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->adj_ready || c->adj_n != c->n_active || !c->panels_loaded)
        return c->fail(BVE_ERR_STATE, "bve_adjoint_backward: call bve_adjoint_forward first");
    if (!lambda_x || !lambda_zeta)
        return c->fail(BVE_ERR_ARGUMENT, "bve_adjoint_backward: null cotangent");
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t P = (size_t)c->adj_pad;
    const int n = c->n_active;
    int rc = ensure_staging(c, 3 * (size_t)n);
    if (rc) return rc;
    HIP_TRY(c, hipMemsetAsync(c->adj_lam, 0, sizeof(double) * 4 * P, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->staging, lambda_x, 3 * sizeof(double) * n, hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, launch_aos_to_soa(c->staging, n, c->adj_lam, c->adj_pad, 0, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->adj_lam + 3 * P, lambda_zeta, sizeof(double) * n,
                              hipMemcpyHostToDevice, c->stream));
    PairPlan fwd, vjp, recompute;
    rc = adjoint_plans(c, fwd, vjp, &recompute);
    if (rc) return rc;
    for (int seg = c->adj_ckpts - 1; seg >= 0; --seg) {
        const int n0 = seg * c->adj_every;
        const int len = std::min(c->adj_every, c->adj_steps - n0);
        if (!(seg == c->adj_ckpts - 1 && c->adj_last_stored)) {
            rc = copy_active_state(c, c->adj_ckpt + (size_t)seg * 4 * P, false);
            if (!rc) rc = run_steps_on(c, c->adj_dt, len, c->n_active, recompute, c->adj_stages, P,
                                       false, 0.0, c->adj_sums);
            if (rc) return rc;
        }
        c->adj_last_stored = false;   // the buffer now goes to earlier segments
        for (int l = len - 1; l >= 0; --l) {
            rc = adjoint_reverse_step(c, c->adj_stages + (size_t)l * 4 * P,
                                      c->adj_sums + (size_t)l * 4 * 3 * P, vjp);
            if (rc) return rc;
        }
    }
    c->downloads.clear();
    rc = download_aos(c, c->adj_lam, c->adj_pad, 0, n, grad_x0);
    if (!rc && grad_zeta0) rc = host_down(c, grad_zeta0, c->adj_lam + 3 * P, sizeof(double) * n);
    if (rc) return rc;
    rc = copy_active_state(c, c->adj_ckpt, false);
    if (!rc) rc = host_sync_checked(c, "bve_adjoint_backward");
    if (rc) return rc;
    collect_profile(c);
    return BVE_OK;
}

int bve_set_profiling(bve_ctx *c, int on)
{
    if (!c) return BVE_ERR_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    collect_profile(c);
    c->profiling = on != 0;
    c->prof_launches = 0;
    c->prof_ms = 0.0;
    return BVE_OK;
}

int bve_get_profile(bve_ctx *c, long long *launches, double *total_ms)
{
    if (!c) return BVE_ERR_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    collect_profile(c);
    if (launches) *launches = c->prof_launches;
    if (total_ms) *total_ms = c->prof_ms;
    return BVE_OK;
}

int bve_get_plan(const bve_ctx *c, int *targets_per_lane, int *tile, int *chunk, int *n_chunks,
                 int *recip)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded) return BVE_ERR_STATE;
    if (targets_per_lane) *targets_per_lane = c->plan.targets_per_lane;
    if (tile) *tile = c->plan.tile;
    if (chunk) *chunk = c->plan.chunk;
    if (n_chunks) *n_chunks = c->plan.segments > 0 ? c->plan.segments : c->plan.n_chunks;
    if (recip) *recip = c->plan.recip;
    return BVE_OK;
}

int bve_get_launch(const bve_ctx *c, int *segments, int *grid, int *n_tiles,
                   long long *partial_bytes)
{
    if (!c) return BVE_ERR_ARGUMENT;
    if (!c->panels_loaded) return BVE_ERR_STATE;
    if (segments) *segments = c->plan.segments;
    if (grid) *grid = c->plan.grid;
    if (n_tiles) *n_tiles = c->plan.n_tiles;
    if (partial_bytes) {
        // rows actually written: chunked, every chunk's rows of all targets
        // (padding included); segmented, one row per (segment, tile) piece,
        // the last piece being row (last segment + last tile)
        const PairPlan &p = c->plan;
        const long long rows = (long long)(p.n_tiles * p.n_src_tiles - 1) / p.seg_len + p.n_tiles;
        *partial_bytes = p.segments > 0
                             ? rows * 3 * kBlock * p.targets_per_lane * (long long)sizeof(double)
                             : (long long)p.n_chunks * 3 * c->t_stride * (long long)sizeof(double);
    }
    return BVE_OK;
}

int bve_pairs_per_step(const bve_ctx *c, double *pairs)
{
    if (!c || !pairs) return BVE_ERR_ARGUMENT;
    const double na = c->n_active, other = (double)c->n_passive + (double)std::max(c->n_particles, 0);
    *pairs = 4.0 * (na * (na - 1.0) + other * na);
    return BVE_OK;
}

}  // extern "C"

// ===========================================================================
// Ensembles: many independent problems advanced by shared launches
// ===========================================================================

struct bve_ensemble {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;
    std::vector<bve_ctx *> members;
    int grid_wg = 0;   // workgroups per persistent pair launch, 0 = occupancy
    int items_per_wg = 16;   // derived segment lengths: work items per resident workgroup
    int n_cu = 256;
    int precision = BVE_PRECISION_FP64;

    // device tables (rebuilt when a member or a parameter changes)
    bool dirty = true;
    EnsItem *d_items = nullptr;
    unsigned *d_queue = nullptr;   // pair-launch queue counters, two per item group
    PairArgs *d_pair = nullptr;
    MixedArgs *d_mixed = nullptr;
    StepState *d_states = nullptr;
    PairPlan *d_plans = nullptr;
    int2 *d_stage_map = nullptr, *d_src_map = nullptr, *d_tgt_map = nullptr;
    unsigned long long *d_status = nullptr;   // numeric status of every member's stages
    EnsLaunch launch{};

    // captured steps
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double g_dt = 0.0;
    int g_steps = 0;

    // profiling: HIP events around pair launches (or graph launches)
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, pool;
    long long prof_launches = 0;
    double prof_ms = 0.0;

    int fail(int code, const char *fmt, ...)
    {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        error = buf;
        return code;
    }
    int hip_fail(hipError_t e, const char *what)
    {
        return fail(e == hipErrorOutOfMemory ? BVE_ERR_MEMORY : BVE_ERR_DEVICE, "%s: %s", what,
                    hipGetErrorString(e));
    }
};

namespace {

void ens_free_tables(bve_ensemble *e)
{
    dev_free(e->d_items);
    dev_free(e->d_queue);
    dev_free(e->d_pair);
    dev_free(e->d_mixed);
    dev_free(e->d_states);
    dev_free(e->d_plans);
    dev_free(e->d_stage_map);
    dev_free(e->d_src_map);
    dev_free(e->d_tgt_map);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->gexec = nullptr;
    e->graph = nullptr;
    e->dirty = true;
}

template <class P>
int ens_upload(bve_ensemble *e, P *&dst, const std::vector<P> &v)
{
    dev_free(dst);
    if (v.empty()) return BVE_OK;
    HIP_TRY(e, dev_alloc(&dst, v.size()));
    HIP_TRY(e, hipMemcpyAsync(dst, v.data(), sizeof(P) * v.size(), hipMemcpyHostToDevice, e->stream));
    return BVE_OK;
}

int ens_build(bve_ensemble *e)
{
    if (!e->dirty) return BVE_OK;
    ens_free_tables(e);
    if (!e->d_status) {
        HIP_TRY(e, dev_alloc(&e->d_status, 1));
        HIP_TRY(e, hipMemsetAsync(e->d_status, 0, sizeof(unsigned long long), e->stream));
    }
    const bool mixed = e->precision == BVE_PRECISION_MIXED ||
                       e->precision == BVE_PRECISION_MIXED_KAHAN;
    std::vector<PairArgs> pairs;
    std::vector<MixedArgs> mixeds;
    std::vector<StepState> states;
    std::vector<PairPlan> plans;
    // (targets-per-lane group, -cost): groups in order, largest items first
    std::vector<std::pair<std::pair<int, long long>, EnsItem>> costed;
    std::vector<int2> stage_map, src_map, tgt_map;
    static const int kT[kEnsGroups] = {1, 2, 4};
    const bool kahan = e->precision == BVE_PRECISION_MIXED_KAHAN;
    const int recip = e->precision == BVE_PRECISION_FP64_FAST ? 4 : 1;
    auto group_of = [](int T) { return T == 1 ? 0 : (T == 2 ? 1 : 2); };
    auto group_slots = [&](int g) {
        const int per_cu = ensemble_pair_blocks_per_cu(kT[g], mixed, kahan, recip);
        return e->grid_wg ? e->grid_wg : (per_cu > 0 ? per_cu : 2) * e->n_cu;
    };
    for (int m = 0; m < (int)e->members.size(); ++m) {
        bve_ctx *c = e->members[m];
        if (!c->panels_loaded || !c->x0)
            return e->fail(BVE_ERR_STATE, "ensemble member %d has no panels loaded", m);
        // the persistent kernel runs 256-source tiles
        if (c->plan_tile != 256 || c->plan.tile != 256) {
            c->plan_tile = 256;
            int rc = update_plan(c);
            if (rc) return e->fail(rc, "member %d: %s", m, c->error.c_str());
        }
    }
    // Members without an explicit chunk share one segment length per
    // targets-per-lane group: items_per_wg segments per resident workgroup
    // of the group's batched launch over all its members' (target tile,
    // source tile) items, which cost the same at every level.  The queue
    // then balances the launch as the single-context plan does, and a
    // member's partial rows number its segments + tiles (level 6 in a
    // 64-member ensemble: a few hundred rows instead of ~2,800 with
    // 1,024-source chunks).
    long long group_items[kEnsGroups] = {0, 0, 0};
    for (bve_ctx *c : e->members)
        if (c->plan_chunk == 0)
            group_items[group_of(c->plan.targets_per_lane)] +=
                (long long)c->plan.n_tiles * c->plan.n_src_tiles;
    for (int m = 0; m < (int)e->members.size(); ++m) {
        bve_ctx *c = e->members[m];
        if (c->plan_chunk != 0) continue;
        const int g = group_of(c->plan.targets_per_lane);
        const long long len =
            std::max(1LL, group_items[g] / ((long long)e->items_per_wg * std::max(1, group_slots(g))));
        c->plan_chunk = (int)std::min<long long>(len, c->plan.n_src_tiles * (long long)c->plan.n_tiles) * 256;
        const PairPlan p = plan_for(c, c->n_targets, c->n_src_pad);
        c->plan_chunk = 0;
        c->plan = p;
        const int rc = ensure_partial(c, partial_doubles(p, c->t_stride));
        if (rc) return e->fail(rc, "member %d: %s", m, c->error.c_str());
    }
    for (int m = 0; m < (int)e->members.size(); ++m) {
        bve_ctx *c = e->members[m];
        if (mixed) {
            int rc = ensure_mixed(c, c->n_src_pad, c->t_stride);
            if (rc) return e->fail(rc, "member %d: %s", m, c->error.c_str());
        }
        const PairArgs pa = pair_args_of(c, c->plan);
        pairs.push_back(pa);
        mixeds.push_back(mixed_args_of(c, pa, c->n_targets, c->n_active));
        states.push_back(step_state(c));
        states.back().status = e->d_status;
        states.back().member = (unsigned)m;
        plans.push_back(c->plan);
        const int T = c->plan.targets_per_lane;
        const int group = group_of(T);
        const long long cost = (long long)kBlock * T * c->plan.chunk;
        for (int sg = 0; sg < c->plan.segments; ++sg)
            costed.push_back({{group, -cost}, EnsItem{m, sg}});
        for (int b = 0; b < ceil_div(c->n_targets, 256); ++b) stage_map.push_back(make_int2(m, b));
        for (int b = 0; b < c->n_src_pad / 256; ++b) src_map.push_back(make_int2(m, b));
        for (int b = 0; b < c->t_stride / 256; ++b) tgt_map.push_back(make_int2(m, b));
    }
    // one launch per targets-per-lane group; in a group, largest items first
    // (the persistent grid takes them from a queue)
    std::stable_sort(costed.begin(), costed.end(),
                     [](const auto &a, const auto &b) { return a.first < b.first; });
    std::vector<EnsItem> items;
    items.reserve(costed.size());
    int group_begin[kEnsGroups] = {0, 0, 0}, group_count[kEnsGroups] = {0, 0, 0};
    for (size_t k = 0; k < costed.size(); ++k) {
        const int g = costed[k].first.first;
        if (group_count[g]++ == 0) group_begin[g] = (int)k;
        items.push_back(costed[k].second);
    }
    int rc = ens_upload(e, e->d_items, items);
    if (!rc && !e->d_queue) {
        HIP_TRY(e, dev_alloc(&e->d_queue, 2 * kEnsGroups));
        HIP_TRY(e, hipMemsetAsync(e->d_queue, 0, 2 * kEnsGroups * sizeof(unsigned), e->stream));
    }
    if (!rc) rc = ens_upload(e, e->d_pair, pairs);
    if (!rc) rc = ens_upload(e, e->d_mixed, mixeds);
    if (!rc) rc = ens_upload(e, e->d_states, states);
    if (!rc) rc = ens_upload(e, e->d_plans, plans);
    if (!rc) rc = ens_upload(e, e->d_stage_map, stage_map);
    if (!rc) rc = ens_upload(e, e->d_src_map, src_map);
    if (!rc) rc = ens_upload(e, e->d_tgt_map, tgt_map);
    if (rc) return rc;
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    EnsLaunch &L = e->launch;
    L = EnsLaunch{};
    L.n_members = (int)e->members.size();
    L.items = e->d_items;
    L.queue = e->d_queue;
    L.pair_members = e->d_pair;
    L.mixed_members = e->d_mixed;
    L.mixed = mixed;
    L.kahan = kahan;
    L.recip = recip;
    for (int g = 0; g < kEnsGroups; ++g) {
        L.group_begin[g] = group_begin[g];
        L.group_count[g] = group_count[g];
        // default grid: every workgroup slot the kernel's occupancy allows
        L.group_grid[g] = std::max(1, std::min(group_count[g], group_slots(g)));
    }
    L.states = e->d_states;
    L.plans = e->d_plans;
    L.stage_map = e->d_stage_map;
    L.n_stage_blocks = (int)stage_map.size();
    L.src_map = e->d_src_map;
    L.n_src_blocks = (int)src_map.size();
    L.tgt_map = e->d_tgt_map;
    L.n_tgt_blocks = (int)tgt_map.size();
    e->dirty = false;
    return BVE_OK;
}

std::pair<hipEvent_t, hipEvent_t> ens_events(bve_ensemble *e)
{
    if (!e->pool.empty()) {
        auto ev = e->pool.back();
        e->pool.pop_back();
        return ev;
    }
    hipEvent_t a = nullptr, b = nullptr;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    return {a, b};
}

void ens_collect(bve_ensemble *e)
{
    for (auto &ev : e->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) {
            e->prof_ms += ms;
            e->prof_launches += 1;
        }
        e->pool.push_back(ev);
    }
    e->pending.clear();
}

// Enqueues n_steps RK4 steps of every member (timed pair launches unless
// the steps are being captured into a graph).
int ens_enqueue(bve_ensemble *e, double dt, int n_steps, bool timed)
{
    static const double a_next[4] = {0.5, 0.5, 1.0, 0.0};
    static const double divisor[4] = {6.0, 3.0, 3.0, 6.0};
    const EnsLaunch &L = e->launch;
    const double omega = e->members.empty() ? 0.0 : e->members[0]->omega;
    HIP_TRY(e, launch_ensemble_prepare(L, e->stream));
    for (int n = 0; n < n_steps; ++n) {
        for (int s = 0; s < 4; ++s) {
            if (L.mixed) HIP_TRY(e, launch_ensemble_mixed_prep(L, e->stream));
            std::pair<hipEvent_t, hipEvent_t> ev{};
            if (timed) {
                ev = ens_events(e);
                HIP_TRY(e, hipEventRecord(ev.first, e->stream));
            }
            HIP_TRY(e, launch_ensemble_pairs(L, e->stream));
            if (timed) {
                HIP_TRY(e, hipEventRecord(ev.second, e->stream));
                e->pending.push_back(ev);
            }
            StageParams sp{};
            sp.stage = s;
            sp.dt = dt;
            sp.omega = omega;
            sp.a_next = a_next[s];
            sp.divisor = divisor[s];
            sp.renormalize = e->members.empty() ? 0 : e->members[0]->renormalize;
            HIP_TRY(e, launch_ensemble_stage(L, sp, e->stream));
        }
    }
    return BVE_OK;
}

}  // namespace

extern "C" {

int bve_ensemble_create(bve_ensemble **out, int device, int n_members)
{
    if (!out || n_members < 1) {
        context_free_error() = "bve_ensemble_create: bad arguments";
        return BVE_ERR_ARGUMENT;
    }
    *out = nullptr;
    bve_ensemble *e = new bve_ensemble;
    e->device = device;
    for (int m = 0; m < n_members; ++m) {
        bve_ctx *c = nullptr;
        const int rc = bve_create(&c, device);
        if (rc) {
            for (bve_ctx *d : e->members) bve_destroy(d);
            delete e;
            return rc;
        }
        c->chunked_only = true;
        e->members.push_back(c);
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        e->n_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        for (bve_ctx *d : e->members) bve_destroy(d);
        delete e;
        context_free_error() = "bve_ensemble_create: hipStreamCreate failed";
        return BVE_ERR_DEVICE;
    }
    *out = e;
    return BVE_OK;
}

int bve_ensemble_destroy(bve_ensemble *e)
{
    if (!e) return BVE_OK;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    ens_free_tables(e);
    dev_free(e->d_status);
    for (auto &ev : e->pending) e->pool.push_back(ev);
    for (auto &ev : e->pool) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    for (bve_ctx *c : e->members) bve_destroy(c);
    (void)hipStreamDestroy(e->stream);
    delete e;
    return BVE_OK;
}

const char *bve_ensemble_last_error(const bve_ensemble *e)
{
    return e ? e->error.c_str() : context_free_error().c_str();
}

int bve_ensemble_set_params(bve_ensemble *e, double omega, double smooth, int precision,
                            double nearfield_radius)
{
    if (!e) return BVE_ERR_ARGUMENT;
    if (precision == BVE_PRECISION_REFERENCE || precision == BVE_PRECISION_MIXED_BOUNDED)
        return e->fail(BVE_ERR_ARGUMENT, "bve_ensemble_set_params: the reference-order and "
                       "bounded mixed modes run per context, not in the batched launch");
    for (bve_ctx *c : e->members) {
        int rc = bve_set_omega(c, omega);
        if (!rc) rc = bve_set_smooth(c, smooth);
        if (!rc) rc = bve_set_precision(c, precision, nearfield_radius);
        if (rc) return e->fail(rc, "%s", c->error.c_str());
    }
    e->precision = precision;
    e->dirty = true;
    ens_free_tables(e);
    return BVE_OK;
}

int bve_ensemble_load_member(bve_ensemble *e, int member, int n_active, const double *x_active,
                             const double *relvort, const double *area, int n_passive,
                             const double *x_passive, int n_particles, const double *x_particles,
                             const double *absvort)
{
    if (!e) return BVE_ERR_ARGUMENT;
    if (member < 0 || member >= (int)e->members.size())
        return e->fail(BVE_ERR_ARGUMENT, "bve_ensemble_load_member: member %d out of range", member);
    bve_ctx *c = e->members[member];
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    ens_free_tables(e);
    int rc = bve_load_panels(c, n_active, x_active, relvort, area, n_passive, x_passive);
    if (!rc) rc = bve_load_particles(c, n_particles, x_particles, absvort);
    // the batched launches run on the ensemble's stream: finish the
    // member's uploads (issued on its own stream) first
    if (!rc) rc = host_sync(c);
    if (rc) return e->fail(rc, "member %d: %s", member, c->error.c_str());
    return BVE_OK;
}

int bve_ensemble_set_member_plan(bve_ensemble *e, int member, int targets_per_lane, int chunk)
{
    if (!e) return BVE_ERR_ARGUMENT;
    if (member < 0 || member >= (int)e->members.size())
        return e->fail(BVE_ERR_ARGUMENT, "bve_ensemble_set_member_plan: member %d out of range", member);
    if (!(targets_per_lane == 0 || targets_per_lane == 1 || targets_per_lane == 2 ||
          targets_per_lane == 4) || chunk < 0)
        return e->fail(BVE_ERR_ARGUMENT, "bve_ensemble_set_member_plan: unsupported plan (T=%d, chunk=%d)",
                       targets_per_lane, chunk);
    bve_ctx *c = e->members[member];
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    ens_free_tables(e);
    c->plan_t = targets_per_lane;
    c->plan_tile = 256;
    c->plan_chunk = chunk;
    if (c->panels_loaded && c->x0) {
        const int rc = update_plan(c);
        if (rc) return e->fail(rc, "member %d: %s", member, c->error.c_str());
    }
    return BVE_OK;
}

int bve_ensemble_get_member_plan(bve_ensemble *e, int member, int *targets_per_lane, int *chunk,
                                 int *work_items)
{
    if (!e) return BVE_ERR_ARGUMENT;
    if (member < 0 || member >= (int)e->members.size())
        return e->fail(BVE_ERR_ARGUMENT, "bve_ensemble_get_member_plan: member %d out of range", member);
    HIP_TRY(e, hipSetDevice(e->device));
    const int rc = ens_build(e);
    if (rc) return rc;
    const PairPlan &p = e->members[member]->plan;
    if (targets_per_lane) *targets_per_lane = p.targets_per_lane;
    if (chunk) *chunk = p.chunk;
    if (work_items) *work_items = p.segments;
    return BVE_OK;
}

int bve_ensemble_set_items_per_workgroup(bve_ensemble *e, int items)
{
    if (!e || items < 0) return BVE_ERR_ARGUMENT;
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    ens_free_tables(e);
    e->items_per_wg = items ? items : 16;
    return BVE_OK;
}

int bve_ensemble_set_grid(bve_ensemble *e, int workgroups)
{
    if (!e || workgroups < 0) return BVE_ERR_ARGUMENT;
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    ens_free_tables(e);
    e->grid_wg = workgroups;
    return BVE_OK;
}

int bve_ensemble_advance(bve_ensemble *e, double dt, int n_steps, int use_graph)
{
    if (!e) return BVE_ERR_ARGUMENT;
    if (n_steps < 0) return e->fail(BVE_ERR_ARGUMENT, "bve_ensemble_advance: n_steps < 0");
    HIP_TRY(e, hipSetDevice(e->device));
    int rc = ens_build(e);
    if (rc) return rc;
    if (!use_graph) return ens_enqueue(e, dt, n_steps, e->profiling);
    if (!e->gexec || e->g_dt != dt || e->g_steps != n_steps) {
        if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
        if (e->graph) (void)hipGraphDestroy(e->graph);
        e->gexec = nullptr;
        e->graph = nullptr;
        HIP_TRY(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        rc = ens_enqueue(e, dt, n_steps, false);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(e->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return e->hip_fail(ec, "hipStreamEndCapture");
        e->graph = g;
        HIP_TRY(e, hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
        e->g_dt = dt;
        e->g_steps = n_steps;
    }
    std::pair<hipEvent_t, hipEvent_t> ev{};
    if (e->profiling) {
        ev = ens_events(e);
        HIP_TRY(e, hipEventRecord(ev.first, e->stream));
    }
    HIP_TRY(e, hipGraphLaunch(e->gexec, e->stream));
    if (e->profiling) {
        HIP_TRY(e, hipEventRecord(ev.second, e->stream));
        e->pending.push_back(ev);
    }
    return BVE_OK;
}

int bve_ensemble_synchronize(bve_ensemble *e)
{
    if (!e) return BVE_ERR_ARGUMENT;
    HIP_TRY(e, hipSetDevice(e->device));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    ens_collect(e);
    if (!e->d_status) return BVE_OK;
    unsigned long long w = 0;
    HIP_TRY(e, hipMemcpy(&w, e->d_status, sizeof w, hipMemcpyDeviceToHost));
    if (w == 0) return BVE_OK;
    HIP_TRY(e, hipMemset(e->d_status, 0, sizeof w));
    const unsigned long long key = ~w;
    const int m = (int)(key >> 32), i = (int)(key & 0xffffffffull);
    const bve_ctx *c = e->members[m];
    const char *cls = i < c->n_active ? "active panel row"
                      : (i < c->n_active + c->n_passive ? "passive panel row" : "particle row");
    return e->fail(BVE_ERR_NUMERIC, "bve_ensemble: member %d: non-finite velocity or position at "
                   "target %d (%s)", m, i, cls);
}

int bve_ensemble_store_member(bve_ensemble *e, int member, double *x_active, double *relvort,
                              double *x_passive, double *x_particles, double *relvort_particles,
                              double *total_ke)
{
    if (!e) return BVE_ERR_ARGUMENT;
    if (member < 0 || member >= (int)e->members.size())
        return e->fail(BVE_ERR_ARGUMENT, "bve_ensemble_store_member: member %d out of range", member);
    int rc = bve_ensemble_synchronize(e);
    if (rc) return rc;
    bve_ctx *c = e->members[member];
    rc = bve_store_panels(c, x_active, relvort, x_passive);
    if (!rc && (x_particles || relvort_particles))
        rc = bve_store_particles(c, x_particles, relvort_particles);
    if (!rc && total_ke) rc = bve_get_kinetic_energy(c, total_ke, nullptr);
    if (rc) return e->fail(rc, "member %d: %s", member, c->error.c_str());
    return BVE_OK;
}

int bve_ensemble_set_profiling(bve_ensemble *e, int on)
{
    if (!e) return BVE_ERR_ARGUMENT;
    int rc = bve_ensemble_synchronize(e);
    if (rc) return rc;
    e->profiling = on != 0;
    e->prof_launches = 0;
    e->prof_ms = 0.0;
    return BVE_OK;
}

int bve_ensemble_get_profile(bve_ensemble *e, long long *launches, double *total_ms)
{
    if (!e) return BVE_ERR_ARGUMENT;
    int rc = bve_ensemble_synchronize(e);
    if (rc) return rc;
    if (launches) *launches = e->prof_launches;
    if (total_ms) *total_ms = e->prof_ms;
    return BVE_OK;
}

int bve_ensemble_pairs_per_step(const bve_ensemble *e, double *pairs)
{
    if (!e || !pairs) return BVE_ERR_ARGUMENT;
    double total = 0.0;
    for (const bve_ctx *c : e->members) {
        double p = 0.0;
        bve_pairs_per_step(c, &p);
        total += p;
    }
    *pairs = total;
    return BVE_OK;
}

}  // extern "C"
