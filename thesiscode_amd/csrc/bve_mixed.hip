// Mixed-precision Biot-Savart pair sum (float32 far field, float64 near
// field and reductions), gfx950.
//
// The sum is the one of bve_kernels.hip, written in difference form:
//
//     S_i = sum_j q_ij (x_j - x_i),  q_ij = w_j / D_ij,
//     D_ij = |x_i - x_j|^2 / 2 + c_i + c_j (+ eps2),  c = (1 - |x|^2) / 2,
//
// which is 1 - x_i.x_j (+ eps2) exactly, without its cancellation for close
// pairs, and x_i x S_i is the same cross product as in the float64 kernel
// (x_i x x_i = 0).  Float32 differences are formed from per-tile float64
// origins: every source tile (256 sources) and every group of 64 targets
// has a float64 centre c; records hold o = fl32(x - c), and a pair's
// difference is (fl32(c_g - c_s) + o_i) - o_j, so its rounding error scales
// with |x_i - x_j| plus the tile sizes, not with |x| = 1.
//
// Pairs closer than the near-field radius are evaluated in float64 from
// the float64 source records.  A tile pair is sent down the careful path
// (per-pair distance test) only when the bounding spheres of the target
// group and the source tile come within the radius, or when the tile holds
// a target's own index; all other tiles run the float32 loop.  Float32
// partial sums are carried into float64 accumulators after every tile, or
// after every pair with Kahan compensation when requested.
#include "bve_common.h"
#include "bve_kernels.h"

namespace bve {

constexpr int kMixedTile = 256;   // sources per LDS tile and per source bounding sphere
constexpr int kGroup = 64;        // targets per bounding sphere (one wave's lanes)

// ---------------------------------------------------------------------------
// Per-stage preparation: bounding spheres and float32 source records
// ---------------------------------------------------------------------------

__device__ double block_sum(double v, double *sh)
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

__device__ double block_max(double v, double *sh)
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

// One block of kMixedTile threads per source tile: centre (mean of the valid
// sources), radius, and the float32 records {x - c, w} and c_j.
__global__ __launch_bounds__(kMixedTile) void mixed_source_prep_kernel(
    const Source *__restrict__ src, int n_sources, float4 *__restrict__ src32,
    float *__restrict__ srcc, double4 *__restrict__ tile_box)
{
    __shared__ double sh[kMixedTile];
    const int j = blockIdx.x * kMixedTile + threadIdx.x;
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
    if (threadIdx.x == 0) tile_box[blockIdx.x] = make_double4(cx, cy, cz, r);
    src32[j] = make_float4((float)dx, (float)dy, (float)dz, (float)s.w);
    srcc[j] = (float)(0.5 * (1.0 - (s.x * s.x + s.y * s.y + s.z * s.z)));
}

// One wave per group of kGroup consecutive targets: centre and radius.
__global__ __launch_bounds__(256) void mixed_target_prep_kernel(
    const double *__restrict__ t, int stride, int n_targets, double4 *__restrict__ group_box,
    int n_groups)
{
    __shared__ double sh[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < n_targets;
    const double x = t[i], y = t[stride + i], z = t[2 * (size_t)stride + i];
    // per-wave sums through LDS, fixed order
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
    const int g = i / kGroup;
    if (lane == 0 && g < n_groups) group_box[g] = make_double4(cx, cy, cz, sh[base]);
}

// ---------------------------------------------------------------------------
// Pair kernel
// ---------------------------------------------------------------------------

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

template <int T, bool KAHAN>
__global__ __launch_bounds__(kBlock) void pair_sum_mixed_kernel(MixedArgs a)
{
    constexpr int TILE = kMixedTile;
    const PairPlan &plan = a.base.plan;
    const int work = (blockIdx.x % 8) * (plan.grid / 8) + blockIdx.x / 8;
    if (work >= plan.n_work) return;
    const int chunk = work / plan.n_tiles;
    const int ttile = work - chunk * plan.n_tiles;
    const int tile_base = ttile * (kBlock * T);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int stride = a.base.t_stride;
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

    __shared__ float4 lds[TILE];
    __shared__ float ldsc[TILE];
    const int j_begin = chunk * plan.chunk;
    const int j_end = min(j_begin + plan.chunk, a.base.n_src_pad);
    for (int jt = j_begin; jt < j_end; jt += TILE) {
        __syncthreads();
        lds[threadIdx.x] = a.src32[jt + threadIdx.x];
        ldsc[threadIdx.x] = a.srcc[jt + threadIdx.x];
        __syncthreads();
        const double4 sb = a.tile_box[jt / TILE];
        float ax[T], ay[T], az[T];
        bool careful = false;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const double dx = gx[t] - sb.x, dy = gy[t] - sb.y, dz = gz[t] - sb.z;
            ax[t] = (float)dx + oix[t];
            ay[t] = (float)dy + oiy[t];
            az[t] = (float)dz + oiz[t];
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            // 1e-6: slack for the float32 rounding of the per-pair test below
            careful |= sb.w >= 0.0 && dist < gr[t] + sb.w + a.rnf + 1e-6;
            const int lo = tile_base + t * kBlock + wave * 64;
            const int hi = min(lo + 64, cls.self_end);
            careful |= (lo < hi) && (lo < jt + TILE) && (jt < hi);
        }
        float Sx[T], Sy[T], Sz[T], Cx[T], Cy[T], Cz[T];
#pragma unroll
        for (int t = 0; t < T; ++t) Sx[t] = Sy[t] = Sz[t] = Cx[t] = Cy[t] = Cz[t] = 0.f;
        if (!careful) {
#pragma unroll 2
            for (int k = 0; k < TILE; ++k) {
                const float4 s = lds[k];
                const float cj = ldsc[k];
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    const float dx = ax[t] - s.x, dy = ay[t] - s.y, dz = az[t] - s.z;
                    const float h = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    const float D = __builtin_fmaf(0.5f, h, cif[t] + cj);
                    const float q = s.w * __builtin_amdgcn_rcpf(D);
                    add_f32<KAHAN>(Sx[t], Cx[t], -q * dx);
                    add_f32<KAHAN>(Sy[t], Cy[t], -q * dy);
                    add_f32<KAHAN>(Sz[t], Cz[t], -q * dz);
                }
            }
        } else {
            for (int k = 0; k < TILE; ++k) {
                const float4 s = lds[k];
                const float cj = ldsc[k];
                const int j = jt + k;
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    const float dx = ax[t] - s.x, dy = ay[t] - s.y, dz = az[t] - s.z;
                    const float h = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    if (j == self[t]) continue;
                    if (h < rnf2) {
                        // near field: float64 from the float64 record
                        const Source sj = a.base.src[j];
                        const double ex = sj.x - xi[t], ey = sj.y - yi[t], ez = sj.z - zi[t];
                        const double cj64 = 0.5 * (1.0 - (sj.x * sj.x + sj.y * sj.y + sj.z * sj.z));
                        const double D = 0.5 * (ex * ex + ey * ey + ez * ez) + (ci[t] + cj64);
                        const double q = sj.w / D;
                        S64x[t] = __builtin_fma(q, ex, S64x[t]);
                        S64y[t] = __builtin_fma(q, ey, S64y[t]);
                        S64z[t] = __builtin_fma(q, ez, S64z[t]);
                    } else {
                        const float D = __builtin_fmaf(0.5f, h, cif[t] + cj);
                        const float q = s.w * __builtin_amdgcn_rcpf(D);
                        add_f32<KAHAN>(Sx[t], Cx[t], -q * dx);
                        add_f32<KAHAN>(Sy[t], Cy[t], -q * dy);
                        add_f32<KAHAN>(Sz[t], Cz[t], -q * dz);
                    }
                }
            }
        }
        // float64 carry of the tile's float32 sums
#pragma unroll
        for (int t = 0; t < T; ++t) {
            S64x[t] += (double)Sx[t] - (double)Cx[t];
            S64y[t] += (double)Sy[t] - (double)Cy[t];
            S64z[t] += (double)Sz[t] - (double)Cz[t];
        }
    }

    double *px = a.base.partial + (size_t)(chunk * 3) * stride;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x;
        px[i] = S64x[t];
        px[stride + i] = S64y[t];
        px[2 * (size_t)stride + i] = S64z[t];
    }
}

template <int T>
static hipError_t launch_mixed_t(const MixedArgs &a, hipStream_t s)
{
    if (a.kahan)
        hipLaunchKernelGGL((pair_sum_mixed_kernel<T, true>), dim3(a.base.plan.grid), dim3(kBlock),
                           0, s, a);
    else
        hipLaunchKernelGGL((pair_sum_mixed_kernel<T, false>), dim3(a.base.plan.grid), dim3(kBlock),
                           0, s, a);
    return hipGetLastError();
}

hipError_t launch_mixed_prep(const MixedArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(mixed_source_prep_kernel, dim3(a.base.n_src_pad / kMixedTile),
                       dim3(kMixedTile), 0, s, a.base.src, a.n_sources, a.src32, a.srcc,
                       a.tile_box);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int n_groups = a.base.t_stride / kGroup;
    hipLaunchKernelGGL(mixed_target_prep_kernel, dim3(a.base.t_stride / 256), dim3(256), 0, s,
                       a.base.tx, a.base.t_stride, a.n_targets, a.group_box, n_groups);
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

}  // namespace bve
