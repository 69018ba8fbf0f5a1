// Adjoint (vector-Jacobian product) of the active-panel Biot-Savart sum and
// the reverse RK4 stage sweep, gfx950.
//
// Forward: U_i = -(1/4 PI) sum_{j != i} w_j (x_i x x_j) / D_ij, D_ij = 1 - x_i.x_j,
// w_j = zeta_j A_j, dzeta_i/dt = -2 Omega U_z,i (BVEActiveRHS,
// PanelVelocityParallel.f90:1287-1334).  For a cotangent (mu_x, mu_zeta) on
// (U, dzeta/dt) let m_i = mu_x,i - 2 Omega mu_zeta,i e_z, n_i = m_i x x_i,
// P_ij = m_i.(x_i x x_j) = x_j.n_i and L = sum_i m_i.U_i.  Then
//
//   dL/dx_k   = -(1/4 PI) (G^T_k + G^S_k)
//   G^T_i     = S_i x m_i + sum_j (w_j P_ij / D_ij^2) x_j,   S_i = sum_j (w_j / D_ij) x_j
//   G^S_j     = w_j sum_i [ n_i / D_ij + P_ij x_i / D_ij^2 ]
//   dL/dzeta_j = -(1/4 PI) A_j sum_i P_ij / D_ij
//
// The target-major sums (S, T) run with a target per lane and sources
// streamed through LDS, the source-major sums (H, g) with a source per lane
// and the cotangent records {x_i, n_i} streamed; both use the forward
// kernel's tiling, chunking, XCD mapping and batched float64 reciprocals,
// and write per-chunk partial sums that the combine kernel adds in a fixed
// order.
#include "bve_common.h"
#include "bve_kernels.h"

namespace bve {

__device__ __forceinline__ double rcp_newton1(double p)
{
    const double r = __builtin_amdgcn_rcp(p);
    const double e = __builtin_fma(-p, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double rcp_newton2(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

// 1/d_r for R denominators: R = 1 each refined twice, R = 4 one shared
// reciprocal of the product refined once (as in the forward kernel).
template <int R>
__device__ __forceinline__ void recips(const double (&d)[R], double (&r)[R])
{
    if constexpr (R == 1) {
        r[0] = rcp_newton2(d[0]);
    } else {
        static_assert(R == 4, "R must be 1 or 4");
        const double p01 = d[0] * d[1], p23 = d[2] * d[3];
        const double rr = rcp_newton1(p01 * p23);
        const double r01 = rr * p23, r23 = rr * p01;
        r[0] = d[1] * r01;
        r[1] = d[0] * r01;
        r[2] = d[3] * r23;
        r[3] = d[2] * r23;
    }
}

__device__ __forceinline__ int xcd_item(int g, int grid)
{
    return (g % 8) * (grid / 8) + g / 8;
}

// Wave-uniform: does source tile [jt, jt + tile) hold the own index of one
// of this wave's lanes (lanes carry consecutive indices from lo)?
__device__ __forceinline__ bool diag_tile(int lo, int self_end, int jt, int tile)
{
    const int hi = min(lo + 64, self_end);
    return (lo < hi) && (lo < jt + tile) && (jt < hi);
}

// ---------------------------------------------------------------------------
// Target-major pass: S_i and T_i
// ---------------------------------------------------------------------------
template <int T, int TILE, int R>
__global__ __launch_bounds__(kBlock) void vjp_target_kernel(
    const AdjRec *__restrict__ rec, int stride, int self_end, const Source *__restrict__ src,
    int n_src_pad, PairPlan plan, double *__restrict__ partial)
{
    const int work = xcd_item(blockIdx.x, plan.grid);
    if (work >= plan.n_work) return;
    const int chunk = work / plan.n_tiles;
    const int ttile = work - chunk * plan.n_tiles;
    const int tile_base = ttile * (kBlock * T);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);

    double nx[T], ny[T], nz[T], ex[T], ey[T], ez[T];
    double Sx[T], Sy[T], Sz[T], Tx[T], Ty[T], Tz[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const AdjRec a = rec[tile_base + t * kBlock + threadIdx.x];
        nx[t] = -a.x;
        ny[t] = -a.y;
        nz[t] = -a.z;
        ex[t] = a.nx;
        ey[t] = a.ny;
        ez[t] = a.nz;
        Sx[t] = Sy[t] = Sz[t] = Tx[t] = Ty[t] = Tz[t] = 0.0;
    }
    __shared__ Source lds[TILE];
    const int j_begin = chunk * plan.chunk;
    const int j_end = min(j_begin + plan.chunk, n_src_pad);
    for (int jt = j_begin; jt < j_end; jt += TILE) {
        __syncthreads();
        for (int k = threadIdx.x; k < TILE; k += kBlock) lds[k] = src[jt + k];
        __syncthreads();
        bool diag = false;
#pragma unroll
        for (int t = 0; t < T; ++t)
            diag |= diag_tile(tile_base + t * kBlock + wave * 64, self_end, jt, TILE);
        for (int k = 0; k < TILE; k += R) {
            Source s[R];
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = lds[k + r];
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int self = tile_base + t * kBlock + threadIdx.x;
                double d[R], rc[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    d[r] = __builtin_fma(nx[t], s[r].x, 1.0);
                    d[r] = __builtin_fma(ny[t], s[r].y, d[r]);
                    d[r] = __builtin_fma(nz[t], s[r].z, d[r]);
                    if (diag && jt + k + r == self) d[r] = 1.0;
                }
                recips<R>(d, rc);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    double q = s[r].w * rc[r];
                    if (diag && jt + k + r == self) q = 0.0;
                    const double p = __builtin_fma(s[r].z, ez[t], __builtin_fma(s[r].y, ey[t], s[r].x * ex[t]));
                    const double tq = q * rc[r] * p;
                    Sx[t] = __builtin_fma(q, s[r].x, Sx[t]);
                    Sy[t] = __builtin_fma(q, s[r].y, Sy[t]);
                    Sz[t] = __builtin_fma(q, s[r].z, Sz[t]);
                    Tx[t] = __builtin_fma(tq, s[r].x, Tx[t]);
                    Ty[t] = __builtin_fma(tq, s[r].y, Ty[t]);
                    Tz[t] = __builtin_fma(tq, s[r].z, Tz[t]);
                }
            }
        }
    }
    double *p = partial + (size_t)(chunk * 6) * stride;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x;
        p[i] = Sx[t];
        p[stride + i] = Sy[t];
        p[2 * (size_t)stride + i] = Sz[t];
        p[3 * (size_t)stride + i] = Tx[t];
        p[4 * (size_t)stride + i] = Ty[t];
        p[5 * (size_t)stride + i] = Tz[t];
    }
}

// ---------------------------------------------------------------------------
// Source-major pass: H_j and g_j (lane = source j, records {x_i, n_i} streamed)
// ---------------------------------------------------------------------------
template <int T, int TILE, int R>
__global__ __launch_bounds__(kBlock) void vjp_source_kernel(
    const Source *__restrict__ src, int stride, int self_end, const AdjRec *__restrict__ rec,
    int n_rec_pad, PairPlan plan, double *__restrict__ partial)
{
    const int work = xcd_item(blockIdx.x, plan.grid);
    if (work >= plan.n_work) return;
    const int chunk = work / plan.n_tiles;
    const int ttile = work - chunk * plan.n_tiles;
    const int tile_base = ttile * (kBlock * T);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);

    double nx[T], ny[T], nz[T], xj[T], yj[T], zj[T];
    double Hx[T], Hy[T], Hz[T], G[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const Source s = src[tile_base + t * kBlock + threadIdx.x];
        xj[t] = s.x;
        yj[t] = s.y;
        zj[t] = s.z;
        nx[t] = -s.x;
        ny[t] = -s.y;
        nz[t] = -s.z;
        Hx[t] = Hy[t] = Hz[t] = G[t] = 0.0;
    }
    __shared__ AdjRec lds[TILE];
    const int i_begin = chunk * plan.chunk;
    const int i_end = min(i_begin + plan.chunk, n_rec_pad);
    for (int it = i_begin; it < i_end; it += TILE) {
        __syncthreads();
        for (int k = threadIdx.x; k < TILE; k += kBlock) lds[k] = rec[it + k];
        __syncthreads();
        bool diag = false;
#pragma unroll
        for (int t = 0; t < T; ++t)
            diag |= diag_tile(tile_base + t * kBlock + wave * 64, self_end, it, TILE);
        for (int k = 0; k < TILE; k += R) {
            AdjRec a[R];
#pragma unroll
            for (int r = 0; r < R; ++r) a[r] = lds[k + r];
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int self = tile_base + t * kBlock + threadIdx.x;
                double d[R], rc[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    d[r] = __builtin_fma(nx[t], a[r].x, 1.0);
                    d[r] = __builtin_fma(ny[t], a[r].y, d[r]);
                    d[r] = __builtin_fma(nz[t], a[r].z, d[r]);
                    if (diag && it + k + r == self) d[r] = 1.0;
                }
                recips<R>(d, rc);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    double rr = rc[r];
                    if (diag && it + k + r == self) rr = 0.0;
                    const double p = __builtin_fma(zj[t], a[r].nz, __builtin_fma(yj[t], a[r].ny, xj[t] * a[r].nx));
                    const double pr = p * rr;
                    const double u = pr * rr;
                    G[t] += pr;
                    Hx[t] = __builtin_fma(rr, a[r].nx, __builtin_fma(u, a[r].x, Hx[t]));
                    Hy[t] = __builtin_fma(rr, a[r].ny, __builtin_fma(u, a[r].y, Hy[t]));
                    Hz[t] = __builtin_fma(rr, a[r].nz, __builtin_fma(u, a[r].z, Hz[t]));
                }
            }
        }
    }
    double *p = partial + (size_t)(chunk * 4) * stride;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x;
        p[i] = Hx[t];
        p[stride + i] = Hy[t];
        p[2 * (size_t)stride + i] = Hz[t];
        p[3 * (size_t)stride + i] = G[t];
    }
}

// ---------------------------------------------------------------------------
// Reverse RK4 stage bookkeeping
// ---------------------------------------------------------------------------

// Cotangent records of stage s: kbar = b lam + c Ybar_{s+1}, mu = h kbar,
// m = mu_x - 2 Omega mu_zeta e_z, n = m x x, from the stage's source records.
__global__ __launch_bounds__(256) void adj_prep_kernel(AdjStage a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n_pad) return;
    if (i >= a.n) {
        a.rec[i] = AdjRec{0, 0, 0, 0, 0, 0};
        return;
    }
    const int st = a.stride;
    const Source y = a.ys[i];
    // the last stage has no successor: its Ybar input is not read
    const bool next = a.c != 0.0;
    const double kx = a.b * a.lam[i] + (next ? a.c * a.ybar[i] : 0.0);
    const double ky = a.b * a.lam[st + i] + (next ? a.c * a.ybar[st + i] : 0.0);
    const double kz = a.b * a.lam[2 * st + i] + (next ? a.c * a.ybar[2 * st + i] : 0.0);
    const double kzeta = a.b * a.lam[3 * st + i] + (next ? a.c * a.ybar[3 * st + i] : 0.0);
    const double mx = a.h * kx, my = a.h * ky;
    const double mz = a.h * kz - 2.0 * a.omega * (a.h * kzeta);
    a.m[i] = mx;
    a.m[st + i] = my;
    a.m[2 * st + i] = mz;
    // n = m x x
    a.rec[i] = AdjRec{y.x, y.y, y.z, my * y.z - mz * y.y, mz * y.x - mx * y.z, mx * y.y - my * y.x};
}

// Ybar of stage s from the two passes' chunk partials; lam_acc += Ybar.
__global__ __launch_bounds__(256) void adj_combine_kernel(AdjStage a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int st = a.stride;
    double S[3] = {0, 0, 0}, Tv[3] = {0, 0, 0}, H[3] = {0, 0, 0}, g = 0.0;
    for (int c = 0; c < a.n_chunks_t; ++c) {
        const double *p = a.partial_t + (size_t)(c * 6) * st;
        for (int q = 0; q < 3; ++q) {
            S[q] += p[(size_t)q * st + i];
            Tv[q] += p[(size_t)(q + 3) * st + i];
        }
    }
    for (int c = 0; c < a.n_chunks_s; ++c) {
        const double *p = a.partial_s + (size_t)(c * 4) * st;
        for (int q = 0; q < 3; ++q) H[q] += p[(size_t)q * st + i];
        g += p[3 * (size_t)st + i];
    }
    const Source y = a.ys[i];
    const double mx = a.m[i], my = a.m[st + i], mz = a.m[2 * st + i];
    const double c4 = -4.0 * kPi;
    // G^T = S x m + T ;  G^S = w H
    const double gx = (S[1] * mz - S[2] * my) + Tv[0] + y.w * H[0];
    const double gy = (S[2] * mx - S[0] * mz) + Tv[1] + y.w * H[1];
    const double gz = (S[0] * my - S[1] * mx) + Tv[2] + y.w * H[2];
    const double bx = gx / c4, by = gy / c4, bz = gz / c4;
    const double bzeta = a.area[i] * g / c4;
    a.ybar[i] = bx;
    a.ybar[st + i] = by;
    a.ybar[2 * st + i] = bz;
    a.ybar[3 * st + i] = bzeta;
    a.lam_acc[i] += bx;
    a.lam_acc[st + i] += by;
    a.lam_acc[2 * st + i] += bz;
    a.lam_acc[3 * st + i] += bzeta;
}

template <int T, int TILE, int R>
static hipError_t launch_vjp3(const AdjStage &a, hipStream_t s)
{
    hipLaunchKernelGGL((vjp_target_kernel<T, TILE, R>), dim3(a.plan.grid), dim3(kBlock), 0, s,
                       a.rec, a.stride, a.n, a.ys, a.n_pad, a.plan, a.partial_t);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((vjp_source_kernel<T, TILE, R>), dim3(a.plan.grid), dim3(kBlock), 0, s,
                       a.ys, a.stride, a.n, a.rec, a.n_pad, a.plan, a.partial_s);
    return hipGetLastError();
}

template <int T>
static hipError_t launch_vjp_t(const AdjStage &a, hipStream_t s)
{
    if (a.plan.recip == 4) return launch_vjp3<T, 256, 4>(a, s);
    return launch_vjp3<T, 256, 1>(a, s);
}

hipError_t launch_adjoint_stage(const AdjStage &a, hipStream_t s)
{
    if (a.plan.tile != 256 || a.n_pad % 256 != 0 || a.stride < a.n_pad) return hipErrorInvalidValue;
    const int blocks = ceil_div(a.n_pad, 256);
    hipLaunchKernelGGL(adj_prep_kernel, dim3(blocks), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (a.plan.targets_per_lane) {
    case 1: e = launch_vjp_t<1>(a, s); break;
    case 2: e = launch_vjp_t<2>(a, s); break;
    case 4: e = launch_vjp_t<4>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(adj_combine_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace bve
