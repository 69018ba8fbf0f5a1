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
// Targets and sources are the same active panels, so one pass evaluates
// all four sums: panel k in a lane, the records {x_l, w_l, n_l} of every
// other panel streamed through LDS; per pair the denominator and its
// reciprocal are shared by the target-side terms (S, T: k as target, l as
// source) and the source-side ones (H, g: k as source, l as target).  The
// pass uses the forward kernel's tiling, chunking, XCD mapping and batched
// float64 reciprocals, and writes per-chunk partial sums that the combine
// kernel adds in a fixed order.
#include "bve_common.h"
#include "bve_device.h"
#include "bve_kernels.h"

namespace bve {

// 1/d_r for R denominators with the forward pass's reciprocals: R = 1 each
// refined twice (recip_f64), R = 4 one shared reciprocal of the product
// refined once (recip1_f64), as in tile_f64.
template <int R>
__device__ __forceinline__ void recips(const double (&d)[R], double (&r)[R])
{
    if constexpr (R == 1) {
        r[0] = recip_f64(d[0]);
    } else {
        static_assert(R == 4, "R must be 1 or 4");
        const double p01 = d[0] * d[1], p23 = d[2] * d[3];
        const double rr = recip1_f64(p01 * p23);
        const double r01 = rr * p23, r23 = rr * p01;
        r[0] = d[1] * r01;
        r[1] = d[0] * r01;
        r[2] = d[3] * r23;
        r[3] = d[2] * r23;
    }
}

// ---------------------------------------------------------------------------
// Pair pass: S, T (panel k as target) and H, g (panel k as source)
// ---------------------------------------------------------------------------
template <int T, int TILE, int R>
__global__ __launch_bounds__(kBlock) void vjp_kernel(const AdjRec *__restrict__ rec, int stride,
                                                     int self_end, int n_pad, PairPlan plan,
                                                     double *__restrict__ partial)
{
    const int work = xcd_work_item(blockIdx.x, plan.grid);
    if (work >= plan.n_work) return;
    const int chunk = work / plan.n_tiles;
    const int ttile = work - chunk * plan.n_tiles;
    const int tile_base = ttile * (kBlock * T);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);

    double xk[T], yk[T], zk[T], ex[T], ey[T], ez[T];
    double Sx[T], Sy[T], Sz[T], Tx[T], Ty[T], Tz[T], Hx[T], Hy[T], Hz[T], G[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const AdjRec a = rec[tile_base + t * kBlock + threadIdx.x];
        xk[t] = a.x;
        yk[t] = a.y;
        zk[t] = a.z;
        ex[t] = a.nx;
        ey[t] = a.ny;
        ez[t] = a.nz;
        Sx[t] = Sy[t] = Sz[t] = Tx[t] = Ty[t] = Tz[t] = 0.0;
        Hx[t] = Hy[t] = Hz[t] = G[t] = 0.0;
    }
    __shared__ AdjRec lds[TILE];
    const int j_begin = chunk * plan.chunk;
    const int j_end = min(j_begin + plan.chunk, n_pad);
    for (int jt = j_begin; jt < j_end; jt += TILE) {
        __syncthreads();
        for (int k = threadIdx.x; k < TILE; k += kBlock) lds[k] = rec[jt + k];
        __syncthreads();
        bool diag = false;
#pragma unroll
        for (int t = 0; t < T; ++t)
            diag |= diag_tile(tile_base + t * kBlock + wave * 64, self_end, jt, TILE);
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
                    d[r] = __builtin_fma(-xk[t], a[r].x, 1.0);
                    d[r] = __builtin_fma(-yk[t], a[r].y, d[r]);
                    d[r] = __builtin_fma(-zk[t], a[r].z, d[r]);
                    if (diag && jt + k + r == self) d[r] = 1.0;
                }
                recips<R>(d, rc);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    double rr = rc[r];
                    if (diag && jt + k + r == self) rr = 0.0;
                    // k as target, l as source: q = w_l / d, P_kl = x_l . n_k
                    const double q = a[r].w * rr;
                    const double pt = __builtin_fma(a[r].z, ez[t], __builtin_fma(a[r].y, ey[t], a[r].x * ex[t]));
                    const double tq = q * rr * pt;
                    Sx[t] = __builtin_fma(q, a[r].x, Sx[t]);
                    Sy[t] = __builtin_fma(q, a[r].y, Sy[t]);
                    Sz[t] = __builtin_fma(q, a[r].z, Sz[t]);
                    Tx[t] = __builtin_fma(tq, a[r].x, Tx[t]);
                    Ty[t] = __builtin_fma(tq, a[r].y, Ty[t]);
                    Tz[t] = __builtin_fma(tq, a[r].z, Tz[t]);
                    // k as source, l as target: P_lk = x_k . n_l
                    const double ps = __builtin_fma(zk[t], a[r].nz, __builtin_fma(yk[t], a[r].ny, xk[t] * a[r].nx));
                    const double pr = ps * rr;
                    const double u = pr * rr;
                    G[t] += pr;
                    Hx[t] = __builtin_fma(rr, a[r].nx, __builtin_fma(u, a[r].x, Hx[t]));
                    Hy[t] = __builtin_fma(rr, a[r].ny, __builtin_fma(u, a[r].y, Hy[t]));
                    Hz[t] = __builtin_fma(rr, a[r].nz, __builtin_fma(u, a[r].z, Hz[t]));
                }
            }
        }
    }
    double *p = partial + (size_t)(chunk * kAdjRows) * stride;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x;
        const double v[kAdjRows] = {Sx[t], Sy[t], Sz[t], Tx[t], Ty[t], Tz[t], Hx[t], Hy[t], Hz[t], G[t]};
#pragma unroll
        for (int q = 0; q < kAdjRows; ++q) p[(size_t)q * stride + i] = v[q];
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
        a.rec[i] = AdjRec{0, 0, 0, 0, 0, 0, 0, 0};
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
    a.rec[i] = AdjRec{y.x, y.y, y.z, y.w, my * y.z - mz * y.y, mz * y.x - mx * y.z,
                      mx * y.y - my * y.x, 0.0};
}

// Ybar of stage s from the pair pass's chunk partials; lam_acc += Ybar.
__global__ __launch_bounds__(256) void adj_combine_kernel(AdjStage a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int st = a.stride;
    double S[3] = {0, 0, 0}, Tv[3] = {0, 0, 0}, H[3] = {0, 0, 0}, g = 0.0;
    for (int c = 0; c < a.n_chunks; ++c) {
        const double *p = a.partial + (size_t)(c * kAdjRows) * st;
        for (int q = 0; q < 3; ++q) {
            S[q] += p[(size_t)q * st + i];
            Tv[q] += p[(size_t)(q + 3) * st + i];
            H[q] += p[(size_t)(q + 6) * st + i];
        }
        g += p[9 * (size_t)st + i];
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
    hipLaunchKernelGGL((vjp_kernel<T, TILE, R>), dim3(a.plan.grid), dim3(kBlock), 0, s, a.rec,
                       a.stride, a.n, a.n_pad, a.plan, a.partial);
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

// ---------------------------------------------------------------------------
// Exact differences of the forward map
// ---------------------------------------------------------------------------
//
// For the trajectory of (x, zeta) and that of (x', zeta') = (x + dx, zeta +
// dzeta) the velocity sums differ by (exact algebra, no linearisation)
//
//   d'_ik - d_ik = Dd = -(dx_i . x'_k + x_i . dx_k),   d = 1 - x_i . x_k
//   w'/d' - w/d  = Dq = (Dw - q Dd) / d',               q = w / d
//   S'_i - S_i   = DS = sum_k (Dq x'_k + q dx_k)
//   U'_i - U_i   = (dx_i x S'_i + x_i x DS_i) / (-4 PI)
//
// Every quantity is formed from differences, never as the difference of two
// nearly equal numbers, so the propagated difference carries rounding
// relative to itself: (J(zeta + e) - J(zeta - e)) / 2e from it is free of the
// cancellation a plain finite difference suffers at small e.

// One (target tile, source chunk) item: base sums S and their differences.
__global__ __launch_bounds__(kBlock) void diff_pair_kernel(DiffState st, PairPlan plan)
{
    const int work = xcd_work_item(blockIdx.x, plan.grid);
    if (work >= plan.n_work) return;
    const int chunk = work / plan.n_tiles;
    const int i = (work - chunk * plan.n_tiles) * kBlock + threadIdx.x;
    __shared__ DiffRec lds[kBlock];
    const DiffRec t = st.rec[i];   // records are zero beyond n, up to the row stride
    double S0 = 0.0, S1 = 0.0, S2 = 0.0, D0 = 0.0, D1 = 0.0, D2 = 0.0;
    const int k0 = chunk * plan.chunk, k1 = min(k0 + plan.chunk, st.n);
    for (int jt = k0; jt < k1; jt += kBlock) {
        __syncthreads();
        if (jt + (int)threadIdx.x < k1) lds[threadIdx.x] = st.rec[jt + threadIdx.x];
        __syncthreads();
        const int m = min(kBlock, k1 - jt);
        for (int kk = 0; kk < m; ++kk) {
            if (jt + kk == i) continue;   // self-exclusion (BVEActiveRHS)
            const DiffRec q = lds[kk];
            const double px = q.x + q.dx, py = q.y + q.dy, pz = q.z + q.dz;
            const double d = 1.0 - (t.x * q.x + t.y * q.y + t.z * q.z);
            const double dd = -((t.dx * px + t.dy * py + t.dz * pz) +
                                (t.x * q.dx + t.y * q.dy + t.z * q.dz));
            const double r = q.w / d;
            const double dr = (q.dw - r * dd) / (d + dd);
            S0 += r * q.x;
            S1 += r * q.y;
            S2 += r * q.z;
            D0 += dr * px + r * q.dx;
            D1 += dr * py + r * q.dy;
            D2 += dr * pz + r * q.dz;
        }
    }
    if (i < st.n) {
        double *p = st.partial + (size_t)6 * chunk * st.stride + i;
        p[0] = S0;
        p[st.stride] = S1;
        p[2 * (size_t)st.stride] = S2;
        p[3 * (size_t)st.stride] = D0;
        p[4 * (size_t)st.stride] = D1;
        p[5 * (size_t)st.stride] = D2;
    }
}

// Stage-1 records from the step-start state.
__global__ void diff_prepare_kernel(DiffState st)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= st.n) return;
    const size_t R = st.stride;
    st.rec[i] = DiffRec{st.x0[i], st.x0[R + i], st.x0[2 * R + i], st.dx0[i], st.dx0[R + i],
                        st.dx0[2 * R + i], st.z0[i] * st.area[i], st.dz0[i] * st.area[i]};
}

// RK4 stage of both trajectories (the same steps as the forward stage body:
// k = dt U, dzeta/dt = -2 Omega U_z, inputs x + a k, update x + sum k_s / c_s).
__global__ void diff_stage_kernel(DiffState st, PairPlan plan, int stage, double dt,
                                  double omega, double a_next, double divisor)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= st.n) return;
    const size_t R = st.stride;
    double S[3] = {0.0, 0.0, 0.0}, D[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < plan.n_chunks; ++c) {
        const double *p = st.partial + (size_t)6 * c * R + i;
        for (int r = 0; r < 3; ++r) {
            S[r] += p[r * R];
            D[r] += p[(3 + r) * R];
        }
    }
    const DiffRec t = st.rec[i];
    const double c4 = -4.0 * kPi;
    const double Sp[3] = {S[0] + D[0], S[1] + D[1], S[2] + D[2]};
    const double u[3] = {(t.y * S[2] - t.z * S[1]) / c4, (t.z * S[0] - t.x * S[2]) / c4,
                         (t.x * S[1] - t.y * S[0]) / c4};
    const double du[3] = {((t.dy * Sp[2] - t.dz * Sp[1]) + (t.y * D[2] - t.z * D[1])) / c4,
                          ((t.dz * Sp[0] - t.dx * Sp[2]) + (t.z * D[0] - t.x * D[2])) / c4,
                          ((t.dx * Sp[1] - t.dy * Sp[0]) + (t.x * D[1] - t.y * D[0])) / c4};
    double nb[4], nd[4];
    for (int r = 0; r < 4; ++r) {
        const double k = r < 3 ? dt * u[r] : dt * (-(2.0 * omega * u[2]));
        const double dk = r < 3 ? dt * du[r] : dt * (-(2.0 * omega * du[2]));
        double *acc = r < 3 ? st.acc + r * R : st.zacc;
        double *dacc = r < 3 ? st.dacc + r * R : st.dzacc;
        const double *b0 = r < 3 ? st.x0 + r * R : st.z0;
        const double *d0 = r < 3 ? st.dx0 + r * R : st.dz0;
        const double a = (stage == 0 ? b0[i] : acc[i]) + k / divisor;
        const double da = (stage == 0 ? d0[i] : dacc[i]) + dk / divisor;
        acc[i] = a;
        dacc[i] = da;
        nb[r] = stage < 3 ? b0[i] + a_next * k : a;
        nd[r] = stage < 3 ? d0[i] + a_next * dk : da;
    }
    if (stage == 3) {
        for (int r = 0; r < 3; ++r) {
            st.x0[r * R + i] = nb[r];
            st.dx0[r * R + i] = nd[r];
        }
        st.z0[i] = nb[3];
        st.dz0[i] = nd[3];
    }
    st.rec[i] = DiffRec{nb[0], nb[1], nb[2], nd[0], nd[1], nd[2], nb[3] * st.area[i],
                        nd[3] * st.area[i]};
}

hipError_t launch_diff_prepare(const DiffState &st, hipStream_t s)
{
    hipLaunchKernelGGL(diff_prepare_kernel, dim3(ceil_div(st.n, 256)), dim3(256), 0, s, st);
    return hipGetLastError();
}

hipError_t launch_diff_pair(const DiffState &st, const PairPlan &plan, hipStream_t s)
{
    if (plan.targets_per_lane != 1 || plan.n_tiles * kBlock > st.stride || plan.grid % 8 != 0)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(diff_pair_kernel, dim3(plan.grid), dim3(kBlock), 0, s, st, plan);
    return hipGetLastError();
}

hipError_t launch_diff_stage(const DiffState &st, const PairPlan &plan, int stage, double dt,
                             double omega, hipStream_t s)
{
    static const double a_next[4] = {0.5, 0.5, 1.0, 0.0};
    static const double divisor[4] = {6.0, 3.0, 3.0, 6.0};
    hipLaunchKernelGGL(diff_stage_kernel, dim3(ceil_div(st.n, 256)), dim3(256), 0, s, st, plan,
                       stage, dt, omega, a_next[stage], divisor[stage]);
    return hipGetLastError();
}

}  // namespace bve
