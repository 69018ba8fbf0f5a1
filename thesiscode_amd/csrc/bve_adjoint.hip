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
// the pair sums: panel k in a lane, the records {x_l, w_l, n_l} of every
// other panel streamed through LDS; per pair the denominator and its
// reciprocal are shared by the target-side terms (T: k as target, l as
// source) and the source-side ones (H: k as source, l as target).  The
// pass runs the forward kernel's segment queue, tiling and batched float64
// reciprocals, and writes per-segment partial sums that the combine kernel
// adds in a fixed order; S comes from the forward sweep.
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
// Pair pass: V and N of panel k (as target and as source)
// ---------------------------------------------------------------------------
//
// With the target-side T_k = sum_l (w_l P_kl / d^2) x_l and the source-side
// H_k = sum_l [n_l / d + P_lk x_l / d^2] both accumulating multiples of x_l,
// the gradient needs only T_k + w_k H_k = V_k + w_k N_k with
//
//     V_k = sum_l (w_l P_kl + w_k P_lk) / d^2  x_l,    N_k = sum_l n_l / d,
//
// (P_kl = x_l . n_k, P_lk = x_k . n_l, d = 1 - x_k . x_l).  The two other
// sums of the gradient come without a pass of their own: S_k is the forward
// sum of the same stage input, kept by the forward sweep, and
// g_k = sum_l P_lk / d = x_k . N_k.  With the records carrying w_l x_l and
// each lane w_k x_k, w_l P_kl + w_k P_lk = (w_l x_l).n_k + (w_k x_k).n_l is one
// chain of six FMAs, and a pair costs about 20 float64 operations with four
// pairs sharing a reciprocal.

// Source tile at jt (in LDS) against the T targets of this lane; MASK: the
// tile holds the own index of some lane's target (self pairs are dropped).
template <int T, int R, bool MASK>
__device__ __forceinline__ void vjp_tile(const AdjRec *lds, int jt, const int (&self)[T],
                                         const double (&xk)[T], const double (&yk)[T],
                                         const double (&zk)[T], const double (&ax)[T],
                                         const double (&ay)[T], const double (&az)[T],
                                         const double (&ex)[T], const double (&ey)[T],
                                         const double (&ez)[T], double (&Vx)[T], double (&Vy)[T],
                                         double (&Vz)[T], double (&Nx)[T], double (&Ny)[T],
                                         double (&Nz)[T])
{
    constexpr int TILE = 256;
    for (int k = 0; k < TILE; k += R) {
        // positions first (denominators and reciprocals of every target),
        // the rest of each record where its pairs are summed
        double px[R], py[R], pz[R], rc[T][R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            px[r] = lds[k + r].x;
            py[r] = lds[k + r].y;
            pz[r] = lds[k + r].z;
        }
#pragma unroll
        for (int t = 0; t < T; ++t) {
            double d[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                d[r] = __builtin_fma(-xk[t], px[r], 1.0);
                d[r] = __builtin_fma(-yk[t], py[r], d[r]);
                d[r] = __builtin_fma(-zk[t], pz[r], d[r]);
                if (MASK && jt + k + r == self[t]) d[r] = 1.0;
            }
            recips<R>(d, rc[t]);
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (MASK && jt + k + r == self[t]) rc[t][r] = 0.0;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const AdjRec &a = lds[k + r];
            const double wx = a.wx, wy = a.wy, wz = a.wz, nx = a.nx, ny = a.ny, nz = a.nz;
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const double rr = rc[t][r];
                // (w_l x_l).n_k + (w_k x_k).n_l
                double sp = az[t] * nz;
                sp = __builtin_fma(ay[t], ny, sp);
                sp = __builtin_fma(ax[t], nx, sp);
                sp = __builtin_fma(wz, ez[t], sp);
                sp = __builtin_fma(wy, ey[t], sp);
                sp = __builtin_fma(wx, ex[t], sp);
                const double gam = (rr * rr) * sp;
                Vx[t] = __builtin_fma(gam, px[r], Vx[t]);
                Vy[t] = __builtin_fma(gam, py[r], Vy[t]);
                Vz[t] = __builtin_fma(gam, pz[r], Vz[t]);
                Nx[t] = __builtin_fma(rr, nx, Nx[t]);
                Ny[t] = __builtin_fma(rr, ny, Ny[t]);
                Nz[t] = __builtin_fma(rr, nz, Nz[t]);
            }
        }
    }
}

// Target tile `ttile` against source tiles [j_begin, j_end) (multiples of
// 256): rows q of target i go to out[q * out_stride + i - out_shift].
template <int T, int R>
__device__ __forceinline__ void vjp_range(const AdjRec *__restrict__ rec, int self_end, int ttile,
                                          int j_begin, int j_end, double *out, int out_stride,
                                          int out_shift, AdjRec *lds)
{
    constexpr int TILE = 256;
    const int tile_base = ttile * (kBlock * T);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    double xk[T], yk[T], zk[T], ax[T], ay[T], az[T], ex[T], ey[T], ez[T];
    double Vx[T], Vy[T], Vz[T], Nx[T], Ny[T], Nz[T];
    int self[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x;
        const AdjRec a = rec[i];
        xk[t] = a.x;
        yk[t] = a.y;
        zk[t] = a.z;
        ax[t] = a.wx;
        ay[t] = a.wy;
        az[t] = a.wz;
        ex[t] = a.nx;
        ey[t] = a.ny;
        ez[t] = a.nz;
        self[t] = i < self_end ? i : -1;
        Vx[t] = Vy[t] = Vz[t] = Nx[t] = Ny[t] = Nz[t] = 0.0;
    }
    for (int jt = j_begin; jt < j_end; jt += TILE) {
        __syncthreads();
        lds[threadIdx.x] = rec[jt + threadIdx.x];
        __syncthreads();
        bool diag = false;
#pragma unroll
        for (int t = 0; t < T; ++t)
            diag |= diag_tile(tile_base + t * kBlock + wave * 64, self_end, jt, TILE);
        if (diag)
            vjp_tile<T, R, true>(lds, jt, self, xk, yk, zk, ax, ay, az, ex, ey, ez, Vx, Vy, Vz, Nx, Ny,
                                 Nz);
        else
            vjp_tile<T, R, false>(lds, jt, self, xk, yk, zk, ax, ay, az, ex, ey, ez, Vx, Vy, Vz, Nx,
                                  Ny, Nz);
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x - out_shift;
        const double v[kAdjRows] = {Vx[t], Vy[t], Vz[t], Nx[t], Ny[t], Nz[t]};
#pragma unroll
        for (int q = 0; q < kAdjRows; ++q) out[(size_t)q * out_stride + i] = v[q];
    }
}

// Segmented persistent pass (the forward pass's queue and row layout:
// segment s writes rows (s + tile) of kAdjRows x kBlock T entries).  The
// register cap keeps 3 waves per SIMD at T = 2 (168 VGPRs, no spills) and 2
// at T = 4.
template <int T, int R>
__global__ __launch_bounds__(kBlock, T == 4 ? 2 : 3) void vjp_kernel(
    const AdjRec *__restrict__ rec, int self_end, PairPlan plan, double *__restrict__ partial,
    unsigned *queue)
{
    constexpr int TT = kBlock * T;
    __shared__ AdjRec lds[256];
    __shared__ int seg;
    run_segments<256>(plan, queue, &seg, [&](int t, int j0, int j1, int s) {
        vjp_range<T, R>(rec, self_end, t, j0, j1, partial + (size_t)(s + t) * kAdjRows * TT, TT,
                        t * TT, lds);
    });
}

// ---------------------------------------------------------------------------
// Mixed pair pass: float32 far field, float64 near field
// ---------------------------------------------------------------------------
//
// A gradient needs far less than float64 accuracy of its pair sums, and the
// forward trajectory it differentiates stays float64.  A (target tile,
// 256-source tile) pair runs the float64 tile above when the source tile
// holds one of the targets' own indices or when its bounding sphere comes
// close to that of a 256-record part of the target tile (r_g + r_s >=
// ratio * |c_g - c_s|); every other pair runs in float32 with the
// tile-relative denominator of the mixed forward pass,
//
//     d = (A_k + B_l) - a_k . o_l,   a_k = x_k - c_s,  o_l = x_l - c_s,
//     A_k = |a_k|^2 / 2 + c_k,       B_l = |o_l|^2 / 2 + c_l,  c = (1 - |x|^2) / 2,
//
// which is 1 - x_k . x_l exactly and, for separated spheres, rounds to
// within about ((1 + ratio) / (1 - ratio))^2 float32 roundings of d
// (|a| + |o| <= (1 + ratio) |c_g - c_s| against |a - o| >= (1 - ratio)
// |c_g - c_s|).  Per pair: an add and three FMAs for d, one v_rcp_f32, six
// FMAs for w_l P_kl + w_k P_lk, two products for gamma and six FMAs into V
// and N; the float32 tile sums are carried into float64 after each tile.

// The float64 tiles of the mixed pass share one reciprocal among four pairs,
// as the float64 pass does.
constexpr int kMixedR = 4;

union AdjMixedLds {
    AdjRec32 r32[256];
    AdjRec r64[256];
};

// Float32 state of one target against one source tile: -a_k, A_k, w_k x_k, n_k.
struct AdjTgt32 {
    float nax, nay, naz, A, wx, wy, wz, nx, ny, nz;
};

struct AdjAcc32 {
    float Vx, Vy, Vz, Nx, Ny, Nz;
};

// Float32 sums of the 256-record tile in LDS for N targets.
template <int N>
__device__ __forceinline__ void vjp_tile32(const AdjRec32 *lds, const AdjTgt32 (&g)[N],
                                           AdjAcc32 (&acc)[N])
{
#pragma unroll 2
    for (int k = 0; k < 256; ++k) {
        const float4 o = lds[k].o, p = lds[k].x, w = lds[k].w, n = lds[k].n;
#pragma unroll
        for (int t = 0; t < N; ++t) {
            const float d = __builtin_fmaf(
                g[t].nax, o.x, __builtin_fmaf(g[t].nay, o.y, __builtin_fmaf(g[t].naz, o.z, g[t].A + o.w)));
            const float r = __builtin_amdgcn_rcpf(d);
            // (w_l x_l).n_k + (w_k x_k).n_l
            float sp = g[t].wz * n.z;
            sp = __builtin_fmaf(g[t].wy, n.y, sp);
            sp = __builtin_fmaf(g[t].wx, n.x, sp);
            sp = __builtin_fmaf(w.z, g[t].nz, sp);
            sp = __builtin_fmaf(w.y, g[t].ny, sp);
            sp = __builtin_fmaf(w.x, g[t].nx, sp);
            const float gam = (r * r) * sp;
            acc[t].Vx = __builtin_fmaf(gam, p.x, acc[t].Vx);
            acc[t].Vy = __builtin_fmaf(gam, p.y, acc[t].Vy);
            acc[t].Vz = __builtin_fmaf(gam, p.z, acc[t].Vz);
            acc[t].Nx = __builtin_fmaf(r, n.x, acc[t].Nx);
            acc[t].Ny = __builtin_fmaf(r, n.y, acc[t].Ny);
            acc[t].Nz = __builtin_fmaf(r, n.z, acc[t].Nz);
        }
    }
}

// Float32 state of a target at x (float64) with float32 w x and n against
// the tile with centre sb.
__device__ __forceinline__ AdjTgt32 adj_tgt32(double x, double y, double z, const float (&wn)[6],
                                              const double4 &sb)
{
    const float ax = (float)(x - sb.x), ay = (float)(y - sb.y), az = (float)(z - sb.z);
    const double ck = 0.5 * (1.0 - (x * x + y * y + z * z));
    const double a2 = (double)ax * ax + (double)ay * ay + (double)az * az;
    return AdjTgt32{-ax, -ay, -az, (float)(0.5 * a2 + ck), wn[0], wn[1], wn[2], wn[3], wn[4], wn[5]};
}

// Target tile `ttile` against source tiles [j_begin, j_end) of a mixed pass
// (rows as vjp_range).  The route is taken per workgroup, from the spheres
// of the T record tiles its targets fill: a source tile runs in float64
// when it holds one of the targets' own indices or lies close to one of
// those spheres, so that the four waves stay in step between the tile
// barriers.  Kept across tiles: the float64 positions and sums and the
// float32 w x and n of each target; a float64 tile re-reads the targets'
// float64 records.  (Routes per wave, with per-wave 64- or 128-target
// spheres, sent fewer pairs to float64 but held three waves at each tile
// barrier behind a fourth on a float64 tile: 400-425 against 336 us per
// level-6 pass.)
template <int T>
__device__ __forceinline__ void vjp_mixed_range(const AdjStage &a,
                                                const AdjRec *__restrict__ rec, int ttile,
                                                int j_begin, int j_end, double *out,
                                                int out_stride, int out_shift, AdjMixedLds &lds)
{
    constexpr int TILE = 256;
    const int tile_base = ttile * (kBlock * T);
    double xk[T], yk[T], zk[T];
    double Vx[T], Vy[T], Vz[T], Nx[T], Ny[T], Nz[T];
    float wn[T][6];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const AdjRec r = rec[tile_base + t * kBlock + threadIdx.x];
        xk[t] = r.x;
        yk[t] = r.y;
        zk[t] = r.z;
        wn[t][0] = (float)r.wx;
        wn[t][1] = (float)r.wy;
        wn[t][2] = (float)r.wz;
        wn[t][3] = (float)r.nx;
        wn[t][4] = (float)r.ny;
        wn[t][5] = (float)r.nz;
        Vx[t] = Vy[t] = Vz[t] = Nx[t] = Ny[t] = Nz[t] = 0.0;
    }
    const int self_hi = min(tile_base + kBlock * T, a.n);
    for (int jt = j_begin; jt < j_end; jt += TILE) {
        const double4 sb = a.tile_box[jt / TILE];
        // workgroup-uniform
        const bool diag = tile_base < self_hi && tile_base < jt + TILE && jt < self_hi;
        bool use64 = diag;
#pragma unroll
        for (int u = 0; u < T; ++u) {
            const double4 gb = a.tile_box[tile_base / TILE + u];
            const double dx = gb.x - sb.x, dy = gb.y - sb.y, dz = gb.z - sb.z;
            use64 |= !(gb.w + sb.w < a.ratio * sqrt(dx * dx + dy * dy + dz * dz));
        }
        __syncthreads();
        if (use64)
            lds.r64[threadIdx.x] = rec[jt + threadIdx.x];
        else
            lds.r32[threadIdx.x] = a.rec32[jt + threadIdx.x];
        __syncthreads();
        if (!use64) {
            AdjTgt32 g[T];
            AdjAcc32 acc[T];
#pragma unroll
            for (int t = 0; t < T; ++t) {
                g[t] = adj_tgt32(xk[t], yk[t], zk[t], wn[t], sb);
                acc[t] = AdjAcc32{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            }
            vjp_tile32<T>(lds.r32, g, acc);
#pragma unroll
            for (int t = 0; t < T; ++t) {
                Vx[t] += (double)acc[t].Vx;
                Vy[t] += (double)acc[t].Vy;
                Vz[t] += (double)acc[t].Vz;
                Nx[t] += (double)acc[t].Nx;
                Ny[t] += (double)acc[t].Ny;
                Nz[t] += (double)acc[t].Nz;
            }
            continue;
        }
        double ax[T], ay[T], az[T], ex[T], ey[T], ez[T];
        int self[T];
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int i = tile_base + t * kBlock + threadIdx.x;
            const AdjRec r = rec[i];
            ax[t] = r.wx;
            ay[t] = r.wy;
            az[t] = r.wz;
            ex[t] = r.nx;
            ey[t] = r.ny;
            ez[t] = r.nz;
            self[t] = i < a.n ? i : -1;
        }
        if (diag)
            vjp_tile<T, kMixedR, true>(lds.r64, jt, self, xk, yk, zk, ax, ay, az, ex, ey, ez, Vx,
                                       Vy, Vz, Nx, Ny, Nz);
        else
            vjp_tile<T, kMixedR, false>(lds.r64, jt, self, xk, yk, zk, ax, ay, az, ex, ey, ez, Vx,
                                        Vy, Vz, Nx, Ny, Nz);
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = tile_base + t * kBlock + threadIdx.x - out_shift;
        const double v[kAdjRows] = {Vx[t], Vy[t], Vz[t], Nx[t], Ny[t], Nz[t]};
#pragma unroll
        for (int q = 0; q < kAdjRows; ++q) out[(size_t)q * out_stride + i] = v[q];
    }
}

template <int T>
__global__ __launch_bounds__(kBlock, T == 4 ? 2 : 3) void vjp_mixed_kernel(
    AdjStage a, const AdjRec *__restrict__ rec, double *__restrict__ partial)
{
    constexpr int TT = kBlock * T;
    __shared__ AdjMixedLds lds;
    __shared__ int seg;
    run_segments<256>(a.plan, a.queue, &seg, [&](int t, int j0, int j1, int s) {
        vjp_mixed_range<T>(a, rec, t, j0, j1, partial + (size_t)(s + t) * kAdjRows * TT, TT,
                           t * TT, lds);
    });
}

template <int T>
static int vjp_mixed_blocks()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, vjp_mixed_kernel<T>, kBlock, 0) !=
        hipSuccess)
        return 0;
    return n;
}

template <int T, int R>
static int vjp_blocks()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, vjp_kernel<T, R>, kBlock, 0) != hipSuccess)
        return 0;
    return n;
}

int vjp_blocks_per_cu(int T, int recip, bool mixed)
{
    if (mixed) return T == 1 ? vjp_mixed_blocks<1>() : (T == 2 ? vjp_mixed_blocks<2>() : vjp_mixed_blocks<4>());
    if (recip == 4) return T == 1 ? vjp_blocks<1, 4>() : (T == 2 ? vjp_blocks<2, 4>() : vjp_blocks<4, 4>());
    return T == 1 ? vjp_blocks<1, 1>() : (T == 2 ? vjp_blocks<2, 1>() : vjp_blocks<4, 1>());
}

// ---------------------------------------------------------------------------
// Reverse RK4 stage bookkeeping
// ---------------------------------------------------------------------------

// Cotangent records of stage s: kbar = b lam + c Ybar_{s+1}, mu = h kbar,
// m = mu_x - 2 Omega mu_zeta e_z, n = m x x, from the stage's source records
// (zero beyond n).  For a mixed pair pass also each 256-record tile's sphere
// (centre the mean of its finite records below n) and the float32 records
// relative to the tile centre.
__global__ __launch_bounds__(256) void adj_prep_kernel(AdjStage a)
{
    __shared__ double sh[256];
    const int i = blockIdx.x * 256 + threadIdx.x;   // n_pad is a multiple of 256
    AdjRec r{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < a.n) {
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
        r = AdjRec{y.x, y.y, y.z, 0.0, y.w * y.x, y.w * y.y, y.w * y.z, my * y.z - mz * y.y,
                   mz * y.x - mx * y.z, mx * y.y - my * y.x};
    }
    a.rec[i] = r;
    if (!a.rec32) return;   // uniform over the launch
    const bool valid = i < a.n && finite3(r.x, r.y, r.z);
    double c[4] = {valid ? 1.0 : 0.0, valid ? r.x : 0.0, valid ? r.y : 0.0, valid ? r.z : 0.0};
    block_sums<4>(c, sh);
    const double inv = c[0] > 0 ? 1.0 / c[0] : 0.0;
    const double cx = c[1] * inv, cy = c[2] * inv, cz = c[3] * inv;
    const double ox = r.x - cx, oy = r.y - cy, oz = r.z - cz;
    const double rad = block_max(valid ? sqrt(ox * ox + oy * oy + oz * oz) : -1.0, sh);
    if (threadIdx.x == 0) a.tile_box[blockIdx.x] = make_double4(cx, cy, cz, rad);
    const float fx = (float)ox, fy = (float)oy, fz = (float)oz;
    const double cl = 0.5 * (1.0 - (r.x * r.x + r.y * r.y + r.z * r.z));
    const double o2 = (double)fx * fx + (double)fy * fy + (double)fz * fz;
    a.rec32[i] = AdjRec32{make_float4(fx, fy, fz, (float)(0.5 * o2 + cl)),
                          make_float4((float)r.x, (float)r.y, (float)r.z, 0.f),
                          make_float4((float)r.wx, (float)r.wy, (float)r.wz, 0.f),
                          make_float4((float)r.nx, (float)r.ny, (float)r.nz, 0.f)};
}

// Ybar of stage s from the pair pass's segment rows and the forward sums S
// of the stage input; lam_acc += Ybar.
__global__ __launch_bounds__(256) void adj_combine_kernel(AdjStage a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int st = a.stride;
    double v[kAdjRows];
#pragma unroll
    for (int q = 0; q < kAdjRows; ++q) v[q] = 0.0;
    // the segments covering i's target tile, in order
    const PairPlan &p = a.plan;
    const int tt = kBlock * p.targets_per_lane;
    const int t = i / tt, loc = i - t * tt;
    const int first = t * p.n_src_tiles;
    const int s_hi = (first + p.n_src_tiles - 1) / p.seg_len;
    for (int s = first / p.seg_len; s <= s_hi; ++s) {
        const double *row = a.partial + (size_t)(s + t) * kAdjRows * tt + loc;
#pragma unroll
        for (int q = 0; q < kAdjRows; ++q) v[q] += row[(size_t)q * tt];
    }
    const Source y = a.ys[i];
    const double Sx = a.sums[i], Sy = a.sums[st + i], Sz = a.sums[2 * st + i];
    const double mx = a.m[i], my = a.m[st + i], mz = a.m[2 * st + i];
    const double c4 = -4.0 * kPi;
    // dL/dx = (S x m + V + w N) / (-4 PI),  dL/dzeta = A (x . N) / (-4 PI)
    const double gx = (Sy * mz - Sz * my) + v[0] + y.w * v[3];
    const double gy = (Sz * mx - Sx * mz) + v[1] + y.w * v[4];
    const double gz = (Sx * my - Sy * mx) + v[2] + y.w * v[5];
    const double g = (y.x * v[3] + y.y * v[4]) + y.z * v[5];
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

template <int T, int R>
static hipError_t launch_vjp2(const AdjStage &a, hipStream_t s)
{
    hipLaunchKernelGGL((vjp_kernel<T, R>), dim3(a.plan.grid), dim3(kBlock), 0, s, a.rec, a.n,
                       a.plan, a.partial, a.queue);
    return hipGetLastError();
}

template <int T>
static hipError_t launch_vjp_t(const AdjStage &a, hipStream_t s)
{
    if (a.rec32) {
        hipLaunchKernelGGL((vjp_mixed_kernel<T>), dim3(a.plan.grid), dim3(kBlock), 0, s, a,
                           (const AdjRec *)a.rec, a.partial);
        return hipGetLastError();
    }
    if (a.plan.recip == 4) return launch_vjp2<T, 4>(a, s);
    return launch_vjp2<T, 1>(a, s);
}

hipError_t launch_adjoint_stage(const AdjStage &a, hipStream_t s)
{
    if (a.plan.tile != 256 || a.n_pad % 256 != 0 || a.stride < a.n_pad || a.plan.segments <= 0 ||
        !a.queue || !a.sums || a.plan.n_tiles * kBlock * a.plan.targets_per_lane > a.n_pad)
        return hipErrorInvalidValue;
    if (a.rec32 && (!a.tile_box || !(a.ratio > 0.0 && a.ratio < 1.0)))
        return hipErrorInvalidValue;
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
