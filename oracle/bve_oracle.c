/*
 * CPU oracle for the Biot-Savart / RK4 path of the barotropic-vorticity
 * particle-panel solver.  TEST INFRASTRUCTURE ONLY: the test suite, the
 * smoke check and the CPU-baseline leg of the benchmark call it as the
 * checker.  The product library never links or calls it.
 *
 * Each function restates one reference routine, in the reference's loop
 * order and rounding points, so that compiled with -O2 -ffp-contract=off it
 * reproduces the reference's float64 results (the golden vectors in
 * tests/golden/ were written by the reference compiled with flang -O2 on
 * x86-64, which has no FMA contraction).
 *
 * Layout: positions are AoS x[3*i+c], the column-major x(3,N) of the
 * reference.  Indices are 0-based; ranges are half-open [i0, i1).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   bvo_active_rhs, bvo_passive_rhs, bvo_smooth_velocity, bvo_rk4_step:
 *       golden vectors from the compiled reference (bitwise or <= 1e-15),
 *       the level-0 cube by hand, tangency u.x = 0, active-block momentum
 *       sum_i w_i u_i = 0, and a __float128 brute force (bvo_velocity_q).
 *   bvo_voronoi_velocity: equals bvo_active_rhs on SoA input (same sum,
 *       the reference's sequential-subtraction denominator), tangency.
 *   bvo_voronoi_rk4_step: golden vectors from the compiled reference's
 *       Voronoi BVERK4 on icosahedral Voronoi grids (bitwise), |x| = 1.
 *   bvo_rk4_step: RK4 polynomial on a solid-body flow, zeta + 2 Omega z
 *       invariant, golden vectors over 12 steps.
 *   bvo_renormalize: |x| = 1 within 1 ulp.
 *   bvo_velocity_q: the level-0 cube in closed form, and bvo_active_rhs.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* PI of NumberKinds.f90:12, rounded to float64 */
static const double BVO_PI = 3.141592653589793238462643383279502797479068098137295573004504331874296718662975536062731407582759857177734375;

/* BVEActiveRHS, PanelVelocityParallel.f90:1287-1334.
 * Targets j in [i0, i1) of the n active panels, sources all active panels
 * except k == j (the loop split at :1319 / :1325).  dX is zeroed for all n
 * (the reference zeroes the whole array, :1316-1317), then divided by
 * (-4 PI) and dVort = -2 Omega dX(3,:) (:1332-1333). */
void bvo_active_rhs(int n, const double *x, const double *vort, const double *area,
                    int i0, int i1, double omega, double *dx, double *dvort)
{
    memset(dx, 0, sizeof(double) * 3 * (size_t)n);
    memset(dvort, 0, sizeof(double) * (size_t)n);
#pragma omp parallel for schedule(dynamic, 16)
    for (int j = i0; j < i1; ++j) {
        const double xj0 = x[3 * j], xj1 = x[3 * j + 1], xj2 = x[3 * j + 2];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int k = 0; k < n; ++k) {
            if (k == j) continue;
            const double xk0 = x[3 * k], xk1 = x[3 * k + 1], xk2 = x[3 * k + 2];
            const double denom = 1.0 - ((xk0 * xj0 + xk1 * xj1) + xk2 * xj2);
            a0 = a0 + ((xj1 * xk2 - xj2 * xk1) * vort[k]) * area[k] / denom;
            a1 = a1 + ((xj2 * xk0 - xj0 * xk2) * vort[k]) * area[k] / denom;
            a2 = a2 + ((xj0 * xk1 - xj1 * xk0) * vort[k]) * area[k] / denom;
        }
        dx[3 * j] = a0;
        dx[3 * j + 1] = a1;
        dx[3 * j + 2] = a2;
    }
    const double c = -4.0 * BVO_PI;
    for (int j = 0; j < n; ++j) {
        dx[3 * j] /= c;
        dx[3 * j + 1] /= c;
        dx[3 * j + 2] /= c;
        dvort[j] = -(2.0 * omega * dx[3 * j + 2]);
    }
}

/* BVEPassiveRHS (PanelVelocityParallel.f90:1337-1373) for eps2 == 0 and
 * BVESmoothVelocity (:1376-1412) for eps2 = smooth*smooth: the sum at m
 * targets X against all n active sources, no exclusion.  The smoothed
 * denominator is (1 - x.y) + smooth^2 (:1405). */
static void bvo_target_sum(int m, const double *X, int n, const double *x,
                           const double *vort, const double *area, int i0, int i1,
                           int smoothed, double eps2, double *dx)
{
    memset(dx, 0, sizeof(double) * 3 * (size_t)m);
#pragma omp parallel for schedule(dynamic, 16)
    for (int j = i0; j < i1; ++j) {
        const double X0 = X[3 * j], X1 = X[3 * j + 1], X2 = X[3 * j + 2];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int k = 0; k < n; ++k) {
            const double xk0 = x[3 * k], xk1 = x[3 * k + 1], xk2 = x[3 * k + 2];
            double denom = 1.0 - ((xk0 * X0 + xk1 * X1) + xk2 * X2);
            if (smoothed) denom = denom + eps2;
            a0 = a0 + ((X1 * xk2 - X2 * xk1) * vort[k]) * area[k] / denom;
            a1 = a1 + ((X2 * xk0 - X0 * xk2) * vort[k]) * area[k] / denom;
            a2 = a2 + ((X0 * xk1 - X1 * xk0) * vort[k]) * area[k] / denom;
        }
        dx[3 * j] = a0;
        dx[3 * j + 1] = a1;
        dx[3 * j + 2] = a2;
    }
    const double c = -4.0 * BVO_PI;
    for (int j = 0; j < 3 * m; ++j) dx[j] /= c;
}

void bvo_passive_rhs(int m, const double *X, int n, const double *x, const double *vort,
                     const double *area, int i0, int i1, double *dx)
{
    bvo_target_sum(m, X, n, x, vort, area, i0, i1, 0, 0.0, dx);
}

void bvo_smooth_velocity(int m, const double *X, int n, const double *x, const double *vort,
                         const double *area, int i0, int i1, double smooth, double *dx)
{
    bvo_target_sum(m, X, n, x, vort, area, i0, i1, 1, smooth * smooth, dx);
}

/* BVEVelocity (Voronoi), VoronoiVelocityParallel.f90:487-515: SoA input,
 * denominator 1 - x_j x_k - y_j y_k - z_j z_k subtracted left to right. */
void bvo_voronoi_velocity(int n, const double *x, const double *y, const double *z,
                          const double *vort, const double *area, int i0, int i1,
                          double omega, double *dx, double *dy, double *dz, double *dvort)
{
    memset(dx, 0, sizeof(double) * (size_t)n);
    memset(dy, 0, sizeof(double) * (size_t)n);
    memset(dz, 0, sizeof(double) * (size_t)n);
#pragma omp parallel for schedule(dynamic, 16)
    for (int j = i0; j < i1; ++j) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int k = 0; k < n; ++k) {
            if (k == j) continue;
            const double denom = 1.0 - x[j] * x[k] - y[j] * y[k] - z[j] * z[k];
            a0 = a0 + ((y[j] * z[k] - z[j] * y[k]) * vort[k]) * area[k] / denom;
            a1 = a1 + ((z[j] * x[k] - x[j] * z[k]) * vort[k]) * area[k] / denom;
            a2 = a2 + ((x[j] * y[k] - y[j] * x[k]) * vort[k]) * area[k] / denom;
        }
        dx[j] = a0;
        dy[j] = a1;
        dz[j] = a2;
    }
    const double c = -4.0 * BVO_PI;
    for (int j = 0; j < n; ++j) {
        dx[j] /= c;
        dy[j] /= c;
        dz[j] /= c;
        dvort[j] = -(2.0 * omega * dz[j]);
    }
}

/* One RK4 step of BVERK4, PanelVelocityParallel.f90:435-658, on one rank.
 *
 * In/out: active x (3*na), active relvort (na); passive x (3*npp);
 * particles x (3*np), particle relvort (np) recomputed from absvort after
 * the step (:651).  area (na) is fixed for the step (:461).
 * Out: ke (na) = |u|^2 at stage 1 (:495-497), *total_ke = 1/2 sum ke*area
 * (:498).  Stage inputs x + a_s k_{s-1} (:458-462, :513-516, :555-558,
 * :597-600), k_s = dt*u_s (:503-506), update x + k1/6 + k2/3 + k3/3 + k4/6
 * evaluated left to right (:637-644). */
void bvo_rk4_step(int na, double *ax, double *avort, const double *area,
                  int npp, double *px, int np, double *qx, double *qrel,
                  const double *qabs, double dt, double omega, double smooth,
                  double *ke, double *total_ke)
{
    const size_t a3 = 3 * (size_t)na, p3 = 3 * (size_t)npp, q3 = 3 * (size_t)np;
    double *ain = malloc(sizeof(double) * a3), *vin = malloc(sizeof(double) * na);
    double *pin = malloc(sizeof(double) * p3), *qin = malloc(sizeof(double) * q3);
    double *ka[4], *kv[4], *kp[4], *kq[4];
    for (int s = 0; s < 4; ++s) {
        ka[s] = malloc(sizeof(double) * a3);
        kv[s] = malloc(sizeof(double) * na);
        kp[s] = malloc(sizeof(double) * p3);
        kq[s] = malloc(sizeof(double) * q3);
    }
    static const double stage_a[4] = {0.0, 0.5, 0.5, 1.0};
    for (int s = 0; s < 4; ++s) {
        if (s == 0) {
            memcpy(ain, ax, sizeof(double) * a3);
            memcpy(vin, avort, sizeof(double) * na);
            memcpy(pin, px, sizeof(double) * p3);
            memcpy(qin, qx, sizeof(double) * q3);
        } else if (s < 3) {
            for (size_t i = 0; i < a3; ++i) ain[i] = ax[i] + stage_a[s] * ka[s - 1][i];
            for (int i = 0; i < na; ++i) vin[i] = avort[i] + stage_a[s] * kv[s - 1][i];
            for (size_t i = 0; i < p3; ++i) pin[i] = px[i] + stage_a[s] * kp[s - 1][i];
            for (size_t i = 0; i < q3; ++i) qin[i] = qx[i] + stage_a[s] * kq[s - 1][i];
        } else {
            for (size_t i = 0; i < a3; ++i) ain[i] = ax[i] + ka[2][i];
            for (int i = 0; i < na; ++i) vin[i] = avort[i] + kv[2][i];
            for (size_t i = 0; i < p3; ++i) pin[i] = px[i] + kp[2][i];
            for (size_t i = 0; i < q3; ++i) qin[i] = qx[i] + kq[2][i];
        }
        bvo_active_rhs(na, ain, vin, area, 0, na, omega, ka[s], kv[s]);
        bvo_passive_rhs(np, qin, na, ain, vin, area, 0, np, kq[s]);
        bvo_smooth_velocity(npp, pin, na, ain, vin, area, 0, npp, smooth, kp[s]);
        if (s == 0) {
            double tot = 0.0;
            for (int j = 0; j < na; ++j) {
                const double *u = ka[0] + 3 * j;
                double e = (u[0] * u[0] + u[1] * u[1]) + u[2] * u[2];
                if (ke) ke[j] = e;
                tot = tot + e * area[j];
            }
            if (total_ke) *total_ke = 0.5 * tot;
        }
        for (size_t i = 0; i < a3; ++i) ka[s][i] = dt * ka[s][i];
        for (size_t i = 0; i < p3; ++i) kp[s][i] = dt * kp[s][i];
        for (size_t i = 0; i < q3; ++i) kq[s][i] = dt * kq[s][i];
        for (int i = 0; i < na; ++i) kv[s][i] = dt * kv[s][i];
    }
    for (size_t i = 0; i < q3; ++i)
        qx[i] = qx[i] + kq[0][i] / 6.0 + kq[1][i] / 3.0 + kq[2][i] / 3.0 + kq[3][i] / 6.0;
    for (size_t i = 0; i < a3; ++i)
        ax[i] = ax[i] + ka[0][i] / 6.0 + ka[1][i] / 3.0 + ka[2][i] / 3.0 + ka[3][i] / 6.0;
    for (size_t i = 0; i < p3; ++i)
        px[i] = px[i] + kp[0][i] / 6.0 + kp[1][i] / 3.0 + kp[2][i] / 3.0 + kp[3][i] / 6.0;
    for (int i = 0; i < na; ++i)
        avort[i] = avort[i] + kv[0][i] / 6.0 + kv[1][i] / 3.0 + kv[2][i] / 3.0 + kv[3][i] / 6.0;
    if (qrel && qabs)
        for (int i = 0; i < np; ++i) qrel[i] = qabs[i] - 2.0 * omega * qx[3 * i + 2];
    free(ain); free(vin); free(pin); free(qin);
    for (int s = 0; s < 4; ++s) { free(ka[s]); free(kv[s]); free(kp[s]); free(kq[s]); }
}

/* One BVERK4 step of the Voronoi module, VoronoiVelocityParallel.f90:304-483,
 * on one rank, on n generators in SoA form (in/out x, y, z, relvort).
 * sum_area (n) weighs the sources (inputArea: the saved or the current
 * areas, :331-335); cur_area (n) is aPanels%area, which only the kinetic
 * energy integral uses (:358-361).  Stage inputs x + a k (:372-375,
 * :399-402, :427-430), k_s = dt u_s (:363-366), update x + k1/6 + k2/3 +
 * k3/3 + k4/6 left to right (:455-467), then NormalizeGrid (:480,
 * VoronoiPanels.f90:395-406): x / sqrt(x x + y y + z z).
 * Out: ke (n) = |u|^2 at stage 1, *total_ke = 1/2 sum ke * cur_area. */
void bvo_voronoi_rk4_step(int n, double *x, double *y, double *z, double *relvort,
                          const double *sum_area, const double *cur_area, double dt,
                          double omega, double *ke, double *total_ke)
{
    const size_t nb = sizeof(double) * (size_t)n;
    double *in[4], *k[4][4];
    for (int c = 0; c < 4; ++c) {
        in[c] = malloc(nb);
        for (int s = 0; s < 4; ++s) k[s][c] = malloc(nb);
    }
    double *state[4] = {x, y, z, relvort};
    static const double stage_a[4] = {0.0, 0.5, 0.5, 1.0};
    for (int s = 0; s < 4; ++s) {
        for (int c = 0; c < 4; ++c) {
            if (s == 0)
                memcpy(in[c], state[c], nb);
            else if (s < 3)
                for (int i = 0; i < n; ++i) in[c][i] = state[c][i] + stage_a[s] * k[s - 1][c][i];
            else
                for (int i = 0; i < n; ++i) in[c][i] = state[c][i] + k[2][c][i];
        }
        bvo_voronoi_velocity(n, in[0], in[1], in[2], in[3], sum_area, 0, n, omega, k[s][0],
                             k[s][1], k[s][2], k[s][3]);
        if (s == 0) {
            double tot = 0.0;
            for (int j = 0; j < n; ++j) {
                const double e = (k[0][0][j] * k[0][0][j] + k[0][1][j] * k[0][1][j]) +
                                 k[0][2][j] * k[0][2][j];
                if (ke) ke[j] = e;
                tot = tot + e * cur_area[j];
            }
            if (total_ke) *total_ke = 0.5 * tot;
        }
        for (int c = 0; c < 4; ++c)
            for (int i = 0; i < n; ++i) k[s][c][i] = dt * k[s][c][i];
    }
    for (int c = 0; c < 4; ++c)
        for (int i = 0; i < n; ++i)
            state[c][i] = state[c][i] + k[0][c][i] / 6.0 + k[1][c][i] / 3.0 + k[2][c][i] / 3.0 +
                          k[3][c][i] / 6.0;
    for (int j = 0; j < n; ++j) {
        const double r = sqrt(x[j] * x[j] + y[j] * y[j] + z[j] * z[j]);
        x[j] = x[j] / r;
        y[j] = y[j] / r;
        z[j] = z[j] / r;
    }
    for (int c = 0; c < 4; ++c) {
        free(in[c]);
        for (int s = 0; s < 4; ++s) free(k[s][c]);
    }
}

/* atan4, SphereGeom.f90:213-248: atan2 with values in [0, 2 PI), by the
 * reference's case split (axes first, then the quadrant of atan2(|y|,|x|)). */
static double bvo_atan4(double y, double x)
{
    if (x == 0.0) {
        if (y > 0.0) return BVO_PI / 2.0;
        if (y < 0.0) return 3.0 * BVO_PI / 2.0;
        return 0.0;
    }
    if (y == 0.0) return x > 0.0 ? 0.0 : BVO_PI;
    const double theta = atan2(fabs(y), fabs(x));
    if (x > 0.0 && y > 0.0) return theta;
    if (x < 0.0 && y > 0.0) return BVO_PI - theta;
    if (x < 0.0 && y < 0.0) return BVO_PI + theta;
    return 2.0 * BVO_PI - theta;
}

/* Juckes_A, TracerAndVorticityDistributions.f90:550-568: forcing ramp,
 * up over [0, 4), 1 on [4, 11), down over [11, 15), 0 outside. */
static double bvo_juckes_a(double t)
{
    const double tfull = 4.0, tend = 15.0, b = BVO_PI / tfull;
    if (0.0 <= t && t < tfull) return 0.5 * (1.0 - cos(b * t));
    if (tfull <= t && t < tend - tfull) return 1.0;
    if (tend - tfull <= t && t < tend) return 0.5 * (1.0 - cos(b * (t - (tend - tfull)) + BVO_PI));
    return 0.0;
}

/* Juckes_Forcing, TracerAndVorticityDistributions.f90:1223-1231, with
 * Juckes_B (:606-617), Latitude (SphereGeom.f90:198-203) and Longitude
 * (:176-181): F0 A(t) B(lat) cos(lon), F0 = 0.3 * 4 PI,
 * B = tan^2(th1)/tan^2(lat) exp(1 - tan^2(th1)/tan^2(lat)) for lat > 0. */
double bvo_juckes_forcing(const double *xyz, double t)
{
    const double f0 = 0.3 * 4.0 * BVO_PI;
    const double lat = atan2(xyz[2], sqrt(xyz[0] * xyz[0] + xyz[1] * xyz[1]));
    const double lon = bvo_atan4(xyz[1], xyz[0]);
    double b = 0.0;
    if (lat > 0.0) {
        const double th1 = BVO_PI / 3.0;
        const double tan1sq = tan(th1) * tan(th1);
        b = tan1sq / (tan(lat) * tan(lat)) * exp(1.0 - tan1sq / (tan(lat) * tan(lat)));
    }
    return f0 * bvo_juckes_a(t) * b * cos(lon);
}

/* StratosphereRK4, PanelVelocityParallel.f90:661-919: the panel RK4 step
 * with forcing.  Every stage's active vorticity is recomputed from the
 * stage positions, absVort - 2 Omega z - F(x, t_s), at t_s = t, t + dt/2,
 * t + dt/2, t + dt (:686-689, :744-747, :792-795, :837-840); the vorticity
 * tendencies are not integrated.  After the update the active panels and
 * the particles get relVort = absVort - 2 Omega z - F(x, t + dt)
 * (:906-915); parent panels keep theirs.  Arrays as in bvo_rk4_step, plus
 * the active panels' absolute vorticity aabs; avort is output only. */
void bvo_stratosphere_rk4_step(int na, double *ax, double *avort, const double *aabs,
                               const double *area, int npp, double *px, int np, double *qx,
                               double *qrel, const double *qabs, double t, double dt,
                               double omega, double smooth, double *ke, double *total_ke)
{
    const size_t a3 = 3 * (size_t)na, p3 = 3 * (size_t)npp, q3 = 3 * (size_t)np;
    double *ain = malloc(sizeof(double) * a3), *vin = malloc(sizeof(double) * na);
    double *pin = malloc(sizeof(double) * p3), *qin = malloc(sizeof(double) * q3);
    double *dv = malloc(sizeof(double) * na);
    double *ka[4], *kp[4], *kq[4];
    for (int s = 0; s < 4; ++s) {
        ka[s] = malloc(sizeof(double) * a3);
        kp[s] = malloc(sizeof(double) * p3);
        kq[s] = malloc(sizeof(double) * q3);
    }
    static const double stage_a[4] = {0.0, 0.5, 0.5, 1.0};
    static const double stage_t[4] = {0.0, 0.5, 0.5, 1.0};
    for (int s = 0; s < 4; ++s) {
        if (s == 0) {
            memcpy(ain, ax, sizeof(double) * a3);
            memcpy(pin, px, sizeof(double) * p3);
            memcpy(qin, qx, sizeof(double) * q3);
        } else if (s < 3) {
            for (size_t i = 0; i < a3; ++i) ain[i] = ax[i] + stage_a[s] * ka[s - 1][i];
            for (size_t i = 0; i < q3; ++i) qin[i] = qx[i] + stage_a[s] * kq[s - 1][i];
            for (size_t i = 0; i < p3; ++i) pin[i] = px[i] + stage_a[s] * kp[s - 1][i];
        } else {
            for (size_t i = 0; i < a3; ++i) ain[i] = ax[i] + ka[2][i];
            for (size_t i = 0; i < q3; ++i) qin[i] = qx[i] + kq[2][i];
            for (size_t i = 0; i < p3; ++i) pin[i] = px[i] + kp[2][i];
        }
        const double ts = s == 0 ? t : t + stage_t[s] * dt;
        for (int j = 0; j < na; ++j)
            vin[j] = aabs[j] - 2.0 * omega * ain[3 * j + 2] - bvo_juckes_forcing(ain + 3 * j, ts);
        bvo_active_rhs(na, ain, vin, area, 0, na, omega, ka[s], dv);
        bvo_passive_rhs(np, qin, na, ain, vin, area, 0, np, kq[s]);
        bvo_smooth_velocity(npp, pin, na, ain, vin, area, 0, npp, smooth, kp[s]);
        if (s == 0) {
            double tot = 0.0;
            for (int j = 0; j < na; ++j) {
                const double *u = ka[0] + 3 * j;
                double e = (u[0] * u[0] + u[1] * u[1]) + u[2] * u[2];
                if (ke) ke[j] = e;
                tot = tot + e * area[j];
            }
            if (total_ke) *total_ke = 0.5 * tot;
        }
        for (size_t i = 0; i < a3; ++i) ka[s][i] = dt * ka[s][i];
        for (size_t i = 0; i < q3; ++i) kq[s][i] = dt * kq[s][i];
        for (size_t i = 0; i < p3; ++i) kp[s][i] = dt * kp[s][i];
    }
    for (size_t i = 0; i < q3; ++i)
        qx[i] = qx[i] + kq[0][i] / 6.0 + kq[1][i] / 3.0 + kq[2][i] / 3.0 + kq[3][i] / 6.0;
    for (size_t i = 0; i < a3; ++i)
        ax[i] = ax[i] + ka[0][i] / 6.0 + ka[1][i] / 3.0 + ka[2][i] / 3.0 + ka[3][i] / 6.0;
    for (size_t i = 0; i < p3; ++i)
        px[i] = px[i] + kp[0][i] / 6.0 + kp[1][i] / 3.0 + kp[2][i] / 3.0 + kp[3][i] / 6.0;
    for (int j = 0; j < na; ++j)
        avort[j] = aabs[j] - 2.0 * omega * ax[3 * j + 2] - bvo_juckes_forcing(ax + 3 * j, t + dt);
    if (qrel && qabs)
        for (int j = 0; j < np; ++j)
            qrel[j] = qabs[j] - 2.0 * omega * qx[3 * j + 2] - bvo_juckes_forcing(qx + 3 * j, t + dt);
    free(ain); free(vin); free(pin); free(qin); free(dv);
    for (int s = 0; s < 4; ++s) { free(ka[s]); free(kp[s]); free(kq[s]); }
}

/* NonDivergentWind (divergent = 0) and DivergentWind (divergent = 1),
 * PanelVelocityParallel.f90:1415-1459: the analytic deformational winds
 * (R = 1, T = 5) at xyz and time t, zero within 1e-14 of the axis. */
void bvo_wind(int divergent, const double *xyz, double t, double *out)
{
    const double RR = 1.0, TT = 5.0, zero_tol = 1e-14;
    const double x = xyz[0], y = xyz[1], z = xyz[2];
    const double lat = atan2(z, sqrt(x * x + y * y));
    const double lon = bvo_atan4(y, x);
    const double raxis = sqrt(x * x + y * y);
    if (!(raxis > zero_tol)) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    double u, v;
    if (!divergent) {
        u = 10.0 * RR / TT * sin(lon - 2.0 * BVO_PI * t / TT) * sin(lon - 2.0 * BVO_PI * t / TT) *
                sin(2.0 * lat) * cos(BVO_PI * t / TT) +
            2.0 * BVO_PI * RR / TT * cos(lat);
        v = 10.0 * RR / TT * sin(2.0 * (lon - 2.0 * BVO_PI * t / TT)) * cos(lat) * cos(BVO_PI * t / TT);
    } else {
        u = -5.0 * RR / TT * sin((lon - 2.0 * BVO_PI * t / TT) / 2.0) *
                sin((lon - 2.0 * BVO_PI * t / TT) / 2.0) * sin(2.0 * lat) * cos(lat) * cos(lat) *
                cos(BVO_PI * t / TT) +
            2.0 * BVO_PI * RR * cos(lat) / TT;
        v = 5.0 * RR / (2.0 * TT) * sin(lon - 2.0 * BVO_PI * t / TT) * cos(lat) * cos(lat) * cos(lat) *
            cos(BVO_PI * t / TT);
    }
    out[0] = -u * y / raxis - v * x * z / raxis;
    out[1] = u * x / raxis - v * y * z / raxis;
    out[2] = v * raxis;
}

/* AdvectionRK4 / DivergentAdvectionRK4, PanelVelocityParallel.f90:922-1099
 * and :1102-1279: classical RK4 of the n points x (AoS, in place) in the
 * analytic wind, stages at t, t + dt/2, t + dt/2, t + dt; every point class
 * is moved the same way and vorticity is untouched. */
void bvo_advection_rk4_step(int divergent, int n, double *x, double t, double dt)
{
    for (int j = 0; j < n; ++j) {
        double *p = x + 3 * (size_t)j;
        double k1[3], k2[3], k3[3], k4[3], in[3];
        bvo_wind(divergent, p, t, k1);
        for (int c = 0; c < 3; ++c) k1[c] = dt * k1[c];
        for (int c = 0; c < 3; ++c) in[c] = p[c] + 0.5 * k1[c];
        bvo_wind(divergent, in, t + 0.5 * dt, k2);
        for (int c = 0; c < 3; ++c) k2[c] = dt * k2[c];
        for (int c = 0; c < 3; ++c) in[c] = p[c] + 0.5 * k2[c];
        bvo_wind(divergent, in, t + 0.5 * dt, k3);
        for (int c = 0; c < 3; ++c) k3[c] = dt * k3[c];
        for (int c = 0; c < 3; ++c) in[c] = p[c] + k3[c];
        bvo_wind(divergent, in, t + dt, k4);
        for (int c = 0; c < 3; ++c) k4[c] = dt * k4[c];
        for (int c = 0; c < 3; ++c)
            p[c] = p[c] + k1[c] / 6.0 + k2[c] / 3.0 + k3[c] / 3.0 + k4[c] / 6.0;
    }
}

/* RenormalizeGrid, ParticlesEdgesPanels2.f90:1690-1702: x / sqrt(x.x). */
void bvo_renormalize(int n, double *x)
{
    for (int j = 0; j < n; ++j) {
        double *p = x + 3 * j;
        const double r = sqrt((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]);
        p[0] /= r; p[1] /= r; p[2] /= r;
    }
}

/* High-precision brute force of the same sum (not the reference's rounding):
 * every quantity in __float128.  Targets X (m), sources x (n); self index
 * for target j is self_off + j when self_off >= 0 (active block), and the
 * denominator is (1 - x.X) + eps2.  Output in float64 after the final
 * division by -4 PI.  For checking float64 kernels on small inputs. */
void bvo_velocity_q(int m, const double *X, int n, const double *x, const double *w,
                    int self_off, double eps2, double *u)
{
    const __float128 pi_q = 3.14159265358979323846264338327950288Q;
#pragma omp parallel for schedule(dynamic, 8)
    for (int j = 0; j < m; ++j) {
        const __float128 X0 = X[3 * j], X1 = X[3 * j + 1], X2 = X[3 * j + 2];
        __float128 a0 = 0, a1 = 0, a2 = 0;
        for (int k = 0; k < n; ++k) {
            if (self_off >= 0 && k == self_off + j) continue;
            const __float128 xk0 = x[3 * k], xk1 = x[3 * k + 1], xk2 = x[3 * k + 2];
            const __float128 d = (1 - (xk0 * X0 + xk1 * X1 + xk2 * X2)) + (__float128)eps2;
            const __float128 s = (__float128)w[k] / d;
            a0 += s * (X1 * xk2 - X2 * xk1);
            a1 += s * (X2 * xk0 - X0 * xk2);
            a2 += s * (X0 * xk1 - X1 * xk0);
        }
        u[3 * j] = (double)(a0 / (-4 * pi_q));
        u[3 * j + 1] = (double)(a1 / (-4 * pi_q));
        u[3 * j + 2] = (double)(a2 / (-4 * pi_q));
    }
}

/* ---- tangent-linear model of the active-panel dynamics ----------------
 *
 * State y = (x_i, zeta_i), i < n, areas A fixed; f(y) = (U, -2 Omega U_z)
 * with U the BVEActiveRHS sum (PanelVelocityParallel.f90:1287-1334) and
 * w_j = zeta_j A_j.  Its Jacobian-vector product, differentiated by hand:
 *
 *   dU_i = -(1/4 PI) sum_{j != i} [ dw_j (x_i x x_j)/D
 *                                 + w_j (dx_i x x_j + x_i x dx_j)/D
 *                                 + w_j (x_i x x_j)(dx_i.x_j + x_i.dx_j)/D^2 ],
 *   D = 1 - x_i.x_j,  dw_j = dzeta_j A_j,  dzeta_dot_i = -2 Omega dU_z,i.
 *
 * Pinned by central differences of bvo_rk4_step (which is bitwise the
 * reference) in tests/test_oracle.py. */
void bvo_jvp_active(int n, const double *x, const double *zeta, const double *area,
                    const double *dx, const double *dzeta, double omega, double *du,
                    double *dzdot)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        const double *xi = x + 3 * i, *dxi = dx + 3 * i;
        double a0 = 0, a1 = 0, a2 = 0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const double *xj = x + 3 * j, *dxj = dx + 3 * j;
            const double D = 1.0 - (xi[0] * xj[0] + xi[1] * xj[1] + xi[2] * xj[2]);
            const double w = zeta[j] * area[j], dw = dzeta[j] * area[j];
            const double c0 = xi[1] * xj[2] - xi[2] * xj[1];
            const double c1 = xi[2] * xj[0] - xi[0] * xj[2];
            const double c2 = xi[0] * xj[1] - xi[1] * xj[0];
            /* dx_i x x_j + x_i x dx_j */
            const double e0 = (dxi[1] * xj[2] - dxi[2] * xj[1]) + (xi[1] * dxj[2] - xi[2] * dxj[1]);
            const double e1 = (dxi[2] * xj[0] - dxi[0] * xj[2]) + (xi[2] * dxj[0] - xi[0] * dxj[2]);
            const double e2 = (dxi[0] * xj[1] - dxi[1] * xj[0]) + (xi[0] * dxj[1] - xi[1] * dxj[0]);
            const double dD = dxi[0] * xj[0] + dxi[1] * xj[1] + dxi[2] * xj[2] +
                              xi[0] * dxj[0] + xi[1] * dxj[1] + xi[2] * dxj[2];
            const double r = 1.0 / D;
            a0 += dw * c0 * r + w * e0 * r + w * c0 * dD * r * r;
            a1 += dw * c1 * r + w * e1 * r + w * c1 * dD * r * r;
            a2 += dw * c2 * r + w * e2 * r + w * c2 * dD * r * r;
        }
        const double c = -4.0 * BVO_PI;
        du[3 * i] = a0 / c;
        du[3 * i + 1] = a1 / c;
        du[3 * i + 2] = a2 / c;
        dzdot[i] = -(2.0 * omega * du[3 * i + 2]);
    }
}

/* Tangent of n_steps RK4 steps (BVERK4's scheme, active panels only):
 * (x, zeta, dx, dzeta) advanced in place; the base state is bvo_rk4_step's
 * up to rounding (the update is accumulated as x + sum b_s k_s). */
void bvo_tlm_rk4(int n, double *x, double *zeta, const double *area, double *dx, double *dzeta,
                 double dt, double omega, int n_steps)
{
    const size_t n3 = 3 * (size_t)n;
    double *X = malloc(sizeof(double) * n3), *Z = malloc(sizeof(double) * n);
    double *dX = malloc(sizeof(double) * n3), *dZ = malloc(sizeof(double) * n);
    double *u = malloc(sizeof(double) * n3), *zd = malloc(sizeof(double) * n);
    double *du = malloc(sizeof(double) * n3), *dzd = malloc(sizeof(double) * n);
    double *kx = malloc(sizeof(double) * n3), *kz = malloc(sizeof(double) * n);
    double *dkx = malloc(sizeof(double) * n3), *dkz = malloc(sizeof(double) * n);
    double *ax = malloc(sizeof(double) * n3), *az = malloc(sizeof(double) * n);
    double *dax = malloc(sizeof(double) * n3), *daz = malloc(sizeof(double) * n);
    static const double a_in[4] = {0.0, 0.5, 0.5, 1.0};
    static const double b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    for (int step = 0; step < n_steps; ++step) {
        memcpy(ax, x, sizeof(double) * n3);
        memcpy(az, zeta, sizeof(double) * n);
        memcpy(dax, dx, sizeof(double) * n3);
        memcpy(daz, dzeta, sizeof(double) * n);
        for (int s = 0; s < 4; ++s) {
            for (size_t k = 0; k < n3; ++k) {
                X[k] = x[k] + a_in[s] * (s ? kx[k] : 0.0);
                dX[k] = dx[k] + a_in[s] * (s ? dkx[k] : 0.0);
            }
            for (int k = 0; k < n; ++k) {
                Z[k] = zeta[k] + a_in[s] * (s ? kz[k] : 0.0);
                dZ[k] = dzeta[k] + a_in[s] * (s ? dkz[k] : 0.0);
            }
            bvo_active_rhs(n, X, Z, area, 0, n, omega, u, zd);
            bvo_jvp_active(n, X, Z, area, dX, dZ, omega, du, dzd);
            for (size_t k = 0; k < n3; ++k) {
                kx[k] = dt * u[k];
                dkx[k] = dt * du[k];
                ax[k] += b[s] * kx[k];
                dax[k] += b[s] * dkx[k];
            }
            for (int k = 0; k < n; ++k) {
                kz[k] = dt * zd[k];
                dkz[k] = dt * dzd[k];
                az[k] += b[s] * kz[k];
                daz[k] += b[s] * dkz[k];
            }
        }
        memcpy(x, ax, sizeof(double) * n3);
        memcpy(zeta, az, sizeof(double) * n);
        memcpy(dx, dax, sizeof(double) * n3);
        memcpy(dzeta, daz, sizeof(double) * n);
    }
    free(X); free(Z); free(dX); free(dZ); free(u); free(zd); free(du); free(dzd);
    free(kx); free(kz); free(dkx); free(dkz); free(ax); free(az); free(dax); free(daz);
}

/* Threads the OpenMP loops above use (1 when built without OpenMP). */
int bvo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void bvo_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
