/*
 * htx_oracle.c -- FP64 CPU restatement of the Hermite-TENO FV residual /
 * time-step path (reference: indrow/Hermite-TENO, C++/Eigen, CPU only).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP product path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it; the product library never links or calls it.
 *
 * The reference cannot be compiled here (it needs Eigen3, absent offline), so
 * this file restates it operation by operation from the source text.  Each
 * function cites the reference lines it follows.  Evaluation order follows the
 * source text left to right (Eigen's own dot/mat-vec order is not pinned by
 * any reference test); the parity build uses -ffp-contract=off.
 *
 * Pinning (tests/test_oracle.py): closed-form known answers (polynomial
 * exactness of teno5/linear/quadrature, Gauss-9 integrals), invariants
 * (L.R = I, R.diag(lambda).L = dF/dq, flux consistency F(q,q) = f(q),
 * Toro HLLC on exact Riemann data with quirks off, conservation with periodic
 * boundaries, free-stream preservation, mirror and x<->y transpose symmetry,
 * SSP-RK3 amplification polynomial) and brute-force CFL on tiny grids.
 *
 * Parity unpinned: the composed trajectories as the reference compiles them
 * (Eigen's evaluation order), which only a build of the reference could
 * pin.
 */
#include "htx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Constants of Schemes::Lagrange (Schemes_Lagrange.cpp:6-142)               */
/* ------------------------------------------------------------------------ */
typedef struct {
    double crj[3][4];                 /* crj(r,c)            :125-136 */
    double dkm[3], dkp[3];            /* 0.3,0.6,0.1 / mirror :140-142 */
    double crjgp3L[3][3], crjgm3L[3][3]; /* :66-84 */
    double linwp[3], linwm[3];        /* :95-105 */
    double linw0p[3], linw0m[3];      /* :86-93 */
    double sumw0p, sumw0m;
    double c1, eps, epsz, ct, machine_eps;
} lag_consts;

static lag_consts LC;
static int LC_ready = 0;

static void lag_init(void) {
    if (LC_ready) return;
    double sq15 = sqrt(15.0);
    LC.c1 = 13.0 / 12.0;
    LC.eps = 1E-6;
    LC.epsz = 1E-40;
    LC.ct = 5E-4;
    LC.machine_eps = 1E-15;

    /* crj: Schemes_Lagrange.cpp:125-136 */
    LC.crj[0][0] = 11.0 / 6.0; LC.crj[1][0] = -7.0 / 6.0; LC.crj[2][0] = 1.0 / 3.0;
    LC.crj[0][1] = 1.0 / 3.0;  LC.crj[1][1] = 5.0 / 6.0;  LC.crj[2][1] = -1.0 / 6.0;
    LC.crj[0][2] = -1.0 / 6.0; LC.crj[1][2] = 5.0 / 6.0;  LC.crj[2][2] = 1.0 / 3.0;
    LC.crj[0][3] = 1.0 / 3.0;  LC.crj[1][3] = -7.0 / 6.0; LC.crj[2][3] = 11.0 / 6.0;
    LC.dkm[0] = 0.3; LC.dkm[1] = 0.6; LC.dkm[2] = 0.1;
    LC.dkp[0] = 0.1; LC.dkp[1] = 0.6; LC.dkp[2] = 0.3;

    /* crjgp3L / crjgm3L: :66-84 */
    LC.crjgp3L[0][0] = 1.0 / 30.0 + sq15 / 20.0;
    LC.crjgp3L[1][0] = -1.0 / 15.0 - sq15 / 5.0;
    LC.crjgp3L[2][0] = 31.0 / 30.0 + 3.0 * sq15 / 20.0;
    LC.crjgp3L[0][1] = 1.0 / 30.0 - sq15 / 20.0;
    LC.crjgp3L[1][1] = 14.0 / 15.0;
    LC.crjgp3L[2][1] = 1.0 / 30.0 + sq15 / 20.0;
    LC.crjgp3L[0][2] = 31.0 / 30.0 - 3.0 * sq15 / 20.0;
    LC.crjgp3L[1][2] = -1.0 / 15.0 + sq15 / 5.0;
    LC.crjgp3L[2][2] = 1.0 / 30.0 - sq15 / 20.0;

    LC.crjgm3L[0][0] = 1.0 / 30.0 - sq15 / 20.0;
    LC.crjgm3L[1][0] = -1.0 / 15.0 + sq15 / 5.0;
    LC.crjgm3L[2][0] = 31.0 / 30.0 - 3.0 * sq15 / 20.0;
    LC.crjgm3L[0][1] = 1.0 / 30.0 + sq15 / 20.0;
    LC.crjgm3L[1][1] = 14.0 / 15.0;
    LC.crjgm3L[2][1] = 1.0 / 30.0 - sq15 / 20.0;
    LC.crjgm3L[0][2] = 31.0 / 30.0 + 3.0 * sq15 / 20.0;
    LC.crjgm3L[1][2] = -1.0 / 15.0 - sq15 / 5.0;
    LC.crjgm3L[2][2] = 1.0 / 30.0 + sq15 / 20.0;

    /* centre-point split into positive / negative groups: :86-93 */
    double lin0[3] = {-9.0 / 80.0, 49.0 / 40.0, -9.0 / 80.0};
    for (int i = 0; i < 3; ++i) {
        LC.linw0p[i] = 0.5 * (lin0[i] + 3.0 * fabs(lin0[i]));
        LC.linw0m[i] = LC.linw0p[i] - lin0[i];
    }
    LC.sumw0p = (LC.linw0p[0] + LC.linw0p[1]) + LC.linw0p[2];
    LC.sumw0m = (LC.linw0m[0] + LC.linw0m[1]) + LC.linw0m[2];
    for (int i = 0; i < 3; ++i) {
        LC.linw0p[i] = LC.linw0p[i] / LC.sumw0p;
        LC.linw0m[i] = LC.linw0m[i] / LC.sumw0m;
    }
    /* +xi / -xi linear weights: :95-105 */
    double denom = 15720.0 * sq15 + 10480.0;
    LC.linwp[0] = (2882.0 * sq15 - 1179.0) / denom;
    LC.linwp[1] = (9672.0 * sq15 + 6448.0) / denom;
    LC.linwp[2] = (3166.0 * sq15 + 5211.0) / denom;
    denom = 15720.0 * sq15 - 10480;
    LC.linwm[0] = (2882.0 * sq15 + 1179.0) / denom;
    LC.linwm[1] = (9672.0 * sq15 - 6448.0) / denom;
    LC.linwm[2] = (3166.0 * sq15 - 5211.0) / denom;
    LC_ready = 1;
}

static inline double dot3(const double a[3], const double b[3]) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static inline double sum3(const double a[3]) { return (a[0] + a[1]) + a[2]; }
static inline double catchnan(double x) { return isnan(x) ? LC.machine_eps : x; }

/* Schemes::Lagrange::smoothnessMeasure, Schemes_Lagrange.cpp:1013-1026 */
static void smoothness5(const double* v, double SI[3]) {
    double c[3], b[3];
    c[0] = v[2] - 2.0 * v[3] + v[4];
    b[0] = 3.0 * v[2] - 4.0 * v[3] + v[4];
    c[1] = v[1] - 2.0 * v[2] + v[3];
    b[1] = v[1] - v[3];
    c[2] = v[0] - 2.0 * v[1] + v[2];
    b[2] = v[0] - 4.0 * v[1] + 3.0 * v[2];
    for (int k = 0; k < 3; ++k) {
        b[k] = b[k] / 2.0;
        SI[k] = b[k] * b[k] + LC.c1 * c[k] * c[k];
    }
}

/* Schemes::Lagrange::fullpointsSmootness, Schemes_Lagrange.cpp:1029-1034 */
static double b35_of(const double* u) {
    double b35 = u[4] * (6908.0 * u[4] - 51001.0 * u[3] + 67923.0 * u[2] - 38947.0 * u[1] + 8209.0 * u[0]) +
                 u[3] * (104963.0 * u[3] - 299076.0 * u[2] + 179098.0 * u[1] - 38947.0 * u[0]) +
                 u[2] * (231153.0 * u[2] - 299076.0 * u[1] + 67923.0 * u[0]) +
                 u[1] * (104963.0 * u[1] - 51001.0 * u[0]) + 6908.0 * u[0] * u[0];
    return b35 / 5040.0;
}

/* One side of teno5: smoothness, tau, gamma^6, cutOff (Schemes_Lagrange.cpp
 * :486-494 and :1141-1149).  v = the 5-point window, p = candidates,
 * d = linear weights.  Returns the reconstructed value; writes delta flags. */
static double teno5_side(const double* v, const double p[3], const double d[3], int sig_out[3]) {
    double SI[3], w[3], sigm[3];
    smoothness5(v, SI);
    double b35 = b35_of(v);
    double tau = fabs(b35 - (SI[0] + SI[2] + 4.0 * SI[1]) / 6.0);
    for (int k = 0; k < 3; ++k) w[k] = pow(1.0 + tau / (SI[k] + LC.epsz), 6);
    /* cutOff(): normalise, then threshold at C_T */
    double s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    for (int k = 0; k < 3; ++k) sigm[k] = w[k] < LC.ct ? 0.0 : 1.0;
    for (int k = 0; k < 3; ++k) w[k] = sigm[k] * d[k];
    s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    if (sig_out)
        for (int k = 0; k < 3; ++k) sig_out[k] = (int)sigm[k];
    return catchnan(dot3(p, w));
}

/* Schemes::Lagrange::teno5, Schemes_Lagrange.cpp:481-511.
 * u = 6 consecutive values u_{i-2..i+3}; out = (u^-_{i+1/2}, u^+_{i+1/2}). */
static void teno5_impl(const double u[6], double out[2], int flags[6]) {
    double p[3], c[3];
    /* left-biased value at x_{i+1/2}: :482-494 */
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][1];
    p[0] = dot3(c, u + 2);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][2];
    p[1] = dot3(c, u + 1);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][3];
    p[2] = dot3(c, u + 0);
    out[0] = teno5_side(u + 0, p, LC.dkm, flags ? flags : NULL);
    /* right-biased value at x_{i+1/2}: :496-508 */
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][0];
    p[0] = dot3(c, u + 3);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][1];
    p[1] = dot3(c, u + 2);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][2];
    p[2] = dot3(c, u + 1);
    out[1] = teno5_side(u + 1, p, LC.dkp, flags ? flags + 3 : NULL);
}

void orc_teno5(const double u[6], double out[2]) {
    lag_init();
    teno5_impl(u, out, NULL);
}

void orc_teno5_flags(const double u[6], int flags[6]) {
    double out[2];
    lag_init();
    teno5_impl(u, out, flags);
}

/* One side of weno5JS (Schemes_Lagrange.cpp:358-381): weights
 * d_k / (eps + SI_k)^2 normalised; or of weno5Z (:383-408): weights
 * d_k (1 + tau / (SI_k + epsz)), tau = |SI_2 - SI_0|, normalised. */
static double weno5_side(const double* v, const double p[3], const double d[3], int z) {
    double SI[3], w[3];
    smoothness5(v, SI);
    if (z) {
        double tau = fabs(SI[2] - SI[0]);
        for (int k = 0; k < 3; ++k) w[k] = d[k] * (1.0 + tau / (SI[k] + LC.epsz));
    } else {
        for (int k = 0; k < 3; ++k) w[k] = d[k] / pow(LC.eps + SI[k], 2);
    }
    double s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    return catchnan(dot3(p, w));
}

/* Schemes::Lagrange::weno5JS / weno5Z: the candidates of teno5_impl. */
static void weno5_impl(const double u[6], double out[2], int z) {
    double p[3], c[3];
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][1];
    p[0] = dot3(c, u + 2);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][2];
    p[1] = dot3(c, u + 1);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][3];
    p[2] = dot3(c, u + 0);
    out[0] = weno5_side(u + 0, p, LC.dkm, z);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][0];
    p[0] = dot3(c, u + 3);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][1];
    p[1] = dot3(c, u + 2);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][2];
    p[2] = dot3(c, u + 1);
    out[1] = weno5_side(u + 1, p, LC.dkp, z);
}

void orc_weno5(const double u[6], int z, double out[2]) {
    lag_init();
    weno5_impl(u, out, z);
}

/* Schemes::Lagrange::linear, Schemes_Lagrange.cpp:1219-1233 */
void orc_linear5(const double u[6], double out[2]) {
    double p[3], c[3];
    lag_init();
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][1];
    p[0] = dot3(c, u + 2);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][2];
    p[1] = dot3(c, u + 1);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][3];
    p[2] = dot3(c, u + 0);
    out[0] = dot3(p, LC.dkm);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][0];
    p[0] = dot3(c, u + 3);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][1];
    p[1] = dot3(c, u + 2);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][2];
    p[2] = dot3(c, u + 1);
    out[1] = dot3(p, LC.dkp);
}

/* Schemes::Lagrange::weno5JS3PointsQuadrature + smoothnessMeasure3L,
 * Schemes_Lagrange.cpp:1155-1217.  u = 5 transverse face values (rows
 * j-2..j+2); out = values at the Gauss points (+xi, 0, -xi), xi = sqrt(15)/10. */
void orc_weno5js3q(const double u[5], double out[3]) {
    double bg[3], cg[3], SI[3], p[3], w[3], c[3], s;
    lag_init();
    bg[0] = u[0] - 4.0 * u[1] + 3.0 * u[2];
    cg[0] = u[0] - 2.0 * u[1] + u[2];
    bg[1] = u[1] - u[3];
    cg[1] = u[1] - 2.0 * u[2] + u[3];
    bg[2] = 3.0 * u[2] - 4.0 * u[3] + u[4];
    cg[2] = u[2] - 2.0 * u[3] + u[4];
    for (int k = 0; k < 3; ++k) {
        bg[k] = 0.25 * (bg[k] * bg[k]);
        cg[k] = LC.c1 * (cg[k] * cg[k]);
        SI[k] = bg[k] + cg[k];
    }
    double den[3];
    for (int k = 0; k < 3; ++k) den[k] = (LC.eps + SI[k]) * (LC.eps + SI[k]);

    /* +xi point: :1178-1186 */
    for (int k = 0; k < 3; ++k) {
        for (int r = 0; r < 3; ++r) c[r] = LC.crjgp3L[r][k];
        p[k] = dot3(c, u + k);
    }
    for (int k = 0; k < 3; ++k) w[k] = LC.linwp[k] / den[k];
    s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    out[0] = dot3(p, w);

    /* centre point with positive/negative weight split: :1188-1200 */
    p[0] = (-u[0] + 2 * u[1] + 23.0 * u[2]) / 24.0;
    p[1] = (-u[1] + 26.0 * u[2] - u[3]) / 24.0;
    p[2] = (23.0 * u[2] + 2.0 * u[3] - u[4]) / 24.0;
    for (int k = 0; k < 3; ++k) w[k] = LC.linw0p[k] / den[k];
    s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    out[1] = LC.sumw0p * dot3(p, w);
    for (int k = 0; k < 3; ++k) w[k] = LC.linw0m[k] / den[k];
    s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    out[1] -= LC.sumw0m * dot3(p, w);

    /* -xi point: :1202-1210 */
    for (int k = 0; k < 3; ++k) {
        for (int r = 0; r < 3; ++r) c[r] = LC.crjgm3L[r][k];
        p[k] = dot3(c, u + k);
    }
    for (int k = 0; k < 3; ++k) w[k] = LC.linwm[k] / den[k];
    s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    out[2] = dot3(p, w);

    for (int k = 0; k < 3; ++k) out[k] = catchnan(out[k]);
}

/* weno5JS3PointsQuadrature's linear limit (Schemes_Lagrange.cpp:1171-1217
 * with every nonlinear weight replaced by its linear weight, :86-105): the
 * candidates of each Gauss point combined with linwp, linw0p / linw0m (the
 * centre's positive and negative groups) and linwm.  No reference routine
 * computes it; the learned path's transverse quadrature on smooth windows is
 * this combination. */
void orc_weno5js3q_linear(const double u[5], double out[3]) {
    double p[3], c[3];
    lag_init();
    for (int k = 0; k < 3; ++k) {
        for (int r = 0; r < 3; ++r) c[r] = LC.crjgp3L[r][k];
        p[k] = dot3(c, u + k);
    }
    out[0] = dot3(p, LC.linwp);
    p[0] = (-u[0] + 2 * u[1] + 23.0 * u[2]) / 24.0;
    p[1] = (-u[1] + 26.0 * u[2] - u[3]) / 24.0;
    p[2] = (23.0 * u[2] + 2.0 * u[3] - u[4]) / 24.0;
    out[1] = LC.sumw0p * dot3(p, LC.linw0p) - LC.sumw0m * dot3(p, LC.linw0m);
    for (int k = 0; k < 3; ++k) {
        for (int r = 0; r < 3; ++r) c[r] = LC.crjgm3L[r][k];
        p[k] = dot3(c, u + k);
    }
    out[2] = dot3(p, LC.linwm);
}

/* ------------------------------------------------------------------------ */
/* The remaining Schemes::Lagrange routines (Schemes_Lagrange.cpp:315-1004,  */
/* 1036-1149, 1265-1294), restated in the member-state style of the          */
/* reference: SI, weights, sigm, tau are locals standing for its members.    */
/* ------------------------------------------------------------------------ */

/* the 5th-order candidates of one side: col c of crj applied to the 3
 * values starting at u[off[c]] (the p(0..2) assignments of teno5, :482-484
 * and :495-497) */
static void cand_side(const double u[6], int plus, double p[3]) {
    double c[3];
    const int cm[3] = {1, 2, 3}, om[3] = {2, 1, 0};
    const int cp[3] = {0, 1, 2}, op[3] = {3, 2, 1};
    for (int k = 0; k < 3; ++k) {
        int col = plus ? cp[k] : cm[k], off = plus ? op[k] : om[k];
        for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][col];
        p[k] = dot3(c, u + off);
    }
}

/* cutOff (:1141-1149) with ct, then sigm.cwiseProduct(d) normalised and
 * the candidates combined, catchnan */
static double cut_and_combine(double weights[3], const double p[3], const double d[3], double ct) {
    double sigm[3], s = sum3(weights);
    for (int k = 0; k < 3; ++k) weights[k] = weights[k] / s;
    for (int k = 0; k < 3; ++k) sigm[k] = weights[k] < ct ? 0.0 : 1.0;
    for (int k = 0; k < 3; ++k) weights[k] = sigm[k] * d[k];
    s = sum3(weights);
    for (int k = 0; k < 3; ++k) weights[k] = weights[k] / s;
    return catchnan(dot3(p, weights));
}

static double normalise_and_combine(double weights[3], const double p[3]) {
    double s = sum3(weights);
    for (int k = 0; k < 3; ++k) weights[k] = weights[k] / s;
    return catchnan(dot3(p, weights));
}

/* the SI and tau of weno5N / teno5N (:445-453, :558-566) on window v */
static void si_n(const double* v, double SI[3], double* fourth) {
    const double theta = 0.1;
    SI[2] = theta * fabs(v[0] - 3.0 * v[1] + 2.0 * v[2]) + fabs(v[0] - 2.0 * v[1] + v[2]);
    SI[1] = theta * fabs(v[3] - v[2]) + fabs(v[1] - 2.0 * v[2] + v[3]);
    SI[0] = theta * fabs(v[3] - v[2]) + fabs(v[2] - 2.0 * v[3] + v[4]);
    *fourth = fabs(v[0] - 4.0 * v[1] + 6.0 * v[2] - 4.0 * v[3] + v[4]);
}

/* Lagrange::weno5N :440-478 */
static void weno5n_impl(const double u[6], double out[2]) {
    double p[3], SI[3], w[3], f;
    for (int side = 0; side < 2; ++side) {
        cand_side(u, side, p);
        si_n(u + side, SI, &f);
        double tau = pow(f, 2.5);
        const double* d = side ? LC.dkp : LC.dkm;
        for (int k = 0; k < 3; ++k) w[k] = d[k] * (1.0 + (tau / pow(SI[k] + LC.eps, 2)));
        out[side] = normalise_and_combine(w, p);
    }
}

/* Lagrange::teno5N :553-593 */
static void teno5n_impl(const double u[6], double out[2]) {
    double p[3], SI[3], w[3], f;
    for (int side = 0; side < 2; ++side) {
        cand_side(u, side, p);
        si_n(u + side, SI, &f);
        double tau = pow(f, 2);
        for (int k = 0; k < 3; ++k) w[k] = pow(1.0 + tau / pow(SI[k] + LC.eps, 2), 6);
        out[side] = cut_and_combine(w, p, side ? LC.dkp : LC.dkm, LC.ct);
    }
}

/* Lagrange::teno5a :1265-1294 */
static void teno5a_impl(const double u[6], double out[2]) {
    double p[3], SI[3], w[3];
    for (int side = 0; side < 2; ++side) {
        cand_side(u, side, p);
        smoothness5(u + side, SI);
        double tau = 1.0;
        for (int k = 0; k < 3; ++k) w[k] = pow(tau / (SI[k] + LC.epsz), 6);
        out[side] = cut_and_combine(w, p, side ? LC.dkp : LC.dkm, LC.ct);
    }
}

/* Lagrange::teno5(u, diff) :514-551 */
static void teno5diff_impl(const double u[6], double diff, double out[2]) {
    double p[3], SI[3], w[3];
    double df = pow(diff, 4);
    for (int side = 0; side < 2; ++side) {
        cand_side(u, side, p);
        smoothness5(u + side, SI);
        double tau = fabs(b35_of(u + side) - (SI[0] + SI[2] + 4.0 * SI[1]) / 6.0);
        for (int k = 0; k < 3; ++k) w[k] = pow(1.0 + pow(tau / (SI[k] + df), 2), 3);
        out[side] = cut_and_combine(w, p, side ? LC.dkp : LC.dkm, LC.ct);
    }
}

/* Lagrange::weno5Z(u, diff) :410-438 (both sides take fullpointsSmootness
 * of u.segment(0, 5), :419 and :431; diff is unused) */
static void weno5zdiff_impl(const double u[6], double out[2]) {
    double p[3], SI[3], w[3];
    for (int side = 0; side < 2; ++side) {
        cand_side(u, side, p);
        smoothness5(u + side, SI);
        double tau = fabs(b35_of(u) - (SI[0] + SI[2] + 4.0 * SI[1]) / 6.0);
        const double* d = side ? LC.dkp : LC.dkm;
        for (int k = 0; k < 3; ++k) w[k] = (1.0 + tau / pow(SI[k] + LC.epsz, 0.75)) * d[k];
        out[side] = normalise_and_combine(w, p);
    }
}

/* smoothnessMeasureIS :1036-1050 (mod = 0) / smoothnessMeasureISmod
 * :1052-1066 (mod = 1, offsets df) -> weightsis */
static void si_is(const double* u, int mod, double df, double weightsis[4]) {
    double SI4is[4], b12, tau;
    SI4is[0] = pow(u[3] - u[2], 2);
    SI4is[1] = pow(u[2] - u[1], 2);
    SI4is[2] = 0.25 * pow(3.0 * u[2] - 4.0 * u[3] + u[4], 2) + LC.c1 * pow(u[2] - 2.0 * u[3] + u[4], 2);
    SI4is[3] = 0.25 * pow(u[0] - 4.0 * u[1] + 3.0 * u[2], 2) + LC.c1 * pow(u[0] - 2.0 * u[1] + u[2], 2);
    b12 = 0.25 * pow(u[1] - u[3], 2) + LC.c1 * pow(u[1] - 2.0 * u[2] + u[3], 2);
    tau = 0.25 * pow(u[4] - 2.0 * (u[3] - u[1]) - u[0], 2) +
          LC.c1 * pow(u[4] - 4.0 * u[3] + 6.0 * u[2] - 4.0 * u[1] + u[0], 2);
    if (!mod) {
        weightsis[0] = 1.0 + pow(tau, 2) / (SI4is[0] + LC.eps) / (b12 + LC.eps);
        weightsis[1] = 1.0 + pow(tau, 2) / (SI4is[1] + LC.eps) / (b12 + LC.eps);
        weightsis[2] = 1.0 + tau / (SI4is[2] + LC.eps);
        weightsis[3] = 1.0 + tau / (SI4is[3] + LC.eps);
    } else {
        weightsis[0] = 1.0 + tau * tau / (SI4is[0] * b12 + df);
        weightsis[1] = 1.0 + tau * tau / (SI4is[1] * b12 + df);
        weightsis[2] = 1.0 + tau / (SI4is[2] + df);
        weightsis[3] = 1.0 + tau / (SI4is[3] + df);
    }
}

/* Lagrange::weno5IS :596-620 and weno5ISmod(u, diff) :622-647 (dkmis /
 * dkpis :144-152) */
static void weno5is_impl(const double u[6], int mod, double diff, double out[2]) {
    const double dkmis[4] = {0.4, 0.2, 0.3, 0.1}, dkpis[4] = {0.2, 0.4, 0.1, 0.3};
    double df = mod ? pow(diff, 3) : 0.0;
    double pis[4], w[4], c[3];
    for (int side = 0; side < 2; ++side) {
        if (!side) {
            pis[0] = 0.5 * (u[2] + u[3]);
            pis[1] = -0.5 * u[1] + 1.5 * u[2];
            for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][1];
            pis[2] = dot3(c, u + 2);
            for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][3];
            pis[3] = dot3(c, u + 0);
        } else {
            pis[0] = 1.5 * u[3] - 0.5 * u[4];
            pis[1] = 0.5 * (u[2] + u[3]);
            for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][0];
            pis[2] = dot3(c, u + 3);
            for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][2];
            pis[3] = dot3(c, u + 1);
        }
        si_is(u + side, mod, df, w);
        const double* d = side ? dkpis : dkmis;
        for (int k = 0; k < 4; ++k) w[k] = d[k] * w[k];
        double s = ((w[0] + w[1]) + w[2]) + w[3];
        for (int k = 0; k < 4; ++k) w[k] = w[k] / s;
        out[side] = catchnan(((pis[0] * w[0] + pis[1] * w[1]) + pis[2] * w[2]) + pis[3] * w[3]);
    }
}

static double sum6(const double a[6]) {
    double s = a[0];
    for (int i = 1; i < 6; ++i) s = s + a[i];
    return s;
}
static double dot6(const double a[6], const double b[6]) {
    double s = a[0] * b[0];
    for (int i = 1; i < 6; ++i) s = s + a[i] * b[i];
    return s;
}

/* Lagrange::weno5IS_Extended :649-717 with smoothnessMeasureISExtended
 * :1068-1084 and the split optimal weights :288-303 */
static void weno5isext_impl(const double u[6], double out[2]) {
    const double dkmise[6] = {-1.0 / 45.0, 11.0 / 90.0, 1.0 / 5.0, 2.0 / 5.0, 1.0 / 5.0, 1.0 / 10.0};
    const double dkpise[6] = {1.0 / 10.0, 1.0 / 5.0, 4.0 / 10.0, 2.0 / 10.0, 11.0 / 90.0, -1.0 / 45.0};
    for (int side = 0; side < 2; ++side) {
        const double* dk = side ? dkpise : dkmise;
        double pos[6], neg[6], pos_sum, neg_sum;
        for (int i = 0; i < 6; ++i) {
            pos[i] = 0.5 * (dk[i] + 3.0 * fabs(dk[i]));
            neg[i] = pos[i] - dk[i];
        }
        pos_sum = sum6(pos);
        neg_sum = sum6(neg);
        for (int i = 0; i < 6; ++i) {
            pos[i] = pos[i] / pos_sum;
            neg[i] = neg[i] / neg_sum;
        }
        double pise[6];
        if (!side) {
            pise[0] = -1.5 * u[0] + 2.5 * u[1];
            pise[1] = -0.5 * u[1] + 1.5 * u[2];
            pise[2] = pise[1];
            pise[3] = 0.5 * (u[2] + u[3]);
            pise[4] = pise[3];
            pise[5] = 1.5 * u[3] - 0.5 * u[4];
        } else {
            pise[0] = -0.5 * u[1] + 1.5 * u[2];
            pise[1] = 0.5 * (u[2] + u[3]);
            pise[2] = pise[1];
            pise[3] = 1.5 * u[3] - 0.5 * u[4];
            pise[4] = pise[3];
            pise[5] = 2.5 * u[4] - 1.5 * u[5];
        }
        const double* v = u + side;
        double c[3], b[3], SI[3], SI4is[4], tau;
        c[2] = v[2] - 2.0 * v[3] + v[4];
        b[2] = 3.0 * v[2] - 4.0 * v[3] + v[4];
        c[1] = v[1] - 2.0 * v[2] + v[3];
        b[1] = v[1] - v[3];
        c[0] = v[0] - 2.0 * v[1] + v[2];
        b[0] = v[0] - 4.0 * v[1] + 3.0 * v[2];
        for (int k = 0; k < 3; ++k) {
            b[k] = b[k] / 2.0;
            SI[k] = b[k] * b[k] + LC.c1 * c[k] * c[k];
        }
        for (int i = 0; i < 4; ++i) SI4is[i] = pow(v[i + 1] - v[i], 2);
        tau = 0.25 * pow(v[4] - 2.0 * (v[3] - v[1]) - v[0], 2) +
              LC.c1 * pow(v[4] - 4.0 * v[3] + 6.0 * v[2] - 4.0 * v[1] + v[0], 2);
        tau *= tau;
        double w[6], wp[6], wn[6];
        w[0] = 1.0 + tau / (SI[0] + LC.eps) / (SI4is[0] + LC.eps);
        w[1] = 1.0 + tau / (SI[0] + LC.eps) / (SI4is[1] + LC.eps);
        w[2] = 1.0 + tau / (SI[1] + LC.eps) / (SI4is[1] + LC.eps);
        w[3] = 1.0 + tau / (SI[1] + LC.eps) / (SI4is[2] + LC.eps);
        w[4] = 1.0 + tau / (SI[2] + LC.eps) / (SI4is[2] + LC.eps);
        w[5] = 1.0 + tau / (SI[2] + LC.eps) / (SI4is[3] + LC.eps);
        for (int i = 0; i < 6; ++i) w[i] = pow(w[i], 4);
        double s = sum6(w);
        for (int i = 0; i < 6; ++i) w[i] = w[i] / s;
        for (int i = 0; i < 6; ++i) w[i] = w[i] < 1E-5 ? 0.0 : 1.0;
        for (int i = 0; i < 6; ++i) {
            wp[i] = pos[i] * w[i];
            wn[i] = neg[i] * w[i];
        }
        double sp = sum6(wp), sn = sum6(wn);
        for (int i = 0; i < 6; ++i) {
            wp[i] = wp[i] / sp;
            wn[i] = wn[i] / sn;
        }
        out[side] = catchnan(pos_sum * dot6(pise, wp) - neg_sum * dot6(pise, wn));
    }
}

static void teno5i_impl(const double u[6], double out[2]);
static void weno5_impl(const double u[6], double out[2], int z);

/* The characteristic reconstruction a solver calls (FV_System1D.cpp:200-201,
 * FV_System2D.cpp:379, :394) with one of the alternatives swapped in:
 * ORC_RECON_* (diff: the cell size normal to the face, for the routines
 * that take one). */
void orc_recon(int recon, const double u[6], double diff, double out[2]) {
    lag_init();
    switch (recon) {
        case ORC_RECON_TENO5I: teno5i_impl(u, out); break;
        case ORC_RECON_WENO5JS: weno5_impl(u, out, 0); break;
        case ORC_RECON_WENO5Z: weno5_impl(u, out, 1); break;
        case ORC_RECON_WENO5N: weno5n_impl(u, out); break;
        case ORC_RECON_TENO5N: teno5n_impl(u, out); break;
        case ORC_RECON_TENO5A: teno5a_impl(u, out); break;
        case ORC_RECON_WENO5IS: weno5is_impl(u, 0, 0.0, out); break;
        case ORC_RECON_WENO5IS_EXT: weno5isext_impl(u, out); break;
        case ORC_RECON_WENO5ISMOD: weno5is_impl(u, 1, diff, out); break;
        case ORC_RECON_WENO5Z_DIFF: weno5zdiff_impl(u, out); break;
        case ORC_RECON_TENO5_DIFF: teno5diff_impl(u, diff, out); break;
        case ORC_RECON_LINEAR: orc_linear5(u, out); break;
        default: teno5_impl(u, out, NULL); break;
    }
}

/* Lagrange::weno3JS :315-335 (its last line stores catchnan(out(0)) in
 * out(1)) and weno3JS2PointsQuadrature :337-356 */
void orc_weno3js(const double u[4], double out[2]) {
    const double eps = 1E-6;
    const double crjm3rd[2][2] = {{-0.5, 0.5}, {1.5, 0.5}}; /* (r, c) :105-108 */
    const double dkm3rd[2] = {1.0 / 3.0, 2.0 / 3.0};
    double crjp3rd[2][2], dkp3rd[2] = {dkm3rd[1], dkm3rd[0]};
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) crjp3rd[r][c] = crjm3rd[1 - r][1 - c];
    double p3rd[2], SI3rd[2], w[2], s;
    p3rd[0] = crjm3rd[0][0] * u[0] + crjm3rd[1][0] * u[1];
    p3rd[1] = crjm3rd[0][1] * u[1] + crjm3rd[1][1] * u[2];
    SI3rd[0] = pow(u[0] - u[1], 2);
    SI3rd[1] = pow(u[1] - u[2], 2);
    for (int k = 0; k < 2; ++k) w[k] = dkm3rd[k] / pow(eps + SI3rd[k], 2);
    s = w[0] + w[1];
    for (int k = 0; k < 2; ++k) w[k] = w[k] / s;
    out[0] = catchnan(p3rd[0] * w[0] + p3rd[1] * w[1]);
    p3rd[0] = crjp3rd[0][0] * u[1] + crjp3rd[1][0] * u[2];
    p3rd[1] = crjp3rd[0][1] * u[2] + crjp3rd[1][1] * u[3];
    SI3rd[0] = pow(u[1] - u[2], 2);
    SI3rd[1] = pow(u[2] - u[3], 2);
    for (int k = 0; k < 2; ++k) w[k] = dkp3rd[k] / pow(eps + SI3rd[k], 2);
    s = w[0] + w[1];
    for (int k = 0; k < 2; ++k) w[k] = w[k] / s;
    out[1] = p3rd[0] * w[0] + p3rd[1] * w[1];
    out[1] = catchnan(out[0]);
}

void orc_weno3js_2q(const double u[3], double out[2]) {
    const double eps = 1E-6, sq3 = sqrt(3.0), c6 = sq3 / 3.0;
    double crjp3rdg[2][2], crjm3rdg[2][2];
    crjp3rdg[0][0] = -sq3 / 6.0;
    crjp3rdg[1][0] = sq3 / 6.0 + 1.0;
    crjp3rdg[0][1] = -sq3 / 6.0 + 1.0;
    crjp3rdg[1][1] = sq3 / 6.0;
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) crjm3rdg[r][c] = crjp3rdg[1 - r][1 - c];
    double SI3rd[2], p3rd[2], w[2], s;
    SI3rd[0] = u[0] - u[1];
    SI3rd[1] = u[1] - u[2];
    for (int k = 0; k < 2; ++k) SI3rd[k] = c6 * pow(SI3rd[k], 2);
    p3rd[0] = crjp3rdg[0][0] * u[0] + crjp3rdg[1][0] * u[1];
    p3rd[1] = crjp3rdg[0][1] * u[1] + crjp3rdg[1][1] * u[2];
    for (int k = 0; k < 2; ++k) w[k] = 0.5 / pow(eps + SI3rd[k], 2);
    s = w[0] + w[1];
    for (int k = 0; k < 2; ++k) w[k] = w[k] / s;
    out[0] = p3rd[0] * w[0] + p3rd[1] * w[1];
    p3rd[0] = crjm3rdg[0][0] * u[0] + crjm3rdg[1][0] * u[1];
    p3rd[1] = crjm3rdg[0][1] * u[1] + crjm3rdg[1][1] * u[2];
    for (int k = 0; k < 2; ++k) w[k] = 0.5 / pow(eps + SI3rd[k], 2);
    s = w[0] + w[1];
    for (int k = 0; k < 2; ++k) w[k] = w[k] / s;
    out[1] = p3rd[0] * w[0] + p3rd[1] * w[1];
}

/* coefficient tables of the 2- and 4-point quadratures (constructor,
 * Schemes_Lagrange.cpp:9-20, :154-286) */
typedef struct {
    double crjgp[3][3], crjgm[3][3], dkpg[3], dkmg[3];
    double crjgpLobatto[3][3], crjgmLobatto[3][3], dkpLobatto[3], dkmLobatto[3];
    double crjgpLegendre_12[3][3], crjgmLegendre_12[3][3], dkpLegendre_12[3], dkmLegendre_12[3];
    double crjgpLegendre_34[3][3], crjgmLegendre_34[3][3], dkpLegendre_34[3], dkmLegendre_34[3];
    double c2, c3, c4, c5, c7, c8, c9, c10;
} quad_tables;

/* M = P with rows 0 and 2 swapped, then columns 0 and 2 */
static void swap02(const double P[3][3], double M[3][3]) {
    double T[3][3];
    for (int c = 0; c < 3; ++c) {
        T[0][c] = P[2][c];
        T[1][c] = P[1][c];
        T[2][c] = P[0][c];
    }
    for (int r = 0; r < 3; ++r) {
        M[r][0] = T[r][2];
        M[r][1] = T[r][1];
        M[r][2] = T[r][0];
    }
}

static void quad_tables_init(quad_tables* Q) {
    double sq3 = sqrt(3.0), sq5 = sqrt(5.0), denom;
    Q->c2 = sq3 / 12.0;
    Q->c3 = 37.0 / 108.0 * sq3;
    Q->c4 = sq5 / 20.0;
    Q->c5 = 61.0 * sq5 / 300.0;
    Q->c7 = sqrt(0.525e3 - 0.70e2 * sqrt(0.30e2)) / 0.140e3;
    Q->c8 = pow(0.525e3 - 0.70e2 * sqrt(0.30e2), 0.3e1 / 0.2e1) / 0.514500e6 +
            sqrt(0.525e3 - 0.70e2 * sqrt(0.30e2)) / 0.35e2;
    Q->c9 = sqrt(0.525e3 + 0.70e2 * sqrt(0.30e2)) / 0.140e3;
    Q->c10 = pow(0.525e3 + 0.70e2 * sqrt(0.30e2), 0.3e1 / 0.2e1) / 0.514500e6 +
             sqrt(0.525e3 + 0.70e2 * sqrt(0.30e2)) / 0.35e2;
    Q->crjgp[0][0] = sq3 / 12.0;
    Q->crjgp[1][0] = -sq3 / 3.0;
    Q->crjgp[2][0] = 1.0 + sq3 / 4.0;
    Q->crjgp[0][1] = -sq3 / 12.0;
    Q->crjgp[1][1] = 1.0;
    Q->crjgp[2][1] = sq3 / 12.0;
    Q->crjgp[0][2] = -sq3 / 4.0 + 1.0;
    Q->crjgp[1][2] = sq3 / 3.0;
    Q->crjgp[2][2] = -sq3 / 12.0;
    swap02(Q->crjgp, Q->crjgm);
    Q->crjgpLobatto[0][0] = sq5 / 20.0 - 1.0 / 60.0;
    Q->crjgpLobatto[1][0] = -sq5 / 5.0 + 1.0 / 30.0;
    Q->crjgpLobatto[2][0] = 3.0 * sq5 / 20.0 + 59 / 60.0;
    Q->crjgpLobatto[0][1] = -sq5 / 20.0 - 1.0 / 60.0;
    Q->crjgpLobatto[1][1] = 31.0 / 30.0;
    Q->crjgpLobatto[2][1] = sq5 / 20.0 - 1.0 / 60.0;
    Q->crjgpLobatto[0][2] = -3.0 * sq5 / 20.0 + 59.0 / 60.0;
    Q->crjgpLobatto[1][2] = sq5 / 5.0 + 1.0 / 30.0;
    Q->crjgpLobatto[2][2] = -sq5 / 20.0 - 1.0 / 60.0;
    swap02(Q->crjgpLobatto, Q->crjgmLobatto);
    denom = 660.0 * sq5 - 220.0;
    Q->dkpLobatto[0] = 132.0 * sq5 + 22.0;
    Q->dkpLobatto[1] = 387.0 * sq5 - 129.0;
    Q->dkpLobatto[2] = 141.0 * sq5 - 113.0;
    for (int k = 0; k < 3; ++k) Q->dkpLobatto[k] = Q->dkpLobatto[k] / denom;
    denom = 660.0 * sq5 + 220.0;
    Q->dkmLobatto[0] = 132.0 * sq5 - 22.0;
    Q->dkmLobatto[1] = 387.0 * sq5 + 129.0;
    Q->dkmLobatto[2] = 141.0 * sq5 + 113.0;
    for (int k = 0; k < 3; ++k) Q->dkmLobatto[k] = Q->dkmLobatto[k] / denom;
    denom = 1080;
    Q->dkpg[0] = (210.0 - sq3) / denom;
    Q->dkpg[1] = 11.0 / 18.0;
    Q->dkpg[2] = (210.0 + sq3) / denom;
    Q->dkmg[0] = Q->dkpg[2];
    Q->dkmg[1] = Q->dkpg[1];
    Q->dkmg[2] = Q->dkpg[0];
    const double r30 = sqrt(0.30e2);
    const double am = sqrt(0.525e3 - 0.70e2 * sqrt(0.30e2)), ap = sqrt(0.525e3 + 0.70e2 * sqrt(0.30e2));
    const double bm = 0.525e3 - 0.70e2 * sqrt(0.30e2), bp = 0.525e3 + 0.70e2 * sqrt(0.30e2);
    Q->crjgpLegendre_12[0][0] = am / 0.140e3 - r30 / 0.140e3 + 0.1e1 / 0.84e2;
    Q->crjgpLegendre_12[1][0] = -am / 0.35e2 + r30 / 0.70e2 - 0.1e1 / 0.42e2;
    Q->crjgpLegendre_12[2][0] = 0.3e1 / 0.140e3 * am - r30 / 0.140e3 + 0.85e2 / 0.84e2;
    Q->crjgpLegendre_12[0][1] = -am / 0.140e3 - r30 / 0.140e3 + 0.1e1 / 0.84e2;
    Q->crjgpLegendre_12[1][1] = r30 / 0.70e2 + 0.41e2 / 0.42e2;
    Q->crjgpLegendre_12[2][1] = am / 0.140e3 - r30 / 0.140e3 + 0.1e1 / 0.84e2;
    Q->crjgpLegendre_12[0][2] = -0.3e1 / 0.140e3 * am - r30 / 0.140e3 + 0.85e2 / 0.84e2;
    Q->crjgpLegendre_12[1][2] = am / 0.35e2 + r30 / 0.70e2 - 0.1e1 / 0.42e2;
    Q->crjgpLegendre_12[2][2] = -am / 0.140e3 - r30 / 0.140e3 + 0.1e1 / 0.84e2;
    swap02(Q->crjgpLegendre_12, Q->crjgmLegendre_12);
    denom = (0.1234800000e10 * r30 - 0.9116940000e10) * am + 0.11174940000e11 * r30 - 0.52238900000e11;
    Q->dkpLegendre_12[2] = ((0.257250000e9 * r30 - 0.1899362500e10) * am +
                            (-0.42000e5 * r30 + 0.310100e6) * pow(bm, 0.3e1 / 0.2e1) - 0.1197315000e10 * r30 +
                            0.5597025000e10) / denom;
    Q->dkpLegendre_12[1] = ((0.1761795000e10 * r30 - 0.13038532500e11) * am +
                            (-0.88200e5 * r30 + 0.16096500e8) * pow(bm, 0.3e1 / 0.2e1) -
                            0.2520e4 * pow(bm, 0.5e1 / 0.2e1) + 0.6757590000e10 * r30 - 0.32331425000e11) / denom;
    Q->dkpLegendre_12[0] = ((0.26092500e8 * r30 - 0.256576250e9) * am +
                            (-0.2100e4 * r30 - 0.3838100e7) * pow(bm, 0.3e1 / 0.2e1) +
                            0.630e3 * pow(bm, 0.5e1 / 0.2e1) + 0.5614665000e10 * r30 - 0.25504500000e11) / denom;
    Q->dkmLegendre_12[0] = Q->dkpLegendre_12[2];
    Q->dkmLegendre_12[1] = Q->dkpLegendre_12[1];
    Q->dkmLegendre_12[2] = Q->dkpLegendre_12[0];
    Q->crjgpLegendre_34[0][0] = ap / 0.140e3 + r30 / 0.140e3 + 0.1e1 / 0.84e2;
    Q->crjgpLegendre_34[1][0] = -ap / 0.35e2 - r30 / 0.70e2 - 0.1e1 / 0.42e2;
    Q->crjgpLegendre_34[2][0] = 0.3e1 / 0.140e3 * ap + r30 / 0.140e3 + 0.85e2 / 0.84e2;
    Q->crjgpLegendre_34[0][1] = -ap / 0.140e3 + r30 / 0.140e3 + 0.1e1 / 0.84e2;
    Q->crjgpLegendre_34[1][1] = -r30 / 0.70e2 + 0.41e2 / 0.42e2;
    Q->crjgpLegendre_34[2][1] = ap / 0.140e3 + r30 / 0.140e3 + 0.1e1 / 0.84e2;
    Q->crjgpLegendre_34[0][2] = -0.3e1 / 0.140e3 * ap + r30 / 0.140e3 + 0.85e2 / 0.84e2;
    Q->crjgpLegendre_34[1][2] = ap / 0.35e2 - r30 / 0.70e2 - 0.1e1 / 0.42e2;
    Q->crjgpLegendre_34[2][2] = -ap / 0.140e3 + r30 / 0.140e3 + 0.1e1 / 0.84e2;
    swap02(Q->crjgpLegendre_34, Q->crjgmLegendre_34);
    denom = (0.1234800000e10 * r30 + 0.9116940000e10) * ap - 0.11174940000e11 * r30 - 0.52238900000e11;
    Q->dkpLegendre_34[2] = ((0.26092500e8 * r30 + 0.256576250e9) * ap +
                            (-0.2100e4 * r30 + 0.3838100e7) * pow(bp, 0.3e1 / 0.2e1) -
                            0.630e3 * pow(bp, 0.5e1 / 0.2e1) - 0.5614665000e10 * r30 - 0.25504500000e11) / denom;
    Q->dkpLegendre_34[1] = ((0.1761795000e10 * r30 + 0.13038532500e11) * ap +
                            (-0.88200e5 * r30 - 0.16096500e8) * pow(bp, 0.3e1 / 0.2e1) +
                            0.2520e4 * pow(bp, 0.5e1 / 0.2e1) - 0.6757590000e10 * r30 - 0.32331425000e11) / denom;
    Q->dkpLegendre_34[0] = ((0.257250000e9 * r30 + 0.1899362500e10) * ap +
                            (-0.42000e5 * r30 - 0.310100e6) * pow(bp, 0.3e1 / 0.2e1) + 0.1197315000e10 * r30 +
                            0.5597025000e10) / denom;
    Q->dkmLegendre_34[0] = Q->dkpLegendre_34[2];
    Q->dkmLegendre_34[1] = Q->dkpLegendre_34[1];
    Q->dkmLegendre_34[2] = Q->dkpLegendre_34[0];
}

/* p(k) = C.col(k).dot(u.segment(k, 3)) */
static void cand3(const double C[3][3], const double u[5], double p[3]) {
    double c[3];
    for (int k = 0; k < 3; ++k) {
        for (int r = 0; r < 3; ++r) c[r] = C[r][k];
        p[k] = dot3(c, u + k);
    }
}

/* bg, cg of smoothnessMeasure2PointsGaussLegendre :1086-1099 (and the 4-point
 * forms :1101-1133), squared */
static void bg_cg2(const double u[5], double bg[3], double cg[3]) {
    bg[0] = u[0] - 4.0 * u[1] + 3.0 * u[2];
    cg[0] = u[0] - 2.0 * u[1] + u[2];
    bg[1] = u[1] - u[3];
    cg[1] = u[1] - 2.0 * u[2] + u[3];
    bg[2] = 3.0 * u[2] - 4.0 * u[3] + u[4];
    cg[2] = u[2] - 2.0 * u[3] + u[4];
    for (int k = 0; k < 3; ++k) {
        bg[k] = pow(bg[k], 2);
        cg[k] = pow(cg[k], 2);
    }
}

/* weights of one quadrature point: 0 JS, 1 Z (SZG), 2 TENO (sigm) */
static double qpoint(int kind, const double p[3], const double d[3], const double SI[3], double tau,
                     const double sigm[3], int nan_guard) {
    double w[3], s;
    for (int k = 0; k < 3; ++k) {
        if (kind == 0)
            w[k] = d[k] / pow(LC.eps + SI[k], 2);
        else if (kind == 1)
            w[k] = d[k] * (1.0 + tau / (SI[k] + LC.epsz));
        else
            w[k] = sigm[k] * d[k];
    }
    s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    double r = dot3(p, w);
    return nan_guard ? catchnan(r) : r;
}

/* cutOff of (1 + tau / (SI + epsz))^6; prenorm: normalised once before
 * cutOff normalises (:962-963) */
static void qsigm(const double SI[3], double tau, int prenorm, double sigm[3]) {
    double w[3], s;
    for (int k = 0; k < 3; ++k) w[k] = pow(1.0 + tau / (SI[k] + LC.epsz), 6);
    if (prenorm) {
        s = sum3(w);
        for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    }
    s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    for (int k = 0; k < 3; ++k) sigm[k] = w[k] < LC.ct ? 0.0 : 1.0;
}

/* weno5JS2PointsQuadrature (kind 0, :719-740), weno5Z2PointsQuadrature (1,
 * :742-765), teno5th2PointsQuadrature (2, :767-791) */
void orc_quad2(int kind, const double u[5], double out[2]) {
    quad_tables Q;
    double bg[3], cg[3], SIGLQ[3], p[3], sigm[3] = {1.0, 1.0, 1.0};
    lag_init();
    quad_tables_init(&Q);
    bg_cg2(u, bg, cg);
    for (int k = 0; k < 3; ++k) SIGLQ[k] = Q.c2 * bg[k] + Q.c3 * cg[k];
    double tau = fabs(SIGLQ[0] - SIGLQ[2]);
    if (kind == 2) qsigm(SIGLQ, tau, 0, sigm);
    cand3(Q.crjgp, u, p);
    out[0] = qpoint(kind, p, Q.dkpg, SIGLQ, tau, sigm, 0);
    cand3(Q.crjgm, u, p);
    out[1] = qpoint(kind, p, Q.dkmg, SIGLQ, tau, sigm, 0);
}

/* weno5JS4PointsGaussLobatto (0, :793-831), weno5Z4PointsGaussLobatto (1,
 * :914-955, no catchnan), teno5th4PointsGaussLobatto (2, :957-1004) */
void orc_quad4_lobatto(int kind, const double u[5], double out[4]) {
    quad_tables Q;
    double bg[3], cg[3], SIA[3], SIB[3], p[3], c[3], sigm[3] = {1.0, 1.0, 1.0};
    lag_init();
    quad_tables_init(&Q);
    bg_cg2(u, bg, cg);
    for (int k = 0; k < 3; ++k) {
        SIA[k] = 0.25 * bg[k] + Q.c3 * cg[k];
        SIB[k] = Q.c4 * bg[k] + Q.c5 * cg[k];
    }
    int ng = kind != 1;
    double tau = fabs(SIA[0] - SIA[2]);
    if (kind == 2) qsigm(SIA, tau, 1, sigm);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][1];
    p[2] = dot3(c, u + 2);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][2];
    p[1] = dot3(c, u + 1);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][3];
    p[0] = dot3(c, u + 0);
    out[0] = qpoint(kind, p, LC.dkm, SIA, tau, sigm, ng);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][0];
    p[2] = dot3(c, u + 2);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][1];
    p[1] = dot3(c, u + 1);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][2];
    p[0] = dot3(c, u + 0);
    out[1] = qpoint(kind, p, LC.dkp, SIA, tau, sigm, ng);
    tau = fabs(SIB[0] - SIB[2]);
    if (kind == 2) qsigm(SIB, tau, 1, sigm);
    cand3(Q.crjgpLobatto, u, p);
    out[2] = qpoint(kind, p, Q.dkpLobatto, SIB, tau, sigm, ng);
    cand3(Q.crjgmLobatto, u, p);
    out[3] = qpoint(kind, p, Q.dkmLobatto, SIB, tau, sigm, ng);
}

/* weno5JS4PointsGaussLegendre (0, :833-869), weno5Z4PointsGaussLegendre (1,
 * :871-912) */
void orc_quad4_legendre(int kind, const double u[5], double out[4]) {
    quad_tables Q;
    double bg[3], cg[3], SIA[3], SIB[3], p[3], sigm[3] = {1.0, 1.0, 1.0};
    lag_init();
    quad_tables_init(&Q);
    bg_cg2(u, bg, cg);
    for (int k = 0; k < 3; ++k) {
        SIA[k] = Q.c7 * bg[k] + Q.c8 * cg[k];
        SIB[k] = Q.c9 * bg[k] + Q.c10 * cg[k];
    }
    double tau = fabs(SIA[0] - SIA[2]);
    cand3(Q.crjgpLegendre_12, u, p);
    out[0] = qpoint(kind, p, Q.dkpLegendre_12, SIA, tau, sigm, 1);
    cand3(Q.crjgmLegendre_12, u, p);
    out[1] = qpoint(kind, p, Q.dkmLegendre_12, SIA, tau, sigm, 1);
    tau = fabs(SIB[0] - SIB[2]);
    cand3(Q.crjgpLegendre_34, u, p);
    out[2] = qpoint(kind, p, Q.dkpLegendre_34, SIB, tau, sigm, 1);
    cand3(Q.crjgmLegendre_34, u, p);
    out[3] = qpoint(kind, p, Q.dkmLegendre_34, SIB, tau, sigm, 1);
}

/* Gauss-Legendre 9-point rule in the reference's node order
 * (Schemes_Lagrange.cpp:22-42): 0, -a1, +a1, -a2, +a2, -a3, +a3, -a4, +a4. */
void orc_gauss9(double nodes[9], double weights[9]) {
    static const double a[4] = {0.8360311073266358, 0.9681602395076261, 0.3242534234038089,
                                0.6133714327005904};
    static const double wa[4] = {0.1806481606948574, 0.08127438836157441, 0.31234707704000286,
                                 0.26061069640293544};
    nodes[0] = 0.0;
    weights[0] = 0.3302393550012598;
    for (int k = 0; k < 4; ++k) {
        nodes[1 + 2 * k] = -a[k];
        nodes[2 + 2 * k] = a[k];
        weights[1 + 2 * k] = wa[k];
        weights[2 + 2 * k] = wa[k];
    }
}

/* ------------------------------------------------------------------------ */
/* Equations::Euler2D (Equations_Euler2D.cpp)                                */
/* ------------------------------------------------------------------------ */

/* primitive: Equations_Euler2D.cpp:20-26 */
void orc_primitive(double gamma, const double q[4], double w[4]) {
    double g1 = gamma - 1.0;
    w[0] = q[0];
    w[1] = q[1] / q[0];
    w[2] = q[2] / q[0];
    w[3] = (q[3] - 0.5 * (q[1] * w[1] + q[2] * w[2])) * g1;
}

/* conservative: Equations_Euler2D.cpp:28-34 */
void orc_conservative(double gamma, const double w[4], double q[4]) {
    double g1 = gamma - 1.0;
    q[0] = w[0];
    q[1] = w[0] * w[1];
    q[2] = w[0] * w[2];
    q[3] = w[3] / g1 + 0.5 * (q[1] * w[1] + q[2] * w[2]);
}

/* fluxx / fluxy: Equations_Euler2D.cpp:36-50 */
static void phys_flux(int dir, const double q[4], const double w[4], double f[4]) {
    if (dir == 1) {
        f[0] = q[1];
        f[1] = q[1] * w[1] + w[3];
        f[2] = q[1] * w[2];
        f[3] = w[1] * (q[3] + w[3]);
    } else {
        f[0] = q[2];
        f[1] = q[2] * w[1];
        f[2] = q[2] * w[2] + w[3];
        f[3] = w[2] * (q[3] + w[3]);
    }
}

/* Roe-averaged eigenvectors of the face between cells a and b in direction
 * dir: FV::System2D::eigen2DEuler (FV_System2D.cpp:265-300 for x, :329-360 for
 * y) with Euler2D::right/leftEigenVectorx/y (Equations_Euler2D.cpp:100-205).
 * The per-cell sqrt(|rho|) and H come from calcPrimitiveVariable
 * (FV_System2D.cpp:226-263).  L and R are row-major 4x4. */
void orc_eigen(double gamma, int dir, const double qa[4], const double qb[4], double L[16],
               double R[16]) {
    double g1 = gamma - 1.0;
    double wa[4], wb[4];
    orc_primitive(gamma, qa, wa);
    orc_primitive(gamma, qb, wb);
    double sa = sqrt(fabs(wa[0])), sb = sqrt(fabs(wb[0]));
    double ha = (qa[3] + wa[3]) / wa[0], hb = (qb[3] + wb[3]) / wb[0];
    double t0 = sa / (sa + sb);
    double t1 = 1.0 - t0;
    double vxm = t0 * wa[1] + t1 * wb[1];
    double vym = t0 * wa[2] + t1 * wb[2];
    double hm = t0 * ha + t1 * hb;
    double qm = 0.5 * (vxm * vxm + vym * vym);
    double cm = sqrt(fabs(g1 * (hm - qm)));
    double rcm = 1.0 / cm;
    double b1 = g1 * rcm * rcm;
    double b2 = qm * b1;
    double t2 = 0.5 * b1;
    if (dir == 1) {
        double tc = vxm * cm;
        double Rx[16] = {1.0, 0.0, 1.0, 1.0,
                         vxm - cm, 0.0, vxm, vxm + cm,
                         vym, 1.0, vym, vym,
                         hm - tc, vym, qm, hm + tc};
        double tr = vxm * rcm, tb = b1 * vxm;
        double Lx[16] = {0.5 * (b2 + tr), -0.5 * (tb + rcm), -t2 * vym, t2,
                         -vym, 0.0, 1.0, 0.0,
                         1.0 - b2, tb, b1 * vym, -b1,
                         0.5 * (b2 - tr), -0.5 * (tb - rcm), -t2 * vym, t2};
        memcpy(R, Rx, sizeof Rx);
        memcpy(L, Lx, sizeof Lx);
    } else {
        double tc = vym * cm;
        double Ry[16] = {1.0, 0.0, 1.0, 1.0,
                         vxm, 1.0, vxm, vxm,
                         vym - cm, 0.0, vym, vym + cm,
                         hm - tc, vxm, qm, hm + tc};
        double tr = vym * rcm, tb = b1 * vym;
        double Ly[16] = {0.5 * (b2 + tr), -t2 * vxm, -0.5 * (tb + rcm), t2,
                         -vxm, 1.0, 0.0, 0.0,
                         1.0 - b2, b1 * vxm, tb, -b1,
                         0.5 * (b2 - tr), -t2 * vxm, -0.5 * (tb - rcm), t2};
        memcpy(R, Ry, sizeof Ry);
        memcpy(L, Ly, sizeof Ly);
    }
}

static inline void matvec4(const double M[16], const double x[4], double y[4]) {
    for (int r = 0; r < 4; ++r)
        y[r] = ((M[4 * r] * x[0] + M[4 * r + 1] * x[1]) + M[4 * r + 2] * x[2]) + M[4 * r + 3] * x[3];
}

/* maxLocalWaveSpeed (Equations_Euler2D.cpp:221-246) + characteristic Rusanov
 * (:310-330): F = 0.5 (f(qL) + f(qR) - R diag(a) L (qR - qL)). */
void orc_rusanov_char(double gamma, int dir, const double ql[4], const double qr[4],
                      const double L[16], const double R[16], double f[4]) {
    double wl[4], wr[4], a[4], fl[4], fr[4], dq[4], ld[4], rd[4];
    orc_primitive(gamma, ql, wl);
    orc_primitive(gamma, qr, wr);
    double cl = sqrt(fabs(gamma * wl[3] / wl[0]));
    double cr = sqrt(fabs(gamma * wr[3] / wr[0]));
    double vl = wl[dir], vr = wr[dir];
    a[0] = fmax(fabs(vl - cl), fabs(vr - cr));
    a[1] = fmax(fabs(vl), fabs(vr));
    a[2] = a[1];
    a[3] = fmax(fabs(vl + cl), fabs(vr + cr));
    /* aM = ... / 1.0 (FV_System2D.cpp:460,485) is a no-op */
    phys_flux(dir, ql, wl, fl);
    phys_flux(dir, qr, wr, fr);
    for (int k = 0; k < 4; ++k) dq[k] = qr[k] - ql[k];
    matvec4(L, dq, ld);
    for (int k = 0; k < 4; ++k) ld[k] = a[k] * ld[k];
    matvec4(R, ld, rd);
    for (int k = 0; k < 4; ++k) f[k] = 0.5 * ((fl[k] + fr[k]) - rd[k]);
}

/* Equations::Euler2D::HLLC, Equations_Euler2D.cpp:372-428.  quirks = 1 keeps
 * the reference's c = sqrt(|p/rho|) (:393-394) and p_L in the right star
 * energy (:423); quirks = 0 is Toro's HLLC with the same Roe-Einfeldt speeds. */
void orc_hllc(double gamma, int dir, int quirks, const double ql[4], const double qr[4],
              double f[4]) {
    double g1 = gamma - 1.0;
    int d1 = dir - 1;
    double wl[4], wr[4], fl[4], fr[4], qs[4], velm[2];
    orc_primitive(gamma, ql, wl);
    orc_primitive(gamma, qr, wr);
    double hl = (ql[3] + wl[3]) / wl[0];
    double hr = (qr[3] + wr[3]) / wr[0];
    double rhoRatio = sqrt(fabs(wr[0] / wl[0]));
    double rr = 1.0 / (1.0 + rhoRatio);
    velm[0] = (wl[1] + wr[1] * rhoRatio) * rr;
    velm[1] = (wl[2] + wr[2] * rhoRatio) * rr;
    double qm = 0.5 * (velm[0] * velm[0] + velm[1] * velm[1]);
    double hm = (hl + hr * rhoRatio) * rr;
    double cm = sqrt(fabs(g1 * (hm - qm)));
    double cl, cr;
    if (quirks) {
        cl = sqrt(fabs(wl[3] / wl[0]));
        cr = sqrt(fabs(wr[3] / wr[0]));
    } else {
        cl = sqrt(fabs(gamma * wl[3] / wl[0]));
        cr = sqrt(fabs(gamma * wr[3] / wr[0]));
    }
    double sl = fmin(wl[dir] - cl, velm[d1] - cm);
    double sr = fmax(wr[dir] + cr, velm[d1] + cm);
    double sm = (qr[dir] * (sr - wr[dir]) - ql[dir] * (sl - wl[dir]) + wl[3] - wr[3]) /
                (wr[0] * (sr - wr[dir]) - wl[0] * (sl - wl[dir]));
    double ps = wl[3] + wl[0] * (wl[dir] - sl) * (wl[dir] - sm);
    phys_flux(dir, ql, wl, fl);
    phys_flux(dir, qr, wr, fr);
    if (sl > 0.0) {
        memcpy(f, fl, sizeof fl);
    } else if (sl <= 0.0 && sm > 0.0) {
        double nom = sl - wl[dir], den = sl - sm;
        qs[0] = wl[0] * nom / den;
        qs[1] = dir == 1 ? (nom * ql[1] + ps - wl[3]) / den : nom * ql[1] / den;
        qs[2] = dir == 2 ? (nom * ql[2] + ps - wl[3]) / den : nom * ql[2] / den;
        qs[3] = (nom * ql[3] - wl[3] * wl[dir] + ps * sm) / den;
        for (int k = 0; k < 4; ++k) f[k] = fl[k] + sl * (qs[k] - ql[k]);
    } else if (sm <= 0.0 && sr >= 0.0) {
        double nom = sr - wr[dir], den = sr - sm;
        double pe = quirks ? wl[3] : wr[3];
        qs[0] = wr[0] * nom / den;
        qs[1] = dir == 1 ? (nom * qr[1] + ps - wr[3]) / den : nom * qr[1] / den;
        qs[2] = dir == 2 ? (nom * qr[2] + ps - wr[3]) / den : nom * qr[2] / den;
        qs[3] = (nom * qr[3] - pe * wr[dir] + ps * sm) / den;
        for (int k = 0; k < 4; ++k) f[k] = fr[k] + sr * (qs[k] - qr[k]);
    } else {
        memcpy(f, fr, sizeof fr);
    }
}

/* Euler2D::maxSpeed, Equations_Euler2D.cpp:207-219 (quirk: no fabs on u-c) */
double orc_max_speed(double gamma, int dir, int quirks, const double w[4]) {
    double c = sqrt(fabs(gamma * w[3] / w[0]));
    if (quirks) return fmax((w[dir] - c), fabs(w[dir] + c));
    return fmax(fabs(w[dir] - c), fabs(w[dir] + c));
}

/* SSPRK::RungeKuttaIII, Schemes_SSPRK.cpp:3-15 */
void orc_rk3_combine(int stage, const double* q0, const double* qs, const double* rhs, double dt,
                     double* out, long n) {
    for (long i = 0; i < n; ++i) {
        if (stage == 1)
            out[i] = q0[i] + dt * rhs[i];
        else if (stage == 2)
            out[i] = 0.75 * q0[i] + 0.25 * (qs[i] + rhs[i] * dt);
        else
            out[i] = (q0[i] + 2.0 * (qs[i] + dt * rhs[i])) / 3.0;
    }
}

/* ------------------------------------------------------------------------ */
/* FV::System2D (FV_System2D.cpp)                                            */
/* ------------------------------------------------------------------------ */

#define FV_K 6 /* ghost width, FV_System2D.cpp:20 */
#define FV_S 2 /* (k-1)/2, :21 */

typedef struct {
    int nx, ny, k, nxx, nyy;
    double *q[12];  /* ghost-padded fields, idx = j*nxx + i */
} padded;

static void pad_alloc(padded* P, int nx, int ny, int k, int nf) {
    P->nx = nx; P->ny = ny; P->k = k;
    P->nxx = nx + 2 * k; P->nyy = ny + 2 * k;
    for (int f = 0; f < nf; ++f) P->q[f] = (double*)calloc((size_t)P->nxx * P->nyy, sizeof(double));
}
static void pad_free(padded* P, int nf) {
    for (int f = 0; f < nf; ++f) free(P->q[f]);
}
#define PIX(P, i, j) ((size_t)(j) * (P)->nxx + (i))

/* zeroGradBc (FV_System2D.cpp:134-152) / periodicBc (:114-132) on one field */
static void fill_ghosts(padded* P, double* a, int bc) {
    int k = P->k, nx = P->nx, ny = P->ny, nxx = P->nxx, nyy = P->nyy;
    for (int j = k; j < ny + k; ++j)
        for (int m = 0; m < k; ++m) {
            if (bc == ORC_BC_ZERO_GRAD) {
                a[PIX(P, m, j)] = a[PIX(P, k, j)];
                a[PIX(P, nx + k + m, j)] = a[PIX(P, nx + k - 1, j)];
            } else {
                a[PIX(P, k - 1 - m, j)] = a[PIX(P, nx + k - 1 - m, j)];
                a[PIX(P, nxx - 1 - m, j)] = a[PIX(P, k + k - 1 - m, j)];
            }
        }
    for (int i = 0; i < nxx; ++i)
        for (int m = 0; m < k; ++m) {
            if (bc == ORC_BC_ZERO_GRAD) {
                a[PIX(P, i, m)] = a[PIX(P, i, k)];
                a[PIX(P, i, ny + k + m)] = a[PIX(P, i, ny + k - 1)];
            } else {
                a[PIX(P, i, k - 1 - m)] = a[PIX(P, i, ny + k - 1 - m)];
                a[PIX(P, i, nyy - 1 - m)] = a[PIX(P, i, k + k - 1 - m)];
            }
        }
}

static void load_padded(padded* P, const double* q, int nf, int bc) {
    size_t n = (size_t)P->nx * P->ny;
    for (int f = 0; f < nf; ++f) {
        for (int j = 0; j < P->ny; ++j)
            memcpy(P->q[f] + PIX(P, P->k, j + P->k), q + f * n + (size_t)j * P->nx,
                   sizeof(double) * P->nx);
        fill_ghosts(P, P->q[f], bc);
    }
}

static inline void cell_q(const padded* P, int i, int j, double out[4]) {
    size_t id = PIX(P, i, j);
    for (int f = 0; f < 4; ++f) out[f] = P->q[f][id];
}

/* FV::System2D::rhs (FV_System2D.cpp:363-573): characteristic TENO5 in x and
 * y (:372-400), 3-point transverse quadrature of the characteristic face
 * values of rows j-2..j+2 (:402-423), back-projection with the face's own R
 * (:425-430, FV_System2D.h:50-58), flux at the 3 Gauss points weighted by
 * (5/18, 4/9, 5/18) (:432-497), divergence (:562-572). */
/* linear = 1: FV::System2D::rhs with Lagrange::linear (Schemes_Lagrange.cpp
 * :1219-1233) in place of teno5 and orc_weno5js3q_linear in place of the
 * transverse WENO quadrature (what the learned path computes where every
 * face is flagged smooth). */
static void fv2d_rhs_impl(const orc_params* p, const double* q, double* rhs, int linear);
static void teno5i_impl(const double u[6], double out[2]);
static void weno5_impl(const double u[6], double out[2], int z);

/* the reconstruction FV_System2D.cpp:379 / :394 calls (lagrangeObj.teno5),
 * or one of the alternatives a user of the reference swaps in there
 * (Schemes_Lagrange.cpp:358-408, :1235-1263) */
static void recon2d(int recon, int linear, const double u[6], double diff, double out[2]) {
    if (linear)
        orc_linear5(u, out);
    else
        orc_recon(recon, u, diff, out);
}

void orc_fv2d_rhs(const orc_params* p, const double* q, double* rhs) { fv2d_rhs_impl(p, q, rhs, 0); }

void orc_fv2d_rhs_linear(const orc_params* p, const double* q, double* rhs) { fv2d_rhs_impl(p, q, rhs, 1); }

static void fv2d_rhs_impl(const orc_params* p, const double* q, double* rhs, int linear) {
    lag_init();
    padded P;
    int nx = p->nx, ny = p->ny, k = FV_K, s = FV_S;
    double gamma = p->gamma;
    double dx = (p->xr - p->xl) / nx, dy = (p->yt - p->yb) / ny, area = dx * dy;
    double gw[3] = {5.0 / 9.0 / 2.0, 8.0 / 9.0 / 2.0, 5.0 / 9.0 / 2.0}; /* :28 */
    pad_alloc(&P, nx, ny, k, 4);
    load_padded(&P, q, 4, p->bc);
    int nxx = P.nxx, nyy = P.nyy;
    size_t nt = (size_t)nxx * nyy;
    /* face arrays (stored at the left/lower cell, like the reference) */
    double* uxl = (double*)calloc(4 * nt, sizeof(double));
    double* uxr = (double*)calloc(4 * nt, sizeof(double));
    double* uyl = (double*)calloc(4 * nt, sizeof(double));
    double* uyr = (double*)calloc(4 * nt, sizeof(double));
    double* fx = (double*)calloc(4 * nt, sizeof(double));
    double* fy = (double*)calloc(4 * nt, sizeof(double));

    /* x reconstruction: FV_System2D.cpp:372-385 */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nyy; ++j) {
        for (int i = s; i < nxx - s - 1; ++i) {
            double qa[4], qb[4], L[16], R[16], C[4][6], cm[4], out[2];
            cell_q(&P, i, j, qa);
            cell_q(&P, i + 1, j, qb);
            orc_eigen(gamma, 1, qa, qb, L, R);
            for (int m = 0; m < 6; ++m) {
                cell_q(&P, i - 2 + m, j, cm);
                double y[4];
                matvec4(L, cm, y);
                for (int e = 0; e < 4; ++e) C[e][m] = y[e];
            }
            for (int e = 0; e < 4; ++e) {
                recon2d(p->recon, linear, C[e], dx, out);
                uxl[e * nt + PIX(&P, i, j)] = out[0];
                uxr[e * nt + PIX(&P, i, j)] = out[1];
            }
        }
    }
    /* y reconstruction: :387-400 */
#pragma omp parallel for schedule(static)
    for (int j = s; j < nyy - s - 1; ++j) {
        for (int i = 0; i < nxx; ++i) {
            double qa[4], qb[4], L[16], R[16], C[4][6], cm[4], out[2];
            cell_q(&P, i, j, qa);
            cell_q(&P, i, j + 1, qb);
            orc_eigen(gamma, 2, qa, qb, L, R);
            for (int m = 0; m < 6; ++m) {
                cell_q(&P, i, j - 2 + m, cm);
                double y[4];
                matvec4(L, cm, y);
                for (int e = 0; e < 4; ++e) C[e][m] = y[e];
            }
            for (int e = 0; e < 4; ++e) {
                recon2d(p->recon, linear, C[e], dy, out);
                uyl[e * nt + PIX(&P, i, j)] = out[0];
                uyr[e * nt + PIX(&P, i, j)] = out[1];
            }
        }
    }
    /* transverse quadrature + back-projection + flux: :402-497.  The
     * quadrature loop runs over [s+2, n-s-3) and the flux loop over
     * [k-1, n-k); only the latter range is consumed, so both are fused. */
#pragma omp parallel for schedule(static)
    for (int j = k - 1; j < nyy - k; ++j) {
        for (int i = k - 1; i < nxx - k; ++i) {
            for (int dir = 1; dir <= 2; ++dir) {
                double qa[4], qb[4], L[16], R[16];
                double gl[4][3], gr[4][3];
                cell_q(&P, i, j, qa);
                if (dir == 1) cell_q(&P, i + 1, j, qb); else cell_q(&P, i, j + 1, qb);
                orc_eigen(gamma, dir, qa, qb, L, R);
                for (int e = 0; e < 4; ++e) {
                    double a5[5], b5[5];
                    for (int m = 0; m < 5; ++m) {
                        size_t id = dir == 1 ? PIX(&P, i, j - 2 + m) : PIX(&P, i - 2 + m, j);
                        a5[m] = (dir == 1 ? uxl : uyl)[e * nt + id];
                        b5[m] = (dir == 1 ? uxr : uyr)[e * nt + id];
                    }
                    if (linear) {
                        orc_weno5js3q_linear(a5, gl[e]);
                        orc_weno5js3q_linear(b5, gr[e]);
                    } else {
                        orc_weno5js3q(a5, gl[e]);
                        orc_weno5js3q(b5, gr[e]);
                    }
                }
                double tf[3][4];
                for (int m = 0; m < 3; ++m) {
                    double cl[4], cr[4], ql[4], qr[4], F[4];
                    for (int e = 0; e < 4; ++e) { cl[e] = gl[e][m]; cr[e] = gr[e][m]; }
                    matvec4(R, cl, ql);
                    matvec4(R, cr, qr);
                    if (p->flux == ORC_FLUX_HLLC)
                        orc_hllc(gamma, dir, p->quirks, ql, qr, F);
                    else
                        orc_rusanov_char(gamma, dir, ql, qr, L, R, F);
                    for (int e = 0; e < 4; ++e) tf[m][e] = gw[m] * F[e];
                }
                double* fd = dir == 1 ? fx : fy;
                for (int e = 0; e < 4; ++e)
                    fd[e * nt + PIX(&P, i, j)] = (tf[0][e] + tf[1][e]) + tf[2][e];
            }
        }
    }
    /* divergence: :562-572 */
    size_t n = (size_t)nx * ny;
    for (int j = k; j < nyy - k; ++j)
        for (int i = k; i < nxx - k; ++i)
            for (int e = 0; e < 4; ++e) {
                double v = ((fx[e * nt + PIX(&P, i - 1, j)] - fx[e * nt + PIX(&P, i, j)]) * dy +
                            (fy[e * nt + PIX(&P, i, j - 1)] - fy[e * nt + PIX(&P, i, j)]) * dx) /
                           area;
                rhs[e * n + (size_t)(j - k) * nx + (i - k)] = v;
            }
    free(uxl); free(uxr); free(uyl); free(uyr); free(fx); free(fy);
    pad_free(&P, 4);
}

/* CFL reduction: FV::System2D::timeStepping (FV_System2D.cpp:623-631) */
/* What the cut-off decisions of FV::System2D::rhs's reconstruction
 * (FV_System2D.cpp:372-400) do to every face of one family: dir 1 = the x
 * face between interior cells (i, j) and (i+1, j), dir 2 = the y face
 * between (i, j) and (i, j+1), labels indexed j * nx + i; label 1 when, for
 * every characteristic field and both sides, teno5 (Schemes_Lagrange.cpp
 * :481-511) gives the value of linear (:1219-1233) to within
 * 1e-12 M + 1e-9 sum|u_{k+1} - u_k| of the field's 6-value stencil, M = the
 * largest |value| of any field on the stencil.  (The learned classifier's
 * training target; the reference has no such label.) */
void orc_face_labels(const orc_params* p, const double* q, int dir, unsigned char* labels) {
    lag_init();
    padded P;
    int nx = p->nx, ny = p->ny, k = FV_K;
    pad_alloc(&P, nx, ny, k, 4);
    load_padded(&P, q, 4, p->bc);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int pi = i + k, pj = j + k;
            double qa[4], qb[4], L[16], R[16], cc[4], y[4], C[4][6];
            cell_q(&P, pi, pj, qa);
            if (dir == 1) cell_q(&P, pi + 1, pj, qb); else cell_q(&P, pi, pj + 1, qb);
            orc_eigen(p->gamma, dir, qa, qb, L, R);
            for (int m = 0; m < 6; ++m) {
                if (dir == 1) cell_q(&P, pi - 2 + m, pj, cc); else cell_q(&P, pi, pj - 2 + m, cc);
                matvec4(L, cc, y);
                for (int e = 0; e < 4; ++e) C[e][m] = y[e];
            }
            int all = 1;
            double mx = 0.0;
            for (int e = 0; e < 4; ++e)
                for (int m = 0; m < 6; ++m) mx = fmax(mx, fabs(C[e][m]));
            for (int e = 0; e < 4; ++e) {
                double tv[2], lv[2], var = 0.0;
                teno5_impl(C[e], tv, NULL);
                orc_linear5(C[e], lv);
                for (int m = 0; m < 5; ++m) var += fabs(C[e][m + 1] - C[e][m]);
                double tol = 1e-12 * mx + 1e-9 * var;
                if (!(fabs(tv[0] - lv[0]) <= tol && fabs(tv[1] - lv[1]) <= tol)) all = 0;
            }
            labels[(size_t)j * nx + i] = (unsigned char)all;
        }
    pad_free(&P, 4);
}

void orc_fv2d_max_speed(const orc_params* p, const double* q, double* sx, double* sy) {
    size_t n = (size_t)p->nx * p->ny;
    double mx = 0.0, my = 0.0;
    for (size_t c = 0; c < n; ++c) {
        double qc[4] = {q[c], q[n + c], q[2 * n + c], q[3 * n + c]}, w[4];
        orc_primitive(p->gamma, qc, w);
        mx = fmax(mx, orc_max_speed(p->gamma, 1, p->quirks, w));
        my = fmax(my, orc_max_speed(p->gamma, 2, p->quirks, w));
    }
    *sx = mx;
    *sy = my;
}

/* dt = cfl * min(dx,dy)^(3/3) / max(sx,sy), clamped to tFinal (:633,645) */
double orc_fv2d_dt(const orc_params* p, const double* q, double cfl, double t, double tfinal) {
    double sx, sy;
    double dx = (p->xr - p->xl) / p->nx, dy = (p->yt - p->yb) / p->ny;
    orc_fv2d_max_speed(p, q, &sx, &sy);
    double dt = cfl * pow(fmin(dx, dy), 3.0 / 3.0) / fmax(sx, sy);
    return t + dt > tfinal ? tfinal - t : dt;
}

/* FV::System2D::RungeKutta3, FV_System2D.cpp:575-612 */
void orc_fv2d_rk3(const orc_params* p, double* q, double dt) {
    size_t n = 4 * (size_t)p->nx * p->ny;
    double* q0 = (double*)malloc(n * sizeof(double));
    double* r = (double*)malloc(n * sizeof(double));
    double* qs = (double*)malloc(n * sizeof(double));
    memcpy(q0, q, n * sizeof(double));
    orc_fv2d_rhs(p, q, r);
    orc_rk3_combine(1, q0, q0, r, dt, q, (long)n);
    memcpy(qs, q, n * sizeof(double));
    orc_fv2d_rhs(p, q, r);
    orc_rk3_combine(2, q0, qs, r, dt, q, (long)n);
    memcpy(qs, q, n * sizeof(double));
    orc_fv2d_rhs(p, q, r);
    orc_rk3_combine(3, q0, qs, r, dt, q, (long)n);
    free(q0); free(r); free(qs);
}

/* FV::System2D::timeStepping, FV_System2D.cpp:614-651 */
int orc_fv2d_advance(const orc_params* p, double* q, double cfl, double t0, double tfinal,
                     int max_steps, double* t_out) {
    double t = t0;
    int nsteps = 0;
    while (t < tfinal && (max_steps <= 0 || nsteps < max_steps)) {
        double dt = orc_fv2d_dt(p, q, cfl, t, tfinal);
        orc_fv2d_rk3(p, q, dt);
        t += dt;
        ++nsteps;
    }
    if (t_out) *t_out = t;
    return nsteps;
}

/* ------------------------------------------------------------------------ */
/* Initial conditions                                                        */
/* ------------------------------------------------------------------------ */

/* conservativeOfIsentropicVortex (Equations_Euler2D.cpp:656-666), centred at
 * (cx, cy), plus an optional density perturbation sum_m a_m sin(2 pi (k_m x +
 * l_m y) / Lx + phi_m) at fixed velocity and pressure.  modes = nmodes rows of
 * (k, l, a, phi, Lx, Ly). */
static void vortex_point(double gamma, double x, double y, double cx, double cy,
                         const double* modes, int nmodes, double q[4]) {
    const double pi = 4.0 * atan(1.0);
    double g1 = gamma - 1.0;
    double xr = x - cx, yr = y - cy;
    double w[4];
    double r = pow(xr, 2) + pow(yr, 2), b = 5.0;
    w[0] = 1.0 - g1 * pow(b, 2) / 8.0 / gamma / pow(pi, 2) * exp(1 - r);
    w[0] = pow(w[0], 1.0 / g1);
    w[1] = 1.0 - b * yr / 2.0 / pi * exp(0.5 * (1 - r));
    w[2] = 1.0 + b * xr / 2.0 / pi * exp(0.5 * (1 - r));
    w[3] = pow(w[0], gamma);
    for (int m = 0; m < nmodes; ++m) {
        const double* md = modes + 6 * m;
        w[0] += md[2] * sin(2.0 * pi * (md[0] * x / md[4] + md[1] * y / md[5]) + md[3]);
    }
    orc_conservative(gamma, w, q);
}

/* Euler2D::isentropicVortex: 9x9 Gauss cell average (Equations_Euler2D.cpp
 * :513-528), cell centres as FV_System2D.cpp:37-38. */
void orc_ic_vortex(const orc_params* p, double cx, double cy, const double* modes, int nmodes,
                   double* q) {
    double a[9], wt[9];
    orc_gauss9(a, wt);
    int nx = p->nx, ny = p->ny;
    double dx = (p->xr - p->xl) / nx, dy = (p->yt - p->yb) / ny;
    size_t n = (size_t)nx * ny;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double x = ((double)i + 0.5) * dx + p->xl;
            double y = ((double)j + 0.5) * dy + p->yb;
            double acc[4] = {0, 0, 0, 0};
            for (int ii = 0; ii < 9; ++ii)
                for (int jj = 0; jj < 9; ++jj) {
                    double qq[4];
                    vortex_point(p->gamma, x + dx * a[ii] / 2.0, y + dy * a[jj] / 2.0, cx, cy, modes,
                                 nmodes, qq);
                    for (int e = 0; e < 4; ++e) acc[e] += wt[ii] * wt[jj] * qq[e] / 4.0;
                }
            for (int e = 0; e < 4; ++e) q[e * n + (size_t)j * nx + i] = acc[e];
        }
}

/* Euler2D::entropyWaves, entropyWavesdx, entropyWavesdy
 * (Equations_Euler2D.cpp:430-511): 9x9 Gauss cell averages of the entropy
 * wave's conserved state (rho + amp sin(pi (x + y)), u, v, p) and, with
 * hermite, of its derivative (fields 4..7 and 8..11, equal: the density
 * depends on x + y only).  The cell centre is shifted by (-sx, -sy) before
 * the Gauss points are spread, as the reference forms its exact solution
 * (FV_System2D.cpp:64-71). */
void orc_ic_entropy(const orc_params* p, double rho, double u, double v, double pr, double amp, double sx,
                    double sy, int hermite, double* q) {
    double a[9], wt[9];
    orc_gauss9(a, wt);
    const double pi = 4.0 * atan(1.0), g1 = p->gamma - 1.0;
    int nx = p->nx, ny = p->ny;
    double dx = (p->xr - p->xl) / nx, dy = (p->yt - p->yb) / ny;
    size_t n = (size_t)nx * ny;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double x = ((double)i + 0.5) * dx + p->xl - sx;
            double y = ((double)j + 0.5) * dy + p->yb - sy;
            double acc[4] = {0, 0, 0, 0}, dacc[4] = {0, 0, 0, 0};
            for (int ii = 0; ii < 9; ++ii)
                for (int jj = 0; jj < 9; ++jj) {
                    double px = x + dx * a[ii] / 2.0, py = y + dy * a[jj] / 2.0;
                    double w[4] = {rho + amp * sin(pi * (px + py)), u, v, pr}, qq[4], d[4];
                    orc_conservative(p->gamma, w, qq);
                    (void)g1;
                    d[0] = amp * cos(pi * (px + py)) * pi;
                    d[1] = u * d[0];
                    d[2] = v * d[0];
                    d[3] = (0.5 * (u * u + v * v)) * d[0];
                    for (int e = 0; e < 4; ++e) {
                        acc[e] += wt[ii] * wt[jj] * qq[e] / 4.0;
                        dacc[e] += wt[ii] * wt[jj] * d[e] / 4.0;
                    }
                }
            for (int e = 0; e < 4; ++e) {
                q[e * n + (size_t)j * nx + i] = acc[e];
                if (hermite) {
                    q[(4 + e) * n + (size_t)j * nx + i] = dacc[e];
                    q[(8 + e) * n + (size_t)j * nx + i] = dacc[e];
                }
            }
        }
}

/* ------------------------------------------------------------------------ */
/* Schemes::Hermite (Schemes_Hermite.cpp)                                    */
/* ------------------------------------------------------------------------ */

/* coefficients of setGridSize (Schemes_Hermite.cpp:31-66) for grid size dif */
typedef struct {
    double crj[3][3], crjd[4][3];
    double dkm[3], dkp[3], dkmd[3], dkpd[3];
    double dif;
} herm_consts;

static void herm_init(herm_consts* H, double dif) {
    H->dif = dif;
    H->crj[0][0] = -7.0 / 6.0; H->crj[1][0] = 13.0 / 6.0; H->crj[2][0] = -2.0 / 3.0 * dif;
    H->crj[0][1] = 1.0 / 6.0;  H->crj[1][1] = 5.0 / 6.0;  H->crj[2][1] = -1.0 / 3.0 * dif;
    H->crj[0][2] = -1.0 / 6.0; H->crj[1][2] = 5.0 / 6.0;  H->crj[2][2] = 1.0 / 3.0;
    H->crjd[0][0] = 4.0 / dif;  H->crjd[1][0] = -4.0 / dif; H->crjd[2][0] = 1.5;        H->crjd[3][0] = 3.5;
    H->crjd[0][1] = -2.0 / dif; H->crjd[1][1] = 2.0 / dif;  H->crjd[2][1] = -0.5;       H->crjd[3][1] = -0.5;
    H->crjd[0][2] = 0.25 / dif; H->crjd[1][2] = -1.0 / dif; H->crjd[2][2] = 0.75 / dif; H->crjd[3][2] = 0.5;
    H->dkm[0] = 9.0 / 80.0; H->dkm[1] = 42.0 / 80.0; H->dkm[2] = 29.0 / 80.0;
    H->dkp[0] = H->dkm[1]; H->dkp[1] = H->dkm[0]; H->dkp[2] = H->dkm[2];
    H->dkmd[0] = 1.0 / 18.0; H->dkmd[1] = 5.0 / 6.0; H->dkmd[2] = 1.0 / 9.0;
    H->dkpd[0] = H->dkmd[1]; H->dkpd[1] = H->dkmd[0]; H->dkpd[2] = H->dkmd[2];
}

/* Hermite::smoothnessMeasure, Schemes_Hermite.cpp:515-526 */
static void herm_si(const double* u, const double* v, double SI[3]) {
    double b[3], c[3];
    b[0] = -2.0 * (u[0] - u[1]) - v[0];
    c[0] = -2.0 * (u[0] - u[1] + v[0]);
    b[1] = -2.0 * (u[1] - u[2]) - v[2];
    c[1] = 2.0 * (u[1] - u[2]) + 2.0 * v[2];
    b[2] = (-u[0] + u[2]) / 2.0;
    c[2] = u[0] - 2.0 * u[1] + u[2];
    for (int k = 0; k < 3; ++k) SI[k] = LC.c1 * (c[k] * c[k]) + b[k] * b[k];
}

/* Hermite::smoothnessMeasureDerivative, :528-541 */
static void herm_sid(const double* u, const double* d, double SI[3]) {
    double b[3], c[3];
    b[0] = 6.0 * (u[0] - u[1]) + (2.0 * d[0] + 4.0 * d[1]);
    c[0] = 12.0 * (u[0] - u[1]) + 6.0 * (d[0] + d[1]);
    b[1] = 6.0 * (u[2] - u[1]) - (4.0 * d[1] + 2.0 * d[2]);
    c[1] = 12.0 * (u[1] - u[2]) + 6.0 * (d[1] + d[2]);
    b[2] = u[0] - 2.0 * u[1] + u[2];
    c[2] = 3.0 * (u[2] - u[0]) - 6.0 * d[1];
    for (int k = 0; k < 3; ++k) SI[k] = LC.c1 * (c[k] * c[k]) + b[k] * b[k];
}

/* Hermite::fullPointsSI, :597-603 */
static double herm_brk(const double* u, const double* d) {
    return fabs((1029270.0 * u[2] - 2129184.0 * u[1] + 1099914.0 * u[0] + 281121.0 * d[0] - 494958.0 * d[2]) * d[0] +
                (-1029270.0 * u[0] + 2129184.0 * u[1] - 1099914.0 * u[2] + 281121.0 * d[2]) * d[2] +
                (1099445.0 * u[2] - 4317056.0 * u[1] + 2118166.0 * u[0]) * u[2] +
                4317056.0 * (u[1] - u[0]) * u[1] + 1099445.0 * u[0] * u[0]) /
           6720.0;
}

/* Hermite::fullPointsSId, :605-612 */
static double herm_brkd(const double* u, const double* d) {
    return fabs((3156862.0 * d[2] + 13189240.0 * d[1] + 1710479.0 * d[0] - 9617754.0 * u[2] - 531552.0 * u[1] +
                 10149306.0 * u[0]) * d[0] +
                (13189240.0 * d[2] + 26511200.0 * d[1] - 39700440.0 * u[2] + 39700440.0 * u[0]) * d[1] +
                (1710479.0 * d[2] - 10149306.0 * u[2] + 531552.0 * u[1] + 9617754.0 * u[0]) * d[2] +
                (15136011.0 * u[2] - 1076544.0 * u[1] - 29195478.0 * u[0]) * u[2] +
                1076544.0 * (u[1] - u[0]) * u[1] + 15136011.0 * u[0] * u[0]) /
           1680.0;
}

/* one HTENO side: tau from the full-stencil indicator, gamma^6, cutOff at
 * C_T = 1e-3 (:614-628), weights, catchnan */
static double hteno_side(double brk, const double SI[3], const double p[3], const double d[3], int* sig_out) {
    double w[3], sigm[3];
    double tau = fabs(brk - 0.25 * (SI[0] + SI[1] + 2.0 * SI[2]));
    for (int k = 0; k < 3; ++k) w[k] = pow(1.0 + tau / (SI[k] + 1E-40), 6);
    double s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    for (int k = 0; k < 3; ++k) sigm[k] = w[k] < 1E-3 ? 0.0 : 1.0;
    for (int k = 0; k < 3; ++k) w[k] = sigm[k] * d[k];
    s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    if (sig_out)
        for (int k = 0; k < 3; ++k) sig_out[k] = (int)sigm[k];
    return catchnan(dot3(p, w));
}

/* Hermite::hteno, Schemes_Hermite.cpp:295-379: cell averages u[0..3] =
 * u_{i-1..i+2} and derivative moments d[0..3] -> (u^-, u^+, d^-, d^+) at
 * x_{i+1/2}. */
static void hteno_impl(const herm_consts* H, const double u[4], const double d[4], double out[4], int* flags) {
    double ud[4], p[3], SI[3];
    for (int k = 0; k < 4; ++k) ud[k] = d[k] * H->dif;
    /* u^- :304-320 */
    p[0] = H->crj[0][0] * u[0] + H->crj[1][0] * u[1] + H->crj[2][0] * d[0];
    p[1] = H->crj[0][1] * u[1] + H->crj[1][1] * u[2] + H->crj[2][1] * d[2];
    p[2] = H->crj[0][2] * u[0] + H->crj[1][2] * u[1] + H->crj[2][2] * u[2];
    herm_si(u, ud, SI);
    out[0] = hteno_side(herm_brk(u, ud), SI, p, H->dkm, flags ? flags : NULL);
    /* d^- :323-339 */
    p[0] = H->crjd[0][0] * u[0] + H->crjd[1][0] * u[1] + H->crjd[2][0] * d[0] + H->crjd[3][0] * d[1];
    p[1] = H->crjd[0][1] * u[1] + H->crjd[1][1] * u[2] + H->crjd[2][1] * d[1] + H->crjd[3][1] * d[2];
    p[2] = H->crjd[0][2] * u[0] + H->crjd[1][2] * u[1] + H->crjd[2][2] * u[2] + H->crjd[3][2] * d[1];
    herm_sid(u, ud, SI);
    out[2] = hteno_side(herm_brkd(u, ud), SI, p, H->dkmd, flags ? flags + 6 : NULL);
    /* u^+ :342-357 */
    p[0] = H->crj[1][1] * u[1] + H->crj[0][1] * u[2] - H->crj[2][1] * d[1];
    p[1] = H->crj[1][0] * u[2] + H->crj[0][0] * u[3] - H->crj[2][0] * d[3];
    p[2] = H->crj[2][2] * u[1] + H->crj[1][2] * u[2] + H->crj[0][2] * u[3];
    herm_si(u + 1, ud + 1, SI);
    out[1] = hteno_side(herm_brk(u + 1, ud + 1), SI, p, H->dkp, flags ? flags + 3 : NULL);
    /* d^+ :360-376 */
    p[0] = H->crjd[0][1] * u[1] + H->crjd[1][1] * u[2] + H->crjd[3][1] * d[1] + H->crjd[2][1] * d[2];
    p[1] = H->crjd[0][0] * u[2] + H->crjd[1][0] * u[3] + H->crjd[3][0] * d[2] + H->crjd[2][0] * d[3];
    p[2] = -(H->crjd[2][2] * u[1] + H->crjd[1][2] * u[2] + H->crjd[0][2] * u[3]) + H->crjd[3][2] * d[2];
    herm_sid(u + 1, ud + 1, SI);
    out[3] = hteno_side(herm_brkd(u + 1, ud + 1), SI, p, H->dkpd, flags ? flags + 9 : NULL);
}

void orc_hteno(double dif, const double u[4], const double d[4], double out[4]) {
    herm_consts H;
    lag_init();
    herm_init(&H, dif);
    hteno_impl(&H, u, d, out, NULL);
}

void orc_hteno_flags(double dif, const double u[4], const double d[4], int flags[12]) {
    herm_consts H;
    double out[4];
    lag_init();
    herm_init(&H, dif);
    hteno_impl(&H, u, d, out, flags);
}

/* Hermite::hweno, Schemes_Hermite.cpp:234-293: hteno's candidates with
 * weights dk / (SI + eps)^2 normalised, catchnan; eps = 1e-6 (:5-6) */
static void hweno_impl(const herm_consts* H, const double u[4], const double d[4], double out[4]) {
    double ud[4], p[3], SI[3], w[3], s;
    const double eps = 1E-6;
    for (int k = 0; k < 4; ++k) ud[k] = d[k] * H->dif;
    for (int o = 0; o < 4; ++o) {
        const double* dk;
        if (o == 0) { /* u^- :238-251 */
            p[0] = H->crj[0][0] * u[0] + H->crj[1][0] * u[1] + H->crj[2][0] * d[0];
            p[1] = H->crj[0][1] * u[1] + H->crj[1][1] * u[2] + H->crj[2][1] * d[2];
            p[2] = H->crj[0][2] * u[0] + H->crj[1][2] * u[1] + H->crj[2][2] * u[2];
            herm_si(u, ud, SI);
            dk = H->dkm;
        } else if (o == 1) { /* u^+ :254-266 */
            p[0] = H->crj[1][1] * u[1] + H->crj[0][1] * u[2] - H->crj[2][1] * d[1];
            p[1] = H->crj[1][0] * u[2] + H->crj[0][0] * u[3] - H->crj[2][0] * d[3];
            p[2] = H->crj[2][2] * u[1] + H->crj[1][2] * u[2] + H->crj[0][2] * u[3];
            herm_si(u + 1, ud + 1, SI);
            dk = H->dkp;
        } else if (o == 2) { /* d^- :269-280 */
            p[0] = H->crjd[0][0] * u[0] + H->crjd[1][0] * u[1] + H->crjd[2][0] * d[0] + H->crjd[3][0] * d[1];
            p[1] = H->crjd[0][1] * u[1] + H->crjd[1][1] * u[2] + H->crjd[2][1] * d[1] + H->crjd[3][1] * d[2];
            p[2] = H->crjd[0][2] * u[0] + H->crjd[1][2] * u[1] + H->crjd[2][2] * u[2] + H->crjd[3][2] * d[1];
            herm_sid(u, ud, SI);
            dk = H->dkmd;
        } else { /* d^+ :283-290 */
            p[0] = H->crjd[0][1] * u[1] + H->crjd[1][1] * u[2] + H->crjd[3][1] * d[1] + H->crjd[2][1] * d[2];
            p[1] = H->crjd[0][0] * u[2] + H->crjd[1][0] * u[3] + H->crjd[3][0] * d[2] + H->crjd[2][0] * d[3];
            p[2] = -(H->crjd[2][2] * u[1] + H->crjd[1][2] * u[2] + H->crjd[0][2] * u[3]) + H->crjd[3][2] * d[2];
            herm_sid(u + 1, ud + 1, SI);
            dk = H->dkpd;
        }
        for (int k = 0; k < 3; ++k) w[k] = dk[k] / pow(SI[k] + eps, 2);
        s = sum3(w);
        for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
        out[o] = catchnan(dot3(p, w));
    }
}

/* Hermite::htenoi, :670-755: tau = SI.maxCoeff(), weights (C + tau / (SI +
 * epsz))^6, C = 1, cutOff / cutOffd at ct = ctd = 1e-3 (:614-628) */
static void htenoi_impl(const herm_consts* H, const double u[4], const double d[4], double out[4]) {
    double ud[4], p[3], SI[3], w[3], sigm[3], s, tau;
    for (int k = 0; k < 4; ++k) ud[k] = d[k] * H->dif;
    const int order[4] = {0, 2, 1, 3}; /* u^-, d^-, u^+, d^+ as the source runs them */
    for (int n = 0; n < 4; ++n) {
        int o = order[n];
        const double* dk;
        if (o == 0) {
            p[0] = H->crj[0][0] * u[0] + H->crj[1][0] * u[1] + H->crj[2][0] * d[0];
            p[1] = H->crj[0][1] * u[1] + H->crj[1][1] * u[2] + H->crj[2][1] * d[2];
            p[2] = H->crj[0][2] * u[0] + H->crj[1][2] * u[1] + H->crj[2][2] * u[2];
            herm_si(u, ud, SI);
            dk = H->dkm;
        } else if (o == 2) {
            p[0] = H->crjd[0][0] * u[0] + H->crjd[1][0] * u[1] + H->crjd[2][0] * d[0] + H->crjd[3][0] * d[1];
            p[1] = H->crjd[0][1] * u[1] + H->crjd[1][1] * u[2] + H->crjd[2][1] * d[1] + H->crjd[3][1] * d[2];
            p[2] = H->crjd[0][2] * u[0] + H->crjd[1][2] * u[1] + H->crjd[2][2] * u[2] + H->crjd[3][2] * d[1];
            herm_sid(u, ud, SI);
            dk = H->dkmd;
        } else if (o == 1) {
            p[0] = H->crj[1][1] * u[1] + H->crj[0][1] * u[2] - H->crj[2][1] * d[1];
            p[1] = H->crj[1][0] * u[2] + H->crj[0][0] * u[3] - H->crj[2][0] * d[3];
            p[2] = H->crj[2][2] * u[1] + H->crj[1][2] * u[2] + H->crj[0][2] * u[3];
            herm_si(u + 1, ud + 1, SI);
            dk = H->dkp;
        } else {
            p[0] = H->crjd[0][1] * u[1] + H->crjd[1][1] * u[2] + H->crjd[3][1] * d[1] + H->crjd[2][1] * d[2];
            p[1] = H->crjd[0][0] * u[2] + H->crjd[1][0] * u[3] + H->crjd[3][0] * d[2] + H->crjd[2][0] * d[3];
            p[2] = -(H->crjd[2][2] * u[1] + H->crjd[1][2] * u[2] + H->crjd[0][2] * u[3]) + H->crjd[3][2] * d[2];
            herm_sid(u + 1, ud + 1, SI);
            dk = H->dkpd;
        }
        tau = SI[0];
        for (int k = 1; k < 3; ++k) tau = SI[k] > tau ? SI[k] : tau;
        for (int k = 0; k < 3; ++k) w[k] = pow(1.0 + tau / (SI[k] + 1E-40), 6);
        s = sum3(w);
        for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
        for (int k = 0; k < 3; ++k) sigm[k] = w[k] < 1E-3 ? 0.0 : 1.0;
        for (int k = 0; k < 3; ++k) w[k] = sigm[k] * dk[k];
        s = sum3(w);
        for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
        out[o] = catchnan(dot3(p, w));
    }
}

/* Hermite::linear, :757-767, with crjlin / crjlind of setGridSize :67-92 */
static void hlinear_impl(double dif, const double u[4], const double d[4], double out[4]) {
    double crjlin[5][2], crjlind[6][2];
    crjlin[0][0] = -23.0 / 120.0;
    crjlin[1][0] = 19.0 / 30.0;
    crjlin[2][0] = 67.0 / 120.0;
    crjlin[3][0] = -3.0 / 40.0 * dif;
    crjlin[4][0] = -7.0 / 40.0 * dif;
    crjlin[0][1] = 67.0 / 120.0;
    crjlin[1][1] = 19.0 / 30.0;
    crjlin[2][1] = -23.0 / 120.0;
    crjlin[3][1] = 7.0 / 40.0 * dif;
    crjlin[4][1] = 3.0 / 40.0 * dif;
    crjlind[0][0] = 0.25 / dif;
    crjlind[1][0] = -2.0 / dif;
    crjlind[2][0] = 1.75 / dif;
    crjlind[3][0] = 1.0 / 12.0;
    crjlind[4][0] = -1.0 / 6.0;
    crjlind[5][0] = -5.0 / 12.0;
    crjlind[0][1] = -7.0 / 4.0 / dif;
    crjlind[1][1] = 2.0 / dif;
    crjlind[2][1] = -0.25 / dif;
    crjlind[3][1] = -5.0 / 12.0;
    crjlind[4][1] = -1.0 / 6.0;
    crjlind[5][1] = 1.0 / 12.0;
    double c[3], e[3];
    for (int r = 0; r < 3; ++r) c[r] = crjlin[r][0];
    out[0] = dot3(c, u) + crjlin[3][0] * d[0] + crjlin[4][0] * d[2];
    for (int r = 0; r < 3; ++r) c[r] = crjlin[r][1];
    out[1] = dot3(c, u + 1) + crjlin[3][1] * d[1] + crjlin[4][1] * d[3];
    for (int r = 0; r < 3; ++r) {
        c[r] = crjlind[r][0];
        e[r] = crjlind[3 + r][0];
    }
    out[2] = dot3(c, u) + dot3(e, d);
    for (int r = 0; r < 3; ++r) {
        c[r] = crjlind[r][1];
        e[r] = crjlind[3 + r][1];
    }
    out[3] = dot3(c, u + 1) + dot3(e, d + 1);
}

/* the Hermite face reconstruction FVHermite_System2D.cpp:459 / :486 calls
 * (hermiteObj.hteno), or an alternative swapped in: ORC_HRECON_* */
static void herm_recon(const herm_consts* H, int kind, const double u[4], const double d[4], double out[4]) {
    if (kind == ORC_HRECON_HWENO)
        hweno_impl(H, u, d, out);
    else if (kind == ORC_HRECON_HTENOI)
        htenoi_impl(H, u, d, out);
    else if (kind == ORC_HRECON_LINEAR)
        hlinear_impl(H->dif, u, d, out);
    else
        hteno_impl(H, u, d, out, NULL);
}

void orc_herm_recon(int kind, double dif, const double u[4], const double d[4], double out[4]) {
    herm_consts H;
    lag_init();
    herm_init(&H, dif);
    herm_recon(&H, kind, u, d, out);
}

/* b, c of smoothnessMeasure2PointsGLQ / 4PointsGLQ (:543-581; deriv = 0)
 * and of their derivative forms (:556-595; deriv = 1), squared */
static void herm_bc(const double u[3], const double udf[3], int deriv, double b2[3], double c2[3]) {
    double b[3], c[3];
    if (!deriv) {
        b[0] = -2.0 * (u[0] - u[1]) - udf[0];
        c[0] = -2.0 * (u[0] - u[1]) - 2.0 * udf[0];
        b[1] = -2.0 * (u[1] - u[2]) - udf[2];
        c[1] = 2.0 * (u[1] - u[2]) + 2.0 * udf[2];
        b[2] = -(u[0] - u[2]) / 2.0;
        c[2] = u[0] - 2.0 * u[1] + u[2];
    } else {
        b[0] = 6.0 * (u[0] - u[1]) + (2.0 * udf[0] + 4.0 * udf[1]);
        c[0] = 12.0 * (u[0] - u[1]) + 6.0 * (udf[0] + udf[1]);
        b[1] = 6.0 * (u[2] - u[1]) - (4.0 * udf[1] + 2.0 * udf[2]);
        c[1] = 12.0 * (u[1] - u[2]) + 6.0 * (udf[1] + udf[2]);
        b[2] = u[0] - 2.0 * u[1] + u[2];
        c[2] = 3.0 * (u[2] - u[0]) - 6.0 * udf[1];
    }
    for (int k = 0; k < 3; ++k) {
        b2[k] = pow(b[k], 2);
        c2[k] = pow(c[k], 2);
    }
}

static double hjs(const double p[3], const double dk[3], const double SI[3]) {
    double w[3], s;
    for (int k = 0; k < 3; ++k) w[k] = dk[k] / pow(SI[k] + 1E-6, 2);
    s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    return dot3(p, w);
}

/* Hermite::hweno2PointsGLQ :381-425 and hweno4PointsGLQ :427-513, with the
 * coefficient tables of setGridSize :93-231 and c2..c7 of the constructor
 * (:5-13) */
void orc_hweno_glq(int npts, double dif, const double u[3], const double d[3], double* out) {
    const double sq3 = sqrt(3.0), sq5 = sqrt(5.0);
    const double c1 = 13.0 / 12.0, c2 = sqrt(3.0) / 3.0, c3 = 37.0 * sqrt(3.0) / 108.0;
    const double c4 = sqrt(3.0) / 3.0, c5 = 37.0 * sqrt(3.0) / 108.0;
    const double c6 = sqrt(5.0) / 5.0, c7 = 61.0 * sqrt(5.0) / 300.0;
    double udglq[3], b2[3], c2v[3], SI[3], SI4[3], p[3], denom;
    herm_consts H;
    lag_init();
    herm_init(&H, dif);
    for (int k = 0; k < 3; ++k) udglq[k] = d[k] * dif;
    if (npts == 2) {
        double crjgp[3][3], crjgm[3][3], crjdgp[4][3], crjdgm[4][3], dkpg[3], dkmg[3], dkpdg[3], dkmdg[3];
        crjgp[0][0] = -sq3 / 3.0; crjgp[1][0] = 1.0 + sq3 / 3.0; crjgp[2][0] = -dif * sq3 / 6.0;
        crjgp[0][1] = 1.0 - sq3 / 3.0; crjgp[1][1] = sq3 / 3.0; crjgp[2][1] = -dif * sq3 / 6.0;
        crjgp[0][2] = -sq3 / 12.0; crjgp[1][2] = 1.0; crjgp[2][2] = sq3 / 12.0;
        crjgm[0][0] = sq3 / 3.0; crjgm[1][0] = 1.0 - sq3 / 3.0; crjgm[2][0] = dif * sq3 / 6.0;
        crjgm[0][1] = 1.0 + sq3 / 3.0; crjgm[1][1] = -sq3 / 3.0; crjgm[2][1] = dif * sq3 / 6.0;
        crjgm[0][2] = sq3 / 12.0; crjgm[1][2] = 1.0; crjgm[2][2] = -sq3 / 12.0;
        denom = 360.0;
        dkpg[0] = (-sq3 + 105.0) / denom; dkpg[1] = (sq3 + 105.0) / denom; dkpg[2] = 5.0 / 12.0;
        dkmg[0] = (sq3 + 105.0) / denom; dkmg[1] = (-sq3 + 105.0) / denom; dkmg[2] = 5.0 / 12.0;
        crjdgp[0][0] = sq3 / dif; crjdgp[1][0] = -sq3 / dif; crjdgp[2][0] = sq3 / 3.0;
        crjdgp[3][0] = (2.0 * sq3 + 3.0) / 3.0;
        crjdgp[0][1] = -sq3 / dif; crjdgp[1][1] = sq3 / dif; crjdgp[2][1] = (-2.0 * sq3 + 3.0) / 3.0;
        crjdgp[3][1] = -sq3 / 3.0;
        crjdgp[0][2] = sq3 / 6.0 / dif; crjdgp[1][2] = -sq3 / 3.0 / dif; crjdgp[2][2] = sq3 / 6.0 / dif;
        crjdgp[3][2] = 1.0;
        crjdgm[0][0] = -sq3 / dif; crjdgm[1][0] = sq3 / dif; crjdgm[2][0] = -sq3 / 3.0;
        crjdgm[3][0] = (-2.0 * sq3 + 3.0) / 3.0;
        crjdgm[0][1] = sq3 / dif; crjdgm[1][1] = -sq3 / dif; crjdgm[2][1] = (2.0 * sq3 + 3.0) / 3.0;
        crjdgm[3][1] = sq3 / 3.0;
        crjdgm[0][2] = -sq3 / 6.0 / dif; crjdgm[1][2] = sq3 / 3.0 / dif; crjdgm[2][2] = -sq3 / 6.0 / dif;
        crjdgm[3][2] = 1.0;
        dkpdg[0] = -sq3 / 144.0 + 1.0 / 3.0; dkpdg[1] = sq3 / 144.0 + 1.0 / 3.0; dkpdg[2] = 1.0 / 3.0;
        dkmdg[0] = sq3 / 144.0 + 1.0 / 3.0; dkmdg[1] = -sq3 / 144.0 + 1.0 / 3.0; dkmdg[2] = 1.0 / 3.0;
        herm_bc(u, udglq, 0, b2, c2v);
        for (int k = 0; k < 3; ++k) SI[k] = c2 * b2[k] + c3 * c2v[k];
        p[0] = crjgp[0][0] * u[0] + crjgp[1][0] * u[1] + crjgp[2][0] * d[0];
        p[1] = crjgp[0][1] * u[1] + crjgp[1][1] * u[2] + crjgp[2][1] * d[2];
        p[2] = crjgp[0][2] * u[0] + crjgp[1][2] * u[1] + crjgp[2][2] * u[2];
        out[0] = hjs(p, dkpg, SI);
        p[0] = crjgm[0][0] * u[0] + crjgm[1][0] * u[1] + crjgm[2][0] * d[0];
        p[1] = crjgm[0][1] * u[1] + crjgm[1][1] * u[2] + crjgm[2][1] * d[2];
        p[2] = crjgm[0][2] * u[0] + crjgm[1][2] * u[1] + crjgm[2][2] * u[2];
        out[1] = hjs(p, dkmg, SI);
        herm_bc(u, udglq, 1, b2, c2v);
        for (int k = 0; k < 3; ++k) SI[k] = c4 * b2[k] + c5 * c2v[k];
        p[0] = crjdgp[0][0] * u[0] + crjdgp[1][0] * u[1] + crjdgp[2][0] * d[0] + crjdgp[3][0] * d[1];
        p[1] = crjdgp[0][1] * u[1] + crjdgp[1][1] * u[2] + crjdgp[2][1] * d[1] + crjdgp[3][1] * d[2];
        p[2] = crjdgp[0][2] * u[0] + crjdgp[1][2] * u[1] + crjdgp[2][2] * u[2] + crjdgp[3][2] * d[1];
        out[2] = hjs(p, dkpdg, SI);
        p[0] = crjdgm[0][0] * u[0] + crjdgm[1][0] * u[1] + crjdgm[2][0] * d[0] + crjdgm[3][0] * d[1];
        p[1] = crjdgm[0][1] * u[1] + crjdgm[1][1] * u[2] + crjdgm[2][1] * d[1] + crjdgm[3][1] * d[2];
        p[2] = crjdgm[0][2] * u[0] + crjdgm[1][2] * u[1] + crjdgm[2][2] * u[2] + crjdgm[3][2] * d[1];
        out[3] = hjs(p, dkmdg, SI);
        return;
    }
    double crjg4p[3][3], crjg4m[3][3], crjdg4p[4][3], crjdg4m[4][3], dkpg4[3], dkmg4[3], dkpdg4[3], dkmdg4[3];
    crjg4p[0][0] = -sq5 / 5.0 + 1.0 / 30.0; crjg4p[1][0] = sq5 / 5.0 + 29.0 / 30.0;
    crjg4p[2][0] = (-sq5 / 10.0 + 1.0 / 30.0) * dif;
    crjg4p[0][1] = -sq5 / 5.0 + 29.0 / 30.0; crjg4p[1][1] = sq5 / 5.0 + 1.0 / 30.0;
    crjg4p[2][1] = -(sq5 / 10.0 + 1.0 / 30.0) * dif;
    crjg4p[0][2] = -(sq5 / 20.0 + 1.0 / 60.0); crjg4p[1][2] = 31.0 / 30.0; crjg4p[2][2] = sq5 / 20.0 - 1.0 / 60.0;
    crjg4m[0][0] = sq5 / 5.0 + 1.0 / 30.0; crjg4m[1][0] = -sq5 / 5.0 + 29.0 / 30.0;
    crjg4m[2][0] = (sq5 / 10.0 + 1.0 / 30.0) * dif;
    crjg4m[0][1] = sq5 / 5.0 + 29.0 / 30.0; crjg4m[1][1] = -sq5 / 5.0 + 1.0 / 30.0;
    crjg4m[2][1] = (sq5 / 10.0 - 1.0 / 30.0) * dif;
    crjg4m[0][2] = sq5 / 20.0 - 1.0 / 60.0; crjg4m[1][2] = 31.0 / 30.0; crjg4m[2][2] = -(sq5 / 20.0 + 1.0 / 60.0);
    denom = 330.0 * sq5 - 110.0;
    dkpg4[0] = (99.0 * sq5 + 33.0) / denom; dkpg4[1] = (108.0 * sq5 - 102.0) / denom;
    dkpg4[2] = (123.0 * sq5 - 41.0) / denom;
    denom = 330.0 * sq5 + 110.0;
    dkmg4[0] = (99.0 * sq5 - 33.0) / denom; dkmg4[1] = (108.0 * sq5 + 102.0) / denom;
    dkmg4[2] = (123.0 * sq5 + 41.0) / denom;
    crjdg4p[0][0] = (6.0 * sq5 - 2.0) / 10.0 / dif; crjdg4p[1][0] = (-6.0 * sq5 + 2.0) / 10.0 / dif;
    crjdg4p[2][0] = 0.2 * sq5 - 0.1; crjdg4p[3][0] = 0.4 * sq5 + 0.9;
    crjdg4p[0][1] = -(6.0 * sq5 + 2.0) / 10.0 / dif; crjdg4p[1][1] = (6.0 * sq5 + 2.0) / 10.0 / dif;
    crjdg4p[2][1] = -0.4 * sq5 + 0.9; crjdg4p[3][1] = -0.2 * sq5 - 0.1;
    crjdg4p[0][2] = (2.0 * sq5 + 1.0) / 20.0 / dif; crjdg4p[1][2] = -sq5 / 5.0 / dif;
    crjdg4p[2][2] = (2.0 * sq5 - 1.0) / 20.0 / dif; crjdg4p[3][2] = 1.1;
    crjdg4m[0][0] = -(6.0 * sq5 + 2.0) / 10.0 / dif; crjdg4m[1][0] = (6.0 * sq5 + 2.0) / 10.0 / dif;
    crjdg4m[2][0] = -0.2 * sq5 - 0.1; crjdg4m[3][0] = -0.4 * sq5 + 0.9;
    crjdg4m[0][1] = (6.0 * sq5 - 2.0) / 10.0 / dif; crjdg4m[1][1] = (-6.0 * sq5 + 2.0) / 10.0 / dif;
    crjdg4m[2][1] = 0.4 * sq5 + 0.9; crjdg4m[3][1] = 0.2 * sq5 - 0.1;
    crjdg4m[0][2] = (-2.0 * sq5 + 1.0) / 20.0 / dif; crjdg4m[1][2] = sq5 / 5.0 / dif;
    crjdg4m[2][2] = (-2.0 * sq5 - 1.0) / 20.0 / dif; crjdg4m[3][2] = 1.1;
    denom = 1140.0 * sq5 - 570.0;
    dkpdg4[0] = (399.0 * sq5 + 190.0) / denom; dkpdg4[1] = (481.0 * sq5 - 630.0) / denom;
    dkpdg4[2] = (260.0 * sq5 - 130.0) / denom;
    denom = 1140.0 * sq5 + 570.0;
    dkmdg4[0] = (399.0 * sq5 - 190.0) / denom; dkmdg4[1] = (481.0 * sq5 + 630.0) / denom;
    dkmdg4[2] = (260.0 * sq5 + 130.0) / denom;
    herm_bc(u, udglq, 0, b2, c2v);
    for (int k = 0; k < 3; ++k) {
        SI[k] = c1 * c2v[k] + b2[k];
        SI4[k] = c6 * b2[k] + c7 * c2v[k];
    }
    p[0] = H.crj[0][0] * u[0] + H.crj[1][0] * u[1] + H.crj[2][0] * d[0];
    p[1] = H.crj[0][1] * u[1] + H.crj[1][1] * u[2] + H.crj[2][1] * d[2];
    p[2] = H.crj[0][2] * u[0] + H.crj[1][2] * u[1] + H.crj[2][2] * u[2];
    out[0] = hjs(p, H.dkm, SI);
    p[0] = H.crj[1][1] * u[0] + H.crj[0][1] * u[1] - H.crj[2][1] * d[0];
    p[1] = H.crj[1][0] * u[1] + H.crj[0][0] * u[2] - H.crj[2][0] * d[2];
    p[2] = H.crj[2][2] * u[0] + H.crj[1][2] * u[1] + H.crj[0][2] * u[2];
    out[1] = hjs(p, H.dkp, SI);
    p[0] = crjg4p[0][0] * u[0] + crjg4p[1][0] * u[1] + crjg4p[2][0] * d[0];
    p[1] = crjg4p[0][1] * u[1] + crjg4p[1][1] * u[2] + crjg4p[2][1] * d[2];
    p[2] = crjg4p[0][2] * u[0] + crjg4p[1][2] * u[1] + crjg4p[2][2] * u[2];
    out[2] = hjs(p, dkpg4, SI4);
    p[0] = crjg4m[0][0] * u[0] + crjg4m[1][0] * u[1] + crjg4m[2][0] * d[0];
    p[1] = crjg4m[0][1] * u[1] + crjg4m[1][1] * u[2] + crjg4m[2][1] * d[2];
    p[2] = crjg4m[0][2] * u[0] + crjg4m[1][2] * u[1] + crjg4m[2][2] * u[2];
    out[3] = hjs(p, dkmg4, SI4);
    herm_bc(u, udglq, 1, b2, c2v);
    for (int k = 0; k < 3; ++k) {
        SI[k] = c1 * c2v[k] + b2[k];
        SI4[k] = c6 * b2[k] + c7 * c2v[k];
    }
    p[0] = H.crjd[0][0] * u[0] + H.crjd[1][0] * u[1] + H.crjd[2][0] * d[0] + H.crjd[3][0] * d[1];
    p[1] = H.crjd[0][1] * u[1] + H.crjd[1][1] * u[2] + H.crjd[2][1] * d[1] + H.crjd[3][1] * d[2];
    p[2] = H.crjd[0][2] * u[0] + H.crjd[1][2] * u[1] + H.crjd[2][2] * u[2] + H.crjd[3][2] * d[1];
    out[4] = hjs(p, H.dkmd, SI);
    p[0] = H.crjd[0][1] * u[0] + H.crjd[1][1] * u[1] + H.crjd[3][1] * d[0] + H.crjd[2][1] * d[1];
    p[1] = H.crjd[0][0] * u[1] + H.crjd[1][0] * u[2] + H.crjd[3][0] * d[1] + H.crjd[2][0] * d[2];
    p[2] = -(H.crjd[2][2] * u[0] + H.crjd[1][2] * u[1] + H.crjd[0][2] * u[2]) + H.crjd[3][2] * d[1];
    out[5] = hjs(p, H.dkpd, SI);
    p[0] = crjdg4p[0][0] * u[0] + crjdg4p[1][0] * u[1] + crjdg4p[2][0] * d[0] + crjdg4p[3][0] * d[1];
    p[1] = crjdg4p[0][1] * u[1] + crjdg4p[1][1] * u[2] + crjdg4p[2][1] * d[1] + crjdg4p[3][1] * d[2];
    p[2] = crjdg4p[0][2] * u[0] + crjdg4p[1][2] * u[1] + crjdg4p[2][2] * u[2] + crjdg4p[3][2] * d[1];
    out[6] = hjs(p, dkpdg4, SI4);
    p[0] = crjdg4m[0][0] * u[0] + crjdg4m[1][0] * u[1] + crjdg4m[2][0] * d[0] + crjdg4m[3][0] * d[1];
    p[1] = crjdg4m[0][1] * u[1] + crjdg4m[1][1] * u[2] + crjdg4m[2][1] * d[1] + crjdg4m[3][1] * d[2];
    p[2] = crjdg4m[0][2] * u[0] + crjdg4m[1][2] * u[1] + crjdg4m[2][2] * u[2] + crjdg4m[3][2] * d[1];
    out[7] = hjs(p, dkmdg4, SI4);
}

/* Hermite::hermiteLimiter, Schemes_Hermite.cpp:634-662 */
static double herm_limiter(double dif, const double u[3], const double d[3]) {
    double p[3], lin[3], sm[3], w[3], ud[3];
    p[0] = 0.75 * (u[2] - u[0]) / dif - 0.25 * (d[0] + d[2]);
    p[1] = (u[1] - u[0]) / dif;
    p[2] = (u[2] - u[1]) / dif;
    lin[0] = 1.0 - 2E-5; lin[1] = 1E-5; lin[2] = 1E-5;
    for (int k = 0; k < 3; ++k) ud[k] = d[k] * dif;
    sm[0] = herm_brk(u, ud);
    sm[1] = (u[0] - u[1]) * (u[0] - u[1]);
    sm[2] = (u[1] - u[2]) * (u[1] - u[2]);
    double a = fabs(sm[0] - sm[1]) + fabs(sm[0] - sm[2]);
    double tau = a * a / 4.0;
    for (int k = 0; k < 3; ++k) w[k] = lin[k] * (1.0 + tau / (sm[k] + 1E-6));
    double s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    return w[0] * (p[0] / lin[0] - ((lin[1] / lin[0]) * p[1] + (lin[2] / lin[0]) * p[2])) + (w[1] * p[1] + w[2] * p[2]);
}

double orc_hermite_limiter(double dif, const double u[3], const double d[3]) {
    lag_init();
    return herm_limiter(dif, u, d);
}

/* Euler2D::dfluxx / dfluxy, Equations_Euler2D.cpp:52-98 (row-major) */
void orc_dflux(double gamma, int dir, const double q[4], const double w[4], double M[16]) {
    double g1 = gamma - 1.0, g3 = gamma - 3.0;
    if (dir == 1) {
        double sq = w[1] * w[1];
        double dyn = 0.5 * g1 * (sq + w[2] * w[2]);
        double h = (q[3] + w[3]) / w[0];
        double m[16] = {0.0, 1.0, 0.0, 0.0,
                        dyn - sq, -g3 * w[1], -g1 * w[2], g1,
                        -w[1] * w[2], w[2], w[1], 0.0,
                        (dyn - h) * w[1], h - g1 * sq, -g1 * w[1] * w[2], gamma * w[1]};
        memcpy(M, m, sizeof m);
    } else {
        double sq = w[2] * w[2];
        double dyn = 0.5 * g1 * (sq + w[1] * w[1]);
        double h = (q[3] + w[3]) / w[0];
        double m[16] = {0.0, 0.0, 1.0, 0.0,
                        -w[1] * w[2], w[2], w[1], 0.0,
                        dyn - sq, -g1 * w[1], -g3 * w[2], g1,
                        (dyn - h) * w[2], -g1 * w[1] * w[2], h - g1 * sq, gamma * w[2]};
        memcpy(M, m, sizeof m);
    }
}

/* Euler2D::dRusanov, Equations_Euler2D.cpp:352-360:
 * 0.5 (Jl dql + Jr dqr - R diag(a) L (dqr - dql)) */
static void drusanov(const double dql[4], const double dqr[4], const double L[16], const double R[16],
                     const double Jl[16], const double Jr[16], const double a[4], double f[4]) {
    double t1[4], t2[4], dd[4], ld[4], rd[4];
    matvec4(Jl, dql, t1);
    matvec4(Jr, dqr, t2);
    for (int k = 0; k < 4; ++k) dd[k] = dqr[k] - dql[k];
    matvec4(L, dd, ld);
    for (int k = 0; k < 4; ++k) ld[k] = a[k] * ld[k];
    matvec4(R, ld, rd);
    for (int k = 0; k < 4; ++k) f[k] = 0.5 * ((t1[k] + t2[k]) - rd[k]);
}

static void max_local_wave(double gamma, int dir, const double wl[4], const double wr[4], double a[4]) {
    double cl = sqrt(fabs(gamma * wl[3] / wl[0]));
    double cr = sqrt(fabs(gamma * wr[3] / wr[0]));
    double vl = wl[dir], vr = wr[dir];
    a[0] = fmax(fabs(vl - cl), fabs(vr - cr));
    a[1] = fmax(fabs(vl), fabs(vr));
    a[2] = a[1];
    a[3] = fmax(fabs(vl + cl), fabs(vr + cr));
}

/* ------------------------------------------------------------------------ */
/* FVHermite::System2D (FVHermite_System2D.cpp)                              */
/* ------------------------------------------------------------------------ */
#define FVH_K 8 /* k = 5 + 3, FVHermite_System2D.cpp:20-23 */
#define FVH_S 2 /* s = (5 - 1) / 2, computed before k += 3 */

/* limitDerivatives, FVHermite_System2D.cpp:393-427, single-threaded order:
 * the x sweep replaces v in place with i ascending (reading the already
 * updated v_{i-1}), the y sweep replaces w with j ascending. */
static void fvh_limit(const padded* P, double** u, double** v, double** w, double dif) {
    int k = P->k, nxx = P->nxx, nyy = P->nyy;
    for (int i = k; i < nxx - k + 1; ++i)
        for (int j = 0; j < nyy; ++j)
            for (int e = 0; e < 4; ++e) {
                double uu[3], vv[3];
                for (int m = 0; m < 3; ++m) {
                    uu[m] = u[e][PIX(P, i - 1 + m, j)];
                    vv[m] = v[e][PIX(P, i - 1 + m, j)];
                }
                v[e][PIX(P, i, j)] = herm_limiter(dif, uu, vv);
            }
    for (int i = 0; i < nxx; ++i)
        for (int j = k; j < nyy - k + 1; ++j)
            for (int e = 0; e < 4; ++e) {
                double uu[3], ww[3];
                for (int m = 0; m < 3; ++m) {
                    uu[m] = u[e][PIX(P, i, j - 1 + m)];
                    ww[m] = w[e][PIX(P, i, j - 1 + m)];
                }
                w[e][PIX(P, i, j)] = herm_limiter(dif, uu, ww);
            }
}

/* FVHermite::System2D::rhs, FVHermite_System2D.cpp:429-731.  q holds 12
 * fields: u (4), v = <dq/dx> (4), w = <dq/dy> (4). */
static void fvh2d_rhs_impl(const orc_params* p, const double* q, double* rhs, double* qlim);

void orc_fvh2d_rhs(const orc_params* p, const double* q, double* rhs) { fvh2d_rhs_impl(p, q, rhs, NULL); }

/* Residual plus the state as rhs() leaves it: limitDerivatives() replaces
 * v and w in place (FVHermite_System2D.cpp:397-426), so after the call the
 * object's state carries the limited moments. */
void orc_fvh2d_rhs_lim(const orc_params* p, const double* q, double* rhs, double* qlim) {
    fvh2d_rhs_impl(p, q, rhs, qlim);
}

static void fvh2d_rhs_impl(const orc_params* p, const double* q, double* rhs, double* qlim) {
    lag_init();
    padded P;
    int nx = p->nx, ny = p->ny, k = FVH_K, s = FVH_S;
    double gamma = p->gamma;
    double dx = (p->xr - p->xl) / nx, dy = (p->yt - p->yb) / ny, area = dx * dy;
    double gw[3] = {5.0 / 9.0 / 2.0, 8.0 / 9.0 / 2.0, 5.0 / 9.0 / 2.0};
    herm_consts H;
    herm_init(&H, dx); /* setGridSize(dx), :28 (used for y as well) */
    pad_alloc(&P, nx, ny, k, 12);
    load_padded(&P, q, 12, p->bc);           /* first BC call, :432 */
    double** U = P.q;
    double** V = P.q + 4;
    double** W = P.q + 8;
    fvh_limit(&P, U, V, W, dx);              /* :435 */
    for (int f = 0; f < 12; ++f) fill_ghosts(&P, P.q[f], p->bc); /* :437 */
    int nxx = P.nxx, nyy = P.nyy;
    if (qlim) {
        size_t nn = (size_t)nx * ny;
        for (int f = 0; f < 12; ++f)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) qlim[f * nn + (size_t)j * nx + i] = P.q[f][PIX(&P, i + k, j + k)];
    }
    size_t nt = (size_t)nxx * nyy;
    /* face states, 6 series (u, v, w left/right) x 4 fields per direction */
    double* FX = (double*)calloc(24 * nt, sizeof(double)); /* uxl uxr vxl vxr wxl wxr */
    double* FY = (double*)calloc(24 * nt, sizeof(double)); /* uyl uyr vyl vyr wyl wyr */
    double* flx = (double*)calloc(12 * nt, sizeof(double)); /* fx gx hx */
    double* fly = (double*)calloc(12 * nt, sizeof(double)); /* fy gy hy */
#define SER(A, s_, e_, id) A[((size_t)(s_) * 4 + (e_)) * nt + (id)]

    /* x reconstruction :443-468 */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nyy; ++j)
        for (int i = s; i < nxx - s - 1; ++i) {
            double qa[4], qb[4], L[16], R[16], A[4][4], B[4][4], C[4][6], cc[4], y[4], o4[4], o2[2];
            cell_q(&P, i, j, qa);
            cell_q(&P, i + 1, j, qb);
            orc_eigen(gamma, 1, qa, qb, L, R);
            for (int m = 0; m < 4; ++m) {
                for (int e = 0; e < 4; ++e) cc[e] = U[e][PIX(&P, i - 1 + m, j)];
                matvec4(L, cc, y);
                for (int e = 0; e < 4; ++e) A[e][m] = y[e];
                for (int e = 0; e < 4; ++e) cc[e] = V[e][PIX(&P, i - 1 + m, j)];
                matvec4(L, cc, y);
                for (int e = 0; e < 4; ++e) B[e][m] = y[e];
            }
            for (int m = 0; m < 6; ++m) {
                for (int e = 0; e < 4; ++e) cc[e] = W[e][PIX(&P, i - 2 + m, j)];
                matvec4(L, cc, y);
                for (int e = 0; e < 4; ++e) C[e][m] = y[e];
            }
            size_t id = PIX(&P, i, j);
            for (int e = 0; e < 4; ++e) {
                herm_recon(&H, p->recon, A[e], B[e], o4);
                SER(FX, 0, e, id) = o4[0];
                SER(FX, 1, e, id) = o4[1];
                SER(FX, 2, e, id) = o4[2];
                SER(FX, 3, e, id) = o4[3];
                teno5_impl(C[e], o2, NULL);
                SER(FX, 4, e, id) = o2[0];
                SER(FX, 5, e, id) = o2[1];
            }
        }
    /* y reconstruction :470-495 (hteno on u and w = d/dy, teno5 on v) */
#pragma omp parallel for schedule(static)
    for (int j = s; j < nyy - s - 1; ++j)
        for (int i = 0; i < nxx; ++i) {
            double qa[4], qb[4], L[16], R[16], A[4][4], B[4][4], C[4][6], cc[4], y[4], o4[4], o2[2];
            cell_q(&P, i, j, qa);
            cell_q(&P, i, j + 1, qb);
            orc_eigen(gamma, 2, qa, qb, L, R);
            for (int m = 0; m < 4; ++m) {
                for (int e = 0; e < 4; ++e) cc[e] = U[e][PIX(&P, i, j - 1 + m)];
                matvec4(L, cc, y);
                for (int e = 0; e < 4; ++e) A[e][m] = y[e];
                for (int e = 0; e < 4; ++e) cc[e] = W[e][PIX(&P, i, j - 1 + m)];
                matvec4(L, cc, y);
                for (int e = 0; e < 4; ++e) B[e][m] = y[e];
            }
            for (int m = 0; m < 6; ++m) {
                for (int e = 0; e < 4; ++e) cc[e] = V[e][PIX(&P, i, j - 2 + m)];
                matvec4(L, cc, y);
                for (int e = 0; e < 4; ++e) C[e][m] = y[e];
            }
            size_t id = PIX(&P, i, j);
            for (int e = 0; e < 4; ++e) {
                herm_recon(&H, p->recon, A[e], B[e], o4);
                SER(FY, 0, e, id) = o4[0];  /* uyl */
                SER(FY, 1, e, id) = o4[1];  /* uyr */
                SER(FY, 4, e, id) = o4[2];  /* wyl */
                SER(FY, 5, e, id) = o4[3];  /* wyr */
                teno5_impl(C[e], o2, NULL);
                SER(FY, 2, e, id) = o2[0];  /* vyl */
                SER(FY, 3, e, id) = o2[1];  /* vyr */
            }
        }
    /* quadrature + transform + fluxes :497-640 (fused over the flux range) */
#pragma omp parallel for schedule(static)
    for (int j = k - 1; j < nyy - k; ++j)
        for (int i = k - 1; i < nxx - k; ++i)
            for (int dir = 1; dir <= 2; ++dir) {
                double qa[4], qb[4], L[16], R[16], G[6][4][3];
                cell_q(&P, i, j, qa);
                if (dir == 1) cell_q(&P, i + 1, j, qb); else cell_q(&P, i, j + 1, qb);
                orc_eigen(gamma, dir, qa, qb, L, R);
                double* F = dir == 1 ? FX : FY;
                for (int sr = 0; sr < 6; ++sr)
                    for (int e = 0; e < 4; ++e) {
                        double a5[5];
                        for (int m = 0; m < 5; ++m)
                            a5[m] = SER(F, sr, e, dir == 1 ? PIX(&P, i, j - 2 + m) : PIX(&P, i - 2 + m, j));
                        orc_weno5js3q(a5, G[sr][e]);
                    }
                double tu[3][4], tv[3][4], tw[3][4];
                for (int m = 0; m < 3; ++m) {
                    double c[6][4], Q[6][4];
                    for (int sr = 0; sr < 6; ++sr) {
                        for (int e = 0; e < 4; ++e) c[sr][e] = G[sr][e][m];
                        matvec4(R, c[sr], Q[sr]);
                    }
                    /* Q: 0 ul, 1 ur, 2 vl, 3 vr, 4 wl, 5 wr (conservative) */
                    double wl[4], wr[4], Jl[16], Jr[16], a[4], Fu[4], Fv[4], Fw[4];
                    orc_primitive(gamma, Q[0], wl);
                    orc_dflux(gamma, dir, Q[0], wl, Jl);
                    orc_primitive(gamma, Q[1], wr);
                    orc_dflux(gamma, dir, Q[1], wr, Jr);
                    max_local_wave(gamma, dir, wl, wr, a);
                    if (p->flux == ORC_FLUX_HLLC)
                        orc_hllc(gamma, dir, p->quirks, Q[0], Q[1], Fu);
                    else
                        orc_rusanov_char(gamma, dir, Q[0], Q[1], L, R, Fu);
                    drusanov(Q[2], Q[3], L, R, Jl, Jr, a, Fv);
                    drusanov(Q[4], Q[5], L, R, Jl, Jr, a, Fw);
                    for (int e = 0; e < 4; ++e) {
                        tu[m][e] = gw[m] * Fu[e];
                        tv[m][e] = gw[m] * Fv[e];
                        tw[m][e] = gw[m] * Fw[e];
                    }
                }
                double* fl = dir == 1 ? flx : fly;
                size_t id = PIX(&P, i, j);
                for (int e = 0; e < 4; ++e) {
                    fl[(0 * 4 + e) * nt + id] = (tu[0][e] + tu[1][e]) + tu[2][e];
                    fl[(1 * 4 + e) * nt + id] = (tv[0][e] + tv[1][e]) + tv[2][e];
                    fl[(2 * 4 + e) * nt + id] = (tw[0][e] + tw[1][e]) + tw[2][e];
                }
            }
    /* divergence :710-730 */
    size_t n = (size_t)nx * ny;
    for (int j = k; j < nyy - k; ++j)
        for (int i = k; i < nxx - k; ++i)
            for (int g = 0; g < 12; ++g) {
                double v = ((flx[g * nt + PIX(&P, i - 1, j)] - flx[g * nt + PIX(&P, i, j)]) * dy +
                            (fly[g * nt + PIX(&P, i, j - 1)] - fly[g * nt + PIX(&P, i, j)]) * dx) /
                           area;
                rhs[g * n + (size_t)(j - k) * nx + (i - k)] = v;
            }
#undef SER
    free(FX); free(FY); free(flx); free(fly);
    pad_free(&P, 12);
}

/* FVHermite::System2D::RungeKutta3, :733-782 (u, v and w).  save() runs
 * after the first rhs() (:738-742), i.e. after limitDerivatives() replaced v
 * and w in place, so u^n of the derivative moments is the limited state;
 * the stage results u1, u2 (:743-764) are stored unlimited. */
void orc_fvh2d_rk3(const orc_params* p, double* q, double dt) {
    size_t n = 12 * (size_t)p->nx * p->ny;
    double* q0 = (double*)malloc(n * sizeof(double));
    double* r = (double*)malloc(n * sizeof(double));
    double* qs = (double*)malloc(n * sizeof(double));
    orc_fvh2d_rhs_lim(p, q, r, q0);
    orc_rk3_combine(1, q0, q0, r, dt, q, (long)n);
    memcpy(qs, q, n * sizeof(double));
    orc_fvh2d_rhs(p, q, r);
    orc_rk3_combine(2, q0, qs, r, dt, q, (long)n);
    memcpy(qs, q, n * sizeof(double));
    orc_fvh2d_rhs(p, q, r);
    orc_rk3_combine(3, q0, qs, r, dt, q, (long)n);
    free(q0); free(r); free(qs);
}

/* FVHermite::System2D::timeStepping, :784-819 (dt from u only) */
int orc_fvh2d_advance(const orc_params* p, double* q, double cfl, double t0, double tfinal, int max_steps,
                      double* t_out) {
    double t = t0;
    int nsteps = 0;
    while (t < tfinal && (max_steps <= 0 || nsteps < max_steps)) {
        double dt = orc_fv2d_dt(p, q, cfl, t, tfinal);
        orc_fvh2d_rk3(p, q, dt);
        t += dt;
        ++nsteps;
    }
    if (t_out) *t_out = t;
    return nsteps;
}

/* Isentropic vortex with derivative moments (FVHermite_System2D.cpp:885-900,
 * Equations_Euler2D.cpp:668-745), centred at (cx, cy), including the
 * reference's /pi where /pi^2 is meant in the x-derivative (:678).  The
 * density modes (k, l, a, phi, Lx, Ly) add a sin(theta) to rho at fixed u, v,
 * p, and their analytic x/y derivatives to the moments. */
static void vortex_point_derivs(double gamma, double x, double y, double cx, double cy, const double* modes,
                                int nmodes, double q[4], double qx[4], double qy[4]) {
    const double pi = 4.0 * atan(1.0);
    double g1 = gamma - 1.0;
    double X = x - cx, Y = y - cy;
    double r = pow(X, 2) + pow(Y, 2);
    double w[4];
    w[0] = 1.0 - g1 * 25.0 / 8.0 / gamma / pow(pi, 2) * exp(1 - r);
    w[0] = pow(w[0], 1.0 / g1);
    w[1] = 1.0 - 5.0 * Y / 2.0 / pi * exp(0.5 * (1 - r));
    w[2] = 1.0 + 5.0 * X / 2.0 / pi * exp(0.5 * (1 - r));
    w[3] = pow(w[0], gamma);
    /* :678-689 */
    double a0 = 1.0 - g1 * 25.0 / 8.0 / gamma / pi * exp(1 - r);
    double rx = 25.0 * X * exp(1 - r) / 4.0 / pow(pi, 2) / gamma * pow(a0, (2.0 - gamma) / g1);
    double ux = 5.0 * X * Y / 2.0 / pi * exp(0.5 * (1 - r));
    double vx = 5.0 * (1 - pow(X, 2)) / 2.0 / pi * exp(0.5 * (1 - r));
    /* :702-713 */
    double b0 = 1.0 - g1 * 25.0 / 8.0 / gamma / pow(pi, 2) * exp(1 - r);
    double ry = 25.0 * Y * exp(1 - r) / 4.0 / pow(pi, 2) / gamma * pow(b0, (2.0 - gamma) / g1);
    double uy = 5.0 * (pow(Y, 2) - 1.0) / 2.0 / pi * exp(0.5 * (1 - r));
    double vy = -5.0 * X * Y / 2.0 / pi * exp(0.5 * (1 - r));
    double pxv = gamma * pow(w[0], g1) * rx; /* dp/dx of the isentropic vortex */
    double pyv = gamma * pow(w[0], g1) * ry;
    /* density modes, accumulated onto rho in the order of vortex_point */
    double rho = w[0], ddx = 0.0, ddy = 0.0;
    for (int m = 0; m < nmodes; ++m) {
        const double* md = modes + 6 * m;
        double th = 2.0 * pi * (md[0] * x / md[4] + md[1] * y / md[5]) + md[3];
        rho += md[2] * sin(th);
        ddx += md[2] * cos(th) * (2.0 * pi * md[0] / md[4]);
        ddy += md[2] * cos(th) * (2.0 * pi * md[1] / md[5]);
    }
    double wp[4] = {rho, w[1], w[2], w[3]};
    orc_conservative(gamma, wp, q);
    double Rx = rx + ddx, Ry = ry + ddy;
    qx[0] = Rx;
    qx[1] = rho * ux + Rx * w[1];
    qx[2] = rho * vx + Rx * w[2];
    qx[3] = pxv / g1 + 0.5 * Rx * (pow(w[1], 2) + pow(w[2], 2)) + 0.5 * rho * (2.0 * w[1] * ux + 2.0 * w[2] * vx);
    qy[0] = Ry;
    qy[1] = rho * uy + Ry * w[1];
    qy[2] = rho * vy + Ry * w[2];
    qy[3] = pyv / g1 + 0.5 * Ry * (pow(w[1], 2) + pow(w[2], 2)) + 0.5 * rho * (2.0 * w[1] * uy + 2.0 * w[2] * vy);
}

void orc_ic_vortex_hermite(const orc_params* p, double cx, double cy, const double* modes, int nmodes,
                           double* q) {
    double a[9], wt[9];
    orc_gauss9(a, wt);
    int nx = p->nx, ny = p->ny;
    double dx = (p->xr - p->xl) / nx, dy = (p->yt - p->yb) / ny;
    size_t n = (size_t)nx * ny;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double x = ((double)i + 0.5) * dx + p->xl;
            double y = ((double)j + 0.5) * dy + p->yb;
            double acc[12] = {0};
            for (int ii = 0; ii < 9; ++ii)
                for (int jj = 0; jj < 9; ++jj) {
                    double qq[4], qx[4], qy[4];
                    vortex_point_derivs(p->gamma, x + dx * a[ii] / 2.0, y + dy * a[jj] / 2.0, cx, cy, modes, nmodes,
                                        qq, qx, qy);
                    for (int e = 0; e < 4; ++e) {
                        acc[e] += wt[ii] * wt[jj] * qq[e] / 4.0;
                        acc[4 + e] += wt[ii] * wt[jj] * qx[e] / 4.0;
                        acc[8 + e] += wt[ii] * wt[jj] * qy[e] / 4.0;
                    }
                }
            for (int g = 0; g < 12; ++g) q[g * n + (size_t)j * nx + i] = acc[g];
        }
}

/* ------------------------------------------------------------------------ */
/* FV::System1D + Equations::Euler1D (FV_System1D.cpp, Equations_Euler1D.cpp) */
/* ------------------------------------------------------------------------ */
#define FV1_K 5 /* FV_System1D.cpp:15 */

static void e1_primitive(double g1, const double q[3], double w[3]) {
    /* Equations_Euler1D.cpp:18-23 */
    w[0] = q[0];
    w[1] = q[1] / q[0];
    w[2] = (q[2] - 0.5 * q[1] * w[1]) * g1;
}

static void e1_conservative(double g1, const double w[3], double q[3]) {
    /* :25-30 */
    q[0] = w[0];
    q[1] = w[0] * w[1];
    q[2] = w[2] / g1 + 0.5 * q[1] * w[1];
}

static void e1_flux(const double q[3], const double w[3], double f[3]) {
    /* :32-37 */
    f[0] = q[1];
    f[1] = q[1] * w[1] + w[2];
    f[2] = w[1] * (q[2] + w[2]);
}

static inline void matvec3(const double M[9], const double x[3], double y[3]) {
    for (int r = 0; r < 3; ++r) y[r] = (M[3 * r] * x[0] + M[3 * r + 1] * x[1]) + M[3 * r + 2] * x[2];
}

/* Euler1D::HLLC, Equations_Euler1D.cpp:156-210 (quirks: c without gamma
 * :174-175, p_L in the right star energy :204) */
static void e1_hllc(double gamma, int quirks, const double ql[3], const double qr[3], double f[3]) {
    double g1 = gamma - 1.0, wl[3], wr[3], fl[3], fr[3], qs[3];
    e1_primitive(g1, ql, wl);
    e1_primitive(g1, qr, wr);
    double hl = (ql[2] + wl[2]) / wl[0], hr = (qr[2] + wr[2]) / wr[0];
    double rr = sqrt(fabs(wr[0] / wl[0]));
    double rrr = 1.0 / (1.0 + rr);
    double vm = (wl[1] + wr[1] * rr) * rrr;
    double qm = 0.5 * vm * vm;
    double hm = (hl + hr * rr) * rrr;
    double cm = sqrt(fabs(g1 * (hm - qm)));
    double cl = quirks ? sqrt(fabs(wl[2] / wl[0])) : sqrt(fabs(gamma * wl[2] / wl[0]));
    double cr = quirks ? sqrt(fabs(wr[2] / wr[0])) : sqrt(fabs(gamma * wr[2] / wr[0]));
    double sl = fmin(wl[1] - cl, vm - cm);
    double sr = fmax(wr[1] + cr, vm + cm);
    double sm = (qr[1] * (sr - wr[1]) - ql[1] * (sl - wl[1]) + wl[2] - wr[2]) /
                (wr[0] * (sr - wr[1]) - wl[0] * (sl - wl[1]));
    double ps = wl[2] + wl[0] * (wl[1] - sl) * (wl[1] - sm);
    e1_flux(ql, wl, fl);
    e1_flux(qr, wr, fr);
    if (sl > 0.0) {
        memcpy(f, fl, sizeof fl);
    } else if (sl <= 0.0 && sm > 0.0) {
        double nom = sl - wl[1], den = sl - sm;
        qs[0] = wl[0] * nom / den;
        qs[1] = (nom * ql[1] + ps - wl[2]) / den;
        qs[2] = (nom * ql[2] - wl[2] * wl[1] + ps * sm) / den;
        for (int k = 0; k < 3; ++k) f[k] = fl[k] + sl * (qs[k] - ql[k]);
    } else if (sm <= 0.0 && sr >= 0.0) {
        double nom = sr - wr[1], den = sr - sm;
        qs[0] = wr[0] * nom / den;
        qs[1] = (nom * qr[1] + ps - wr[2]) / den;
        qs[2] = (nom * qr[2] - (quirks ? wl[2] : wr[2]) * wr[1] + ps * sm) / den;
        for (int k = 0; k < 3; ++k) f[k] = fr[k] + sr * (qs[k] - qr[k]);
    } else {
        memcpy(f, fr, sizeof fr);
    }
}

void orc_e1_hllc(double gamma, int quirks, const double ql[3], const double qr[3], double f[3]) {
    e1_hllc(gamma, quirks, ql, qr, f);
}

/* Euler1D::localWaveSpeed, Equations_Euler1D.cpp:108-119 */
static void e1_local_speed(double gamma, const double w[3], double s[3]) {
    double c = sqrt(fabs(gamma * w[2] / w[0]));
    s[0] = fabs(w[1] - c);
    s[1] = fabs(w[1]);
    s[2] = fabs(w[1] + c);
}

static inline double cmax(double a, double b) { return (a < b) ? b : a; } /* cwiseMax */
static inline double cmin(double a, double b) { return (b < a) ? b : a; } /* cwiseMin */

/* Euler1D::CUFlux, Equations_Euler1D.cpp:137-154 (L, R row-major) */
static void e1_cuflux(double gamma, const double L[9], const double R[9], const double ql[3], const double qr[3],
                      double f[3]) {
    double g1 = gamma - 1.0, qmax[3], qmin[3], w[3], spdmax[3], spdmin[3], aplus[3], aminus[3];
    for (int e = 0; e < 3; ++e) {
        qmax[e] = cmax(ql[e], qr[e]);
        qmin[e] = cmin(ql[e], qr[e]);
    }
    e1_primitive(g1, qmax, w);
    e1_local_speed(gamma, w, spdmax);
    e1_primitive(g1, qmin, w);
    e1_local_speed(gamma, w, spdmin);
    for (int e = 0; e < 3; ++e) {
        aplus[e] = cmax(cmax(spdmax[e], spdmin[e]), 0.0);
        aminus[e] = cmin(cmin(spdmax[e], spdmin[e]), 0.0);
    }
    double wl[3], wr[3], fl[3], fr[3], inv[3], dq[3], Ldq[3], t[3], Rt[3];
    e1_primitive(g1, ql, wl);
    e1_primitive(g1, qr, wr);
    e1_flux(ql, wl, fl);
    e1_flux(qr, wr, fr);
    for (int e = 0; e < 3; ++e) inv[e] = 1.0 / (aplus[e] - aminus[e]);
    for (int e = 0; e < 3; ++e) f[e] = (aplus[e] * fl[e] - aminus[e] * fr[e]) * inv[e];
    for (int e = 0; e < 3; ++e) dq[e] = qr[e] - ql[e];
    matvec3(L, dq, Ldq);
    for (int e = 0; e < 3; ++e) t[e] = ((aplus[e] * aminus[e]) * inv[e]) * Ldq[e];
    matvec3(R, t, Rt);
    for (int e = 0; e < 3; ++e) f[e] += Rt[e];
}

void orc_e1_cuflux(double gamma, const double L[9], const double R[9], const double ql[3], const double qr[3],
                   double f[3]) {
    e1_cuflux(gamma, L, R, ql, qr, f);
}

/* Lagrange::teno5i, Schemes_Lagrange.cpp:1235-1263: tau = max(beta) */
static double teno5i_side(const double* v, const double p[3], const double d[3]) {
    double SI[3], w[3], sigm[3];
    smoothness5(v, SI);
    double tau = SI[0];
    if (SI[1] > tau) tau = SI[1];
    if (SI[2] > tau) tau = SI[2];
    for (int k = 0; k < 3; ++k) w[k] = pow(1.0 + tau / (SI[k] + LC.epsz), 6);
    double s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    for (int k = 0; k < 3; ++k) sigm[k] = w[k] < LC.ct ? 0.0 : 1.0;
    for (int k = 0; k < 3; ++k) w[k] = sigm[k] * d[k];
    s = sum3(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    return catchnan(dot3(p, w));
}

static void teno5i_impl(const double u[6], double out[2]) {
    double p[3], c[3];
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][1];
    p[0] = dot3(c, u + 2);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][2];
    p[1] = dot3(c, u + 1);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][3];
    p[2] = dot3(c, u + 0);
    out[0] = teno5i_side(u, p, LC.dkm);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][0];
    p[0] = dot3(c, u + 3);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][1];
    p[1] = dot3(c, u + 2);
    for (int r = 0; r < 3; ++r) c[r] = LC.crj[r][2];
    p[2] = dot3(c, u + 1);
    out[1] = teno5i_side(u + 1, p, LC.dkp);
}

void orc_teno5i(const double u[6], double out[2]) {
    lag_init();
    teno5i_impl(u, out);
}

/* FV::System1D::rhs, FV_System1D.cpp:158-241 (zero-gradient BC :74-80,
 * eigen1DEuler :121-156; the KXRCF flags :166-192 feed nothing).
 * q = 3 fields x nx.  recon: 0 teno5, 1 teno5i (shipped), 2 weno5JS, 3 weno5Z;
 * flux: ORC_FLUX_*. */
void orc_fv1d_rhs(int nx, double xl, double xr, double gamma, int recon, int flux, int quirks, const double* q,
                  double* rhs) {
    lag_init();
    int k = FV1_K, nxx = nx + 2 * k;
    double g1 = gamma - 1.0, dx = (xr - xl) / nx;
    double* U = (double*)calloc(3 * (size_t)nxx, sizeof(double));
    double* fx = (double*)calloc(3 * (size_t)nxx, sizeof(double));
    for (int e = 0; e < 3; ++e) {
        for (int i = 0; i < nx; ++i) U[e * nxx + k + i] = q[e * nx + i];
        for (int m = 0; m < k; ++m) {
            U[e * nxx + k - 1 - m] = U[e * nxx + k];
            U[e * nxx + nxx - 1 - m] = U[e * nxx + nxx - k - 1];
        }
    }
    for (int i = k - 1; i < nxx - k; ++i) {
        double wa[3], wb[3], qa[3] = {U[i], U[nxx + i], U[2 * nxx + i]};
        double qb[3] = {U[i + 1], U[nxx + i + 1], U[2 * nxx + i + 1]};
        e1_primitive(g1, qa, wa);
        e1_primitive(g1, qb, wb);
        double sa = sqrt(fabs(wa[0])), sb = sqrt(fabs(wb[0]));
        double ha = (qa[2] + wa[2]) / wa[0], hb = (qb[2] + wb[2]) / wb[0];
        double t0 = sa / (sa + sb), t1 = 1.0 - t0;
        double vxm = t0 * wa[1] + t1 * wb[1];
        double hm = t0 * ha + t1 * hb;
        double qm = 0.5 * vxm * vxm;
        double cm = sqrt(fabs(g1 * (hm - qm)));
        double tc = vxm * cm;
        double R[9] = {1.0, 1.0, 1.0, vxm - cm, vxm, vxm + cm, hm - tc, qm, hm + tc};
        double rcm = 1.0 / cm, b1 = g1 * rcm * rcm, b2 = qm * b1, tr = vxm * rcm, tb = b1 * vxm, t2 = 0.5 * b1;
        double L[9] = {0.5 * (b2 + tr), -0.5 * (tb + rcm), t2, 1.0 - b2, tb, -b1, 0.5 * (b2 - tr), -0.5 * (tb - rcm), t2};
        double C[3][6], cc[3], y[3], ul[3], ur[3], o[2];
        for (int m = 0; m < 6; ++m) {
            for (int e = 0; e < 3; ++e) cc[e] = U[e * nxx + i - 2 + m];
            matvec3(L, cc, y);
            for (int e = 0; e < 3; ++e) C[e][m] = y[e];
        }
        for (int e = 0; e < 3; ++e) {
            orc_recon(recon, C[e], dx, o);
            ul[e] = o[0];
            ur[e] = o[1];
        }
        double ql[3], qr[3], F[3];
        matvec3(R, ul, ql);  /* transformToConservative, FV_System1D.h:48-51 */
        matvec3(R, ur, qr);
        if (flux == ORC_FLUX_HLLC) {
            e1_hllc(gamma, quirks, ql, qr, F);
        } else if (flux == ORC_FLUX_CU) { /* the commented alternative, FV_System1D.cpp:234 */
            e1_cuflux(gamma, L, R, ql, qr, F);
        } else {
            double wl[3], wr[3], fl[3], fr[3], a[3], dd[3], ld[3], rd[3];
            e1_primitive(g1, ql, wl);
            e1_primitive(g1, qr, wr);
            double cl = sqrt(fabs(gamma * wl[2] / wl[0])), cr = sqrt(fabs(gamma * wr[2] / wr[0]));
            a[0] = fmax(fabs(wl[1] - cl), fabs(wr[1] - cr));
            a[1] = fmax(fabs(wl[1]), fabs(wr[1]));
            a[2] = fmax(fabs(wl[1] + cl), fabs(wr[1] + cr));
            e1_flux(ql, wl, fl);
            e1_flux(qr, wr, fr);
            for (int e = 0; e < 3; ++e) dd[e] = qr[e] - ql[e];
            matvec3(L, dd, ld);
            for (int e = 0; e < 3; ++e) ld[e] = a[e] * ld[e];
            matvec3(R, ld, rd);
            for (int e = 0; e < 3; ++e) F[e] = 0.5 * ((fl[e] + fr[e]) - rd[e]);
        }
        for (int e = 0; e < 3; ++e) fx[e * nxx + i] = F[e];
    }
    for (int i = k; i < nxx - k; ++i)
        for (int e = 0; e < 3; ++e) rhs[e * nx + i - k] = (fx[e * nxx + i - 1] - fx[e * nxx + i]) / dx;
    free(U);
    free(fx);
}

/* The KXRCF troubled-cell flags of FV::System1D::rhs (FV_System1D.cpp:166-192,
 * radius circleRadMod = (dx/2)^1.5 from :19): the reference's loop runs over
 * i = s+1 .. nxx-s-2 with s = 2 (:15, :166), i.e. every interior cell and
 * two ghost cells on each side, of the zero-gradient ghost-padded density;
 * for each cell the value at the face upwind of the momentum's sign is
 * reconstructed from the stencils of the two cells sharing it, and the cell
 * is flagged when |h1 - h2| / (radius |rho_i|) > 1.  The interior cells'
 * flags are returned (the first and last cells read the ghosts). */
void orc_fv1d_kxrcf(int nx, double xl, double xr, const double* q, unsigned char* flags) {
    int k = FV1_K, nxx = nx + 2 * k;
    double dx = (xr - xl) / nx, radius = pow(dx / 2.0, 1.5);
    double* r = (double*)calloc((size_t)nxx, sizeof(double));
    double* m = (double*)calloc((size_t)nxx, sizeof(double));
    for (int i = 0; i < nx; ++i) {
        r[k + i] = q[i];
        m[k + i] = q[nx + i];
    }
    for (int g = 0; g < k; ++g) { /* zeroGradBc, FV_System1D.cpp:74-80 */
        r[k - 1 - g] = r[k];
        r[nxx - 1 - g] = r[nxx - k - 1];
        m[k - 1 - g] = m[k];
        m[nxx - 1 - g] = m[nxx - k - 1];
    }
    const int s = 2; /* FV_System1D.cpp:15 */
    for (int i = s + 1; i < nxx - s - 1; ++i) {
        double h1, h2;
        if (m[i] < 0.0) {
            h1 = (2.0 * r[i - 1] + 5.0 * r[i] - r[i + 1]) / 6.0;
            h2 = (-r[i - 2] + 5.0 * r[i - 1] + 2.0 * r[i]) / 6.0;
        } else {
            h2 = (2.0 * r[i] + 5.0 * r[i + 1] - r[i + 2]) / 6.0;
            h1 = (-r[i - 1] + 5.0 * r[i] + 2.0 * r[i + 1]) / 6.0;
        }
        const int fl = fabs(h1 - h2) / (radius * fabs(r[i])) > 1.0;
        if (i >= k && i < nxx - k) flags[i - k] = (unsigned char)fl;
    }
    free(r);
    free(m);
}

/* The LMPE troubled-cell detector of FVHermite::System1D::rhs
 * (FVHermite_System1D.cpp:239-257; k = 4 ghost cells, s = 1, :20-21), from
 * a reset state (all flags clear; the reference's reset loop at :214-216 is
 * commented out, so there the flags accumulate over steps).  For each field
 * eq and each cell i = s+1 .. nxx-s-2 in order (two ghost cells on each
 * side, filled by the boundary condition: periodicBc :77-86 or zero
 * gradient), when cell i is not yet flagged:
 *     A  = |-u_{i-1} + 2 u_i - u_{i+1}| / 2,
 *     xc = (-2 u_{i-1} + 3 u_i - u_{i+1}) / (-u_{i-1} + 2 u_i - u_{i+1}) dx;
 * A > dx / 2 and 0 <= xc <= 3 dx flag cells i-1, i, i+1 (so a cell flagged
 * by its left neighbour is not tested itself).  flags[eq nx + c] receives
 * the interior cells' flags.  bc: 1 periodic, 0 zero gradient. */
void orc_fv1d_lmpe(int nx, double xl, double xr, int bc, const double* q, unsigned char* flags) {
    const int k = 4, s = 1, nxx = nx + 2 * k;
    const double dx = (xr - xl) / nx;
    double* u = (double*)calloc((size_t)nxx, sizeof(double));
    unsigned char* f = (unsigned char*)calloc((size_t)nxx, 1);
    for (int eq = 0; eq < 3; ++eq) {
        for (int i = 0; i < nxx; ++i) {
            int g = i - k;
            if (bc == 1) g = ((g % nx) + nx) % nx;
            else g = g < 0 ? 0 : (g >= nx ? nx - 1 : g);
            u[i] = q[eq * nx + g];
            f[i] = 0;
        }
        for (int i = s + 1; i < nxx - s - 1; ++i) {
            if (f[i]) continue;
            double A = fabs(-u[i - 1] + 2.0 * u[i] - u[i + 1]) / 2.0;
            if (A > 0.5 * dx) {
                double xc = (-2. * u[i - 1] + 3. * u[i] - u[i + 1]) / (-u[i - 1] + 2.0 * u[i] - u[i + 1]) * dx;
                if (0.0 <= xc && xc <= 3. * dx) f[i] = f[i - 1] = f[i + 1] = 1;
            }
        }
        for (int c = 0; c < nx; ++c) flags[eq * nx + c] = f[k + c];
    }
    free(u);
    free(f);
}

/* FV::System1D::timeStepping + RungeKutta3, FV_System1D.cpp:243-292 */
int orc_fv1d_advance(int nx, double xl, double xr, double gamma, int recon, int flux, int quirks, double* q,
                     double cfl, double t0, double tfinal, int max_steps, double* t_out) {
    double g1 = gamma - 1.0, dx = (xr - xl) / nx, t = t0;
    size_t n = 3 * (size_t)nx;
    double* q0 = (double*)malloc(n * sizeof(double));
    double* qs = (double*)malloc(n * sizeof(double));
    double* r = (double*)malloc(n * sizeof(double));
    int nsteps = 0;
    while (t < tfinal && (max_steps <= 0 || nsteps < max_steps)) {
        double smax = 0.0;
        for (int i = 0; i < nx; ++i) {
            double qq[3] = {q[i], q[nx + i], q[2 * nx + i]}, w[3];
            e1_primitive(g1, qq, w);
            double c = sqrt(fabs(gamma * w[2] / w[0]));
            smax = fmax(smax, fmax(fabs(w[1] - c), fabs(w[1] + c))); /* Equations_Euler1D.cpp:91-94 */
        }
        double dt = cfl * pow(dx, 3.0 / 3.0) / smax;
        dt = t + dt > tfinal ? tfinal - t : dt;
        memcpy(q0, q, n * sizeof(double));
        orc_fv1d_rhs(nx, xl, xr, gamma, recon, flux, quirks, q, r);
        orc_rk3_combine(1, q0, q0, r, dt, q, (long)n);
        memcpy(qs, q, n * sizeof(double));
        orc_fv1d_rhs(nx, xl, xr, gamma, recon, flux, quirks, q, r);
        orc_rk3_combine(2, q0, qs, r, dt, q, (long)n);
        memcpy(qs, q, n * sizeof(double));
        orc_fv1d_rhs(nx, xl, xr, gamma, recon, flux, quirks, q, r);
        orc_rk3_combine(3, q0, qs, r, dt, q, (long)n);
        t += dt;
        ++nsteps;
    }
    free(q0); free(qs); free(r);
    if (t_out) *t_out = t;
    return nsteps;
}

/* Euler1D::sodShock (Equations_Euler1D.cpp:317-332) point-sampled at the
 * cell centres of FV_System1D.cpp:25. */
void orc_ic_sod(int nx, double xl, double xr, double gamma, double* q) {
    double dx = (xr - xl) / nx, g1 = gamma - 1.0;
    for (int i = 0; i < nx; ++i) {
        double x = ((double)(i + FV1_K) - (double)FV1_K + 0.5) * dx + xl;
        double w[3], qq[3];
        if (x < 0.5) { w[0] = 1.0; w[1] = 0.0; w[2] = 1.0; }
        else { w[0] = 0.125; w[1] = 0.0; w[2] = 0.1; }
        e1_conservative(g1, w, qq);
        for (int e = 0; e < 3; ++e) q[e * nx + i] = qq[e];
    }
}

/* ---- test hooks: smoothness indicators, exported for independent pinning ---- */
void orc_teno5_indicators(const double u[6], double SI[6], double b35[2]) {
    lag_init();
    smoothness5(u, SI);
    smoothness5(u + 1, SI + 3);
    b35[0] = b35_of(u);
    b35[1] = b35_of(u + 1);
}

void orc_hteno_indicators(const double u[3], const double ud[3], double SI[3], double SId[3], double brk[2]) {
    lag_init();
    herm_si(u, ud, SI);
    herm_sid(u, ud, SId);
    brk[0] = herm_brk(u, ud);
    brk[1] = herm_brkd(u, ud);
}
