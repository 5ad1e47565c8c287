// htx_variants.h -- the reference's remaining reconstruction and quadrature
// routines as host+device functions of one value set:
//   Schemes::Lagrange  weno5N, teno5N, teno5a, weno5IS, weno5ISmod,
//                      weno5IS_Extended, weno5Z(u, diff), teno5(u, diff),
//                      weno3JS, weno3JS2PointsQuadrature, the 2- and 4-point
//                      Gauss quadratures (Schemes_Lagrange.cpp:315-1004,
//                      1086-1149, 1265-1294)
//   Schemes::Hermite   hweno, linear, htenoi, hweno2PointsGLQ,
//                      hweno4PointsGLQ (Schemes_Hermite.cpp:93-293, 381-513,
//                      543-595, 670-777)
// Each follows its reference routine operation by operation (left-to-right
// sums, divisions kept, no contraction) in every build, so the GPU solvers
// that select one (the 1D and 2D reconstructions, the 2D Hermite face
// reconstruction) and the host classes compute the same values.  Integer
// powers x^n (Eigen .pow(n), std::pow) are formed exactly in double-double
// arithmetic and rounded once (pow_n); the fractional powers of weno5N
// (2.5) and weno5Z(u, diff) (0.75) call pow.
#pragma once
#include "htx_numerics.h"

namespace HTX_NS {

// x^n (n = 2, 3, 4, 6) rounded once to the nearest double: the power is
// carried as an unevaluated sum hi + lo of two doubles (products split with
// an FMA), whose relative error is below 2^-100, and the final addition
// rounds it; this equals a correctly rounded pow(x, n) except for powers
// within 2^-100 relative of a rounding midpoint.
HTX_HD double pow_n(double x, int n) {
    HTX_NO_CONTRACT
    if (n == 2) return x * x;
    const double h2 = x * x, l2 = fma(x, x, -h2);
    double h, l;
    if (n == 4 || n == 6) {
        double b = h2, bl = l2;
        if (n == 6) {  // x^3 = x^2 x
            const double h3 = h2 * x;
            bl = fma(h2, x, -h3) + l2 * x;
            b = h3;
        }
        h = b * b;  // (b + bl)^2
        l = fma(b, b, -h) + 2.0 * (b * bl);
    } else {  // n == 3
        h = h2 * x;
        l = fma(h2, x, -h) + l2 * x;
    }
    if (!(fabs(h) <= 1.7976931348623157e308)) return h;  // inf or NaN
    return h + l;
}

// ---------------------------------------------------------------------------
// Schemes::Lagrange, 5-point reconstructions on u_{i-2..i+3}.  The candidate
// values of the three 3-cell stencils at x_{i+1/2} (crj, Schemes_Lagrange.cpp
// :125-136): minus side p0 = crj.col(1) u_{i..i+2}, p1 = crj.col(2)
// u_{i-1..i+1}, p2 = crj.col(3) u_{i-2..i}; plus side q0 = crj.col(0)
// u_{i+1..i+3}, q1 = crj.col(1) u_{i..i+2}, q2 = crj.col(2) u_{i-1..i+1}.
HTX_HD void cand5(const double u[6], double p[3], double q[3]) {
    HTX_NO_CONTRACT
    const double a = 1.0 / 3.0, b = 5.0 / 6.0, c = -1.0 / 6.0, d = 11.0 / 6.0, e = -7.0 / 6.0;
    p[0] = (a * u[2] + b * u[3]) + c * u[4];
    p[1] = (c * u[1] + b * u[2]) + a * u[3];
    p[2] = (a * u[0] + e * u[1]) + d * u[2];
    q[0] = (d * u[3] + e * u[4]) + a * u[5];
    q[1] = (a * u[2] + b * u[3]) + c * u[4];
    q[2] = (c * u[1] + b * u[2]) + a * u[3];
}

HTX_HD double dot3v(const double a[3], const double b[3]) {
    HTX_NO_CONTRACT
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

HTX_HD double catchnan(double x) { return x != x ? 1e-15 : x; }

// smoothnessMeasure (Schemes_Lagrange.cpp:1013-1026) on the window v0..v4
HTX_HD void si5(const double* v, double s[3]) {
    HTX_NO_CONTRACT
    const double c1 = 13.0 / 12.0;
    const double c0 = v[2] - 2.0 * v[3] + v[4], b0 = (3.0 * v[2] - 4.0 * v[3] + v[4]) / 2.0;
    const double cc1 = v[1] - 2.0 * v[2] + v[3], b1 = (v[1] - v[3]) / 2.0;
    const double c2 = v[0] - 2.0 * v[1] + v[2], b2 = (v[0] - 4.0 * v[1] + 3.0 * v[2]) / 2.0;
    s[0] = b0 * b0 + c1 * c0 * c0;
    s[1] = b1 * b1 + c1 * cc1 * cc1;
    s[2] = b2 * b2 + c1 * c2 * c2;
}

// fullpointsSmootness (:1029-1034)
HTX_HD double b35_5(const double* v) {
    HTX_NO_CONTRACT
    double b = v[4] * (6908.0 * v[4] - 51001.0 * v[3] + 67923.0 * v[2] - 38947.0 * v[1] + 8209.0 * v[0]) +
               v[3] * (104963.0 * v[3] - 299076.0 * v[2] + 179098.0 * v[1] - 38947.0 * v[0]) +
               v[2] * (231153.0 * v[2] - 299076.0 * v[1] + 67923.0 * v[0]) +
               v[1] * (104963.0 * v[1] - 51001.0 * v[0]) + 6908.0 * v[0] * v[0];
    return b / 5040.0;
}

// cutOff (:1141-1149) on the unnormalised weights g, then the kept linear
// weights normalised and the candidates combined, catchnan (the tail every
// TENO variant shares, e.g. :490-497)
HTX_HD double teno_tail(const double g[3], const double p[3], const double d[3], double ct) {
    HTX_NO_CONTRACT
    const double gs = (g[0] + g[1]) + g[2];
    double w[3];
    for (int k = 0; k < 3; ++k) w[k] = (g[k] / gs < ct ? 0.0 : 1.0) * d[k];
    const double ws = (w[0] + w[1]) + w[2];
    for (int k = 0; k < 3; ++k) w[k] = w[k] / ws;
    return catchnan(dot3v(p, w));
}

// weights w normalised and combined with the candidates, catchnan
HTX_HD double weno_tail(double w[3], const double p[3]) {
    HTX_NO_CONTRACT
    const double s = (w[0] + w[1]) + w[2];
    for (int k = 0; k < 3; ++k) w[k] = w[k] / s;
    return catchnan(dot3v(p, w));
}

// the N-type smoothness measures and tau of weno5N / teno5N (:440-478,
// :553-593) on the window v0..v4
HTX_HD void si5_n(const double* v, double s[3], double& t4) {
    HTX_NO_CONTRACT
    const double th = 0.1;
    s[2] = th * fabs(v[0] - 3.0 * v[1] + 2.0 * v[2]) + fabs(v[0] - 2.0 * v[1] + v[2]);
    s[1] = th * fabs(v[3] - v[2]) + fabs(v[1] - 2.0 * v[2] + v[3]);
    s[0] = th * fabs(v[3] - v[2]) + fabs(v[2] - 2.0 * v[3] + v[4]);
    t4 = fabs(v[0] - 4.0 * v[1] + 6.0 * v[2] - 4.0 * v[3] + v[4]);
}

// Lagrange::weno5N (Schemes_Lagrange.cpp:440-478)
HTX_HD void weno5n(const double u[6], double& um, double& up) {
    HTX_NO_CONTRACT
    const double dm[3] = {0.3, 0.6, 0.1}, dp[3] = {0.1, 0.6, 0.3};
    double p[3], q[3], s[3], t4, w[3];
    cand5(u, p, q);
    for (int side = 0; side < 2; ++side) {
        si5_n(u + side, s, t4);
        const double tau = pow(t4, 2.5);
        const double* d = side ? dp : dm;
        for (int k = 0; k < 3; ++k) w[k] = d[k] * (1.0 + tau / pow_n(s[k] + 1e-6, 2));
        (side ? up : um) = weno_tail(w, side ? q : p);
    }
}

// Lagrange::teno5N (:553-593): tau = (fourth difference)^2, weights
// (1 + tau / (s + 1e-6)^2)^6, cut-off at 5e-4
HTX_HD void teno5n(const double u[6], double& um, double& up) {
    HTX_NO_CONTRACT
    const double dm[3] = {0.3, 0.6, 0.1}, dp[3] = {0.1, 0.6, 0.3};
    double p[3], q[3], s[3], t4, g[3];
    cand5(u, p, q);
    for (int side = 0; side < 2; ++side) {
        si5_n(u + side, s, t4);
        const double tau = pow_n(t4, 2);
        for (int k = 0; k < 3; ++k) g[k] = pow_n(1.0 + tau / pow_n(s[k] + 1e-6, 2), 6);
        (side ? up : um) = teno_tail(g, side ? q : p, side ? dp : dm, 5e-4);
    }
}

// Lagrange::teno5a (:1265-1294): weights (1 / (s + 1e-40))^6, cut-off at 5e-4
HTX_HD void teno5a(const double u[6], double& um, double& up) {
    HTX_NO_CONTRACT
    const double dm[3] = {0.3, 0.6, 0.1}, dp[3] = {0.1, 0.6, 0.3};
    double p[3], q[3], s[3], g[3];
    cand5(u, p, q);
    for (int side = 0; side < 2; ++side) {
        si5(u + side, s);
        for (int k = 0; k < 3; ++k) g[k] = pow_n(1.0 / (s[k] + 1e-40), 6);
        (side ? up : um) = teno_tail(g, side ? q : p, side ? dp : dm, 5e-4);
    }
}

// Lagrange::teno5(u, diff) (:514-551): weights (1 + (tau / (s + diff^4))^2)^3
// with TENO5's tau, cut-off at 5e-4
HTX_HD void teno5_diff(const double u[6], double diff, double& um, double& up) {
    HTX_NO_CONTRACT
    const double dm[3] = {0.3, 0.6, 0.1}, dp[3] = {0.1, 0.6, 0.3};
    const double df = pow_n(diff, 4);
    double p[3], q[3], s[3], g[3];
    cand5(u, p, q);
    for (int side = 0; side < 2; ++side) {
        si5(u + side, s);
        const double tau = fabs(b35_5(u + side) - (s[0] + s[2] + 4.0 * s[1]) / 6.0);
        for (int k = 0; k < 3; ++k) g[k] = pow_n(1.0 + pow_n(tau / (s[k] + df), 2), 3);
        (side ? up : um) = teno_tail(g, side ? q : p, side ? dp : dm, 5e-4);
    }
}

// Lagrange::weno5Z(u, diff) (:410-438): weights d (1 + tau / (s + 1e-40)^0.75)
// with TENO5's tau; the plus side takes the full-stencil indicator of the
// minus side's window (fullpointsSmootness(u.segment(0, 5)) at :431), and
// diff is not used
HTX_HD void weno5z_diff(const double u[6], double& um, double& up) {
    HTX_NO_CONTRACT
    const double dm[3] = {0.3, 0.6, 0.1}, dp[3] = {0.1, 0.6, 0.3};
    double p[3], q[3], s[3], w[3];
    cand5(u, p, q);
    const double b35 = b35_5(u);
    for (int side = 0; side < 2; ++side) {
        si5(u + side, s);
        const double tau = fabs(b35 - (s[0] + s[2] + 4.0 * s[1]) / 6.0);
        const double* d = side ? dp : dm;
        for (int k = 0; k < 3; ++k) w[k] = (1.0 + tau / pow(s[k] + 1e-40, 0.75)) * d[k];
        (side ? up : um) = weno_tail(w, side ? q : p);
    }
}

// smoothnessMeasureIS (:1036-1050) / smoothnessMeasureISmod (:1052-1066) on
// v0..v4: the 4 nonlinear factors (mod: divisors s + df, s b12 + df)
HTX_HD void si5_is(const double* v, bool mod, double df, double w[4]) {
    HTX_NO_CONTRACT
    const double c1 = 13.0 / 12.0, eps = 1e-6;
    const double s0 = pow_n(v[3] - v[2], 2), s1 = pow_n(v[2] - v[1], 2);
    const double s2 = 0.25 * pow_n(3.0 * v[2] - 4.0 * v[3] + v[4], 2) + c1 * pow_n(v[2] - 2.0 * v[3] + v[4], 2);
    const double s3 = 0.25 * pow_n(v[0] - 4.0 * v[1] + 3.0 * v[2], 2) + c1 * pow_n(v[0] - 2.0 * v[1] + v[2], 2);
    const double b12 = 0.25 * pow_n(v[1] - v[3], 2) + c1 * pow_n(v[1] - 2.0 * v[2] + v[3], 2);
    const double tau = 0.25 * pow_n(v[4] - 2.0 * (v[3] - v[1]) - v[0], 2) +
                       c1 * pow_n(v[4] - 4.0 * v[3] + 6.0 * v[2] - 4.0 * v[1] + v[0], 2);
    if (mod) {
        w[0] = 1.0 + tau * tau / (s0 * b12 + df);
        w[1] = 1.0 + tau * tau / (s1 * b12 + df);
        w[2] = 1.0 + tau / (s2 + df);
        w[3] = 1.0 + tau / (s3 + df);
    } else {
        w[0] = 1.0 + pow_n(tau, 2) / (s0 + eps) / (b12 + eps);
        w[1] = 1.0 + pow_n(tau, 2) / (s1 + eps) / (b12 + eps);
        w[2] = 1.0 + tau / (s2 + eps);
        w[3] = 1.0 + tau / (s3 + eps);
    }
}

// Lagrange::weno5IS (:596-620) and weno5ISmod(u, diff) (:622-647, divisor
// offsets diff^3): four candidates (two linear, two quadratic) with the
// optimal weights (0.4, 0.2, 0.3, 0.1) / (0.2, 0.4, 0.1, 0.3)
HTX_HD void weno5is(const double u[6], bool mod, double diff, double& um, double& up) {
    HTX_NO_CONTRACT
    const double dm[4] = {0.4, 0.2, 0.3, 0.1}, dp[4] = {0.2, 0.4, 0.1, 0.3};
    const double df = mod ? pow_n(diff, 3) : 0.0;
    double p3[3], q3[3], w[4];
    cand5(u, p3, q3);
    for (int side = 0; side < 2; ++side) {
        double c[4];
        if (side == 0) {
            c[0] = 0.5 * (u[2] + u[3]);
            c[1] = -0.5 * u[1] + 1.5 * u[2];
            c[2] = p3[0];  // crj.col(1) u_{i..i+2}
            c[3] = p3[2];  // crj.col(3) u_{i-2..i}
        } else {
            c[0] = 1.5 * u[3] - 0.5 * u[4];
            c[1] = 0.5 * (u[2] + u[3]);
            c[2] = q3[0];  // crj.col(0) u_{i+1..i+3}
            c[3] = q3[2];  // crj.col(2) u_{i-1..i+1}
        }
        si5_is(u + side, mod, df, w);
        const double* d = side ? dp : dm;
        for (int k = 0; k < 4; ++k) w[k] = d[k] * w[k];
        const double s = ((w[0] + w[1]) + w[2]) + w[3];
        for (int k = 0; k < 4; ++k) w[k] = w[k] / s;
        (side ? up : um) = catchnan(((c[0] * w[0] + c[1] * w[1]) + c[2] * w[2]) + c[3] * w[3]);
    }
}

// Lagrange::weno5IS_Extended (:649-717 with smoothnessMeasureISExtended
// :1068-1084): six linear candidates, the optimal weights (dkmise, dkpise
// :288-303) split into positive and negative groups, each candidate kept
// when its normalised fourth-power weight reaches 1e-5
HTX_HD void weno5is_ext(const double u[6], double& um, double& up) {
    HTX_NO_CONTRACT
    const double c1 = 13.0 / 12.0, eps = 1e-6;
    const double dmi[6] = {-1.0 / 45.0, 11.0 / 90.0, 1.0 / 5.0, 2.0 / 5.0, 1.0 / 5.0, 1.0 / 10.0};
    const double dpi[6] = {1.0 / 10.0, 1.0 / 5.0, 4.0 / 10.0, 2.0 / 10.0, 11.0 / 90.0, -1.0 / 45.0};
    for (int side = 0; side < 2; ++side) {
        const double* v = u + side;
        // candidates
        double c[6];
        if (side == 0) {
            c[0] = -1.5 * u[0] + 2.5 * u[1];
            c[1] = -0.5 * u[1] + 1.5 * u[2];
            c[2] = c[1];
            c[3] = 0.5 * (u[2] + u[3]);
            c[4] = c[3];
            c[5] = 1.5 * u[3] - 0.5 * u[4];
        } else {
            c[0] = -0.5 * u[1] + 1.5 * u[2];
            c[1] = 0.5 * (u[2] + u[3]);
            c[2] = c[1];
            c[3] = 1.5 * u[3] - 0.5 * u[4];
            c[4] = c[3];
            c[5] = 2.5 * u[4] - 1.5 * u[5];
        }
        // smoothness (this routine orders the 3-cell measures left to right)
        const double cc2 = v[2] - 2.0 * v[3] + v[4], b2 = (3.0 * v[2] - 4.0 * v[3] + v[4]) / 2.0;
        const double cc1 = v[1] - 2.0 * v[2] + v[3], b1 = (v[1] - v[3]) / 2.0;
        const double cc0 = v[0] - 2.0 * v[1] + v[2], b0 = (v[0] - 4.0 * v[1] + 3.0 * v[2]) / 2.0;
        const double S0 = b0 * b0 + c1 * cc0 * cc0, S1 = b1 * b1 + c1 * cc1 * cc1, S2 = b2 * b2 + c1 * cc2 * cc2;
        double s4[4];
        for (int i = 0; i < 4; ++i) s4[i] = pow_n(v[i + 1] - v[i], 2);
        double tau = 0.25 * pow_n(v[4] - 2.0 * (v[3] - v[1]) - v[0], 2) +
                     c1 * pow_n(v[4] - 4.0 * v[3] + 6.0 * v[2] - 4.0 * v[1] + v[0], 2);
        tau *= tau;
        double w[6];
        w[0] = 1.0 + tau / (S0 + eps) / (s4[0] + eps);
        w[1] = 1.0 + tau / (S0 + eps) / (s4[1] + eps);
        w[2] = 1.0 + tau / (S1 + eps) / (s4[1] + eps);
        w[3] = 1.0 + tau / (S1 + eps) / (s4[2] + eps);
        w[4] = 1.0 + tau / (S2 + eps) / (s4[2] + eps);
        w[5] = 1.0 + tau / (S2 + eps) / (s4[3] + eps);
        double sum = 0.0;
        for (int i = 0; i < 6; ++i) {
            w[i] = pow_n(w[i], 4);
            sum = i ? sum + w[i] : w[i];
        }
        for (int i = 0; i < 6; ++i) w[i] = (w[i] / sum < 1e-5) ? 0.0 : 1.0;
        // optimal weights split into positive / negative groups (:291-303)
        const double* d = side ? dpi : dmi;
        double pos[6], neg[6], ps = 0.0, ns = 0.0;
        for (int i = 0; i < 6; ++i) {
            pos[i] = 0.5 * (d[i] + 3.0 * fabs(d[i]));
            neg[i] = pos[i] - d[i];
            ps = i ? ps + pos[i] : pos[i];
            ns = i ? ns + neg[i] : neg[i];
        }
        double wp[6], wn[6], sp = 0.0, sn = 0.0;
        for (int i = 0; i < 6; ++i) {
            wp[i] = (pos[i] / ps) * w[i];
            wn[i] = (neg[i] / ns) * w[i];
            sp = i ? sp + wp[i] : wp[i];
            sn = i ? sn + wn[i] : wn[i];
        }
        double dp_ = 0.0, dn_ = 0.0;
        for (int i = 0; i < 6; ++i) {
            const double a = c[i] * (wp[i] / sp), b = c[i] * (wn[i] / sn);
            dp_ = i ? dp_ + a : a;
            dn_ = i ? dn_ + b : b;
        }
        (side ? up : um) = catchnan(ps * dp_ - ns * dn_);
    }
}

// The selectable 5-point characteristic reconstructions of the GPU solvers
// (FV_System1D.cpp:200-201, FV_System2D.cpp:379, :394, where a user of the
// reference swaps the call): kind 4 weno5N, 5 teno5N, 6 teno5a, 7 weno5IS,
// 8 weno5IS_Extended, 9 weno5ISmod(u, diff), 10 weno5Z(u, diff), 11
// teno5(u, diff), 12 linear; diff = the cell size normal to the face.
// Kinds 0-3 (teno5, teno5i, weno5JS, weno5Z) have their own paths.
enum { RECON_WENO5N = 4, RECON_TENO5N, RECON_TENO5A, RECON_WENO5IS, RECON_WENO5IS_EXT, RECON_WENO5ISMOD,
       RECON_WENO5Z_DIFF, RECON_TENO5_DIFF, RECON_LINEAR, RECON_LAST = RECON_LINEAR };

HTX_HD void recon5_variant(int kind, const double u[6], double diff, double& um, double& up) {
    switch (kind) {
        case RECON_WENO5N: weno5n(u, um, up); break;
        case RECON_TENO5N: teno5n(u, um, up); break;
        case RECON_TENO5A: teno5a(u, um, up); break;
        case RECON_WENO5IS: weno5is(u, false, 0.0, um, up); break;
        case RECON_WENO5IS_EXT: weno5is_ext(u, um, up); break;
        case RECON_WENO5ISMOD: weno5is(u, true, diff, um, up); break;
        case RECON_WENO5Z_DIFF: weno5z_diff(u, um, up); break;
        case RECON_TENO5_DIFF: teno5_diff(u, diff, um, up); break;
        default: linear5(u, um, up); break;
    }
}

// ---------------------------------------------------------------------------
// Third-order WENO (Schemes_Lagrange.cpp:315-356)

// Lagrange::weno3JS on u_{i-1..i+2}.  As the reference writes it (:333), the
// second output is catchnan of the first: both components are u^-_{i+1/2}.
HTX_HD void weno3js(const double u[4], double& o0, double& o1) {
    HTX_NO_CONTRACT
    const double eps = 1e-6;
    // crjm3rd columns (-0.5, 1.5), (0.5, 0.5); dkm3rd = (1/3, 2/3)
    double p0 = -0.5 * u[0] + 1.5 * u[1], p1 = 0.5 * u[1] + 0.5 * u[2];
    double s0 = pow_n(u[0] - u[1], 2), s1 = pow_n(u[1] - u[2], 2);
    double w0 = (1.0 / 3.0) / pow_n(eps + s0, 2), w1 = (2.0 / 3.0) / pow_n(eps + s1, 2);
    double s = w0 + w1;
    w0 = w0 / s; w1 = w1 / s;
    o0 = catchnan(p0 * w0 + p1 * w1);
    // the plus side (crjp3rd, dkp3rd) is evaluated and then overwritten
    o1 = o0;
}

// Lagrange::weno3JS2PointsQuadrature (:337-356) on u_{i-1..i+1}: values at
// the two Gauss-Legendre points x_i + xi dx, xi = +-sqrt(3)/6, order (+, -)
HTX_HD void weno3js_2q(const double u[3], double& o0, double& o1) {
    HTX_NO_CONTRACT
    const double sq3 = 0x1.bb67ae8584caap+0, eps = 1e-6;
    const double c6 = sq3 / 3.0;
    const double s0 = c6 * pow_n(u[0] - u[1], 2), s1 = c6 * pow_n(u[1] - u[2], 2);
    const double e0 = pow_n(eps + s0, 2), e1 = pow_n(eps + s1, 2);
    // crjp3rdg (:168-171): (0,0) -sq3/6, (1,0) sq3/6 + 1, (0,1) -sq3/6 + 1, (1,1) sq3/6
    double p0 = (-sq3 / 6.0) * u[0] + (sq3 / 6.0 + 1.0) * u[1];
    double p1 = (-sq3 / 6.0 + 1.0) * u[1] + (sq3 / 6.0) * u[2];
    double w0 = 0.5 / e0, w1 = 0.5 / e1, s = w0 + w1;
    w0 = w0 / s; w1 = w1 / s;
    o0 = p0 * w0 + p1 * w1;
    // crjm3rdg: rows and columns of crjp3rdg swapped (:172-174)
    p0 = (sq3 / 6.0) * u[0] + (-sq3 / 6.0 + 1.0) * u[1];
    p1 = (sq3 / 6.0 + 1.0) * u[1] + (-sq3 / 6.0) * u[2];
    w0 = 0.5 / e0; w1 = 0.5 / e1; s = w0 + w1;
    w0 = w0 / s; w1 = w1 / s;
    o1 = p0 * w0 + p1 * w1;
}

// ---------------------------------------------------------------------------
// 2- and 4-point Gauss quadratures of the 5th-order WENO / TENO
// (Schemes_Lagrange.cpp:719-1004).  Their coefficient tables
// (Schemes_Lagrange.cpp:156-286), formed with the reference's expressions;
// the Gauss-Legendre ones need pow(., 1.5) and pow(., 2.5) and are filled on
// the host (fill_quad_var).
struct QuadVar {
    double crjgp[3][3], crjgm[3][3], dkpg[3], dkmg[3];            // 2-point (+-sqrt(3)/6)
    double crjgpLo[3][3], crjgmLo[3][3], dkpLo[3], dkmLo[3];      // Gauss-Lobatto (+-sqrt(5)/10)
    double crjgpL12[3][3], crjgmL12[3][3], dkpL12[3], dkmL12[3];  // Gauss-Legendre inner pair
    double crjgpL34[3][3], crjgmL34[3][3], dkpL34[3], dkmL34[3];  // Gauss-Legendre outer pair
    double c2, c3, c4, c5, c7, c8, c9, c10;
};

inline void fill_quad_var(QuadVar& Q) {
    const double sq3 = sqrt(3.0), sq5 = sqrt(5.0);
    Q.c2 = sq3 / 12.0;
    Q.c3 = 37.0 / 108.0 * sq3;
    Q.c4 = sq5 / 20.0;
    Q.c5 = 61.0 * sq5 / 300.0;
    Q.c7 = sqrt(0.525e3 - 0.70e2 * sqrt(0.30e2)) / 0.140e3;
    Q.c8 = pow(0.525e3 - 0.70e2 * sqrt(0.30e2), 0.3e1 / 0.2e1) / 0.514500e6 +
           sqrt(0.525e3 - 0.70e2 * sqrt(0.30e2)) / 0.35e2;
    Q.c9 = sqrt(0.525e3 + 0.70e2 * sqrt(0.30e2)) / 0.140e3;
    Q.c10 = pow(0.525e3 + 0.70e2 * sqrt(0.30e2), 0.3e1 / 0.2e1) / 0.514500e6 +
            sqrt(0.525e3 + 0.70e2 * sqrt(0.30e2)) / 0.35e2;
    // mirror: M(r, c) = P(2 - r, 2 - c) (rows 0, 2 and columns 0, 2 swapped)
    auto mirror = [](const double P[3][3], double M[3][3]) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) M[r][c] = P[2 - r][2 - c];
    };
    // crjgp (:156-166)
    Q.crjgp[0][0] = sq3 / 12.0; Q.crjgp[1][0] = -sq3 / 3.0; Q.crjgp[2][0] = 1.0 + sq3 / 4.0;
    Q.crjgp[0][1] = -sq3 / 12.0; Q.crjgp[1][1] = 1.0; Q.crjgp[2][1] = sq3 / 12.0;
    Q.crjgp[0][2] = -sq3 / 4.0 + 1.0; Q.crjgp[1][2] = sq3 / 3.0; Q.crjgp[2][2] = -sq3 / 12.0;
    mirror(Q.crjgp, Q.crjgm);
    // dkpg, dkmg (:223-230)
    Q.dkpg[0] = (210.0 - sq3) / 1080; Q.dkpg[1] = 11.0 / 18.0; Q.dkpg[2] = (210.0 + sq3) / 1080;
    Q.dkmg[0] = Q.dkpg[2]; Q.dkmg[1] = Q.dkpg[1]; Q.dkmg[2] = Q.dkpg[0];
    // Gauss-Lobatto (:185-221)
    Q.crjgpLo[0][0] = sq5 / 20.0 - 1.0 / 60.0; Q.crjgpLo[1][0] = -sq5 / 5.0 + 1.0 / 30.0;
    Q.crjgpLo[2][0] = 3.0 * sq5 / 20.0 + 59 / 60.0;
    Q.crjgpLo[0][1] = -sq5 / 20.0 - 1.0 / 60.0; Q.crjgpLo[1][1] = 31.0 / 30.0; Q.crjgpLo[2][1] = sq5 / 20.0 - 1.0 / 60.0;
    Q.crjgpLo[0][2] = -3.0 * sq5 / 20.0 + 59.0 / 60.0; Q.crjgpLo[1][2] = sq5 / 5.0 + 1.0 / 30.0;
    Q.crjgpLo[2][2] = -sq5 / 20.0 - 1.0 / 60.0;
    mirror(Q.crjgpLo, Q.crjgmLo);
    double den = 660.0 * sq5 - 220.0;
    Q.dkpLo[0] = (132.0 * sq5 + 22.0) / den; Q.dkpLo[1] = (387.0 * sq5 - 129.0) / den;
    Q.dkpLo[2] = (141.0 * sq5 - 113.0) / den;
    den = 660.0 * sq5 + 220.0;
    Q.dkmLo[0] = (132.0 * sq5 - 22.0) / den; Q.dkmLo[1] = (387.0 * sq5 + 129.0) / den;
    Q.dkmLo[2] = (141.0 * sq5 + 113.0) / den;
    // Gauss-Legendre (:232-286)
    const double s30 = sqrt(0.30e2);
    const double A = sqrt(0.525e3 - 0.70e2 * s30), B = sqrt(0.525e3 + 0.70e2 * s30);
    const double Ap = 0.525e3 - 0.70e2 * s30, Bp = 0.525e3 + 0.70e2 * s30;
    Q.crjgpL12[0][0] = A / 0.140e3 - s30 / 0.140e3 + 0.1e1 / 0.84e2;
    Q.crjgpL12[1][0] = -A / 0.35e2 + s30 / 0.70e2 - 0.1e1 / 0.42e2;
    Q.crjgpL12[2][0] = 0.3e1 / 0.140e3 * A - s30 / 0.140e3 + 0.85e2 / 0.84e2;
    Q.crjgpL12[0][1] = -A / 0.140e3 - s30 / 0.140e3 + 0.1e1 / 0.84e2;
    Q.crjgpL12[1][1] = s30 / 0.70e2 + 0.41e2 / 0.42e2;
    Q.crjgpL12[2][1] = A / 0.140e3 - s30 / 0.140e3 + 0.1e1 / 0.84e2;
    Q.crjgpL12[0][2] = -0.3e1 / 0.140e3 * A - s30 / 0.140e3 + 0.85e2 / 0.84e2;
    Q.crjgpL12[1][2] = A / 0.35e2 + s30 / 0.70e2 - 0.1e1 / 0.42e2;
    Q.crjgpL12[2][2] = -A / 0.140e3 - s30 / 0.140e3 + 0.1e1 / 0.84e2;
    mirror(Q.crjgpL12, Q.crjgmL12);
    den = (0.1234800000e10 * s30 - 0.9116940000e10) * A + 0.11174940000e11 * s30 - 0.52238900000e11;
    Q.dkpL12[2] = ((0.257250000e9 * s30 - 0.1899362500e10) * A + (-0.42000e5 * s30 + 0.310100e6) * pow(Ap, 0.3e1 / 0.2e1) -
                   0.1197315000e10 * s30 + 0.5597025000e10) / den;
    Q.dkpL12[1] = ((0.1761795000e10 * s30 - 0.13038532500e11) * A + (-0.88200e5 * s30 + 0.16096500e8) * pow(Ap, 0.3e1 / 0.2e1) -
                   0.2520e4 * pow(Ap, 0.5e1 / 0.2e1) + 0.6757590000e10 * s30 - 0.32331425000e11) / den;
    Q.dkpL12[0] = ((0.26092500e8 * s30 - 0.256576250e9) * A + (-0.2100e4 * s30 - 0.3838100e7) * pow(Ap, 0.3e1 / 0.2e1) +
                   0.630e3 * pow(Ap, 0.5e1 / 0.2e1) + 0.5614665000e10 * s30 - 0.25504500000e11) / den;
    Q.dkmL12[0] = Q.dkpL12[2]; Q.dkmL12[1] = Q.dkpL12[1]; Q.dkmL12[2] = Q.dkpL12[0];
    Q.crjgpL34[0][0] = B / 0.140e3 + s30 / 0.140e3 + 0.1e1 / 0.84e2;
    Q.crjgpL34[1][0] = -B / 0.35e2 - s30 / 0.70e2 - 0.1e1 / 0.42e2;
    Q.crjgpL34[2][0] = 0.3e1 / 0.140e3 * B + s30 / 0.140e3 + 0.85e2 / 0.84e2;
    Q.crjgpL34[0][1] = -B / 0.140e3 + s30 / 0.140e3 + 0.1e1 / 0.84e2;
    Q.crjgpL34[1][1] = -s30 / 0.70e2 + 0.41e2 / 0.42e2;
    Q.crjgpL34[2][1] = B / 0.140e3 + s30 / 0.140e3 + 0.1e1 / 0.84e2;
    Q.crjgpL34[0][2] = -0.3e1 / 0.140e3 * B + s30 / 0.140e3 + 0.85e2 / 0.84e2;
    Q.crjgpL34[1][2] = B / 0.35e2 - s30 / 0.70e2 - 0.1e1 / 0.42e2;
    Q.crjgpL34[2][2] = -B / 0.140e3 + s30 / 0.140e3 + 0.1e1 / 0.84e2;
    mirror(Q.crjgpL34, Q.crjgmL34);
    den = (0.1234800000e10 * s30 + 0.9116940000e10) * B - 0.11174940000e11 * s30 - 0.52238900000e11;
    Q.dkpL34[2] = ((0.26092500e8 * s30 + 0.256576250e9) * B + (-0.2100e4 * s30 + 0.3838100e7) * pow(Bp, 0.3e1 / 0.2e1) -
                   0.630e3 * pow(Bp, 0.5e1 / 0.2e1) - 0.5614665000e10 * s30 - 0.25504500000e11) / den;
    Q.dkpL34[1] = ((0.1761795000e10 * s30 + 0.13038532500e11) * B + (-0.88200e5 * s30 - 0.16096500e8) * pow(Bp, 0.3e1 / 0.2e1) +
                   0.2520e4 * pow(Bp, 0.5e1 / 0.2e1) - 0.6757590000e10 * s30 - 0.32331425000e11) / den;
    Q.dkpL34[0] = ((0.257250000e9 * s30 + 0.1899362500e10) * B + (-0.42000e5 * s30 - 0.310100e6) * pow(Bp, 0.3e1 / 0.2e1) +
                   0.1197315000e10 * s30 + 0.5597025000e10) / den;
    Q.dkmL34[0] = Q.dkpL34[2]; Q.dkmL34[1] = Q.dkpL34[1]; Q.dkmL34[2] = Q.dkpL34[0];
}

// the 3-cell candidate values at one point: p_k = C.col(k) u_{k..k+2}
HTX_HD void cand3_at(const double C[3][3], const double u[5], double p[3]) {
    HTX_NO_CONTRACT
    for (int k = 0; k < 3; ++k) p[k] = (C[0][k] * u[k] + C[1][k] * u[k + 1]) + C[2][k] * u[k + 2];
}

// squared first and second differences of the three stencils of u0..u4
// (smoothnessMeasure2PointsGaussLegendre :1086-1099 and the 4-point forms):
// bg_k^2, cg_k^2, stencil k = cells k..k+2
HTX_HD void bgcg(const double u[5], double b2[3], double c2[3]) {
    HTX_NO_CONTRACT
    const double b0 = u[0] - 4.0 * u[1] + 3.0 * u[2], g0 = u[0] - 2.0 * u[1] + u[2];
    const double b1 = u[1] - u[3], g1 = u[1] - 2.0 * u[2] + u[3];
    const double bb = 3.0 * u[2] - 4.0 * u[3] + u[4], gb = u[2] - 2.0 * u[3] + u[4];
    b2[0] = pow_n(b0, 2); b2[1] = pow_n(b1, 2); b2[2] = pow_n(bb, 2);
    c2[0] = pow_n(g0, 2); c2[1] = pow_n(g1, 2); c2[2] = pow_n(gb, 2);
}

// weights of one point: kind 0 JS  d / (1e-6 + s)^2;  1 Z  d (1 + tau / (s + 1e-40));
// 2 TENO  the kept d (mask), normalised; then the candidates combined
HTX_HD double quad_point(int kind, const double p[3], const double d[3], const double s[3], double tau,
                         unsigned keep, bool nan_guard) {
    HTX_NO_CONTRACT
    double w[3];
    for (int k = 0; k < 3; ++k) {
        if (kind == 0)
            w[k] = d[k] / pow_n(1e-6 + s[k], 2);
        else if (kind == 1)
            w[k] = d[k] * (1.0 + tau / (s[k] + 1e-40));
        else
            w[k] = d[k] * ((keep >> k) & 1u ? 1.0 : 0.0);
    }
    const double ws = (w[0] + w[1]) + w[2];
    for (int k = 0; k < 3; ++k) w[k] = w[k] / ws;
    const double r = dot3v(p, w);
    return nan_guard ? catchnan(r) : r;
}

// cutOff on g = (1 + tau / (s + 1e-40))^6: candidate k kept unless
// g_k / sum g < 5e-4; prenorm: the weights are normalised once more before
// cutOff normalises them (teno5th4PointsGaussLobatto :962-963, :987-988)
HTX_HD unsigned quad_keep(const double s[3], double tau, bool prenorm) {
    HTX_NO_CONTRACT
    double g[3];
    for (int k = 0; k < 3; ++k) g[k] = pow_n(1.0 + tau / (s[k] + 1e-40), 6);
    if (prenorm) {
        const double gs = (g[0] + g[1]) + g[2];
        for (int k = 0; k < 3; ++k) g[k] = g[k] / gs;
    }
    const double gs = (g[0] + g[1]) + g[2];
    unsigned keep = 0;
    for (int k = 0; k < 3; ++k) keep |= (g[k] / gs < 5e-4 ? 0u : 1u) << k;
    return keep;
}

// weno5JS2PointsQuadrature (kind 0), weno5Z2PointsQuadrature (1),
// teno5th2PointsQuadrature (2) (:719-791): values at x_i +- sqrt(3)/6 dx
HTX_HD void weno5_2q(const QuadVar& Q, int kind, const double u[5], double& o0, double& o1) {
    HTX_NO_CONTRACT
    double b2[3], c2[3], s[3], p[3];
    bgcg(u, b2, c2);
    for (int k = 0; k < 3; ++k) s[k] = Q.c2 * b2[k] + Q.c3 * c2[k];
    const double tau = fabs(s[0] - s[2]);
    const unsigned keep = kind == 2 ? quad_keep(s, tau, false) : 7u;
    cand3_at(Q.crjgp, u, p);
    o0 = quad_point(kind, p, Q.dkpg, s, tau, keep, false);
    cand3_at(Q.crjgm, u, p);
    o1 = quad_point(kind, p, Q.dkmg, s, tau, keep, false);
}

// the Gauss-Lobatto end points x_{i+-1/2} take the 5th-order candidates
// (crj) ordered with the left stencil first (:796-811)
HTX_HD void cand5_lobatto(const double u[5], double pm[3], double pp[3]) {
    HTX_NO_CONTRACT
    const double a = 1.0 / 3.0, b = 5.0 / 6.0, c = -1.0 / 6.0, d = 11.0 / 6.0, e = -7.0 / 6.0;
    pm[2] = (a * u[2] + b * u[3]) + c * u[4];  // crj.col(1) u2..4
    pm[1] = (c * u[1] + b * u[2]) + a * u[3];  // crj.col(2) u1..3
    pm[0] = (a * u[0] + e * u[1]) + d * u[2];  // crj.col(3) u0..2
    pp[2] = (d * u[2] + e * u[3]) + a * u[4];  // crj.col(0) u2..4
    pp[1] = (a * u[1] + b * u[2]) + c * u[3];  // crj.col(1) u1..3
    pp[0] = (c * u[0] + b * u[1]) + a * u[2];  // crj.col(2) u0..2
}

// weno5JS4PointsGaussLobatto (kind 0), weno5Z4PointsGaussLobatto (1),
// teno5th4PointsGaussLobatto (2) (:793-1004): values at x_{i+1/2},
// x_{i-1/2}, x_i + sqrt(5)/10 dx, x_i - sqrt(5)/10 dx.  The Z form applies
// no catchnan (:914-955).
HTX_HD void weno5_4lobatto(const QuadVar& Q, int kind, const double u[5], double o[4]) {
    HTX_NO_CONTRACT
    const double dkm[3] = {0.3, 0.6, 0.1}, dkp[3] = {0.1, 0.6, 0.3};
    double b2[3], c2[3], sa[3], sb[3], pm[3], pp[3], p[3];
    bgcg(u, b2, c2);
    for (int k = 0; k < 3; ++k) {
        sa[k] = 0.25 * b2[k] + Q.c3 * c2[k];
        sb[k] = Q.c4 * b2[k] + Q.c5 * c2[k];
    }
    const bool ng = kind != 1;
    double tau = fabs(sa[0] - sa[2]);
    unsigned keep = kind == 2 ? quad_keep(sa, tau, true) : 7u;
    cand5_lobatto(u, pm, pp);
    o[0] = quad_point(kind, pm, dkm, sa, tau, keep, ng);
    o[1] = quad_point(kind, pp, dkp, sa, tau, keep, ng);
    tau = fabs(sb[0] - sb[2]);
    keep = kind == 2 ? quad_keep(sb, tau, true) : 7u;
    cand3_at(Q.crjgpLo, u, p);
    o[2] = quad_point(kind, p, Q.dkpLo, sb, tau, keep, ng);
    cand3_at(Q.crjgmLo, u, p);
    o[3] = quad_point(kind, p, Q.dkmLo, sb, tau, keep, ng);
}

// weno5JS4PointsGaussLegendre (kind 0), weno5Z4PointsGaussLegendre (1)
// (:833-912): values at the inner (+, -) then outer (+, -) Gauss-Legendre
// points of the cell
HTX_HD void weno5_4legendre(const QuadVar& Q, int kind, const double u[5], double o[4]) {
    HTX_NO_CONTRACT
    double b2[3], c2[3], sa[3], sb[3], p[3];
    bgcg(u, b2, c2);
    for (int k = 0; k < 3; ++k) {
        sa[k] = Q.c7 * b2[k] + Q.c8 * c2[k];
        sb[k] = Q.c9 * b2[k] + Q.c10 * c2[k];
    }
    double tau = fabs(sa[0] - sa[2]);
    cand3_at(Q.crjgpL12, u, p);
    o[0] = quad_point(kind, p, Q.dkpL12, sa, tau, 7u, true);
    cand3_at(Q.crjgmL12, u, p);
    o[1] = quad_point(kind, p, Q.dkmL12, sa, tau, 7u, true);
    tau = fabs(sb[0] - sb[2]);
    cand3_at(Q.crjgpL34, u, p);
    o[2] = quad_point(kind, p, Q.dkpL34, sb, tau, 7u, true);
    cand3_at(Q.crjgmL34, u, p);
    o[3] = quad_point(kind, p, Q.dkmL34, sb, tau, 7u, true);
}

}  // namespace HTX_NS
