// htx_hermite.h -- device numerics of the Hermite-TENO scheme
// (Schemes::Hermite, Schemes_Hermite.cpp) and the per-face / per-line
// bodies of the FVHermite 2D kernels (FVHermite_System2D.cpp:393-731),
// written as host+device functions of one work item so that the GPU kernels
// and a CPU build of them run the same code.
#pragma once
#include "htx_internal.h"
#include "htx_variants.h"

namespace HTX_NS {

// Coefficients of Hermite::setGridSize (Schemes_Hermite.cpp:31-66) for the
// grid size dif, filled on the host with the reference's expressions.
struct HermConsts {
    double crj[3][3], crjd[4][3];
    double dkm[3], dkp[3], dkmd[3], dkpd[3];
    double dif;
};

inline void fill_herm_consts(HermConsts& H, double dif) {
    H.dif = dif;
    H.crj[0][0] = -7.0 / 6.0; H.crj[1][0] = 13.0 / 6.0; H.crj[2][0] = -2.0 / 3.0 * dif;
    H.crj[0][1] = 1.0 / 6.0;  H.crj[1][1] = 5.0 / 6.0;  H.crj[2][1] = -1.0 / 3.0 * dif;
    H.crj[0][2] = -1.0 / 6.0; H.crj[1][2] = 5.0 / 6.0;  H.crj[2][2] = 1.0 / 3.0;
    H.crjd[0][0] = 4.0 / dif;  H.crjd[1][0] = -4.0 / dif; H.crjd[2][0] = 1.5;        H.crjd[3][0] = 3.5;
    H.crjd[0][1] = -2.0 / dif; H.crjd[1][1] = 2.0 / dif;  H.crjd[2][1] = -0.5;       H.crjd[3][1] = -0.5;
    H.crjd[0][2] = 0.25 / dif; H.crjd[1][2] = -1.0 / dif; H.crjd[2][2] = 0.75 / dif; H.crjd[3][2] = 0.5;
    H.dkm[0] = 9.0 / 80.0; H.dkm[1] = 42.0 / 80.0; H.dkm[2] = 29.0 / 80.0;
    H.dkp[0] = H.dkm[1]; H.dkp[1] = H.dkm[0]; H.dkp[2] = H.dkm[2];
    H.dkmd[0] = 1.0 / 18.0; H.dkmd[1] = 5.0 / 6.0; H.dkmd[2] = 1.0 / 9.0;
    H.dkpd[0] = H.dkmd[1]; H.dkpd[1] = H.dkmd[0]; H.dkpd[2] = H.dkmd[2];
}

// Hermite::fullPointsSI (Schemes_Hermite.cpp:597-603); u, d are 3 consecutive
// cell averages and scaled derivative moments.
HTX_HD double herm_brk(double u0, double u1, double u2, double d0, double d2) {
    HTX_NO_CONTRACT
    const double v = fabs((1029270.0 * u2 - 2129184.0 * u1 + 1099914.0 * u0 + 281121.0 * d0 - 494958.0 * d2) * d0 +
                (-1029270.0 * u0 + 2129184.0 * u1 - 1099914.0 * u2 + 281121.0 * d2) * d2 +
                (1099445.0 * u2 - 4317056.0 * u1 + 2118166.0 * u0) * u2 + 4317056.0 * (u1 - u0) * u1 +
                1099445.0 * u0 * u0);
    return quot_const(v, 6720.0, 1.0 / 6720.0);
}

// Hermite::fullPointsSId (:605-612)
HTX_HD double herm_brkd(double u0, double u1, double u2, double d0, double d1, double d2) {
    HTX_NO_CONTRACT
    const double v = fabs((3156862.0 * d2 + 13189240.0 * d1 + 1710479.0 * d0 - 9617754.0 * u2 - 531552.0 * u1 +
                 10149306.0 * u0) * d0 +
                (13189240.0 * d2 + 26511200.0 * d1 - 39700440.0 * u2 + 39700440.0 * u0) * d1 +
                (1710479.0 * d2 - 10149306.0 * u2 + 531552.0 * u1 + 9617754.0 * u0) * d2 +
                (15136011.0 * u2 - 1076544.0 * u1 - 29195478.0 * u0) * u2 + 1076544.0 * (u1 - u0) * u1 +
                15136011.0 * u0 * u0);
    return quot_const(v, 1680.0, 1.0 / 1680.0);
}

// one HTENO side (Schemes_Hermite.cpp:311-320 pattern): tau from the full
// indicator, gamma^6, cutOff at C_T = 1e-3 (:614-628), weights, catchnan
HTX_HD double hteno_side(double brk, double s0, double s1, double s2, double p0, double p1, double p2, double d0,
                         double d1, double d2) {
    HTX_NO_CONTRACT
    double tau = fabs(brk - 0.25 * (s0 + s1 + 2.0 * s2));
    return teno_combine(teno_keep(tau, s0, s1, s2, 1e-3), p0, p1, p2, d0, d1, d2);
}

// Mixed-precision HTENO side: the indicators brk, s_k come in FP64 (formed
// exactly as for hteno_side); tau, a_k = 1 + tau / (s_k + eps) and the cut-off
// (C_T = 1e-3 on a_k^6 / sum_j a_j^6) are evaluated in FP32 on the ratios
// r_k = a_k / max_j a_j as in teno5_side_mixed; the value is the FP64
// combination of the selected linear weights.  The 1e-40 guard becomes 1e-30
// (FP32's normal range); both act only on exactly flat stencils.
HTX_HD double hteno_side_mixed(double brk, double s0, double s1, double s2, double p0, double p1, double p2,
                               double d0, double d1, double d2) {
    const float tau = (float)fabs(brk - 0.25 * (s0 + s1 + 2.0 * s2));
    const float a0 = 1.0f + tau * frcp((float)s0 + 1e-30f);
    const float a1 = 1.0f + tau * frcp((float)s1 + 1e-30f);
    const float a2 = 1.0f + tau * frcp((float)s2 + 1e-30f);
    const float am = fmaxf(fmaxf(a0, a1), a2);
    bool k0, k1, k2;
    if (am > 3.40282347e38f) {  // am is infinite (am >= 1; NaN fails the test)
        k0 = a0 == am;
        k1 = a1 == am;
        k2 = a2 == am;
    } else {
        const float im = frcp(am);
        float r0 = a0 * im, r1 = a1 * im, r2 = a2 * im;
        r0 = r0 * r0 * r0; r0 = r0 * r0;
        r1 = r1 * r1 * r1; r1 = r1 * r1;
        r2 = r2 * r2 * r2; r2 = r2 * r2;
        const float lim = 1e-3f * ((r0 + r1) + r2);
        k0 = !(r0 < lim);
        k1 = !(r1 < lim);
        k2 = !(r2 < lim);
    }
    double w0 = k0 ? d0 : 0.0, w1 = k1 ? d1 : 0.0, w2 = k2 ? d2 : 0.0;
    double ws = (w0 + w1) + w2;
    double out = ((p0 * w0 + p1 * w1) + p2 * w2) * drcp(ws);
    return out != out ? 1e-15 : out;
}

// Hermite::smoothnessMeasure (:515-526) and smoothnessMeasureDerivative (:528-541)
HTX_HD void herm_si(double u0, double u1, double u2, double v0, double v2, double& s0, double& s1, double& s2) {
    HTX_NO_CONTRACT
    const double c1 = 13.0 / 12.0;
    double b0 = -2.0 * (u0 - u1) - v0, c0 = -2.0 * (u0 - u1 + v0);
    double b1 = -2.0 * (u1 - u2) - v2, cc1 = 2.0 * (u1 - u2) + 2.0 * v2;
    double b2 = (-u0 + u2) / 2.0, c2 = u0 - 2.0 * u1 + u2;
    s0 = c1 * (c0 * c0) + b0 * b0;
    s1 = c1 * (cc1 * cc1) + b1 * b1;
    s2 = c1 * (c2 * c2) + b2 * b2;
}

HTX_HD void herm_sid(double u0, double u1, double u2, double d0, double d1, double d2, double& s0, double& s1,
                     double& s2) {
    HTX_NO_CONTRACT
    const double c1 = 13.0 / 12.0;
    double b0 = 6.0 * (u0 - u1) + (2.0 * d0 + 4.0 * d1), c0 = 12.0 * (u0 - u1) + 6.0 * (d0 + d1);
    double b1 = 6.0 * (u2 - u1) - (4.0 * d1 + 2.0 * d2), cc1 = 12.0 * (u1 - u2) + 6.0 * (d1 + d2);
    double b2 = u0 - 2.0 * u1 + u2, c2 = 3.0 * (u2 - u0) - 6.0 * d1;
    s0 = c1 * (c0 * c0) + b0 * b0;
    s1 = c1 * (cc1 * cc1) + b1 * b1;
    s2 = c1 * (c2 * c2) + b2 * b2;
}

// Hermite::hteno (Schemes_Hermite.cpp:295-379): cell averages u_{i-1..i+2} and
// derivative moments d -> (u^-, u^+, d^-, d^+) at x_{i+1/2}.
template <bool MIXED = false>
HTX_HD void hteno(const HermConsts& H, const double u[4], const double d[4], double out[4]) {
    HTX_NO_CONTRACT
    auto side = [](double brk, double s0, double s1, double s2, double p0, double p1, double p2, double d0, double d1,
                   double d2) {
        return MIXED ? hteno_side_mixed(brk, s0, s1, s2, p0, p1, p2, d0, d1, d2)
                     : hteno_side(brk, s0, s1, s2, p0, p1, p2, d0, d1, d2);
    };
    const double dif = H.dif;
    double ud0 = d[0] * dif, ud1 = d[1] * dif, ud2 = d[2] * dif, ud3 = d[3] * dif;
    double s0, s1, s2, p0, p1, p2;
    // u^-
    p0 = H.crj[0][0] * u[0] + H.crj[1][0] * u[1] + H.crj[2][0] * d[0];
    p1 = H.crj[0][1] * u[1] + H.crj[1][1] * u[2] + H.crj[2][1] * d[2];
    p2 = H.crj[0][2] * u[0] + H.crj[1][2] * u[1] + H.crj[2][2] * u[2];
    herm_si(u[0], u[1], u[2], ud0, ud2, s0, s1, s2);
    out[0] = side(herm_brk(u[0], u[1], u[2], ud0, ud2), s0, s1, s2, p0, p1, p2, H.dkm[0], H.dkm[1], H.dkm[2]);
    // d^-
    p0 = H.crjd[0][0] * u[0] + H.crjd[1][0] * u[1] + H.crjd[2][0] * d[0] + H.crjd[3][0] * d[1];
    p1 = H.crjd[0][1] * u[1] + H.crjd[1][1] * u[2] + H.crjd[2][1] * d[1] + H.crjd[3][1] * d[2];
    p2 = H.crjd[0][2] * u[0] + H.crjd[1][2] * u[1] + H.crjd[2][2] * u[2] + H.crjd[3][2] * d[1];
    herm_sid(u[0], u[1], u[2], ud0, ud1, ud2, s0, s1, s2);
    out[2] = side(herm_brkd(u[0], u[1], u[2], ud0, ud1, ud2), s0, s1, s2, p0, p1, p2, H.dkmd[0], H.dkmd[1],
                        H.dkmd[2]);
    // u^+
    p0 = H.crj[1][1] * u[1] + H.crj[0][1] * u[2] - H.crj[2][1] * d[1];
    p1 = H.crj[1][0] * u[2] + H.crj[0][0] * u[3] - H.crj[2][0] * d[3];
    p2 = H.crj[2][2] * u[1] + H.crj[1][2] * u[2] + H.crj[0][2] * u[3];
    herm_si(u[1], u[2], u[3], ud1, ud3, s0, s1, s2);
    out[1] = side(herm_brk(u[1], u[2], u[3], ud1, ud3), s0, s1, s2, p0, p1, p2, H.dkp[0], H.dkp[1], H.dkp[2]);
    // d^+
    p0 = H.crjd[0][1] * u[1] + H.crjd[1][1] * u[2] + H.crjd[3][1] * d[1] + H.crjd[2][1] * d[2];
    p1 = H.crjd[0][0] * u[2] + H.crjd[1][0] * u[3] + H.crjd[3][0] * d[2] + H.crjd[2][0] * d[3];
    p2 = -(H.crjd[2][2] * u[1] + H.crjd[1][2] * u[2] + H.crjd[0][2] * u[3]) + H.crjd[3][2] * d[2];
    herm_sid(u[1], u[2], u[3], ud1, ud2, ud3, s0, s1, s2);
    out[3] = side(herm_brkd(u[1], u[2], u[3], ud1, ud2, ud3), s0, s1, s2, p0, p1, p2, H.dkpd[0], H.dkpd[1],
                        H.dkpd[2]);
}

// The four candidate sets of a Hermite face (the same for hteno, hweno and
// htenoi): u^- (p[0]), u^+ (p[1]), d^- (p[2]), d^+ (p[3]) at x_{i+1/2} from
// u[0..3] = u_{i-1..i+2} and their derivative moments d
// (Schemes_Hermite.cpp:240-290)
HTX_HD void herm_cands(const HermConsts& H, const double u[4], const double d[4], double p[4][3]) {
    HTX_NO_CONTRACT
    p[0][0] = H.crj[0][0] * u[0] + H.crj[1][0] * u[1] + H.crj[2][0] * d[0];
    p[0][1] = H.crj[0][1] * u[1] + H.crj[1][1] * u[2] + H.crj[2][1] * d[2];
    p[0][2] = H.crj[0][2] * u[0] + H.crj[1][2] * u[1] + H.crj[2][2] * u[2];
    p[1][0] = H.crj[1][1] * u[1] + H.crj[0][1] * u[2] - H.crj[2][1] * d[1];
    p[1][1] = H.crj[1][0] * u[2] + H.crj[0][0] * u[3] - H.crj[2][0] * d[3];
    p[1][2] = H.crj[2][2] * u[1] + H.crj[1][2] * u[2] + H.crj[0][2] * u[3];
    p[2][0] = H.crjd[0][0] * u[0] + H.crjd[1][0] * u[1] + H.crjd[2][0] * d[0] + H.crjd[3][0] * d[1];
    p[2][1] = H.crjd[0][1] * u[1] + H.crjd[1][1] * u[2] + H.crjd[2][1] * d[1] + H.crjd[3][1] * d[2];
    p[2][2] = H.crjd[0][2] * u[0] + H.crjd[1][2] * u[1] + H.crjd[2][2] * u[2] + H.crjd[3][2] * d[1];
    p[3][0] = H.crjd[0][1] * u[1] + H.crjd[1][1] * u[2] + H.crjd[3][1] * d[1] + H.crjd[2][1] * d[2];
    p[3][1] = H.crjd[0][0] * u[2] + H.crjd[1][0] * u[3] + H.crjd[3][0] * d[2] + H.crjd[2][0] * d[3];
    p[3][2] = -(H.crjd[2][2] * u[1] + H.crjd[1][2] * u[2] + H.crjd[0][2] * u[3]) + H.crjd[3][2] * d[2];
}

// Hermite::hweno (Schemes_Hermite.cpp:234-293): the candidates of hteno with
// WENO-JS weights d_k / (s_k + 1e-6)^2, normalised, then catchnan.
HTX_HD void hweno(const HermConsts& H, const double u[4], const double d[4], double out[4]) {
    HTX_NO_CONTRACT
    const double dif = H.dif;
    const double ud0 = d[0] * dif, ud1 = d[1] * dif, ud2 = d[2] * dif, ud3 = d[3] * dif;
    double p[4][3], s[3], w[3];
    herm_cands(H, u, d, p);
    for (int o = 0; o < 4; ++o) {
        const int b = o & 1;  // 0: cells i-1..i+1, 1: cells i..i+2
        if (o < 2)
            herm_si(u[b], u[b + 1], u[b + 2], b ? ud1 : ud0, b ? ud3 : ud2, s[0], s[1], s[2]);
        else
            herm_sid(u[b], u[b + 1], u[b + 2], b ? ud1 : ud0, b ? ud2 : ud1, b ? ud3 : ud2, s[0], s[1], s[2]);
        const double* dk = o == 0 ? H.dkm : o == 1 ? H.dkp : o == 2 ? H.dkmd : H.dkpd;
        for (int k = 0; k < 3; ++k) {
            const double e = s[k] + 1e-6;
            w[k] = dk[k] / (e * e);
        }
        const double ws = (w[0] + w[1]) + w[2];
        for (int k = 0; k < 3; ++k) w[k] = w[k] / ws;
        const double r = (p[o][0] * w[0] + p[o][1] * w[1]) + p[o][2] * w[2];
        out[o] = r != r ? 1e-15 : r;
    }
}

// Hermite::htenoi (Schemes_Hermite.cpp:670-755): hteno's candidates with tau
// the largest smoothness measure, weights (1 + tau / (s_k + 1e-40))^6 cut off
// at 1e-3 (cutOff / cutOffd, :614-628), kept linear weights normalised,
// catchnan.  The sixth power is rounded once (pow_n).
HTX_HD void htenoi(const HermConsts& H, const double u[4], const double d[4], double out[4]) {
    HTX_NO_CONTRACT
    const double dif = H.dif;
    const double ud0 = d[0] * dif, ud1 = d[1] * dif, ud2 = d[2] * dif, ud3 = d[3] * dif;
    double p[4][3], s[3], g[3], w[3];
    herm_cands(H, u, d, p);
    for (int o = 0; o < 4; ++o) {
        const int b = o & 1;
        if (o < 2)
            herm_si(u[b], u[b + 1], u[b + 2], b ? ud1 : ud0, b ? ud3 : ud2, s[0], s[1], s[2]);
        else
            herm_sid(u[b], u[b + 1], u[b + 2], b ? ud1 : ud0, b ? ud2 : ud1, b ? ud3 : ud2, s[0], s[1], s[2]);
        double tau = s[0];
        if (s[1] > tau) tau = s[1];
        if (s[2] > tau) tau = s[2];
        for (int k = 0; k < 3; ++k) g[k] = pow_n(1.0 + tau / (s[k] + 1e-40), 6);
        const double gs = (g[0] + g[1]) + g[2];
        const double* dk = o == 0 ? H.dkm : o == 1 ? H.dkp : o == 2 ? H.dkmd : H.dkpd;
        for (int k = 0; k < 3; ++k) w[k] = (g[k] / gs < 1e-3 ? 0.0 : 1.0) * dk[k];
        const double ws = (w[0] + w[1]) + w[2];
        for (int k = 0; k < 3; ++k) w[k] = w[k] / ws;
        const double r = (p[o][0] * w[0] + p[o][1] * w[1]) + p[o][2] * w[2];
        out[o] = r != r ? 1e-15 : r;
    }
}

// Hermite::linear (Schemes_Hermite.cpp:757-767 with crjlin / crjlind of
// setGridSize :67-92): the optimal 5th-order Hermite interpolants
// (u^-, u^+, d^-, d^+).
HTX_HD void herm_linear(const HermConsts& H, const double u[4], const double d[4], double out[4]) {
    HTX_NO_CONTRACT
    const double dif = H.dif;
    out[0] = ((-23.0 / 120.0) * u[0] + (19.0 / 30.0) * u[1]) + (67.0 / 120.0) * u[2] + (-3.0 / 40.0 * dif) * d[0] +
             (-7.0 / 40.0 * dif) * d[2];
    out[1] = ((67.0 / 120.0) * u[1] + (19.0 / 30.0) * u[2]) + (-23.0 / 120.0) * u[3] + (7.0 / 40.0 * dif) * d[1] +
             (3.0 / 40.0 * dif) * d[3];
    out[2] = (((0.25 / dif) * u[0] + (-2.0 / dif) * u[1]) + (1.75 / dif) * u[2]) +
             (((1.0 / 12.0) * d[0] + (-1.0 / 6.0) * d[1]) + (-5.0 / 12.0) * d[2]);
    out[3] = (((-7.0 / 4.0 / dif) * u[1] + (2.0 / dif) * u[2]) + (-0.25 / dif) * u[3]) +
             (((-5.0 / 12.0) * d[1] + (-1.0 / 6.0) * d[2]) + (1.0 / 12.0) * d[3]);
}

// The Hermite face reconstruction of the 2D Hermite solver
// (FVHermite_System2D.cpp:447-470 calls hteno): 0 hteno, 1 hweno, 2 htenoi,
// 3 linear, each on (L u, L d_normal) of a characteristic field.
HTX_HD void herm_recon(int kind, const HermConsts& H, const double u[4], const double d[4], double out[4]) {
    if (kind == 1)
        hweno(H, u, d, out);
    else if (kind == 2)
        htenoi(H, u, d, out);
    else if (kind == 3)
        herm_linear(H, u, d, out);
    else
        hteno<false>(H, u, d, out);
}

// Coefficients of the Hermite 2- and 4-point Gauss-Legendre quadratures
// (setGridSize, Schemes_Hermite.cpp:93-231) for grid size dif, with the
// reference's expressions.
struct HermGLQ {
    double crjgp[3][3], crjgm[3][3], crjdgp[4][3], crjdgm[4][3];
    double dkpg[3], dkmg[3], dkpdg[3], dkmdg[3];
    double crjg4p[3][3], crjg4m[3][3], crjdg4p[4][3], crjdg4m[4][3];
    double dkpg4[3], dkmg4[3], dkpdg4[3], dkmdg4[3];
    double c2, c3, c4, c5, c6, c7;
};

inline void fill_herm_glq(HermGLQ& G, double dif) {
    const double sq3 = sqrt(3.0), sq5 = sqrt(5.0);
    G.c2 = sqrt(3.0) / 3.0;
    G.c3 = 37.0 * sqrt(3.0) / 108.0;
    G.c4 = sqrt(3.0) / 3.0;
    G.c5 = 37.0 * sqrt(3.0) / 108.0;
    G.c6 = sqrt(5.0) / 5.0;
    G.c7 = 61.0 * sqrt(5.0) / 300.0;
    G.crjgp[0][0] = -sq3 / 3.0; G.crjgp[1][0] = 1.0 + sq3 / 3.0; G.crjgp[2][0] = -dif * sq3 / 6.0;
    G.crjgp[0][1] = 1.0 - sq3 / 3.0; G.crjgp[1][1] = sq3 / 3.0; G.crjgp[2][1] = -dif * sq3 / 6.0;
    G.crjgp[0][2] = -sq3 / 12.0; G.crjgp[1][2] = 1.0; G.crjgp[2][2] = sq3 / 12.0;
    G.crjgm[0][0] = sq3 / 3.0; G.crjgm[1][0] = 1.0 - sq3 / 3.0; G.crjgm[2][0] = dif * sq3 / 6.0;
    G.crjgm[0][1] = 1.0 + sq3 / 3.0; G.crjgm[1][1] = -sq3 / 3.0; G.crjgm[2][1] = dif * sq3 / 6.0;
    G.crjgm[0][2] = sq3 / 12.0; G.crjgm[1][2] = 1.0; G.crjgm[2][2] = -sq3 / 12.0;
    G.dkpg[0] = (-sq3 + 105.0) / 360.0; G.dkpg[1] = (sq3 + 105.0) / 360.0; G.dkpg[2] = 5.0 / 12.0;
    G.dkmg[0] = (sq3 + 105.0) / 360.0; G.dkmg[1] = (-sq3 + 105.0) / 360.0; G.dkmg[2] = 5.0 / 12.0;
    G.crjdgp[0][0] = sq3 / dif; G.crjdgp[1][0] = -sq3 / dif; G.crjdgp[2][0] = sq3 / 3.0;
    G.crjdgp[3][0] = (2.0 * sq3 + 3.0) / 3.0;
    G.crjdgp[0][1] = -sq3 / dif; G.crjdgp[1][1] = sq3 / dif; G.crjdgp[2][1] = (-2.0 * sq3 + 3.0) / 3.0;
    G.crjdgp[3][1] = -sq3 / 3.0;
    G.crjdgp[0][2] = sq3 / 6.0 / dif; G.crjdgp[1][2] = -sq3 / 3.0 / dif; G.crjdgp[2][2] = sq3 / 6.0 / dif;
    G.crjdgp[3][2] = 1.0;
    G.crjdgm[0][0] = -sq3 / dif; G.crjdgm[1][0] = sq3 / dif; G.crjdgm[2][0] = -sq3 / 3.0;
    G.crjdgm[3][0] = (-2.0 * sq3 + 3.0) / 3.0;
    G.crjdgm[0][1] = sq3 / dif; G.crjdgm[1][1] = -sq3 / dif; G.crjdgm[2][1] = (2.0 * sq3 + 3.0) / 3.0;
    G.crjdgm[3][1] = sq3 / 3.0;
    G.crjdgm[0][2] = -sq3 / 6.0 / dif; G.crjdgm[1][2] = sq3 / 3.0 / dif; G.crjdgm[2][2] = -sq3 / 6.0 / dif;
    G.crjdgm[3][2] = 1.0;
    G.dkpdg[0] = -sq3 / 144.0 + 1.0 / 3.0; G.dkpdg[1] = sq3 / 144.0 + 1.0 / 3.0; G.dkpdg[2] = 1.0 / 3.0;
    G.dkmdg[0] = sq3 / 144.0 + 1.0 / 3.0; G.dkmdg[1] = -sq3 / 144.0 + 1.0 / 3.0; G.dkmdg[2] = 1.0 / 3.0;
    G.crjg4p[0][0] = -sq5 / 5.0 + 1.0 / 30.0; G.crjg4p[1][0] = sq5 / 5.0 + 29.0 / 30.0;
    G.crjg4p[2][0] = (-sq5 / 10.0 + 1.0 / 30.0) * dif;
    G.crjg4p[0][1] = -sq5 / 5.0 + 29.0 / 30.0; G.crjg4p[1][1] = sq5 / 5.0 + 1.0 / 30.0;
    G.crjg4p[2][1] = -(sq5 / 10.0 + 1.0 / 30.0) * dif;
    G.crjg4p[0][2] = -(sq5 / 20.0 + 1.0 / 60.0); G.crjg4p[1][2] = 31.0 / 30.0; G.crjg4p[2][2] = sq5 / 20.0 - 1.0 / 60.0;
    G.crjg4m[0][0] = sq5 / 5.0 + 1.0 / 30.0; G.crjg4m[1][0] = -sq5 / 5.0 + 29.0 / 30.0;
    G.crjg4m[2][0] = (sq5 / 10.0 + 1.0 / 30.0) * dif;
    G.crjg4m[0][1] = sq5 / 5.0 + 29.0 / 30.0; G.crjg4m[1][1] = -sq5 / 5.0 + 1.0 / 30.0;
    G.crjg4m[2][1] = (sq5 / 10.0 - 1.0 / 30.0) * dif;
    G.crjg4m[0][2] = sq5 / 20.0 - 1.0 / 60.0; G.crjg4m[1][2] = 31.0 / 30.0; G.crjg4m[2][2] = -(sq5 / 20.0 + 1.0 / 60.0);
    double den = 330.0 * sq5 - 110.0;
    G.dkpg4[0] = (99.0 * sq5 + 33.0) / den; G.dkpg4[1] = (108.0 * sq5 - 102.0) / den;
    G.dkpg4[2] = (123.0 * sq5 - 41.0) / den;
    den = 330.0 * sq5 + 110.0;
    G.dkmg4[0] = (99.0 * sq5 - 33.0) / den; G.dkmg4[1] = (108.0 * sq5 + 102.0) / den;
    G.dkmg4[2] = (123.0 * sq5 + 41.0) / den;
    G.crjdg4p[0][0] = (6.0 * sq5 - 2.0) / 10.0 / dif; G.crjdg4p[1][0] = (-6.0 * sq5 + 2.0) / 10.0 / dif;
    G.crjdg4p[2][0] = 0.2 * sq5 - 0.1; G.crjdg4p[3][0] = 0.4 * sq5 + 0.9;
    G.crjdg4p[0][1] = -(6.0 * sq5 + 2.0) / 10.0 / dif; G.crjdg4p[1][1] = (6.0 * sq5 + 2.0) / 10.0 / dif;
    G.crjdg4p[2][1] = -0.4 * sq5 + 0.9; G.crjdg4p[3][1] = -0.2 * sq5 - 0.1;
    G.crjdg4p[0][2] = (2.0 * sq5 + 1.0) / 20.0 / dif; G.crjdg4p[1][2] = -sq5 / 5.0 / dif;
    G.crjdg4p[2][2] = (2.0 * sq5 - 1.0) / 20.0 / dif; G.crjdg4p[3][2] = 1.1;
    G.crjdg4m[0][0] = -(6.0 * sq5 + 2.0) / 10.0 / dif; G.crjdg4m[1][0] = (6.0 * sq5 + 2.0) / 10.0 / dif;
    G.crjdg4m[2][0] = -0.2 * sq5 - 0.1; G.crjdg4m[3][0] = -0.4 * sq5 + 0.9;
    G.crjdg4m[0][1] = (6.0 * sq5 - 2.0) / 10.0 / dif; G.crjdg4m[1][1] = (-6.0 * sq5 + 2.0) / 10.0 / dif;
    G.crjdg4m[2][1] = 0.4 * sq5 + 0.9; G.crjdg4m[3][1] = 0.2 * sq5 - 0.1;
    G.crjdg4m[0][2] = (-2.0 * sq5 + 1.0) / 20.0 / dif; G.crjdg4m[1][2] = sq5 / 5.0 / dif;
    G.crjdg4m[2][2] = (-2.0 * sq5 - 1.0) / 20.0 / dif; G.crjdg4m[3][2] = 1.1;
    den = 1140.0 * sq5 - 570.0;
    G.dkpdg4[0] = (399.0 * sq5 + 190.0) / den; G.dkpdg4[1] = (481.0 * sq5 - 630.0) / den;
    G.dkpdg4[2] = (260.0 * sq5 - 130.0) / den;
    den = 1140.0 * sq5 + 570.0;
    G.dkmdg4[0] = (399.0 * sq5 - 190.0) / den; G.dkmdg4[1] = (481.0 * sq5 + 630.0) / den;
    G.dkmdg4[2] = (260.0 * sq5 + 130.0) / den;
}

// smoothnessMeasure2PointsGLQ / 4PointsGLQ (Schemes_Hermite.cpp:543-581) and
// their derivative forms (:556-595): the squared slope and curvature terms
// b_k^2, c_k^2 of the three Hermite stencils of u0..u2 with moments ud
HTX_HD void herm_bc2(const double u[3], const double ud[3], bool deriv, double b2[3], double c2[3]) {
    HTX_NO_CONTRACT
    double b[3], c[3];
    if (!deriv) {
        b[0] = -2.0 * (u[0] - u[1]) - ud[0];
        c[0] = -2.0 * (u[0] - u[1]) - 2.0 * ud[0];
        b[1] = -2.0 * (u[1] - u[2]) - ud[2];
        c[1] = 2.0 * (u[1] - u[2]) + 2.0 * ud[2];
        b[2] = -(u[0] - u[2]) / 2.0;
        c[2] = u[0] - 2.0 * u[1] + u[2];
    } else {
        b[0] = 6.0 * (u[0] - u[1]) + (2.0 * ud[0] + 4.0 * ud[1]);
        c[0] = 12.0 * (u[0] - u[1]) + 6.0 * (ud[0] + ud[1]);
        b[1] = 6.0 * (u[2] - u[1]) - (4.0 * ud[1] + 2.0 * ud[2]);
        c[1] = 12.0 * (u[1] - u[2]) + 6.0 * (ud[1] + ud[2]);
        b[2] = u[0] - 2.0 * u[1] + u[2];
        c[2] = 3.0 * (u[2] - u[0]) - 6.0 * ud[1];
    }
    for (int k = 0; k < 3; ++k) {
        b2[k] = b[k] * b[k];
        c2[k] = c[k] * c[k];
    }
}

// WENO-JS value of one point: weights d_k / (s_k + 1e-6)^2 normalised, no
// catchnan (the quadrature routines apply none)
HTX_HD double herm_js_point(const double p[3], const double dk[3], const double s[3]) {
    HTX_NO_CONTRACT
    double w[3];
    for (int k = 0; k < 3; ++k) {
        const double e = s[k] + 1e-6;
        w[k] = dk[k] / (e * e);
    }
    const double ws = (w[0] + w[1]) + w[2];
    for (int k = 0; k < 3; ++k) w[k] = w[k] / ws;
    return (p[0] * w[0] + p[1] * w[1]) + p[2] * w[2];
}

// u-type candidates at one point: C(0, k) u_{a} + C(1, k) u_{b} + C(2, k) m
HTX_HD void herm_ucand(const double C[3][3], const double u[3], const double d[3], double p[3]) {
    HTX_NO_CONTRACT
    p[0] = C[0][0] * u[0] + C[1][0] * u[1] + C[2][0] * d[0];
    p[1] = C[0][1] * u[1] + C[1][1] * u[2] + C[2][1] * d[2];
    p[2] = C[0][2] * u[0] + C[1][2] * u[1] + C[2][2] * u[2];
}

// derivative-type candidates at one point (the crjd pattern of :270-276)
HTX_HD void herm_dcand(const double C[4][3], const double u[3], const double d[3], double p[3]) {
    HTX_NO_CONTRACT
    p[0] = C[0][0] * u[0] + C[1][0] * u[1] + C[2][0] * d[0] + C[3][0] * d[1];
    p[1] = C[0][1] * u[1] + C[1][1] * u[2] + C[2][1] * d[1] + C[3][1] * d[2];
    p[2] = C[0][2] * u[0] + C[1][2] * u[1] + C[2][2] * u[2] + C[3][2] * d[1];
}

// Hermite::hweno2PointsGLQ (Schemes_Hermite.cpp:381-425): u at x_i + sqrt(3)/6
// dx and x_i - sqrt(3)/6 dx, then the derivative moment at the same points,
// from cells i-1..i+1
HTX_HD void hweno_2q(const HermGLQ& G, double dif, const double u[3], const double d[3], double out[4]) {
    HTX_NO_CONTRACT
    const double ud[3] = {d[0] * dif, d[1] * dif, d[2] * dif};
    double b2[3], c2[3], s[3], p[3];
    herm_bc2(u, ud, false, b2, c2);
    for (int k = 0; k < 3; ++k) s[k] = G.c2 * b2[k] + G.c3 * c2[k];
    herm_ucand(G.crjgp, u, d, p);
    out[0] = herm_js_point(p, G.dkpg, s);
    herm_ucand(G.crjgm, u, d, p);
    out[1] = herm_js_point(p, G.dkmg, s);
    herm_bc2(u, ud, true, b2, c2);
    for (int k = 0; k < 3; ++k) s[k] = G.c4 * b2[k] + G.c5 * c2[k];
    herm_dcand(G.crjdgp, u, d, p);
    out[2] = herm_js_point(p, G.dkpdg, s);
    herm_dcand(G.crjdgm, u, d, p);
    out[3] = herm_js_point(p, G.dkmdg, s);
}

// Hermite::hweno4PointsGLQ (Schemes_Hermite.cpp:427-513): u at x_{i+1/2},
// x_{i-1/2}, x_i + sqrt(5)/10 dx, x_i - sqrt(5)/10 dx (out[0..3]), then the
// derivative moment at the same four points (out[4..7])
HTX_HD void hweno_4q(const HermConsts& H, const HermGLQ& G, const double u[3], const double d[3], double out[8]) {
    HTX_NO_CONTRACT
    const double dif = H.dif;
    const double c1 = 13.0 / 12.0;
    const double ud[3] = {d[0] * dif, d[1] * dif, d[2] * dif};
    double b2[3], c2[3], s[3], s4[3], p[3];
    herm_bc2(u, ud, false, b2, c2);
    for (int k = 0; k < 3; ++k) {
        s[k] = c1 * c2[k] + b2[k];
        s4[k] = G.c6 * b2[k] + G.c7 * c2[k];
    }
    herm_ucand(H.crj, u, d, p);
    out[0] = herm_js_point(p, H.dkm, s);
    p[0] = H.crj[1][1] * u[0] + H.crj[0][1] * u[1] - H.crj[2][1] * d[0];
    p[1] = H.crj[1][0] * u[1] + H.crj[0][0] * u[2] - H.crj[2][0] * d[2];
    p[2] = H.crj[2][2] * u[0] + H.crj[1][2] * u[1] + H.crj[0][2] * u[2];
    out[1] = herm_js_point(p, H.dkp, s);
    herm_ucand(G.crjg4p, u, d, p);
    out[2] = herm_js_point(p, G.dkpg4, s4);
    herm_ucand(G.crjg4m, u, d, p);
    out[3] = herm_js_point(p, G.dkmg4, s4);
    herm_bc2(u, ud, true, b2, c2);
    for (int k = 0; k < 3; ++k) {
        s[k] = c1 * c2[k] + b2[k];
        s4[k] = G.c6 * b2[k] + G.c7 * c2[k];
    }
    herm_dcand(H.crjd, u, d, p);
    out[4] = herm_js_point(p, H.dkmd, s);
    p[0] = H.crjd[0][1] * u[0] + H.crjd[1][1] * u[1] + H.crjd[3][1] * d[0] + H.crjd[2][1] * d[1];
    p[1] = H.crjd[0][0] * u[1] + H.crjd[1][0] * u[2] + H.crjd[3][0] * d[1] + H.crjd[2][0] * d[2];
    p[2] = -(H.crjd[2][2] * u[0] + H.crjd[1][2] * u[1] + H.crjd[0][2] * u[2]) + H.crjd[3][2] * d[1];
    out[5] = herm_js_point(p, H.dkpd, s);
    herm_dcand(G.crjdg4p, u, d, p);
    out[6] = herm_js_point(p, G.dkpdg4, s4);
    herm_dcand(G.crjdg4m, u, d, p);
    out[7] = herm_js_point(p, G.dkmdg4, s4);
}

// Hermite::hermiteLimiter (Schemes_Hermite.cpp:634-662): u and derivative
// moments of cells i-1, i, i+1 -> the limited moment of cell i.
HTX_HD double herm_limiter(double dif, double u0, double u1, double u2, double d0, double d2) {
    HTX_NO_CONTRACT
    double p0 = 0.75 * (u2 - u0) / dif - 0.25 * (d0 + d2);
    double p1 = (u1 - u0) / dif;
    double p2 = (u2 - u1) / dif;
    const double l0 = 1.0 - 2E-5, l1 = 1E-5, l2 = 1E-5;
    double sm0 = herm_brk(u0, u1, u2, d0 * dif, d2 * dif);
    double sm1 = (u0 - u1) * (u0 - u1);
    double sm2 = (u1 - u2) * (u1 - u2);
    double a = fabs(sm0 - sm1) + fabs(sm0 - sm2);
    double tau = a * a / 4.0;
    double w0 = l0 * (1.0 + tau / (sm0 + 1E-6));
    double w1 = l1 * (1.0 + tau / (sm1 + 1E-6));
    double w2 = l2 * (1.0 + tau / (sm2 + 1E-6));
    double ws = (w0 + w1) + w2;
    w0 = w0 / ws; w1 = w1 / ws; w2 = w2 / ws;
    return w0 * (p0 / l0 - ((l1 / l0) * p1 + (l2 / l0) * p2)) + (w1 * p1 + w2 * p2);
}

// Flux Jacobian times a moment vector, Euler2D::dfluxx/dfluxy
// (Equations_Euler2D.cpp:52-98), in the natural (rho, x, y, E) order: each
// row is summed left to right over the Jacobian's columns, as the 4x4
// product of the reference does (its zero entries drop out exactly).
template <bool YDIR>
HTX_HD void dflux_mul(double gam, double g1, double q3, double rho, double u, double v, double p,
                      const double d[4], double o[4]) {
    double g3 = gam - 3.0;
    double h = (q3 + p) / rho;
    if (!YDIR) {
        double sq = u * u;
        double dyn = 0.5 * g1 * (sq + v * v);
        o[0] = d[1];
        o[1] = (((dyn - sq) * d[0] + (-g3 * u) * d[1]) + (-g1 * v) * d[2]) + g1 * d[3];
        o[2] = ((-u * v) * d[0] + v * d[1]) + u * d[2];
        o[3] = ((((dyn - h) * u) * d[0] + (h - g1 * sq) * d[1]) + (-g1 * u * v) * d[2]) + (gam * u) * d[3];
    } else {
        double sq = v * v;
        double dyn = 0.5 * g1 * (sq + u * u);
        o[0] = d[2];
        o[1] = ((-u * v) * d[0] + v * d[1]) + u * d[2];
        o[2] = (((dyn - sq) * d[0] + (-g1 * u) * d[1]) + (-g3 * v) * d[2]) + g1 * d[3];
        o[3] = ((((dyn - h) * v) * d[0] + (-g1 * u * v) * d[1]) + (h - g1 * sq) * d[2]) + (gam * v) * d[3];
    }
}

// Hermite state layout on the device: 12 fields, u (4), v = <dq/dx> (4),
// w = <dq/dy> (4), each nx*ny with pitch; face buffers hold 24 values per face
// (series-major: u^-, u^+, v^-, v^+, w^-, w^+ for each of 4 characteristic
// fields); flux buffers 12 per face (u, v, w fluxes).
struct FvhArgs {
    const double* q;     // current 12-field state (u, v, w)
    const double* lv;    // limited moments: 8 fields (v', w')
    double* lv_out;      // limiter output
    double* fx;          // x-face states, 24 fields of ny x (nx+1)
    double* fy;          // y-face states, 24 fields of (ny+1) x nx
    double* gx;          // x-face fluxes, 12 fields of ny x (nx+1)
    double* gy;          // y-face fluxes, 12 fields of (ny+1) x nx
    const double* q0;    // u^n (stages 2, 3)
    double* qout;        // residual (mode 0) or stage result
    Ctl* ctl;
    int nx, ny, pitch;
    long long fstride;   // per field of the state
    int bc, mode, quirks, flux;
    int mixed;           // 1: FP32 cut-off rules and transverse weights (fast build only)
    int bulk32;          // 1: FP32 face values and fluxes (with mixed; fast build only)
    int hrecon;          // face reconstruction (herm_recon): 0 hteno, 1 hweno, 2 htenoi, 3 linear
    double gam, g1, dx, dy, area;
    QuadConsts qc;
    HermConsts hc;
};

HTX_HD long long fvh_off(const FvhArgs& a, int i, int j) {
    return (long long)map_index(j, a.ny, a.bc) * a.pitch + map_index(i, a.nx, a.bc);
}

// herm_limiter split at its Gauss-Seidel dependence.  Along a limiter chain
// only d0, the limited left neighbour, comes from the previous step; every
// subexpression of u0, u1, u2 and d2 alone is computed ahead of the chain
// (herm_limiter_pre, 15 values) and the chain step (herm_limiter_step) does
// the rest.  In the strict build the two halves perform exactly the
// operations of herm_limiter in its order, so the result is bit-identical to
// it; the fast build reassociates (p0, the smoothness sum, the weighted
// sum over the candidates, normalised once at the end), turns divisions by
// chain-independent or constant denominators into products and the two
// left on the chain into refined hardware reciprocals.
constexpr int LIM_NV = 15;
enum LimPreSlot {
    LP_X = 0,   // 0.75 (u2 - u0) / dif          (fast: minus 0.25 d2)
    LP_D2,      // d2
    LP_P,       // u part of the first smoothness factor  (fast: minus LP_Q)
    LP_Q,       // 494958 d2 dif
    LP_T2,      // remaining smoothness terms            (fast: their sum)
    LP_T3,
    LP_T4,
    LP_T5,
    LP_SM1,     // (u0 - u1)^2
    LP_SM2,     // (u1 - u2)^2
    LP_S1,      // sm1 + 1e-6                 (fast: K = l1 / (sm1 + 1e-6) + l2 / (sm2 + 1e-6))
    LP_S2,      // sm2 + 1e-6                 (fast: M = l1 p1 / (sm1 + 1e-6) + l2 p2 / (sm2 + 1e-6))
    LP_P1,      // (u1 - u0) / dif
    LP_P2,      // (u2 - u1) / dif
    LP_C        // (l1 / l0) p1 + (l2 / l0) p2  (fast: l1 p1 + l2 p2)
};

// Slots the chain step reads (the fast build folds the others away): their
// number LIM_NS and each slot's position among them (-1: not read), the
// layout of the values handed from the precompute to the chain.
#if HTX_EXACT_DIVISION
constexpr int LIM_NS = LIM_NV;
HTX_HD constexpr int lim_slot_pos(int v) { return v; }
#else
constexpr int LIM_NS = 8;
HTX_HD constexpr int lim_slot_pos(int v) {
    return v == LP_X ? 0 : v == LP_P ? 1 : v == LP_T2 ? 2 : v == LP_SM1 ? 3 : v == LP_SM2 ? 4 : v == LP_S1 ? 5
         : v == LP_S2 ? 6 : v == LP_C ? 7 : -1;
}
#endif

HTX_HD void herm_limiter_pre(double dif, double u0, double u1, double u2, double d2, double o[LIM_NV]) {
#if HTX_EXACT_DIVISION
    HTX_NO_CONTRACT
    const double l0 = 1.0 - 2E-5, l1 = 1E-5, l2 = 1E-5;
    const double D2 = d2 * dif;
    const double p1 = (u1 - u0) / dif, p2 = (u2 - u1) / dif;
    const double sm1 = (u0 - u1) * (u0 - u1), sm2 = (u1 - u2) * (u1 - u2);
    o[LP_X] = 0.75 * (u2 - u0) / dif;
    o[LP_D2] = d2;
    o[LP_P] = 1029270.0 * u2 - 2129184.0 * u1 + 1099914.0 * u0;
    o[LP_Q] = 494958.0 * D2;
    o[LP_T2] = (-1029270.0 * u0 + 2129184.0 * u1 - 1099914.0 * u2 + 281121.0 * D2) * D2;
    o[LP_T3] = (1099445.0 * u2 - 4317056.0 * u1 + 2118166.0 * u0) * u2;
    o[LP_T4] = 4317056.0 * (u1 - u0) * u1;
    o[LP_T5] = 1099445.0 * u0 * u0;
    o[LP_SM1] = sm1;
    o[LP_SM2] = sm2;
    o[LP_S1] = sm1 + 1E-6;
    o[LP_S2] = sm2 + 1E-6;
    o[LP_P1] = p1;
    o[LP_P2] = p2;
    o[LP_C] = (l1 / l0) * p1 + (l2 / l0) * p2;
#else
    // w_k = l_k (1 + tau R_k) with R_k = 1 / (sm_k + 1e-6) and l0 + l1 + l2 = 1:
    //   sum_k w_k = 1 + tau (l0 R0 + K),  K = l1 R1 + l2 R2
    //   w0 (p0 / l0 - C) + w1 p1 + w2 p2 = (1 + tau R0)(p0 - C') + C' + tau M,
    //   C' = l1 p1 + l2 p2,  M = l1 R1 p1 + l2 R2 p2
    // (only the slots lim_slot_pos places are set).  Contraction is off and
    // the fused operations are written out, so that every kernel built from
    // these bodies (single problem, batch) rounds the same way.
    HTX_NO_CONTRACT
    const double l1 = 1E-5, l2 = 1E-5;
    const double idif = 1.0 / dif;  // uniform: hoisted out of the callers' loops
    const double D2 = d2 * dif;
    const double p1 = (u1 - u0) * idif, p2 = (u2 - u1) * idif;
    const double sm1 = (u0 - u1) * (u0 - u1), sm2 = (u1 - u2) * (u1 - u2);
    const double r1 = drcp(sm1 + 1E-6), r2 = drcp(sm2 + 1E-6);
    o[LP_X] = 0.75 * (u2 - u0) * idif - 0.25 * d2;
    o[LP_P] = ((1029270.0 * u2 - 2129184.0 * u1) + 1099914.0 * u0) - 494958.0 * D2;
    o[LP_T2] = (((-1029270.0 * u0 + 2129184.0 * u1) - 1099914.0 * u2) + 281121.0 * D2) * D2 +
               ((1099445.0 * u2 - 4317056.0 * u1) + 2118166.0 * u0) * u2 + 4317056.0 * (u1 - u0) * u1 +
               1099445.0 * u0 * u0;
    o[LP_SM1] = sm1;
    o[LP_SM2] = sm2;
    o[LP_S1] = l1 * r1 + l2 * r2;
    o[LP_S2] = l1 * r1 * p1 + l2 * r2 * p2;
    o[LP_C] = l1 * p1 + l2 * p2;
#endif
}

// one chain step: the limited moment from the precomputed values and d0
HTX_HD double herm_limiter_step(double dif, const double o[LIM_NV], double d0) {
#if HTX_EXACT_DIVISION
    HTX_NO_CONTRACT
    const double l0 = 1.0 - 2E-5, l1 = 1E-5, l2 = 1E-5;
    const double D0 = d0 * dif;
    double p0 = o[LP_X] - 0.25 * (d0 + o[LP_D2]);
    const double e1 = (o[LP_P] + 281121.0 * D0) - o[LP_Q];
    double sm0 = fabs((((e1 * D0 + o[LP_T2]) + o[LP_T3]) + o[LP_T4]) + o[LP_T5]) / 6720.0;
    double a = fabs(sm0 - o[LP_SM1]) + fabs(sm0 - o[LP_SM2]);
    double tau = a * a / 4.0;
    double w0 = l0 * (1.0 + tau / (sm0 + 1E-6));
    double w1 = l1 * (1.0 + tau / o[LP_S1]);
    double w2 = l2 * (1.0 + tau / o[LP_S2]);
    double ws = (w0 + w1) + w2;
    w0 = w0 / ws; w1 = w1 / ws; w2 = w2 / ws;
    return w0 * (p0 / l0 - o[LP_C]) + (w1 * o[LP_P1] + w2 * o[LP_P2]);
#else
    HTX_NO_CONTRACT
    const double l0 = 1.0 - 2E-5;
    const double D0 = d0 * dif;
    const double p0 = fma(-0.25, d0, o[LP_X]);
    const double sm0 = fabs(fma(fma(281121.0, D0, o[LP_P]), D0, o[LP_T2])) * (1.0 / 6720.0);
    const double a = fabs(sm0 - o[LP_SM1]) + fabs(sm0 - o[LP_SM2]);
    const double tau = a * a * 0.25;
    const double r0 = drcp(sm0 + 1E-6);
    const double ws = fma(tau, fma(l0, r0, o[LP_S1]), 1.0);
    const double num = fma(fma(tau, r0, 1.0), p0 - o[LP_C], fma(tau, o[LP_S2], o[LP_C]));
    return num * drcp(ws);
#endif
}

// limitDerivatives, FVHermite_System2D.cpp:393-427: one sequential chain.
// dir 0: row j, field e, v replaced with i ascending (Gauss-Seidel: the new
// v_{i-1} is used, ghost neighbours keep their pre-limiting values); dir 1:
// column i, w replaced with j ascending.  This is the CPU form (one chain at
// a time, the test harnesses and the batched launch); the single-problem GPU
// kernel computes the same pre/step halves on two wavefronts.
HTX_HD void fvh_limit_line(const FvhArgs& a, int dir, int line, int e) {
    const long long fs = a.fstride;
    const double* U = a.q + e * fs;
    const double* D = a.q + (dir == 0 ? 4 + e : 8 + e) * fs;
    double* O = a.lv_out + (dir == 0 ? e : 4 + e) * fs;
    int n = dir == 0 ? a.nx : a.ny;
    auto at = [&](int k) { return dir == 0 ? fvh_off(a, k, line) : fvh_off(a, line, k); };
    double prev = D[at(-1)];  // ghost: pre-limiting value
    double um = U[at(-1)], uc = U[at(0)];
    for (int k = 0; k < n; ++k) {
        // indices past the line map through the boundary condition
        const double up = U[at(k + 1)];
        double o[LIM_NV] = {};
        herm_limiter_pre(a.hc.dif, um, uc, up, D[at(k + 1)], o);
        double v = herm_limiter_step(a.hc.dif, o, prev);
        O[at(k)] = v;
        prev = v;
        um = uc;
        uc = up;
    }
}

// The limiter: 64 chains per workgroup, one consumer and LIM_NP producer
// wavefronts.  A chain step's cost is the latency of its dependent
// arithmetic, and one wavefront per SIMD runs it, so everything that does not
// depend on the chain moves off the consumer: the producers load u and the
// old moments, compute herm_limiter_pre for the next chunk of LIM_B steps of
// the 64 chains into LDS (LIM_B x 64 items shared among the producer lanes
// in the order that keeps their loads and stores coalesced) and store the
// chunk the consumer finished before, while the consumer runs
// herm_limiter_step over the current chunk.  The consumer touches LDS only (a
// global store would make a later step wait for the store's completion
// before reusing its registers).  Two LDS buffers alternate, one barrier per
// chunk, and the producers' loads run one chunk ahead of their arithmetic.
//   LDS: pre[buffer][slot][step][chain], 2 x LIM_NS x LIM_B x 65 doubles
//        (15 slots strict, 8 fast), res[buffer][step][chain], 2 x LIM_B x 65
//   blocks: blk < g0 = ceil(4 ny / 64) take the row chains (v), the next
//           ceil(4 nx / 64) the column chains (w); LIM_THREADS threads each
constexpr int LIM_B = HTX_EXACT_DIVISION ? 8 : 16, LIM_NP = 8, LIM_SP = LIM_B / LIM_NP;
constexpr int LIM_THREADS = 64 * (1 + LIM_NP);
static_assert(LIM_B % LIM_NP == 0, "every producer takes the same number of steps");

// workgroups of one problem's limiter launch
HTX_HD int lim_blocks(int nx, int ny) { return (4 * ny + 63) / 64 + (4 * nx + 63) / 64; }

__device__ inline void fvh_limit_block(const FvhArgs& a, int blk, int g0) {
    // [.][65]: the padding spreads a producer's writes of consecutive steps
    // of one chain over the LDS banks
    __shared__ double pre[2][LIM_NS][LIM_B][65];
    __shared__ double res[2][LIM_B][65];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int dir = blk < g0 ? 0 : 1;
    const int grp = dir == 0 ? blk : blk - g0;
    const int nl = dir == 0 ? a.ny : a.nx;  // lines per field
    const int n = dir == 0 ? a.nx : a.ny;   // chain length
    const long long fs = a.fstride;
    const long long sk = dir == 0 ? 1 : a.pitch;
    // chain c of the workgroup: its U, old-moment and output fields and the
    // offset of its first cell; idle chains (past 4 nl) shadow the last one
    // and store nothing
    struct Chain {
        const double *U, *D;
        double* O;
        long long base;
        bool live;
    };
    auto chain = [&](int c) {
        const int t = grp * 64 + c;
        const bool live = t < 4 * nl;
        const int tt = live ? t : 4 * nl - 1;
        const int line = tt % nl, e = tt / nl;
        Chain ch;
        ch.U = a.q + e * fs;
        ch.D = a.q + (dir == 0 ? 4 + e : 8 + e) * fs;
        ch.O = a.lv_out + (dir == 0 ? e : 4 + e) * fs;
        ch.base = dir == 0 ? (long long)line * a.pitch : line;
        ch.live = live;
        return ch;
    };
    // step index -> cell along the chain; off the line through the boundary
    // condition (one wrap or a clamp covers k in [-1, n + LIM_B] when
    // n >= LIM_B + 2)
    auto kmap = [&](int k) {
        if (n >= LIM_B + 2) return a.bc == 1 ? (k < 0 ? k + n : (k >= n ? k - n : k)) : min(max(k, 0), n - 1);
        return map_index(k, n, a.bc);
    };
    const double dif = a.hc.dif;
    const int nch = (n + LIM_B - 1) / LIM_B;

    // Producer items: LIM_SP (chain, step) pairs per lane and chunk.  Row
    // chains (dir 0) are contiguous along the chain, so consecutive lanes
    // take consecutive steps of one chain (coalesced loads and stores);
    // column chains are contiguous across chains, so consecutive lanes take
    // consecutive chains at one step.
    const int q = (wave > 0 ? wave - 1 : 0) * 64 + lane;
    int ic[LIM_SP], is[LIM_SP];
    Chain pc[LIM_SP];
#pragma unroll
    for (int j = 0; j < LIM_SP; ++j) {
        const int idx = q + 64 * LIM_NP * j;
        ic[j] = dir == 0 ? idx / LIM_B : idx % 64;
        is[j] = dir == 0 ? idx % LIM_B : idx / 64;
        pc[j] = chain(ic[j]);
    }
    // u_{k-1}, u_k, u_{k+1} and the old moment_{k+1} of the items of the
    // chunk the producer computes next
    double r0[LIM_SP], r1[LIM_SP], r2[LIM_SP], rd[LIM_SP];
    auto fetch = [&](int chunk) {
#pragma unroll
        for (int j = 0; j < LIM_SP; ++j) {
            const int k = chunk * LIM_B + is[j];
            r0[j] = pc[j].U[pc[j].base + kmap(k - 1) * sk];
            r1[j] = pc[j].U[pc[j].base + kmap(k) * sk];
            r2[j] = pc[j].U[pc[j].base + kmap(k + 1) * sk];
            rd[j] = pc[j].D[pc[j].base + kmap(k + 1) * sk];
        }
    };
    auto produce = [&](int chunk) {
        const int buf = chunk & 1;
#pragma unroll
        for (int j = 0; j < LIM_SP; ++j) {
            double o[LIM_NV] = {};
            herm_limiter_pre(dif, r0[j], r1[j], r2[j], rd[j], o);
#pragma unroll
            for (int v = 0; v < LIM_NV; ++v)
                if (lim_slot_pos(v) >= 0) pre[buf][lim_slot_pos(v)][is[j]][ic[j]] = o[v];
        }
    };
    auto store = [&](int chunk) {  // the consumer's results of a finished chunk
        const int buf = chunk & 1;
#pragma unroll
        for (int j = 0; j < LIM_SP; ++j) {
            const int k = chunk * LIM_B + is[j];
            if (k < n && pc[j].live) pc[j].O[pc[j].base + (long long)k * sk] = res[buf][is[j]][ic[j]];
        }
    };
    double prev = 0.0;
    if (wave > 0) {
        fetch(0);
        produce(0);
        if (nch > 1) fetch(1);
    } else {
        const Chain me = chain(lane);
        prev = me.D[me.base + kmap(-1) * sk];  // ghost: pre-limiting value
    }
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
        if (wave == 0) {
            const int buf = ch & 1;
            // the values of step s + 1 are read while step s computes
            double o[LIM_NV], on[LIM_NV];
#pragma unroll
            for (int v = 0; v < LIM_NV; ++v) o[v] = lim_slot_pos(v) >= 0 ? pre[buf][lim_slot_pos(v)][0][lane] : 0.0;
#pragma unroll
            for (int s = 0; s < LIM_B; ++s) {
                if (s + 1 < LIM_B) {
#pragma unroll
                    for (int v = 0; v < LIM_NV; ++v)
                        on[v] = lim_slot_pos(v) >= 0 ? pre[buf][lim_slot_pos(v)][s + 1][lane] : 0.0;
                }
                if (ch * LIM_B + s < n) {
                    prev = herm_limiter_step(dif, o, prev);
                    res[buf][s][lane] = prev;
                }
#pragma unroll
                for (int v = 0; v < LIM_NV; ++v) o[v] = on[v];
            }
        } else {
            if (ch + 1 < nch) {
                produce(ch + 1);
                if (ch + 2 < nch) fetch(ch + 2);
            }
            if (ch > 0) store(ch - 1);
        }
        __syncthreads();
    }
    if (wave > 0) store(nch - 1);
}

// limited moment of the state at (i, j): the limited interior value at the
// boundary-mapped cell (the reference re-applies the boundary condition
// after limiting, FVHermite_System2D.cpp:437)
HTX_HD double fvh_lim(const FvhArgs& a, int comp, int i, int j) { return a.lv[comp * a.fstride + fvh_off(a, i, j)]; }

// x or y face reconstruction (FVHermite_System2D.cpp:443-495) for the face
// between cell (i, j) and its +1 neighbour in direction YDIR; writes the 24
// characteristic face values at face index fid of the face buffer.  MIXED:
// the HTENO and TENO5 cut-off rules in FP32 (hteno_side_mixed,
// teno5_side_mixed), every value and projection in FP64.
template <bool YDIR, bool MIXED = false, bool VARIANT = false, typename FT = double>
HTX_HD void fvh_recon_face(const FvhArgs& a, int i, int j, FT* F, long long fid, long long fstr) {
    const long long fs = a.fstride;
    auto cellq = [&](int k, double q[4]) {  // k: offset along the normal
        long long o = YDIR ? fvh_off(a, i, j + k) : fvh_off(a, i + k, j);
        q[0] = a.q[o]; q[1] = a.q[o + fs]; q[2] = a.q[o + 2 * fs]; q[3] = a.q[o + 3 * fs];
    };
    double qa[4], qb[4];
    cellq(0, qa);
    cellq(1, qb);
    Eig E = roe_eig(a.g1, YDIR, cell_prim(a.g1, qa[0], qa[1], qa[2], qa[3]), cell_prim(a.g1, qb[0], qb[1], qb[2], qb[3]));
    // hteno on (L u, L d_normal) over cells -1..2, teno5 on L d_tangential over -2..3
    double A[4][4], B[4][4], C[4][6];
    const int nc = YDIR ? 4 : 0;  // normal-derivative block: v (x) or w (y) -> limited comp offset
    // FP32-mixed (and FP32 bulk): the loops are unrolled so that A, B, C stay
    // in registers (236 VGPRs, two waves per SIMD; rolled, 129 VGPRs and
    // three waves, the Hermite stage ran 6 % slower, and with a four-wave
    // launch bound 13 % slower); FP64: they stay loops over A, B, C in
    // scratch, 164 VGPRs and three waves per SIMD instead of 262 and one
    // (+13 % for fp64_fast; the strict build's compiler makes that choice by
    // itself)
#pragma unroll(MIXED ? 6 : 1)
    for (int m = 0; m < 6; ++m) {
        int k = m - 2;
        int ci = YDIR ? i : i + k, cj = YDIR ? j + k : j;
        double q[4];
        cellq(k, q);
        double dn[4], dt[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            dn[e] = fvh_lim(a, nc + e, ci, cj);        // normal derivative moment
            dt[e] = fvh_lim(a, (4 - nc) + e, ci, cj);  // tangential derivative moment
        }
        double c[4];
        if (m >= 1 && m <= 4) {
            eig_left<YDIR>(E, q[0], q[1], q[2], q[3], c[0], c[1], c[2], c[3]);
#pragma unroll
            for (int e = 0; e < 4; ++e) A[e][m - 1] = c[e];
            eig_left<YDIR>(E, dn[0], dn[1], dn[2], dn[3], c[0], c[1], c[2], c[3]);
#pragma unroll
            for (int e = 0; e < 4; ++e) B[e][m - 1] = c[e];
        }
        eig_left<YDIR>(E, dt[0], dt[1], dt[2], dt[3], c[0], c[1], c[2], c[3]);
#pragma unroll
        for (int e = 0; e < 4; ++e) C[e][m] = c[e];
    }
#pragma unroll(MIXED ? 4 : 1)
    for (int e = 0; e < 4; ++e) {
        double o[4], um, up;
        if (VARIANT)  // hweno, htenoi or linear in hteno's place (FVHermite_System2D.cpp:459, :486)
            herm_recon(a.hrecon, a.hc, A[e], B[e], o);
        else
            hteno<MIXED>(a.hc, A[e], B[e], o);
        if (MIXED)
            teno5_mixed(C[e], um, up);
        else
            teno5(C[e], um, up);
        // series: 0 u^-, 1 u^+, 2 vn^-, 3 vn^+, 4 vt^-, 5 vt^+ (vn: normal moment)
        F[(0 * 4 + e) * fstr + fid] = (FT)o[0];
        F[(1 * 4 + e) * fstr + fid] = (FT)o[1];
        F[(2 * 4 + e) * fstr + fid] = (FT)o[2];
        F[(3 * 4 + e) * fstr + fid] = (FT)o[3];
        F[(4 * 4 + e) * fstr + fid] = (FT)um;
        F[(5 * 4 + e) * fstr + fid] = (FT)up;
    }
}

}  // namespace HTX_NS

namespace HTX_NS {

template <bool YDIR>
HTX_HD void rot_to_nat(const double r[4], double& q0, double& qx, double& qy, double& q3) {
    q0 = r[0];
    qx = YDIR ? r[2] : r[1];
    qy = YDIR ? r[1] : r[2];
    q3 = r[3];
}

// Euler2D::dRusanov (Equations_Euler2D.cpp:352-360):
// 0.5 (Jl dql + Jr dqr - R diag(a) L (dqr - dql)).  States and moments come
// in (rho, n, t, E) order, the result goes out in the same order; the
// products run in the natural order.
template <bool YDIR>
HTX_HD void drusanov(double gam, double g1, const Eig& E, const double ql[4], const double qr[4], double pl,
                     double pr, const double a[4], const double dql[4], const double dqr[4], double f[4]) {
    double dl[4], dr[4], jl[4], jr[4], c0, c1, c2, c3, r0, rn, rt, r3;
    rot_to_nat<YDIR>(dql, dl[0], dl[1], dl[2], dl[3]);
    rot_to_nat<YDIR>(dqr, dr[0], dr[1], dr[2], dr[3]);
#if HTX_EXACT_DIVISION
    double unl = ql[1] / ql[0], utl = ql[2] / ql[0], unr = qr[1] / qr[0], utr = qr[2] / qr[0];
#else
    double il = drcp(ql[0]), ir = drcp(qr[0]);
    double unl = ql[1] * il, utl = ql[2] * il, unr = qr[1] * ir, utr = qr[2] * ir;
#endif
    dflux_mul<YDIR>(gam, g1, ql[3], ql[0], YDIR ? utl : unl, YDIR ? unl : utl, pl, dl, jl);
    dflux_mul<YDIR>(gam, g1, qr[3], qr[0], YDIR ? utr : unr, YDIR ? unr : utr, pr, dr, jr);
    eig_left<YDIR>(E, dr[0] - dl[0], dr[1] - dl[1], dr[2] - dl[2], dr[3] - dl[3], c0, c1, c2, c3);
    eig_right(E, a[0] * c0, a[1] * c1, a[1] * c2, a[3] * c3, r0, rn, rt, r3);
    const int n = YDIR ? 2 : 1, t = YDIR ? 1 : 2;
    f[0] = 0.5 * ((jl[0] + jr[0]) - r0);
    f[1] = 0.5 * ((jl[n] + jr[n]) - rn);
    f[2] = 0.5 * ((jl[t] + jr[t]) - rt);
    f[3] = 0.5 * ((jl[3] + jr[3]) - r3);
}

// Transverse quadrature, back-projection and the u, v, w fluxes of a face
// (FVHermite_System2D.cpp:497-640).  The face lies between cell (i, j) and its
// +1 neighbour in direction YDIR.  In two parts:
//   fvh_series_gp3  the three Gauss-point values of one of the 24 face-state
//                   series (6 series x 4 characteristic fields), from its 5
//                   transverse rows read at boundary-mapped positions fid
//   fvh_flux_at     from the 24 series values at Gauss point m: back-projection
//                   and the 12 fluxes (u, v, w; each in (rho, x, y, E) order),
//                   weighted; the caller sums the three Gauss points in order.
template <bool YDIR>
HTX_HD void fvh_flux_fids(const FvhArgs& a, int i, int j, long long fid[5]) {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (YDIR)
            fid[k] = (long long)(j + 1) * a.nx + map_index(i - 2 + k, a.nx, a.bc);
        else
            fid[k] = (long long)map_index(j - 2 + k, a.ny, a.bc) * (a.nx + 1) + (i + 1);
    }
}

template <bool MIXED = false>
HTX_HD void fvh_series_gp3(const QuadConsts& qc, const double* F, long long fstr, int series, const long long fid[5],
                           double g[3]) {
    const double* S = F + (long long)series * fstr;
    if (MIXED)  // nonlinear weights in FP32, candidate values in FP64
        weno3q_mixed(qc, qc.lw32, S[fid[0]], S[fid[1]], S[fid[2]], S[fid[3]], S[fid[4]], g[0], g[1], g[2]);
    else
        weno3q_gp3(qc, S[fid[0]], S[fid[1]], S[fid[2]], S[fid[3]], S[fid[4]], g);
}

// the face's Roe eigensystem from its two cells
template <bool YDIR>
HTX_HD Eig fvh_face_eig(const FvhArgs& a, int i, int j) {
    double qa[4], qb[4];
    const long long fs = a.fstride;
    long long oa = fvh_off(a, i, j), ob = YDIR ? fvh_off(a, i, j + 1) : fvh_off(a, i + 1, j);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        qa[e] = a.q[oa + e * fs];
        qb[e] = a.q[ob + e * fs];
    }
    return roe_eig(a.g1, YDIR, cell_prim(a.g1, qa[0], qa[1], qa[2], qa[3]), cell_prim(a.g1, qb[0], qb[1], qb[2], qb[3]));
}

template <bool YDIR, int FLUX>
HTX_HD void fvh_flux_at_e(const FvhArgs& a, const Eig& E, int m, const double c[6][4], double out[12]) {
    double Q[6][4];  // unrolled loops keep Q in registers
#pragma unroll
    for (int sr = 0; sr < 6; ++sr) eig_right(E, c[sr][0], c[sr][1], c[sr][2], c[sr][3], Q[sr][0], Q[sr][1], Q[sr][2], Q[sr][3]);
    // primitive states at the Gauss point and maxLocalWaveSpeed (Equations_Euler2D.cpp:221-246)
    double fl0, fln, flt, fl3, wnl, pl, rl, fr0, frn, frt, fr3, wnr, pr, rr;
    phys_flux(a.g1, Q[0][0], Q[0][1], Q[0][2], Q[0][3], fl0, fln, flt, fl3, wnl, pl, rl);
    phys_flux(a.g1, Q[1][0], Q[1][1], Q[1][2], Q[1][3], fr0, frn, frt, fr3, wnr, pr, rr);
#if HTX_EXACT_DIVISION
    double cl = sqrt(fabs(a.gam * pl / rl)), cr = sqrt(fabs(a.gam * pr / rr));
#else
    double cl = sqrt(fabs(a.gam * pl * drcp(rl))), cr = sqrt(fabs(a.gam * pr * drcp(rr)));
#endif
    double am[4];
    am[0] = fmax(fabs(wnl - cl), fabs(wnr - cr));
    am[1] = fmax(fabs(wnl), fabs(wnr));
    am[2] = am[1];
    am[3] = fmax(fabs(wnl + cl), fabs(wnr + cr));
    double fu[4], fv[4], fw[4];
    if (FLUX == FLUX_RUSANOV_CHAR)
        rusanov_char<YDIR>(a.gam, a.g1, E, Q[0], Q[1], fu);
    else if (FLUX == FLUX_HLLC)
        hllc<true>(a.gam, a.g1, Q[0], Q[1], fu);
    else
        hllc<false>(a.gam, a.g1, Q[0], Q[1], fu);
    // series 2,3 carry the normal moment (v in x, w in y), 4,5 the tangential one
    const int sv = YDIR ? 4 : 2, sw = YDIR ? 2 : 4;
    drusanov<YDIR>(a.gam, a.g1, E, Q[0], Q[1], pl, pr, am, Q[sv], Q[sv + 1], fv);
    drusanov<YDIR>(a.gam, a.g1, E, Q[0], Q[1], pl, pr, am, Q[sw], Q[sw + 1], fw);
    const double gw = (m == 1) ? 8.0 / 9.0 / 2.0 : 5.0 / 9.0 / 2.0;
    auto put = [&](int g, const double f[4]) {
        out[4 * g + 0] = gw * f[0];
        out[4 * g + 1] = gw * (YDIR ? f[2] : f[1]);
        out[4 * g + 2] = gw * (YDIR ? f[1] : f[2]);
        out[4 * g + 3] = gw * f[3];
    };
    put(0, fu);
    put(1, fv);
    put(2, fw);
}

template <bool YDIR, int FLUX>
HTX_HD void fvh_flux_at(const FvhArgs& a, int i, int j, int m, const double c[6][4], double out[12]) {
    fvh_flux_at_e<YDIR, FLUX>(a, fvh_face_eig<YDIR>(a, i, j), m, c, out);
}

// ---------------------------------------------------------------------------
// FP32 bulk precision of the fluxes: the face values come in FP32, the
// transverse quadrature (weno3q_f32), the back-projection, the u flux and the
// moment fluxes run in FP32 from the face's FP64 Roe eigensystem rounded
// once; the Gauss-weighted sums are FP64.  A Gauss point whose FP32 u states
// fail state_fp32_ok is recomputed in FP64 from the same Gauss-point values.

template <bool YDIR>
HTX_HD void dflux_mul_f32(float gam, float g1, float q3, float rho, float u, float v, float p, const float d[4],
                          float o[4]) {
    const float g3 = gam - 3.0f;
    const float h = (q3 + p) * frcp(rho);
    if (!YDIR) {
        const float sq = u * u, dyn = 0.5f * g1 * (sq + v * v);
        o[0] = d[1];
        o[1] = (((dyn - sq) * d[0] + (-g3 * u) * d[1]) + (-g1 * v) * d[2]) + g1 * d[3];
        o[2] = ((-u * v) * d[0] + v * d[1]) + u * d[2];
        o[3] = ((((dyn - h) * u) * d[0] + (h - g1 * sq) * d[1]) + (-g1 * u * v) * d[2]) + (gam * u) * d[3];
    } else {
        const float sq = v * v, dyn = 0.5f * g1 * (sq + u * u);
        o[0] = d[2];
        o[1] = ((-u * v) * d[0] + v * d[1]) + u * d[2];
        o[2] = (((dyn - sq) * d[0] + (-g1 * u) * d[1]) + (-g3 * v) * d[2]) + g1 * d[3];
        o[3] = ((((dyn - h) * v) * d[0] + (-g1 * u * v) * d[1]) + (h - g1 * sq) * d[2]) + (gam * v) * d[3];
    }
}

// drusanov in FP32 (Equations_Euler2D.cpp:352-360), the same (rho, n, t, E)
// conventions
template <bool YDIR>
HTX_HD void drusanov_f32(float gam, float g1, const Eig32& E, const float ql[4], const float qr[4], float pl,
                         float pr, const float a[4], const float dql[4], const float dqr[4], float f[4]) {
    const int n = YDIR ? 2 : 1, t = YDIR ? 1 : 2;
    float dl[4], dr[4], jl[4], jr[4];
    // natural (rho, x, y, E) order for the Jacobian products
    dl[0] = dql[0], dl[n] = dql[1], dl[t] = dql[2], dl[3] = dql[3];
    dr[0] = dqr[0], dr[n] = dqr[1], dr[t] = dqr[2], dr[3] = dqr[3];
    const float il = frcp(ql[0]), ir = frcp(qr[0]);
    const float unl = ql[1] * il, utl = ql[2] * il, unr = qr[1] * ir, utr = qr[2] * ir;
    dflux_mul_f32<YDIR>(gam, g1, ql[3], ql[0], YDIR ? utl : unl, YDIR ? unl : utl, pl, dl, jl);
    dflux_mul_f32<YDIR>(gam, g1, qr[3], qr[0], YDIR ? utr : unr, YDIR ? unr : utr, pr, dr, jr);
    float c0, c1, c2, c3, r0, rn, rt, r3;
    eig_left_nt_f32(E, dr[0] - dl[0], dr[n] - dl[n], dr[t] - dl[t], dr[3] - dl[3], c0, c1, c2, c3);
    eig_right_f32(E, a[0] * c0, a[1] * c1, a[1] * c2, a[3] * c3, r0, rn, rt, r3);
    f[0] = 0.5f * ((jl[0] + jr[0]) - r0);
    f[1] = 0.5f * ((jl[n] + jr[n]) - rn);
    f[2] = 0.5f * ((jl[t] + jr[t]) - rt);
    f[3] = 0.5f * ((jl[3] + jr[3]) - r3);
}

HTX_HD void fvh_series_gp3_f32(const QuadConsts& qc, const float* F, long long fstr, int series, const long long fid[5],
                               float g[3]) {
    const float* S = F + (long long)series * fstr;
    weno3q_f32(qc, S[fid[0]], S[fid[1]], S[fid[2]], S[fid[3]], S[fid[4]], g[0], g[1], g[2]);
}

template <bool YDIR, int FLUX>
HTX_HD void fvh_flux_at_f32(const FvhArgs& a, int i, int j, int m, const float c[6][4], double out[12]) {
    const Eig E = fvh_face_eig<YDIR>(a, i, j);
    const Eig32 E32 = eig32(E);
    float Q[6][4];
#pragma unroll
    for (int sr = 0; sr < 6; ++sr)
        eig_right_f32(E32, c[sr][0], c[sr][1], c[sr][2], c[sr][3], Q[sr][0], Q[sr][1], Q[sr][2], Q[sr][3]);
    if (!(state_fp32_ok(Q[0], c[0][0], c[0][2], c[0][3]) && state_fp32_ok(Q[1], c[1][0], c[1][2], c[1][3]))) {
        double cd[6][4];
#pragma unroll
        for (int sr = 0; sr < 6; ++sr)
#pragma unroll
            for (int f = 0; f < 4; ++f) cd[sr][f] = (double)c[sr][f];
        fvh_flux_at_e<YDIR, FLUX>(a, E, m, cd, out);
        return;
    }
    const float gam = (float)a.gam, g1 = (float)a.g1;
    const float il = frcp(Q[0][0]), ir = frcp(Q[1][0]);
    const float wnl = Q[0][1] * il, wtl = Q[0][2] * il, wnr = Q[1][1] * ir, wtr = Q[1][2] * ir;
    const float pl = (Q[0][3] - 0.5f * (Q[0][1] * wnl + Q[0][2] * wtl)) * g1;
    const float pr = (Q[1][3] - 0.5f * (Q[1][1] * wnr + Q[1][2] * wtr)) * g1;
    const float cl = fsqrt(fabsf(gam * pl * il)), cr = fsqrt(fabsf(gam * pr * ir));
    float am[4];
    am[0] = fmaxf(fabsf(wnl - cl), fabsf(wnr - cr));
    am[1] = fmaxf(fabsf(wnl), fabsf(wnr));
    am[2] = am[1];
    am[3] = fmaxf(fabsf(wnl + cl), fabsf(wnr + cr));
    float fu[4], fv[4], fw[4];
    if (FLUX == FLUX_RUSANOV_CHAR)
        rusanov_char_f32(gam, g1, E32, Q[0], Q[1], fu);
    else if (FLUX == FLUX_HLLC)
        hllc_f32<true>(gam, g1, Q[0], Q[1], fu);
    else
        hllc_f32<false>(gam, g1, Q[0], Q[1], fu);
    const int sv = YDIR ? 4 : 2, sw = YDIR ? 2 : 4;
    drusanov_f32<YDIR>(gam, g1, E32, Q[0], Q[1], pl, pr, am, Q[sv], Q[sv + 1], fv);
    drusanov_f32<YDIR>(gam, g1, E32, Q[0], Q[1], pl, pr, am, Q[sw], Q[sw + 1], fw);
    const double gw = (m == 1) ? 8.0 / 9.0 / 2.0 : 5.0 / 9.0 / 2.0;
    auto put = [&](int g, const float f[4]) {
        out[4 * g + 0] = gw * (double)f[0];
        out[4 * g + 1] = gw * (double)(YDIR ? f[2] : f[1]);
        out[4 * g + 2] = gw * (double)(YDIR ? f[1] : f[2]);
        out[4 * g + 3] = gw * (double)f[3];
    };
    put(0, fu);
    put(1, fv);
    put(2, fw);
}

// All of one face on one work item (the CPU form): out[m] = weighted fluxes
// at Gauss point m.
template <bool YDIR, int FLUX, bool MIXED = false>
HTX_HD void fvh_flux_face(const FvhArgs& a, int i, int j, const double* F, long long fstr, double out[3][12]) {
    long long fid[5];
    fvh_flux_fids<YDIR>(a, i, j, fid);
    double c[3][6][4];
    for (int s = 0; s < 24; ++s) {
        double g[3];
        fvh_series_gp3<MIXED>(kQuad, F, fstr, s, fid, g);
        for (int m = 0; m < 3; ++m) c[m][s / 4][s % 4] = g[m];
    }
    for (int m = 0; m < 3; ++m) fvh_flux_at<YDIR, FLUX>(a, i, j, m, c[m], out[m]);
}

// One workgroup of 64 faces x 3 Gauss points (threadIdx.x face t0 + x,
// threadIdx.y = m), faces t < nf of the x (YDIR false) or y family.  The
// thread of Gauss point m forms the three Gauss-point values of series
// 8m .. 8m+7 (smoothness shared by the three points) into LDS; after a
// barrier each thread takes its point's 24 values, computes the fluxes, and
// the three points are summed in order through the same LDS.
//   lds: 24 x 3 x 64 doubles (series values), then 3 x 12 x 64 (fluxes)
template <bool YDIR, int FLUX, bool MIXED = false>
__device__ void fvh_flux_block(const FvhArgs& a, long long t0, long long nf, double* lds) {
    const int x = threadIdx.x, m = threadIdx.y;
    const long long t = t0 + x;
    const bool valid = t < nf;
    int i = 0, j = 0;
    if (YDIR) {
        i = (int)(t % a.nx);
        j = (int)(t / a.nx) - 1;
    } else {
        i = (int)(t % (a.nx + 1)) - 1;
        j = (int)(t / (a.nx + 1));
    }
    const double* F = YDIR ? a.fy : a.fx;
    if (valid) {
        long long fid[5];
        fvh_flux_fids<YDIR>(a, i, j, fid);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int s = 8 * m + q;
            double g[3];
            fvh_series_gp3<MIXED>(kQuad, F, nf, s, fid, g);
#pragma unroll
            for (int p = 0; p < 3; ++p) lds[(s * 3 + p) * 64 + x] = g[p];
        }
    }
    __syncthreads();
    double c[6][4];
#pragma unroll
    for (int s = 0; s < 24; ++s) c[s / 4][s % 4] = valid ? lds[(s * 3 + m) * 64 + x] : 0.0;
    __syncthreads();
    if (valid) {
        double o[12];
        fvh_flux_at<YDIR, FLUX>(a, i, j, m, c, o);
#pragma unroll
        for (int k = 0; k < 12; ++k) lds[(m * 12 + k) * 64 + x] = o[k];
    }
    __syncthreads();
    if (valid && m == 0) {
        double* G = YDIR ? a.gy : a.gx;
#pragma unroll
        for (int k = 0; k < 12; ++k) G[k * nf + t] = (lds[k * 64 + x] + lds[(12 + k) * 64 + x]) + lds[(24 + k) * 64 + x];
    }
}

// FP32 bulk form of fvh_flux_block: the series values go through LDS in FP32
// (24 x 3 x 64 floats), the weighted fluxes in FP64 (12 x 3 x 64 doubles) in
// the same 18 KB.
template <bool YDIR, int FLUX>
__device__ void fvh_flux_block_f32(const FvhArgs& a, long long t0, long long nf, double* lds) {
    float* ldsf = reinterpret_cast<float*>(lds);
    const int x = threadIdx.x, m = threadIdx.y;
    const long long t = t0 + x;
    const bool valid = t < nf;
    int i = 0, j = 0;
    if (YDIR) {
        i = (int)(t % a.nx);
        j = (int)(t / a.nx) - 1;
    } else {
        i = (int)(t % (a.nx + 1)) - 1;
        j = (int)(t / (a.nx + 1));
    }
    const float* F = reinterpret_cast<const float*>(YDIR ? a.fy : a.fx);
    if (valid) {
        long long fid[5];
        fvh_flux_fids<YDIR>(a, i, j, fid);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int s = 8 * m + q;
            float g[3];
            fvh_series_gp3_f32(kQuad, F, nf, s, fid, g);
#pragma unroll
            for (int p = 0; p < 3; ++p) ldsf[(s * 3 + p) * 64 + x] = g[p];
        }
    }
    __syncthreads();
    float c[6][4];
#pragma unroll
    for (int s = 0; s < 24; ++s) c[s / 4][s % 4] = valid ? ldsf[(s * 3 + m) * 64 + x] : 0.0f;
    __syncthreads();
    if (valid) {
        double o[12];
        fvh_flux_at_f32<YDIR, FLUX>(a, i, j, m, c, o);
#pragma unroll
        for (int k = 0; k < 12; ++k) lds[(m * 12 + k) * 64 + x] = o[k];
    }
    __syncthreads();
    if (valid && m == 0) {
        double* G = YDIR ? a.gy : a.gx;
#pragma unroll
        for (int k = 0; k < 12; ++k) G[k * nf + t] = (lds[k * 64 + x] + lds[(12 + k) * 64 + x]) + lds[(24 + k) * 64 + x];
    }
}

// All of one face on one work item in the FP32 bulk precision (the CPU form)
template <bool YDIR, int FLUX>
HTX_HD void fvh_flux_face_f32(const FvhArgs& a, int i, int j, const float* F, long long fstr, double out[3][12]) {
    long long fid[5];
    fvh_flux_fids<YDIR>(a, i, j, fid);
    float c[3][6][4];
    for (int s = 0; s < 24; ++s) {
        float g[3];
        fvh_series_gp3_f32(kQuad, F, fstr, s, fid, g);
        for (int m = 0; m < 3; ++m) c[m][s / 4][s % 4] = g[m];
    }
    for (int m = 0; m < 3; ++m) fvh_flux_at_f32<YDIR, FLUX>(a, i, j, m, c[m], out[m]);
}

// divergence of the 12 fluxes (FVHermite_System2D.cpp:710-730), SSP-RK3
// stage (:733-782) and, in stage 3, the CFL maxima of u.  In stage 1 the
// base state u^n of v and w is the limited one (save() runs after rhs()
// limited them in place), so it is also written back into q0.
HTX_HD void fvh_update_cell(const FvhArgs& a, int i, int j, double dt, double& smx, double& smy,
                            bool* bad = nullptr) {
    const long long fs = a.fstride;
    long long off = (long long)j * a.pitch + i;
    long long gxs = (long long)a.ny * (a.nx + 1), gys = (long long)(a.ny + 1) * a.nx;
    long long bx = (long long)j * (a.nx + 1) + i, by = (long long)j * a.nx + i;
    double unew[4];
    bool fin = true;
    for (int c = 0; c < 12; ++c) {
        double rhs = ((a.gx[c * gxs + bx] - a.gx[c * gxs + bx + 1]) * a.dy +
                      (a.gy[c * gys + by] - a.gy[c * gys + by + a.nx]) * a.dx) /
                     a.area;
        double us = a.q[off + c * fs];
        double v;
        if (a.mode == 0) {
            v = rhs;
        } else if (a.mode == 1) {
            double base = c < 4 ? us : a.lv[(c - 4) * fs + off];
            v = base + dt * rhs;
            if (c >= 4) const_cast<double*>(a.q0)[off + c * fs] = base;
        } else if (a.mode == 2) {
            v = 0.75 * a.q0[off + c * fs] + 0.25 * (us + rhs * dt);
        } else {
            v = (a.q0[off + c * fs] + 2.0 * (us + dt * rhs)) / 3.0;
        }
        if (c < 4) unew[c] = v;
        fin = fin && fabs(v) <= 1.7976931348623157e308;
        a.qout[off + c * fs] = v;
    }
    if (a.mode == 3) {
        if (bad && !fin) *bad = true;
        double wu = unew[1] / unew[0], wv = unew[2] / unew[0];
        double p = (unew[3] - 0.5 * (unew[1] * wu + unew[2] * wv)) * a.g1;
        if (a.quirks) {
            smx = fmax(smx, max_speed<true>(a.gam, wu, p, unew[0]));
            smy = fmax(smy, max_speed<true>(a.gam, wv, p, unew[0]));
        } else {
            smx = fmax(smx, max_speed<false>(a.gam, wu, p, unew[0]));
            smy = fmax(smy, max_speed<false>(a.gam, wv, p, unew[0]));
        }
    }
}

}  // namespace HTX_NS
