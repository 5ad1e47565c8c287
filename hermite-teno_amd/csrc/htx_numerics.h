// htx_numerics.h -- device numerics of the 2D Euler FV path (gfx950).
//
// Reconstruction, transverse quadrature, Roe eigensystem and Riemann fluxes,
// written for registers: no Eigen temporaries, no per-object scratch.  Each
// routine follows one reference routine; the citation is at the definition.
// Divisions by a common denominator are kept as divisions (IEEE, correctly
// rounded) so the FP64 path matches the FP64 CPU restatement to rounding
// of fused multiply-adds only.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

// The kernels are compiled twice into libhtx.so: once as the fast FP64 /
// mixed-precision build (namespace htx) and once as the strict IEEE build
// (htx_strict.hip: no contraction, division by division), each in its own
// namespace so the two builds of every inline function never meet.
#ifndef HTX_NS
#define HTX_NS htx
#endif

// Host+device so that the same numerics can be exercised by CPU builds.
#define HTX_HD __host__ __device__ __forceinline__

// The reconstruction chain (cell primitives, Roe average, characteristic
// projection, TENO5) is evaluated without FMA contraction, in the reference's
// operation order.  TENO5's full-stencil indicator sums terms of size ~1e5 u^2
// down to ~beta, so a one-ulp change in any input moves the cut-off decision
// near its threshold; keeping this chain uncontracted makes the FP64 face
// states bit-identical to an IEEE evaluation of the reference formulas.
#define HTX_NO_CONTRACT _Pragma("clang fp contract(off)")

// Divisions that feed no discontinuous decision (normalisations, the
// quadrature's nonlinear weights, the HLLC star state) are strength-reduced
// to one reciprocal and multiplies, which moves results by a few ulp.  The
// CPU tiling harness builds with HTX_EXACT_DIVISION=1 to reproduce the
// reference's division-by-division evaluation bit for bit.
#ifndef HTX_EXACT_DIVISION
#define HTX_EXACT_DIVISION 0
#endif

namespace HTX_NS {

// Coefficients of Schemes::Lagrange that involve sqrt(15)
// (Schemes_Lagrange.cpp:64-107), formed with the reference's own expressions
// from sqrt(15).  The kernels read the compile-time copy kQuad (literal
// operands the compiler rematerialises, instead of kernel arguments held in
// scalar registers across the whole tile); the host forms the same table at
// run time from sqrt(15.0) and the library refuses to start if the two differ
// in any bit (check_quad_consts).
struct QuadConsts {
    double gp[3][3];   // crjgp3L(r, k)
    double gm[3][3];   // crjgm3L(r, k)
    double linwp[3], linwm[3], linw0p[3], linw0m[3];
    double sumw0p, sumw0m;
    float lw32[12];  // the four linear weight sets in FP32 (+xi, centre+, centre-, -xi)
    // weno3q with every nonlinear weight equal to its linear weight, folded into
    // one 5-point filter per Gauss point (+xi, centre, -xi): the learned path's
    // transverse quadrature on smooth windows
    double lq[3][5];
    // FP32 bulk precision: the candidate coefficients and linear filters in
    // FP32, the centre's group sums
    float gp32[3][3], gm32[3][3], lq32[3][5], sumw0p32, sumw0m32;
};

constexpr QuadConsts make_quad_consts(double sq15) {
    // Schemes_Lagrange.cpp:64-105, same expressions
    QuadConsts C{};
    C.gp[0][0] = 1.0 / 30.0 + sq15 / 20.0;
    C.gp[1][0] = -1.0 / 15.0 - sq15 / 5.0;
    C.gp[2][0] = 31.0 / 30.0 + 3.0 * sq15 / 20.0;
    C.gp[0][1] = 1.0 / 30.0 - sq15 / 20.0;
    C.gp[1][1] = 14.0 / 15.0;
    C.gp[2][1] = 1.0 / 30.0 + sq15 / 20.0;
    C.gp[0][2] = 31.0 / 30.0 - 3.0 * sq15 / 20.0;
    C.gp[1][2] = -1.0 / 15.0 + sq15 / 5.0;
    C.gp[2][2] = 1.0 / 30.0 - sq15 / 20.0;
    C.gm[0][0] = 1.0 / 30.0 - sq15 / 20.0;
    C.gm[1][0] = -1.0 / 15.0 + sq15 / 5.0;
    C.gm[2][0] = 31.0 / 30.0 - 3.0 * sq15 / 20.0;
    C.gm[0][1] = 1.0 / 30.0 + sq15 / 20.0;
    C.gm[1][1] = 14.0 / 15.0;
    C.gm[2][1] = 1.0 / 30.0 - sq15 / 20.0;
    C.gm[0][2] = 31.0 / 30.0 + 3.0 * sq15 / 20.0;
    C.gm[1][2] = -1.0 / 15.0 - sq15 / 5.0;
    C.gm[2][2] = 1.0 / 30.0 + sq15 / 20.0;
    const double lin0[3] = {-9.0 / 80.0, 49.0 / 40.0, -9.0 / 80.0};
    double p[3] = {0.0, 0.0, 0.0}, m[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; ++i) {
        p[i] = 0.5 * (lin0[i] + 3.0 * (lin0[i] < 0.0 ? -lin0[i] : lin0[i]));
        m[i] = p[i] - lin0[i];
    }
    C.sumw0p = (p[0] + p[1]) + p[2];
    C.sumw0m = (m[0] + m[1]) + m[2];
    for (int i = 0; i < 3; ++i) {
        C.linw0p[i] = p[i] / C.sumw0p;
        C.linw0m[i] = m[i] / C.sumw0m;
    }
    double den = 15720.0 * sq15 + 10480.0;
    C.linwp[0] = (2882.0 * sq15 - 1179.0) / den;
    C.linwp[1] = (9672.0 * sq15 + 6448.0) / den;
    C.linwp[2] = (3166.0 * sq15 + 5211.0) / den;
    den = 15720.0 * sq15 - 10480;
    C.linwm[0] = (2882.0 * sq15 + 1179.0) / den;
    C.linwm[1] = (9672.0 * sq15 - 6448.0) / den;
    C.linwm[2] = (3166.0 * sq15 - 5211.0) / den;
    for (int i = 0; i < 3; ++i) {
        C.lw32[i] = (float)C.linwp[i];
        C.lw32[3 + i] = (float)C.linw0p[i];
        C.lw32[6 + i] = (float)C.linw0m[i];
        C.lw32[9 + i] = (float)C.linwm[i];
    }
    // candidate k at a Gauss point is sum_j c[j][k] u_{k+j}; its linear weight
    // multiplies each tap.  The centre's split groups recombine to the plain
    // centre weights lin0 (sumw0p linw0p - sumw0m linw0m = p - m = lin0).
    const double cc[3][3] = {{-1.0 / 24.0, -1.0 / 24.0, 23.0 / 24.0},
                             {2.0 / 24.0, 26.0 / 24.0, 2.0 / 24.0},
                             {23.0 / 24.0, -1.0 / 24.0, -1.0 / 24.0}};
    for (int m = 0; m < 3; ++m)
        for (int i = 0; i < 5; ++i) C.lq[m][i] = 0.0;
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) {
            C.lq[0][k + j] += C.linwp[k] * C.gp[j][k];
            C.lq[1][k + j] += lin0[k] * cc[j][k];
            C.lq[2][k + j] += C.linwm[k] * C.gm[j][k];
        }
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            C.gp32[r][k] = (float)C.gp[r][k];
            C.gm32[r][k] = (float)C.gm[r][k];
        }
    for (int m = 0; m < 3; ++m)
        for (int i = 0; i < 5; ++i) C.lq32[m][i] = (float)C.lq[m][i];
    C.sumw0p32 = (float)C.sumw0p;
    C.sumw0m32 = (float)C.sumw0m;
    return C;
}

// sqrt(15) correctly rounded (IEEE sqrt is, so the host's sqrt(15.0) equals it)
constexpr double kSqrt15 = 0x1.efbdeb14f4edap+1;
constexpr QuadConsts kQuad = make_quad_consts(kSqrt15);

inline void fill_quad_consts(QuadConsts& C) { C = make_quad_consts(sqrt(15.0)); }

// true when the run-time table equals the compile-time one in every bit
inline bool check_quad_consts() {
    QuadConsts C;
    fill_quad_consts(C);
    const unsigned char* a = reinterpret_cast<const unsigned char*>(&C);
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&kQuad);
    // the members are contiguous; the tail padding is not compared
    static_assert(offsetof(QuadConsts, sumw0m32) + sizeof(float) + 4 == sizeof(QuadConsts), "layout");
    for (unsigned k = 0; k < offsetof(QuadConsts, sumw0m32) + sizeof(float); ++k)
        if (a[k] != b[k]) return false;
    return true;
}

// Fast FP32 reciprocal and base-2 logarithm on the device (one hardware
// instruction each, about 1 ulp), used only by the FP32 decision paths;
// exact forms on the host.
HTX_HD float frcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// FP64 reciprocal for the fast build's latency-bound chains: the hardware
// estimate refined by two Newton steps (within about 1 ulp, five dependent
// instructions instead of a correctly rounded division's eleven); the exact
// quotient on the host.  For finite x != 0.
HTX_HD double drcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(x);
    y = fma(fma(-x, y, 1.0), y, y);
    return fma(fma(-x, y, 1.0), y, y);
#else
    return 1.0 / x;
#endif
}

// Correctly rounded FP64 quotients x / b that share the divisor's reciprocal
// (strict build).  gfx950 divides as follows: v_div_scale pre-scales both
// operands, y = v_rcp_f64(b) is refined by two Newton steps
// (e = 1 - b y, y = y + y e, twice), q0 = x y, r = x - b q0 (exact with an
// FMA), q = q0 + r y (v_div_fmas), and v_div_fixup handles special values.
// For a positive divisor b in [2^-300, 2^300] and a numerator x = +-0 or
// 2^-400 <= |x| <= 2^400 the scaling and the fix-up leave every value
// unchanged, so the same operations without them give the same bits, and
// the reciprocal steps are shared by all quotients of one divisor: three
// instructions per quotient instead of eleven.  The residual is formed
// negated (b q0 - x, exact as well) so that x = -0 gives -0 as the division
// does.  Callers guarantee the ranges (or test them and divide otherwise).
// On the host every quotient is the plain division, the same correctly
// rounded value.
struct DivBy {
    double b, y;
};

HTX_HD DivBy div_by(double b) {
    DivBy d;
    d.b = b;
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    d.y = fma(y, e, y);
#else
    d.y = 1.0 / b;
#endif
    return d;
}

HTX_HD double quot(const DivBy& d, double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double q0 = x * d.y;
    const double nr = fma(d.b, q0, -x);
    return fma(-nr, d.y, q0);
#else
    return x / d.b;
#endif
}

// x / c for a constant divisor c given with yc = 1 / c correctly rounded
// (a compile-time constant): with y the correctly rounded reciprocal, one
// correction of q0 = x y by the exact residual gives the correctly rounded
// quotient (Markstein's theorem), for numerators far from the exponent
// limits; others (and zero, whose sign the correction may drop) are divided.
HTX_HD double quot_const(double x, double c, double yc) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double ax = fabs(x);
    if (!(ax >= 0x1p-400 && ax <= 0x1p400)) return x / c;
    const double q0 = x * yc;
    const double nr = fma(c, q0, -x);
    return fma(-nr, yc, q0);
#else
    (void)yc;
    return x / c;
#endif
}

// TENO cut-off decisions of one side (cutOff, Schemes_Lagrange.cpp:1141-1149;
// Schemes_Hermite.cpp:614-628): candidate k is kept unless g_k / sum g <
// ct, g_k = (1 + tau / (s_k + 1e-40))^6.  The sixth power is formed as a
// cube squared; in the strict build, where a ratio falls within 1e-13 of ct
// (there the power's last-bit rounding could decide), the powers are formed
// again with pow(., 6), as the reference's Eigen .pow(6) does.  NaN ratios
// (inf / inf) keep the candidate.  Returns the kept mask (bit k).
HTX_HD unsigned teno_keep(double tau, double s0, double s1, double s2, double ct) {
    HTX_NO_CONTRACT
    double a0 = 1.0 + tau / (s0 + 1e-40);
    double a1 = 1.0 + tau / (s1 + 1e-40);
    double a2 = 1.0 + tau / (s2 + 1e-40);
    double g0 = a0 * a0 * a0; g0 = g0 * g0;
    double g1 = a1 * a1 * a1; g1 = g1 * g1;
    double g2 = a2 * a2 * a2; g2 = g2 * g2;
    double gs = (g0 + g1) + g2;
    // g_k >= 1, so with gs <= 2^300 every quotient is in the shared-reciprocal
    // range (the correctly rounded quotient, in both builds)
    double r0, r1, r2;
    {
        const DivBy D = div_by(gs);
        r0 = quot(D, g0);
        r1 = quot(D, g1);
        r2 = quot(D, g2);
    }
    if (!(gs <= 0x1p300)) {
        r0 = g0 / gs;
        r1 = g1 / gs;
        r2 = g2 / gs;
    }
#if HTX_EXACT_DIVISION
    const double band = ct * 1e-13;
    if (fabs(r0 - ct) <= band || fabs(r1 - ct) <= band || fabs(r2 - ct) <= band) {
        g0 = pow(a0, 6.0);
        g1 = pow(a1, 6.0);
        g2 = pow(a2, 6.0);
        gs = (g0 + g1) + g2;
        r0 = g0 / gs;
        r1 = g1 / gs;
        r2 = g2 / gs;
    }
#endif
    return (r0 < ct ? 0u : 1u) | (r1 < ct ? 0u : 2u) | (r2 < ct ? 0u : 4u);
}

// p0 w0 + p1 w1 + p2 w2 with the kept linear weights normalised,
// w_k = d_k / ((d'_0 + d'_1) + d'_2), d'_k = d_k if kept else 0, as
// Schemes_Lagrange.cpp:496-498 evaluates it, then catchnan.  At least one
// candidate is always kept, so the divisor lies in [min d, 1] and the
// quotients share its reciprocal in the strict build.
HTX_HD double teno_combine(unsigned keep, double p0, double p1, double p2, double d0, double d1, double d2) {
    HTX_NO_CONTRACT
    double w0 = (keep & 1u) ? d0 : 0.0, w1 = (keep & 2u) ? d1 : 0.0, w2 = (keep & 4u) ? d2 : 0.0;
    double ws = (w0 + w1) + w2;
#if HTX_EXACT_DIVISION
    const DivBy D = div_by(ws);
    w0 = quot(D, w0);
    w1 = quot(D, w1);
    w2 = quot(D, w2);
    double out = (p0 * w0 + p1 * w1) + p2 * w2;
#else
    double out = ((p0 * w0 + p1 * w1) + p2 * w2) * drcp(ws);
#endif
    return out != out ? 1e-15 : out;  // catchnan
}

// 1 / x where the reference divides: the correctly rounded quotient in the
// strict build, drcp in the fast one
HTX_HD double inv(double x) {
#if HTX_EXACT_DIVISION
    return 1.0 / x;
#else
    return drcp(x);
#endif
}

HTX_HD float flog2(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_logf(x);
#else
    return log2f(x);
#endif
}

struct Prim {  // per-cell quantities used by the Roe average
    double sr, u, v, h;
};

// Schemes::Lagrange::teno5, one side (Schemes_Lagrange.cpp:481-511 with
// smoothnessMeasure :1013-1026, fullpointsSmootness :1029-1034, cutOff
// :1141-1149, catchnan :1151-1153).  v0..v4 = 5-point window, p = the three
// candidate values, d = linear weights.  Returns the TENO value; *cut gets a
// bitmask of discarded stencils.
HTX_HD double teno5_side(double v0, double v1, double v2, double v3, double v4,
                                             double p0, double p1, double p2, double d0, double d1,
                                             double d2, unsigned* cut) {
    HTX_NO_CONTRACT
    const double c1 = 13.0 / 12.0;
    double c_0 = v2 - 2.0 * v3 + v4, b_0 = (3.0 * v2 - 4.0 * v3 + v4) / 2.0;
    double c_1 = v1 - 2.0 * v2 + v3, b_1 = (v1 - v3) / 2.0;
    double c_2 = v0 - 2.0 * v1 + v2, b_2 = (v0 - 4.0 * v1 + 3.0 * v2) / 2.0;
    double s0 = b_0 * b_0 + c1 * c_0 * c_0;
    double s1 = b_1 * b_1 + c1 * c_1 * c_1;
    double s2 = b_2 * b_2 + c1 * c_2 * c_2;
    double b35 = v4 * (6908.0 * v4 - 51001.0 * v3 + 67923.0 * v2 - 38947.0 * v1 + 8209.0 * v0) +
                 v3 * (104963.0 * v3 - 299076.0 * v2 + 179098.0 * v1 - 38947.0 * v0) +
                 v2 * (231153.0 * v2 - 299076.0 * v1 + 67923.0 * v0) +
                 v1 * (104963.0 * v1 - 51001.0 * v0) + 6908.0 * v0 * v0;
    b35 = quot_const(b35, 5040.0, 1.0 / 5040.0);
    double tau = fabs(b35 - quot_const((s0 + s2 + 4.0 * s1), 6.0, 1.0 / 6.0));
    const unsigned keep = teno_keep(tau, s0, s1, s2, 5e-4);
    if (cut) *cut = keep ^ 7u;
    return teno_combine(keep, p0, p1, p2, d0, d1, d2);
}

// Mixed-precision TENO5 side: the same cut-off rule evaluated in FP32 on
// differences formed in FP64; the reconstructed value is the FP64 candidate
// combination with the selected linear weights, exactly as in teno5_side.
// beta5 uses its translation-invariant form in first differences
// d_k = v_k - v_{k-1}: beta5 = d^T M d with the exact rational M of the
// reference polynomial (Schemes_Lagrange.cpp:1030-1032), which avoids the
// 1e5-sized cancellation of the expanded form.  The 1e-40 guard becomes
// 1e-30 (below FP32's normal range it would flush); both only act on
// exactly flat stencils, where every candidate coincides.
HTX_HD double teno5_side_mixed(double v0, double v1, double v2, double v3, double v4, double p0, double p1,
                               double p2, double d0, double d1, double d2) {
    // smoothness measures from the first differences e_k = v_k - v_{k-1}
    // (formed in FP64, then FP32): v2 - 2 v3 + v4 = e4 - e3,
    // (3 v2 - 4 v3 + v4)/2 = (e4 - 3 e3)/2, (v1 - v3)/2 = -(e2 + e3)/2,
    // (v0 - 4 v1 + 3 v2)/2 = (3 e2 - e1)/2
    const float c1 = 13.0f / 12.0f;
    float e1 = (float)(v1 - v0), e2 = (float)(v2 - v1), e3 = (float)(v3 - v2), e4 = (float)(v4 - v3);
    float c_0 = e4 - e3, b_0 = (e4 - 3.0f * e3) * 0.5f;
    float c_1 = e3 - e2, b_1 = -(e2 + e3) * 0.5f;
    float c_2 = e2 - e1, b_2 = (3.0f * e2 - e1) * 0.5f;
    float s0 = b_0 * b_0 + c1 * c_0 * c_0;
    float s1 = b_1 * b_1 + c1 * c_1 * c_1;
    float s2 = b_2 * b_2 + c1 * c_2 * c_2;
    const float m11 = 1727.0f / 1260.0f, m22 = 2029.0f / 168.0f, m12 = -2479.0f / 672.0f;
    const float m13 = 5123.0f / 1680.0f, m14 = -8209.0f / 10080.0f, m23 = -12157.0f / 1120.0f;
    float t1 = m11 * e1 + m12 * e2 + m13 * e3 + m14 * e4;
    float t2 = m12 * e1 + m22 * e2 + m23 * e3 + m13 * e4;
    float t3 = m13 * e1 + m23 * e2 + m22 * e3 + m12 * e4;
    float t4 = m14 * e1 + m13 * e2 + m12 * e3 + m11 * e4;
    float b5 = e1 * t1 + e2 * t2 + e3 * t3 + e4 * t4;
    float tau = fabsf(b5 - (s0 + s2 + 4.0f * s1) * (1.0f / 6.0f));
    // a_k = 1 + tau / (beta_k + eps); the cut-off compares a_k^6 / sum_j a_j^6
    // with C_T.  a_k^6 overflows FP32 long before FP64, so the rule is
    // evaluated on r_k = a_k / max_j a_j <= 1: a_k^6 / sum a_j^6 = r_k^6 / sum r_j^6.
    // Candidates whose a_k is infinite (tau / eps overflows: a flat candidate
    // next to a jump) dominate every finite one; NaN keeps every stencil.
    float a0 = 1.0f + tau * frcp(s0 + 1e-30f);
    float a1 = 1.0f + tau * frcp(s1 + 1e-30f);
    float a2 = 1.0f + tau * frcp(s2 + 1e-30f);
    float am = fmaxf(fmaxf(a0, a1), a2);
    bool k0, k1, k2;
    if (am > 3.40282347e38f) {  // am is infinite (am >= 1; NaN fails the test)
        k0 = a0 == am;
        k1 = a1 == am;
        k2 = a2 == am;
    } else {
        float im = frcp(am);
        float r0 = a0 * im, r1 = a1 * im, r2 = a2 * im;
        r0 = r0 * r0 * r0; r0 = r0 * r0;
        r1 = r1 * r1 * r1; r1 = r1 * r1;
        r2 = r2 * r2 * r2; r2 = r2 * r2;
        float lim = 5e-4f * ((r0 + r1) + r2);
        k0 = !(r0 < lim);
        k1 = !(r1 < lim);
        k2 = !(r2 < lim);
    }
    double w0 = k0 ? d0 : 0.0, w1 = k1 ? d1 : 0.0, w2 = k2 ? d2 : 0.0;
    double ws = (w0 + w1) + w2;
    double out = ((p0 * w0 + p1 * w1) + p2 * w2) * drcp(ws);
    return out != out ? 1e-15 : out;
}

// Schemes::Lagrange::teno5 (Schemes_Lagrange.cpp:481-511): u[0..5] =
// u_{i-2..i+3} -> (u^-_{i+1/2}, u^+_{i+1/2}).
HTX_HD void teno5(const double u[6], double& um, double& up, unsigned* cut = nullptr) {
    HTX_NO_CONTRACT
    const double a = 1.0 / 3.0, b = 5.0 / 6.0, c = -1.0 / 6.0, d = 11.0 / 6.0, e = -7.0 / 6.0;
    // crj columns: col0 = (d, e, a), col1 = (a, b, c), col2 = (c, b, a), col3 = (a, e, d)
    double p0 = (a * u[2] + b * u[3]) + c * u[4];
    double p1 = (c * u[1] + b * u[2]) + a * u[3];
    double p2 = (a * u[0] + e * u[1]) + d * u[2];
    unsigned cm = 0, cp = 0;
    um = teno5_side(u[0], u[1], u[2], u[3], u[4], p0, p1, p2, 0.3, 0.6, 0.1, cut ? &cm : nullptr);
    double q0 = (d * u[3] + e * u[4]) + a * u[5];
    double q1 = (a * u[2] + b * u[3]) + c * u[4];
    double q2 = (c * u[1] + b * u[2]) + a * u[3];
    up = teno5_side(u[1], u[2], u[3], u[4], u[5], q0, q1, q2, 0.1, 0.6, 0.3, cut ? &cp : nullptr);
    if (cut) *cut = cm | (cp << 3);
}

// Schemes::Lagrange::linear (Schemes_Lagrange.cpp:1219-1233): the same three
// candidates combined with the optimal weights (0.3, 0.6, 0.1) / mirrored.
HTX_HD void linear5(const double u[6], double& um, double& up) {
    HTX_NO_CONTRACT
    const double a = 1.0 / 3.0, b = 5.0 / 6.0, c = -1.0 / 6.0, d = 11.0 / 6.0, e = -7.0 / 6.0;
    double p0 = (a * u[2] + b * u[3]) + c * u[4];
    double p1 = (c * u[1] + b * u[2]) + a * u[3];
    double p2 = (a * u[0] + e * u[1]) + d * u[2];
    um = (p0 * 0.3 + p1 * 0.6) + p2 * 0.1;
    double q0 = (d * u[3] + e * u[4]) + a * u[5];
    double q1 = (a * u[2] + b * u[3]) + c * u[4];
    double q2 = (c * u[1] + b * u[2]) + a * u[3];
    up = (q0 * 0.1 + q1 * 0.6) + q2 * 0.3;
}

// TENO5 with the FP32 cut-off rule of teno5_side_mixed.
HTX_HD void teno5_mixed(const double u[6], double& um, double& up) {
    const double a = 1.0 / 3.0, b = 5.0 / 6.0, c = -1.0 / 6.0, d = 11.0 / 6.0, e = -7.0 / 6.0;
    // crj columns: col0 = (d, e, a), col1 = (a, b, c), col2 = (c, b, a), col3 = (a, e, d)
    double p0 = (a * u[2] + b * u[3]) + c * u[4];
    double p1 = (c * u[1] + b * u[2]) + a * u[3];
    double p2 = (a * u[0] + e * u[1]) + d * u[2];
    um = teno5_side_mixed(u[0], u[1], u[2], u[3], u[4], p0, p1, p2, 0.3, 0.6, 0.1);
    double q0 = (d * u[3] + e * u[4]) + a * u[5];
    double q1 = (a * u[2] + b * u[3]) + c * u[4];
    double q2 = (c * u[1] + b * u[2]) + a * u[3];
    up = teno5_side_mixed(u[1], u[2], u[3], u[4], u[5], q0, q1, q2, 0.1, 0.6, 0.3);
}

// Normalised WENO weights lin_k / e_k / sum_j (lin_j / e_j) with the
// divisions of weno5JS3PointsQuadrature (Schemes_Lagrange.cpp:1178-1210):
// each lin_k / e_k, then each weight by their sum; the e_k reciprocals come
// shared across the quadrature's four weight sets.  Valid (the correctly
// rounded quotients) for lin_k in (0, 1] and e_k in [1e-12, 2^200]: the
// weights lie in [2^-210, 2^40], their sum in the same range and the
// normalised weights in [2^-250, 1].
HTX_HD void weno_weights_shared(const DivBy E[3], const double lin[3], double& w0, double& w1, double& w2) {
    HTX_NO_CONTRACT
    w0 = quot(E[0], lin[0]);
    w1 = quot(E[1], lin[1]);
    w2 = quot(E[2], lin[2]);
    const DivBy S = div_by((w0 + w1) + w2);
    w0 = quot(S, w0);
    w1 = quot(S, w1);
    w2 = quot(S, w2);
}

// Schemes::Lagrange::weno5JS3PointsQuadrature + smoothnessMeasure3L
// (Schemes_Lagrange.cpp:1155-1217): 5 transverse values -> values at the
// Gauss points (+xi, 0, -xi).
HTX_HD void weno3q(const QuadConsts& C, double u0, double u1, double u2, double u3,
                                       double u4, double& o0, double& o1, double& o2) {
    const double c1 = 13.0 / 12.0;
    double bg0 = u0 - 4.0 * u1 + 3.0 * u2, cg0 = u0 - 2.0 * u1 + u2;
    double bg1 = u1 - u3, cg1 = u1 - 2.0 * u2 + u3;
    double bg2 = 3.0 * u2 - 4.0 * u3 + u4, cg2 = u2 - 2.0 * u3 + u4;
    double s0 = 0.25 * (bg0 * bg0) + c1 * (cg0 * cg0);
    double s1 = 0.25 * (bg1 * bg1) + c1 * (cg1 * cg1);
    double s2 = 0.25 * (bg2 * bg2) + c1 * (cg2 * cg2);
    double e0 = 1e-6 + s0, e1 = 1e-6 + s1, e2 = 1e-6 + s2;
    e0 = e0 * e0; e1 = e1 * e1; e2 = e2 * e2;
    double w0, w1, w2, ws, p0, p1, p2;
#if HTX_EXACT_DIVISION
    if (e0 <= 0x1p200 && e1 <= 0x1p200 && e2 <= 0x1p200) {
        // every division in the shared-reciprocal range (NaN fails the test)
        const DivBy E[3] = {div_by(e0), div_by(e1), div_by(e2)};
        p0 = (C.gp[0][0] * u0 + C.gp[1][0] * u1) + C.gp[2][0] * u2;
        p1 = (C.gp[0][1] * u1 + C.gp[1][1] * u2) + C.gp[2][1] * u3;
        p2 = (C.gp[0][2] * u2 + C.gp[1][2] * u3) + C.gp[2][2] * u4;
        weno_weights_shared(E, C.linwp, w0, w1, w2);
        o0 = (p0 * w0 + p1 * w1) + p2 * w2;
        p0 = quot_const(-u0 + 2.0 * u1 + 23.0 * u2, 24.0, 1.0 / 24.0);
        p1 = quot_const(-u1 + 26.0 * u2 - u3, 24.0, 1.0 / 24.0);
        p2 = quot_const(23.0 * u2 + 2.0 * u3 - u4, 24.0, 1.0 / 24.0);
        weno_weights_shared(E, C.linw0p, w0, w1, w2);
        o1 = C.sumw0p * ((p0 * w0 + p1 * w1) + p2 * w2);
        weno_weights_shared(E, C.linw0m, w0, w1, w2);
        o1 -= C.sumw0m * ((p0 * w0 + p1 * w1) + p2 * w2);
        p0 = (C.gm[0][0] * u0 + C.gm[1][0] * u1) + C.gm[2][0] * u2;
        p1 = (C.gm[0][1] * u1 + C.gm[1][1] * u2) + C.gm[2][1] * u3;
        p2 = (C.gm[0][2] * u2 + C.gm[1][2] * u3) + C.gm[2][2] * u4;
        weno_weights_shared(E, C.linwm, w0, w1, w2);
        o2 = (p0 * w0 + p1 * w1) + p2 * w2;
    } else {
    // +xi
    p0 = (C.gp[0][0] * u0 + C.gp[1][0] * u1) + C.gp[2][0] * u2;
    p1 = (C.gp[0][1] * u1 + C.gp[1][1] * u2) + C.gp[2][1] * u3;
    p2 = (C.gp[0][2] * u2 + C.gp[1][2] * u3) + C.gp[2][2] * u4;
    w0 = C.linwp[0] / e0; w1 = C.linwp[1] / e1; w2 = C.linwp[2] / e2;
    ws = (w0 + w1) + w2;
    w0 = w0 / ws; w1 = w1 / ws; w2 = w2 / ws;
    o0 = (p0 * w0 + p1 * w1) + p2 * w2;
    // centre, positive/negative split
    p0 = (-u0 + 2.0 * u1 + 23.0 * u2) / 24.0;
    p1 = (-u1 + 26.0 * u2 - u3) / 24.0;
    p2 = (23.0 * u2 + 2.0 * u3 - u4) / 24.0;
    w0 = C.linw0p[0] / e0; w1 = C.linw0p[1] / e1; w2 = C.linw0p[2] / e2;
    ws = (w0 + w1) + w2;
    w0 = w0 / ws; w1 = w1 / ws; w2 = w2 / ws;
    o1 = C.sumw0p * ((p0 * w0 + p1 * w1) + p2 * w2);
    w0 = C.linw0m[0] / e0; w1 = C.linw0m[1] / e1; w2 = C.linw0m[2] / e2;
    ws = (w0 + w1) + w2;
    w0 = w0 / ws; w1 = w1 / ws; w2 = w2 / ws;
    o1 -= C.sumw0m * ((p0 * w0 + p1 * w1) + p2 * w2);
    // -xi
    p0 = (C.gm[0][0] * u0 + C.gm[1][0] * u1) + C.gm[2][0] * u2;
    p1 = (C.gm[0][1] * u1 + C.gm[1][1] * u2) + C.gm[2][1] * u3;
    p2 = (C.gm[0][2] * u2 + C.gm[1][2] * u3) + C.gm[2][2] * u4;
    w0 = C.linwm[0] / e0; w1 = C.linwm[1] / e1; w2 = C.linwm[2] / e2;
    ws = (w0 + w1) + w2;
    w0 = w0 / ws; w1 = w1 / ws; w2 = w2 / ws;
    o2 = (p0 * w0 + p1 * w1) + p2 * w2;
    }
#else
    // 3 reciprocals of (eps + beta_k)^2 shared by the four weight sets, one
    // reciprocal per normalisation
    double r0 = drcp(e0), r1 = drcp(e1), r2 = drcp(e2);
    // +xi
    p0 = (C.gp[0][0] * u0 + C.gp[1][0] * u1) + C.gp[2][0] * u2;
    p1 = (C.gp[0][1] * u1 + C.gp[1][1] * u2) + C.gp[2][1] * u3;
    p2 = (C.gp[0][2] * u2 + C.gp[1][2] * u3) + C.gp[2][2] * u4;
    w0 = C.linwp[0] * r0; w1 = C.linwp[1] * r1; w2 = C.linwp[2] * r2;
    ws = (w0 + w1) + w2;
    o0 = ((p0 * w0 + p1 * w1) + p2 * w2) / ws;
    // centre, positive/negative split
    const double r24 = 1.0 / 24.0;
    p0 = (-u0 + 2.0 * u1 + 23.0 * u2) * r24;
    p1 = (-u1 + 26.0 * u2 - u3) * r24;
    p2 = (23.0 * u2 + 2.0 * u3 - u4) * r24;
    w0 = C.linw0p[0] * r0; w1 = C.linw0p[1] * r1; w2 = C.linw0p[2] * r2;
    ws = (w0 + w1) + w2;
    o1 = C.sumw0p * (((p0 * w0 + p1 * w1) + p2 * w2) / ws);
    w0 = C.linw0m[0] * r0; w1 = C.linw0m[1] * r1; w2 = C.linw0m[2] * r2;
    ws = (w0 + w1) + w2;
    o1 -= C.sumw0m * (((p0 * w0 + p1 * w1) + p2 * w2) / ws);
    // -xi
    p0 = (C.gm[0][0] * u0 + C.gm[1][0] * u1) + C.gm[2][0] * u2;
    p1 = (C.gm[0][1] * u1 + C.gm[1][1] * u2) + C.gm[2][1] * u3;
    p2 = (C.gm[0][2] * u2 + C.gm[1][2] * u3) + C.gm[2][2] * u4;
    w0 = C.linwm[0] * r0; w1 = C.linwm[1] * r1; w2 = C.linwm[2] * r2;
    ws = (w0 + w1) + w2;
    o2 = ((p0 * w0 + p1 * w1) + p2 * w2) / ws;
#endif
    o0 = o0 != o0 ? 1e-15 : o0;  // catchnan
    o1 = o1 != o1 ? 1e-15 : o1;
    o2 = o2 != o2 ? 1e-15 : o2;
}

// Mixed-precision weno3q: smoothness and nonlinear weights in FP32 from
// differences formed in FP64, candidates and their combination in FP64.  The
// last weight of each group is 1 minus the others (in FP64), so the
// combination stays exact for constant data.
HTX_HD void weno3q_mixed(const QuadConsts& C, const float* lw, double u0, double u1, double u2, double u3,
                         double u4, double& o0, double& o1, double& o2) {
    // smoothness measures from the first differences d_k = u_k - u_{k-1}
    // (formed in FP64, then FP32): u0 - 4 u1 + 3 u2 = 3 d2 - d1,
    // u0 - 2 u1 + u2 = d2 - d1, u1 - u3 = -(d2 + d3), u1 - 2 u2 + u3 = d3 - d2,
    // 3 u2 - 4 u3 + u4 = d4 - 3 d3, u2 - 2 u3 + u4 = d4 - d3
    const float c1 = 13.0f / 12.0f;
    float d1 = (float)(u1 - u0), d2 = (float)(u2 - u1), d3 = (float)(u3 - u2), d4 = (float)(u4 - u3);
    float bg0 = 3.0f * d2 - d1, cg0 = d2 - d1;
    float bg1 = -(d2 + d3), cg1 = d3 - d2;
    float bg2 = d4 - 3.0f * d3, cg2 = d4 - d3;
    float e0 = 1e-6f + (0.25f * (bg0 * bg0) + c1 * (cg0 * cg0));
    float e1 = 1e-6f + (0.25f * (bg1 * bg1) + c1 * (cg1 * cg1));
    float e2 = 1e-6f + (0.25f * (bg2 * bg2) + c1 * (cg2 * cg2));
    float r0 = frcp(e0 * e0), r1 = frcp(e1 * e1), r2 = frcp(e2 * e2);
    // lw: 4 sets of 3 FP32 linear weights (+xi, centre+, centre-, -xi)
    double W[4][2];
    for (int k = 0; k < 4; ++k) {
        float w0 = lw[3 * k] * r0, w1 = lw[3 * k + 1] * r1, w2 = lw[3 * k + 2] * r2;
        float is = frcp((w0 + w1) + w2);
        W[k][0] = (double)(w0 * is);
        W[k][1] = (double)(w1 * is);
    }
    double p0, p1, p2;
    p0 = (C.gp[0][0] * u0 + C.gp[1][0] * u1) + C.gp[2][0] * u2;
    p1 = (C.gp[0][1] * u1 + C.gp[1][1] * u2) + C.gp[2][1] * u3;
    p2 = (C.gp[0][2] * u2 + C.gp[1][2] * u3) + C.gp[2][2] * u4;
    o0 = p2 + W[0][0] * (p0 - p2) + W[0][1] * (p1 - p2);
    const double r24 = 1.0 / 24.0;
    p0 = (-u0 + 2.0 * u1 + 23.0 * u2) * r24;
    p1 = (-u1 + 26.0 * u2 - u3) * r24;
    p2 = (23.0 * u2 + 2.0 * u3 - u4) * r24;
    o1 = C.sumw0p * (p2 + W[1][0] * (p0 - p2) + W[1][1] * (p1 - p2)) -
         C.sumw0m * (p2 + W[2][0] * (p0 - p2) + W[2][1] * (p1 - p2));
    p0 = (C.gm[0][0] * u0 + C.gm[1][0] * u1) + C.gm[2][0] * u2;
    p1 = (C.gm[0][1] * u1 + C.gm[1][1] * u2) + C.gm[2][1] * u3;
    p2 = (C.gm[0][2] * u2 + C.gm[1][2] * u3) + C.gm[2][2] * u4;
    o2 = p2 + W[3][0] * (p0 - p2) + W[3][1] * (p1 - p2);
    o0 = o0 != o0 ? 1e-15 : o0;
    o1 = o1 != o1 ? 1e-15 : o1;
    o2 = o2 != o2 ? 1e-15 : o2;
}

// weno3q's linear limit: the 5th-order central interpolation of 5 transverse
// values to the Gauss points (+xi, 0, -xi), exact for cell averages of a
// quartic.  Used by the learned path where the window is smooth.
HTX_HD void weno3q_linear(const QuadConsts& C, double u0, double u1, double u2, double u3, double u4, double& o0,
                          double& o1, double& o2) {
    o0 = C.lq[0][0] * u0 + C.lq[0][1] * u1 + C.lq[0][2] * u2 + C.lq[0][3] * u3 + C.lq[0][4] * u4;
    o1 = C.lq[1][0] * u0 + C.lq[1][1] * u1 + C.lq[1][2] * u2 + C.lq[1][3] * u3 + C.lq[1][4] * u4;
    o2 = C.lq[2][0] * u0 + C.lq[2][1] * u1 + C.lq[2][2] * u2 + C.lq[2][3] * u3 + C.lq[2][4] * u4;
}

// The three Gauss points of weno3q (o[0]: +xi, o[1]: centre, o[2]: -xi),
// each evaluated with exactly the operations weno3q uses for that output;
// the smoothness indicators (and in the fast build their reciprocals) are
// formed once for the three.
HTX_HD void weno3q_gp3(const QuadConsts& C, double u0, double u1, double u2, double u3, double u4, double o[3]) {
    const double c1 = 13.0 / 12.0;
    double bg0 = u0 - 4.0 * u1 + 3.0 * u2, cg0 = u0 - 2.0 * u1 + u2;
    double bg1 = u1 - u3, cg1 = u1 - 2.0 * u2 + u3;
    double bg2 = 3.0 * u2 - 4.0 * u3 + u4, cg2 = u2 - 2.0 * u3 + u4;
    double s0 = 0.25 * (bg0 * bg0) + c1 * (cg0 * cg0);
    double s1 = 0.25 * (bg1 * bg1) + c1 * (cg1 * cg1);
    double s2 = 0.25 * (bg2 * bg2) + c1 * (cg2 * cg2);
    double e0 = 1e-6 + s0, e1 = 1e-6 + s1, e2 = 1e-6 + s2;
    e0 = e0 * e0; e1 = e1 * e1; e2 = e2 * e2;
#if !HTX_EXACT_DIVISION
    const double r0 = drcp(e0), r1 = drcp(e1), r2 = drcp(e2);
#else
    // shared reciprocals when every division is in their range (see
    // weno_weights_shared; NaN fails the test)
    const bool shared = e0 <= 0x1p200 && e1 <= 0x1p200 && e2 <= 0x1p200;
    const DivBy E[3] = {div_by(e0), div_by(e1), div_by(e2)};
#endif
    double w0, w1, w2, ws, p0, p1, p2, v;
    // centre
    {
#if HTX_EXACT_DIVISION
        if (shared) {
            p0 = quot_const(-u0 + 2.0 * u1 + 23.0 * u2, 24.0, 1.0 / 24.0);
            p1 = quot_const(-u1 + 26.0 * u2 - u3, 24.0, 1.0 / 24.0);
            p2 = quot_const(23.0 * u2 + 2.0 * u3 - u4, 24.0, 1.0 / 24.0);
            weno_weights_shared(E, C.linw0p, w0, w1, w2);
            v = C.sumw0p * ((p0 * w0 + p1 * w1) + p2 * w2);
            weno_weights_shared(E, C.linw0m, w0, w1, w2);
            v -= C.sumw0m * ((p0 * w0 + p1 * w1) + p2 * w2);
        } else {
        p0 = (-u0 + 2.0 * u1 + 23.0 * u2) / 24.0;
        p1 = (-u1 + 26.0 * u2 - u3) / 24.0;
        p2 = (23.0 * u2 + 2.0 * u3 - u4) / 24.0;
        w0 = C.linw0p[0] / e0; w1 = C.linw0p[1] / e1; w2 = C.linw0p[2] / e2;
        ws = (w0 + w1) + w2;
        w0 = w0 / ws; w1 = w1 / ws; w2 = w2 / ws;
        v = C.sumw0p * ((p0 * w0 + p1 * w1) + p2 * w2);
        w0 = C.linw0m[0] / e0; w1 = C.linw0m[1] / e1; w2 = C.linw0m[2] / e2;
        ws = (w0 + w1) + w2;
        w0 = w0 / ws; w1 = w1 / ws; w2 = w2 / ws;
        v -= C.sumw0m * ((p0 * w0 + p1 * w1) + p2 * w2);
        }
#else
        const double r24 = 1.0 / 24.0;
        p0 = (-u0 + 2.0 * u1 + 23.0 * u2) * r24;
        p1 = (-u1 + 26.0 * u2 - u3) * r24;
        p2 = (23.0 * u2 + 2.0 * u3 - u4) * r24;
        w0 = C.linw0p[0] * r0; w1 = C.linw0p[1] * r1; w2 = C.linw0p[2] * r2;
        ws = (w0 + w1) + w2;
        v = C.sumw0p * (((p0 * w0 + p1 * w1) + p2 * w2) / ws);
        w0 = C.linw0m[0] * r0; w1 = C.linw0m[1] * r1; w2 = C.linw0m[2] * r2;
        ws = (w0 + w1) + w2;
        v -= C.sumw0m * (((p0 * w0 + p1 * w1) + p2 * w2) / ws);
#endif
        o[1] = v != v ? 1e-15 : v;
    }
#pragma unroll
    for (int m = 0; m <= 2; m += 2) {
        const double(*cc)[3] = m == 0 ? C.gp : C.gm;
        const double* lw = m == 0 ? C.linwp : C.linwm;
        p0 = (cc[0][0] * u0 + cc[1][0] * u1) + cc[2][0] * u2;
        p1 = (cc[0][1] * u1 + cc[1][1] * u2) + cc[2][1] * u3;
        p2 = (cc[0][2] * u2 + cc[1][2] * u3) + cc[2][2] * u4;
#if HTX_EXACT_DIVISION
        if (shared) {
            weno_weights_shared(E, lw, w0, w1, w2);
        } else {
            w0 = lw[0] / e0; w1 = lw[1] / e1; w2 = lw[2] / e2;
            ws = (w0 + w1) + w2;
            w0 = w0 / ws; w1 = w1 / ws; w2 = w2 / ws;
        }
        v = (p0 * w0 + p1 * w1) + p2 * w2;
#else
        w0 = lw[0] * r0; w1 = lw[1] * r1; w2 = lw[2] * r2;
        ws = (w0 + w1) + w2;
        v = ((p0 * w0 + p1 * w1) + p2 * w2) / ws;
#endif
        o[m] = v != v ? 1e-15 : v;
    }
}

// Per-cell quantities of calcPrimitiveVariable (FV_System2D.cpp:226-263).
HTX_HD Prim cell_prim(double g1, double q0, double q1, double q2, double q3) {
    HTX_NO_CONTRACT
    Prim P;
    // exact in both builds: these feed the eigenvectors, and through them the
    // TENO cut-off decisions
    double u = q1 / q0, v = q2 / q0;
    double p = (q3 - 0.5 * (q1 * u + q2 * v)) * g1;
    P.sr = sqrt(fabs(q0));
    P.u = u;
    P.v = v;
    P.h = (q3 + p) / q0;
    return P;
}

// Roe-averaged eigensystem of one face (FV_System2D.cpp:271-298 / :332-358,
// Equations_Euler2D.cpp:100-205), kept as the 9 scalars the 4x4 matrices
// are built from.
struct Eig {
    double vn, vt, hm, qm, cm, rcm, b1, b2, t2;  // vn: normal, vt: tangential velocity
};

HTX_HD Eig roe_eig(double g1, bool ydir, const Prim& a, const Prim& b) {
    HTX_NO_CONTRACT
    Eig E;
    double t0 = a.sr / (a.sr + b.sr);
    double t1 = 1.0 - t0;
    double vxm = t0 * a.u + t1 * b.u;
    double vym = t0 * a.v + t1 * b.v;
    E.hm = t0 * a.h + t1 * b.h;
    E.qm = 0.5 * (vxm * vxm + vym * vym);
    E.cm = sqrt(fabs(g1 * (E.hm - E.qm)));
    E.rcm = 1.0 / E.cm;  // exact in both builds: the eigenvectors feed the TENO decisions
    E.b1 = g1 * E.rcm * E.rcm;
    E.b2 = E.qm * E.b1;
    E.t2 = 0.5 * E.b1;
    E.vn = ydir ? vym : vxm;
    E.vt = ydir ? vxm : vym;
    return E;
}

// Characteristic projection L*q of a state given in (rho, rho u, rho v, E)
// order.  The x and y left eigenvector matrices are one matrix once the
// momenta are read as (normal, tangential); the sums run over the natural
// component order, as Eigen's row-times-vector does.  Products with the
// matrices' exact 0 and 1 entries are not formed.
template <bool YDIR>
HTX_HD void eig_left(const Eig& E, double q0, double qx, double qy, double q3, double& c0, double& c1,
                     double& c2, double& c3) {
    HTX_NO_CONTRACT
    double tr = E.vn * E.rcm, tb = E.b1 * E.vn;
    double ln0 = -0.5 * (tb + E.rcm), lt0 = -E.t2 * E.vt;  // row 0: normal / tangential entries
    double ln3 = -0.5 * (tb - E.rcm);                      // row 3 normal entry (tangential = lt0)
    double lt2 = E.b1 * E.vt;                              // row 2 tangential entry (normal = tb)
    if (!YDIR) {
        c0 = ((0.5 * (E.b2 + tr) * q0 + ln0 * qx) + lt0 * qy) + E.t2 * q3;
        c1 = -E.vt * q0 + qy;
        c2 = (((1.0 - E.b2) * q0 + tb * qx) + lt2 * qy) + (-E.b1) * q3;
        c3 = ((0.5 * (E.b2 - tr) * q0 + ln3 * qx) + lt0 * qy) + E.t2 * q3;
    } else {
        c0 = ((0.5 * (E.b2 + tr) * q0 + lt0 * qx) + ln0 * qy) + E.t2 * q3;
        c1 = -E.vt * q0 + qx;
        c2 = (((1.0 - E.b2) * q0 + lt2 * qx) + tb * qy) + (-E.b1) * q3;
        c3 = ((0.5 * (E.b2 - tr) * q0 + lt0 * qx) + ln3 * qy) + E.t2 * q3;
    }
}

// Back-projection R*c, returned in (rho, normal, tangential, E) ordering.
HTX_HD void eig_right(const Eig& E, double c0, double c1, double c2, double c3,
                                          double& q0, double& qn, double& qt, double& q3) {
    double tc = E.vn * E.cm;
    q0 = (c0 + c2) + c3;
    qn = ((E.vn - E.cm) * c0 + E.vn * c2) + (E.vn + E.cm) * c3;
    qt = ((E.vt * c0 + c1) + E.vt * c2) + E.vt * c3;
    q3 = (((E.hm - tc) * c0 + E.vt * c1) + E.qm * c2) + (E.hm + tc) * c3;
}

// Physical normal flux (Equations_Euler2D.cpp:36-50) in (rho, n, t, E) order.
HTX_HD void phys_flux(double g1, double q0, double qn, double qt, double q3,
                                          double& f0, double& fn, double& ft, double& f3, double& wn,
                                          double& p, double& rho) {
    rho = q0;
#if HTX_EXACT_DIVISION
    wn = qn / q0;
    double wt = qt / q0;
#else
    double iq = drcp(q0);
    wn = qn * iq;
    double wt = qt * iq;
#endif
    p = (q3 - 0.5 * (qn * wn + qt * wt)) * g1;
    f0 = qn;
    fn = qn * wn + p;
    ft = qn * wt;
    f3 = wn * (q3 + p);
}

// maxLocalWaveSpeed (Equations_Euler2D.cpp:221-246) + characteristic Rusanov
// (:310-330); states and flux in (rho, n, t, E) ordering.
template <bool YDIR>
HTX_HD void rusanov_char(double gam, double g1, const Eig& E, const double ql[4],
                                             const double qr[4], double f[4]) {
    double fl0, fln, flt, fl3, wnl, pl, rl;
    double fr0, frn, frt, fr3, wnr, pr, rr;
    phys_flux(g1, ql[0], ql[1], ql[2], ql[3], fl0, fln, flt, fl3, wnl, pl, rl);
    phys_flux(g1, qr[0], qr[1], qr[2], qr[3], fr0, frn, frt, fr3, wnr, pr, rr);
#if HTX_EXACT_DIVISION
    double cl = sqrt(fabs(gam * pl / rl));
    double cr = sqrt(fabs(gam * pr / rr));
#else
    double cl = sqrt(fabs(gam * pl * drcp(rl)));
    double cr = sqrt(fabs(gam * pr * drcp(rr)));
#endif
    double a0 = fmax(fabs(wnl - cl), fabs(wnr - cr));
    double a1 = fmax(fabs(wnl), fabs(wnr));
    double a3 = fmax(fabs(wnl + cl), fabs(wnr + cr));
    double d0, d1, d2, d3;
    if (YDIR)
        eig_left<true>(E, qr[0] - ql[0], qr[2] - ql[2], qr[1] - ql[1], qr[3] - ql[3], d0, d1, d2, d3);
    else
        eig_left<false>(E, qr[0] - ql[0], qr[1] - ql[1], qr[2] - ql[2], qr[3] - ql[3], d0, d1, d2, d3);
    d0 *= a0; d1 *= a1; d2 *= a1; d3 *= a3;
    double r0, rn, rt, r3;
    eig_right(E, d0, d1, d2, d3, r0, rn, rt, r3);
    f[0] = 0.5 * ((fl0 + fr0) - r0);
    f[1] = 0.5 * ((fln + frn) - rn);
    f[2] = 0.5 * ((flt + frt) - rt);
    f[3] = 0.5 * ((fl3 + fr3) - r3);
}

// Equations::Euler2D::HLLC (Equations_Euler2D.cpp:372-428) in (rho, n, t, E)
// ordering.  QUIRKS keeps c = sqrt(|p/rho|) (:393-394) and p_L in the right
// star energy (:423).
template <bool QUIRKS>
HTX_HD void hllc(double gam, double g1, const double ql[4], const double qr[4],
                                     double f[4]) {
#if HTX_EXACT_DIVISION
    double wl0 = ql[0], wln = ql[1] / ql[0], wlt = ql[2] / ql[0];
    double wr0 = qr[0], wrn = qr[1] / qr[0], wrt = qr[2] / qr[0];
    double pl = (ql[3] - 0.5 * (ql[1] * wln + ql[2] * wlt)) * g1;
    double pr = (qr[3] - 0.5 * (qr[1] * wrn + qr[2] * wrt)) * g1;
    double hl = (ql[3] + pl) / wl0;
    double hr = (qr[3] + pr) / wr0;
    double rho_ratio = sqrt(fabs(wr0 / wl0));
    double cl = QUIRKS ? sqrt(fabs(pl / wl0)) : sqrt(fabs(gam * pl / wl0));
    double cr = QUIRKS ? sqrt(fabs(pr / wr0)) : sqrt(fabs(gam * pr / wr0));
#else
    double il = drcp(ql[0]), ir = drcp(qr[0]);
    double wl0 = ql[0], wln = ql[1] * il, wlt = ql[2] * il;
    double wr0 = qr[0], wrn = qr[1] * ir, wrt = qr[2] * ir;
    double pl = (ql[3] - 0.5 * (ql[1] * wln + ql[2] * wlt)) * g1;
    double pr = (qr[3] - 0.5 * (qr[1] * wrn + qr[2] * wrt)) * g1;
    double hl = (ql[3] + pl) * il;
    double hr = (qr[3] + pr) * ir;
    double rho_ratio = sqrt(fabs(wr0 * il));
    double cl = QUIRKS ? sqrt(fabs(pl * il)) : sqrt(fabs(gam * pl * il));
    double cr = QUIRKS ? sqrt(fabs(pr * ir)) : sqrt(fabs(gam * pr * ir));
#endif
    double rrr = inv(1.0 + rho_ratio);
    double vn = (wln + wrn * rho_ratio) * rrr;
    double vt = (wlt + wrt * rho_ratio) * rrr;
    double qm = 0.5 * (vn * vn + vt * vt);
    double hm = (hl + hr * rho_ratio) * rrr;
    double cm = sqrt(fabs(g1 * (hm - qm)));
    double sl = fmin(wln - cl, vn - cm);
    double sr = fmax(wrn + cr, vn + cm);
    double sm = (qr[1] * (sr - wrn) - ql[1] * (sl - wln) + pl - pr) /
                (wr0 * (sr - wrn) - wl0 * (sl - wln));
    double ps = pl + wl0 * (wln - sl) * (wln - sm);
    if (sl > 0.0) {
        f[0] = ql[1];
        f[1] = ql[1] * wln + pl;
        f[2] = ql[1] * wlt;
        f[3] = wln * (ql[3] + pl);
    } else if (sl <= 0.0 && sm > 0.0) {
        double nom = sl - wln, den = sl - sm;
#if HTX_EXACT_DIVISION
        double s0 = wl0 * nom / den;
        double sn = (nom * ql[1] + ps - pl) / den;
        double st = nom * ql[2] / den;
        double s3 = (nom * ql[3] - pl * wln + ps * sm) / den;
#else
        double id = drcp(den);
        double s0 = wl0 * nom * id;
        double sn = (nom * ql[1] + ps - pl) * id;
        double st = nom * ql[2] * id;
        double s3 = (nom * ql[3] - pl * wln + ps * sm) * id;
#endif
        f[0] = ql[1] + sl * (s0 - ql[0]);
        f[1] = (ql[1] * wln + pl) + sl * (sn - ql[1]);
        f[2] = ql[1] * wlt + sl * (st - ql[2]);
        f[3] = wln * (ql[3] + pl) + sl * (s3 - ql[3]);
    } else if (sm <= 0.0 && sr >= 0.0) {
        double nom = sr - wrn, den = sr - sm;
        double pe = QUIRKS ? pl : pr;
#if HTX_EXACT_DIVISION
        double s0 = wr0 * nom / den;
        double sn = (nom * qr[1] + ps - pr) / den;
        double st = nom * qr[2] / den;
        double s3 = (nom * qr[3] - pe * wrn + ps * sm) / den;
#else
        double id = drcp(den);
        double s0 = wr0 * nom * id;
        double sn = (nom * qr[1] + ps - pr) * id;
        double st = nom * qr[2] * id;
        double s3 = (nom * qr[3] - pe * wrn + ps * sm) * id;
#endif
        f[0] = qr[1] + sr * (s0 - qr[0]);
        f[1] = (qr[1] * wrn + pr) + sr * (sn - qr[1]);
        f[2] = qr[1] * wrt + sr * (st - qr[2]);
        f[3] = wrn * (qr[3] + pr) + sr * (s3 - qr[3]);
    } else {
        f[0] = qr[1];
        f[1] = qr[1] * wrn + pr;
        f[2] = qr[1] * wrt;
        f[3] = wrn * (qr[3] + pr);
    }
}

// ---------------------------------------------------------------------------
// FP32 bulk precision (transverse quadrature, back-projection and Riemann
// flux in FP32).  On CDNA4 a non-packed FP32 FMA issues at the FP64 rate, so
// FP32 pays through single-instruction reciprocals and square roots, half
// the registers and half the LDS bytes of the face values, not through FMA
// throughput.  Every formula is written relative to a reference value of its
// stencil so that constant data stay exact in FP32 (free streams are
// preserved to the last bit).
// ---------------------------------------------------------------------------

HTX_HD float fsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}

// FP32 weno3q: values at the Gauss points (+xi, 0, -xi) of the 5 transverse
// values (Schemes_Lagrange.cpp:1155-1217).  Smoothness and weights as
// weno3q_mixed; each candidate as u2 plus its combination of the differences
// u_j - u2 (the candidates' coefficients sum to one), each weighted sum as
// u2 plus the weighted candidate differences, the centre's split groups
// recombined as u2 + sumw0p (...) - sumw0m (...) (sumw0p - sumw0m = 1).
HTX_HD void weno3q_f32(const QuadConsts& C, float u0, float u1, float u2, float u3, float u4, float& o0, float& o1,
                       float& o2) {
    const float c1 = 13.0f / 12.0f;
    const float d1 = u1 - u0, d2 = u2 - u1, d3 = u3 - u2, d4 = u4 - u3;
    const float bg0 = 3.0f * d2 - d1, cg0 = d2 - d1;
    const float bg1 = -(d2 + d3), cg1 = d3 - d2;
    const float bg2 = d4 - 3.0f * d3, cg2 = d4 - d3;
    const float e0 = 1e-6f + (0.25f * (bg0 * bg0) + c1 * (cg0 * cg0));
    const float e1 = 1e-6f + (0.25f * (bg1 * bg1) + c1 * (cg1 * cg1));
    const float e2 = 1e-6f + (0.25f * (bg2 * bg2) + c1 * (cg2 * cg2));
    const float r0 = frcp(e0 * e0), r1 = frcp(e1 * e1), r2 = frcp(e2 * e2);
    // differences to the centre value
    const float a0 = u0 - u2, a1 = u1 - u2, a3 = u3 - u2, a4 = u4 - u2;
    const float* lw = C.lw32;
    auto mix = [&](int k, float q0, float q1, float q2) {
        const float w0 = lw[3 * k] * r0, w1 = lw[3 * k + 1] * r1, w2 = lw[3 * k + 2] * r2;
        return ((w0 * q0 + w1 * q1) + w2 * q2) * frcp((w0 + w1) + w2);
    };
    float q0 = C.gp32[0][0] * a0 + C.gp32[1][0] * a1;
    float q1 = C.gp32[0][1] * a1 + C.gp32[2][1] * a3;
    float q2 = C.gp32[1][2] * a3 + C.gp32[2][2] * a4;
    o0 = u2 + mix(0, q0, q1, q2);
    const float r24 = 1.0f / 24.0f;
    q0 = (2.0f * a1 - a0) * r24;
    q1 = -(a1 + a3) * r24;
    q2 = (2.0f * a3 - a4) * r24;
    o1 = u2 + (C.sumw0p32 * mix(1, q0, q1, q2) - C.sumw0m32 * mix(2, q0, q1, q2));
    q0 = C.gm32[0][0] * a0 + C.gm32[1][0] * a1;
    q1 = C.gm32[0][1] * a1 + C.gm32[2][1] * a3;
    q2 = C.gm32[1][2] * a3 + C.gm32[2][2] * a4;
    o2 = u2 + mix(3, q0, q1, q2);
    o0 = o0 != o0 ? 1e-15f : o0;  // catchnan
    o1 = o1 != o1 ? 1e-15f : o1;
    o2 = o2 != o2 ? 1e-15f : o2;
}

// FP32 linear limit of the quadrature (the learned path's smooth windows):
// u2 plus each Gauss point's filter applied to the differences u_j - u2
HTX_HD void weno3q_linear_f32(const QuadConsts& C, float u0, float u1, float u2, float u3, float u4, float& o0,
                              float& o1, float& o2) {
    const float a0 = u0 - u2, a1 = u1 - u2, a3 = u3 - u2, a4 = u4 - u2;
    o0 = u2 + (((C.lq32[0][0] * a0 + C.lq32[0][1] * a1) + C.lq32[0][3] * a3) + C.lq32[0][4] * a4);
    o1 = u2 + (((C.lq32[1][0] * a0 + C.lq32[1][1] * a1) + C.lq32[1][3] * a3) + C.lq32[1][4] * a4);
    o2 = u2 + (((C.lq32[2][0] * a0 + C.lq32[2][1] * a1) + C.lq32[2][3] * a3) + C.lq32[2][4] * a4);
}

// The face eigensystem in FP32 (formed in FP64 by roe_eig, rounded once)
struct Eig32 {
    float vn, vt, hm, qm, cm, rcm, b1, b2, t2;
};

HTX_HD Eig32 eig32(const Eig& E) {
    return Eig32{(float)E.vn, (float)E.vt, (float)E.hm, (float)E.qm, (float)E.cm,
                 (float)E.rcm, (float)E.b1, (float)E.b2, (float)E.t2};
}

// R c in (rho, normal, tangential, E) order, FP32
HTX_HD void eig_right_f32(const Eig32& E, float c0, float c1, float c2, float c3, float& q0, float& qn, float& qt,
                          float& q3) {
    const float tc = E.vn * E.cm;
    q0 = (c0 + c2) + c3;
    qn = ((E.vn - E.cm) * c0 + E.vn * c2) + (E.vn + E.cm) * c3;
    qt = ((E.vt * c0 + c1) + E.vt * c2) + E.vt * c3;
    q3 = (((E.hm - tc) * c0 + E.vt * c1) + E.qm * c2) + (E.hm + tc) * c3;
}

// L q (normal, tangential momenta order), FP32
HTX_HD void eig_left_nt_f32(const Eig32& E, float q0, float qn, float qt, float q3, float& c0, float& c1, float& c2,
                            float& c3) {
    const float tr = E.vn * E.rcm, tb = E.b1 * E.vn;
    const float ln0 = -0.5f * (tb + E.rcm), lt0 = -E.t2 * E.vt, ln3 = -0.5f * (tb - E.rcm), lt2 = E.b1 * E.vt;
    c0 = ((0.5f * (E.b2 + tr) * q0 + ln0 * qn) + lt0 * qt) + E.t2 * q3;
    c1 = -E.vt * q0 + qt;
    c2 = (((1.0f - E.b2) * q0 + tb * qn) + lt2 * qt) + (-E.b1) * q3;
    c3 = ((0.5f * (E.b2 - tr) * q0 + ln3 * qn) + lt0 * qt) + E.t2 * q3;
}

// An FP32 back-projected state (rho, m_n, m_t, E) from characteristic values
// c is usable when its density is positive and well above the cancellation
// of the back-projection (rho > 2^-10 (|c0| + |c2| + |c3|)) and its pressure
// is positive and at least 2^-10 of its energy (rho E - m^2 / 2 > 2^-10
// rho E); else the FP32 bulk paths recompute in FP64.
HTX_HD bool state_fp32_ok(const float q[4], float c0, float c2, float c3) {
    const float re = q[3] * q[0];
    return q[0] > 0x1p-10f * (fabsf(c0) + fabsf(c2) + fabsf(c3)) &&
           re - 0.5f * (q[1] * q[1] + q[2] * q[2]) > 0x1p-10f * re;
}

// density and pressure (gamma - 1)(E - m^2 / 2 rho) of a state positive,
// tested as rho E - m^2 / 2 > 0 (same sign for rho > 0)
HTX_HD bool state_positive_f32(const float q[4]) {
    return q[0] > 0.0f && q[3] * q[0] - 0.5f * (q[1] * q[1] + q[2] * q[2]) > 0.0f;
}

// HLLC (Equations_Euler2D.cpp:372-428) in FP32, (rho, n, t, E) order; QUIRKS
// as hllc<>
template <bool QUIRKS>
HTX_HD void hllc_f32(float gam, float g1, const float ql[4], const float qr[4], float f[4]) {
    const float il = frcp(ql[0]), ir = frcp(qr[0]);
    const float wl0 = ql[0], wln = ql[1] * il, wlt = ql[2] * il;
    const float wr0 = qr[0], wrn = qr[1] * ir, wrt = qr[2] * ir;
    const float pl = (ql[3] - 0.5f * (ql[1] * wln + ql[2] * wlt)) * g1;
    const float pr = (qr[3] - 0.5f * (qr[1] * wrn + qr[2] * wrt)) * g1;
    const float hl = (ql[3] + pl) * il, hr = (qr[3] + pr) * ir;
    const float rho_ratio = fsqrt(fabsf(wr0 * il));
    const float cl = QUIRKS ? fsqrt(fabsf(pl * il)) : fsqrt(fabsf(gam * pl * il));
    const float cr = QUIRKS ? fsqrt(fabsf(pr * ir)) : fsqrt(fabsf(gam * pr * ir));
    const float rrr = frcp(1.0f + rho_ratio);
    const float vn = (wln + wrn * rho_ratio) * rrr, vt = (wlt + wrt * rho_ratio) * rrr;
    const float qm = 0.5f * (vn * vn + vt * vt);
    const float hm = (hl + hr * rho_ratio) * rrr;
    const float cm = fsqrt(fabsf(g1 * (hm - qm)));
    const float sl = fminf(wln - cl, vn - cm), sr = fmaxf(wrn + cr, vn + cm);
    const float sm = (qr[1] * (sr - wrn) - ql[1] * (sl - wln) + pl - pr) * frcp(wr0 * (sr - wrn) - wl0 * (sl - wln));
    const float ps = pl + wl0 * (wln - sl) * (wln - sm);
    if (sl > 0.0f) {
        f[0] = ql[1];
        f[1] = ql[1] * wln + pl;
        f[2] = ql[1] * wlt;
        f[3] = wln * (ql[3] + pl);
    } else if (sl <= 0.0f && sm > 0.0f) {
        const float nom = sl - wln, id = frcp(sl - sm);
        f[0] = ql[1] + sl * (wl0 * nom * id - ql[0]);
        f[1] = (ql[1] * wln + pl) + sl * ((nom * ql[1] + ps - pl) * id - ql[1]);
        f[2] = ql[1] * wlt + sl * (nom * ql[2] * id - ql[2]);
        f[3] = wln * (ql[3] + pl) + sl * ((nom * ql[3] - pl * wln + ps * sm) * id - ql[3]);
    } else if (sm <= 0.0f && sr >= 0.0f) {
        const float nom = sr - wrn, id = frcp(sr - sm), pe = QUIRKS ? pl : pr;
        f[0] = qr[1] + sr * (wr0 * nom * id - qr[0]);
        f[1] = (qr[1] * wrn + pr) + sr * ((nom * qr[1] + ps - pr) * id - qr[1]);
        f[2] = qr[1] * wrt + sr * (nom * qr[2] * id - qr[2]);
        f[3] = wrn * (qr[3] + pr) + sr * ((nom * qr[3] - pe * wrn + ps * sm) * id - qr[3]);
    } else {
        f[0] = qr[1];
        f[1] = qr[1] * wrn + pr;
        f[2] = qr[1] * wrt;
        f[3] = wrn * (qr[3] + pr);
    }
}

// characteristic Rusanov (Equations_Euler2D.cpp:221-246, 310-330) in FP32,
// (rho, n, t, E) order
HTX_HD void rusanov_char_f32(float gam, float g1, const Eig32& E, const float ql[4], const float qr[4], float f[4]) {
    const float il = frcp(ql[0]), ir = frcp(qr[0]);
    const float wnl = ql[1] * il, wtl = ql[2] * il, wnr = qr[1] * ir, wtr = qr[2] * ir;
    const float pl = (ql[3] - 0.5f * (ql[1] * wnl + ql[2] * wtl)) * g1;
    const float pr = (qr[3] - 0.5f * (qr[1] * wnr + qr[2] * wtr)) * g1;
    const float cl = fsqrt(fabsf(gam * pl * il)), cr = fsqrt(fabsf(gam * pr * ir));
    const float a0 = fmaxf(fabsf(wnl - cl), fabsf(wnr - cr));
    const float a1 = fmaxf(fabsf(wnl), fabsf(wnr));
    const float a3 = fmaxf(fabsf(wnl + cl), fabsf(wnr + cr));
    float d0, d1, d2, d3;
    eig_left_nt_f32(E, qr[0] - ql[0], qr[1] - ql[1], qr[2] - ql[2], qr[3] - ql[3], d0, d1, d2, d3);
    d0 *= a0;
    d1 *= a1;
    d2 *= a1;
    d3 *= a3;
    float r0, rn, rt, r3;
    eig_right_f32(E, d0, d1, d2, d3, r0, rn, rt, r3);
    f[0] = 0.5f * ((ql[1] + qr[1]) - r0);
    f[1] = 0.5f * (((ql[1] * wnl + pl) + (qr[1] * wnr + pr)) - rn);
    f[2] = 0.5f * ((ql[1] * wtl + qr[1] * wtr) - rt);
    f[3] = 0.5f * ((wnl * (ql[3] + pl) + wnr * (qr[3] + pr)) - r3);
}

// Euler2D::maxSpeed (Equations_Euler2D.cpp:207-219); QUIRKS drops the fabs
// on u - c (:218).  Negative/NaN results are clamped to 0, which is what the
// reference's fmax(0, .) accumulation does with them.
template <bool QUIRKS>
HTX_HD double max_speed(double gam, double wn, double p, double rho) {
#if HTX_EXACT_DIVISION
    double c = sqrt(fabs(gam * p / rho));
#else
    double c = sqrt(fabs(gam * p * drcp(rho)));
#endif
    double s = QUIRKS ? fmax(wn - c, fabs(wn + c)) : fmax(fabs(wn - c), fabs(wn + c));
    return s > 0.0 ? s : 0.0;
}

// The 1D solver's reconstruction alternatives, also selectable for the 2D
// TENO5 path (HTX_SCHEME_TENO5I / WENO5JS / WENO5Z).
// Lagrange::teno5i (Schemes_Lagrange.cpp:1235-1263): teno5 with tau the
// largest smoothness measure of the three candidates
HTX_HD double teno5i_side(double v0, double v1, double v2, double v3, double v4, double p0, double p1, double p2,
                          double d0, double d1, double d2) {
    HTX_NO_CONTRACT
    const double c1 = 13.0 / 12.0;
    double c_0 = v2 - 2.0 * v3 + v4, b_0 = (3.0 * v2 - 4.0 * v3 + v4) / 2.0;
    double c_1 = v1 - 2.0 * v2 + v3, b_1 = (v1 - v3) / 2.0;
    double c_2 = v0 - 2.0 * v1 + v2, b_2 = (v0 - 4.0 * v1 + 3.0 * v2) / 2.0;
    double s0 = b_0 * b_0 + c1 * c_0 * c_0;
    double s1 = b_1 * b_1 + c1 * c_1 * c_1;
    double s2 = b_2 * b_2 + c1 * c_2 * c_2;
    double tau = s0;
    if (s1 > tau) tau = s1;
    if (s2 > tau) tau = s2;
    double g0 = 1.0 + tau / (s0 + 1e-40);
    double g1 = 1.0 + tau / (s1 + 1e-40);
    double g2 = 1.0 + tau / (s2 + 1e-40);
    g0 = g0 * g0 * g0; g0 = g0 * g0;
    g1 = g1 * g1 * g1; g1 = g1 * g1;
    g2 = g2 * g2 * g2; g2 = g2 * g2;
    double gs = (g0 + g1) + g2;
    bool k0 = !(g0 / gs < 5e-4), k1 = !(g1 / gs < 5e-4), k2 = !(g2 / gs < 5e-4);
    double w0 = k0 ? d0 : 0.0, w1 = k1 ? d1 : 0.0, w2 = k2 ? d2 : 0.0;
    double ws = (w0 + w1) + w2;
    w0 = w0 / ws; w1 = w1 / ws; w2 = w2 / ws;
    double out = (p0 * w0 + p1 * w1) + p2 * w2;
    return out != out ? 1e-15 : out;
}

HTX_HD void teno5i(const double u[6], double& um, double& up) {
    HTX_NO_CONTRACT
    const double a = 1.0 / 3.0, b = 5.0 / 6.0, c = -1.0 / 6.0, d = 11.0 / 6.0, e = -7.0 / 6.0;
    double p0 = (a * u[2] + b * u[3]) + c * u[4];
    double p1 = (c * u[1] + b * u[2]) + a * u[3];
    double p2 = (a * u[0] + e * u[1]) + d * u[2];
    um = teno5i_side(u[0], u[1], u[2], u[3], u[4], p0, p1, p2, 0.3, 0.6, 0.1);
    double q0 = (d * u[3] + e * u[4]) + a * u[5];
    double q1 = (a * u[2] + b * u[3]) + c * u[4];
    double q2 = (c * u[1] + b * u[2]) + a * u[3];
    up = teno5i_side(u[1], u[2], u[3], u[4], u[5], q0, q1, q2, 0.1, 0.6, 0.3);
}

// One side of Lagrange::weno5JS (Schemes_Lagrange.cpp:358-381; weights
// d_k / (1e-6 + SI_k)^2) or weno5Z (:383-408; weights d_k (1 + tau / (SI_k
// + 1e-40)), tau = |SI_2 - SI_0|), normalised, on the window v0..v4 with
// candidates p0..p2; the smoothness measures of teno5_side
// (Schemes_Lagrange.cpp:1013-1026).  A NaN result becomes 1e-15 (catchnan).
HTX_HD double weno5_side(double v0, double v1, double v2, double v3, double v4, double p0, double p1, double p2,
                         double d0, double d1, double d2, bool z) {
    HTX_NO_CONTRACT
    const double c1 = 13.0 / 12.0;
    double c_0 = v2 - 2.0 * v3 + v4, b_0 = (3.0 * v2 - 4.0 * v3 + v4) / 2.0;
    double c_1 = v1 - 2.0 * v2 + v3, b_1 = (v1 - v3) / 2.0;
    double c_2 = v0 - 2.0 * v1 + v2, b_2 = (v0 - 4.0 * v1 + 3.0 * v2) / 2.0;
    double s0 = b_0 * b_0 + c1 * c_0 * c_0;
    double s1 = b_1 * b_1 + c1 * c_1 * c_1;
    double s2 = b_2 * b_2 + c1 * c_2 * c_2;
    double w0, w1, w2;
    if (z) {
        const double tau = fabs(s2 - s0);
        w0 = d0 * (1.0 + tau / (s0 + 1e-40));
        w1 = d1 * (1.0 + tau / (s1 + 1e-40));
        w2 = d2 * (1.0 + tau / (s2 + 1e-40));
    } else {
        const double e0 = 1e-6 + s0, e1 = 1e-6 + s1, e2 = 1e-6 + s2;
        w0 = d0 / (e0 * e0);
        w1 = d1 / (e1 * e1);
        w2 = d2 / (e2 * e2);
    }
    const double sum = (w0 + w1) + w2;
    w0 = w0 / sum;
    w1 = w1 / sum;
    w2 = w2 / sum;
    const double r = (p0 * w0 + p1 * w1) + p2 * w2;
    return r != r ? 1e-15 : r;
}

// Lagrange::weno5JS (z false) / weno5Z on u_{i-2..i+3}: (u-, u+) at x_{i+1/2},
// with the candidates and linear weights of teno5.
HTX_HD void weno5(const double u[6], bool z, double& um, double& up) {
    HTX_NO_CONTRACT
    const double a = 1.0 / 3.0, b = 5.0 / 6.0, c = -1.0 / 6.0, d = 11.0 / 6.0, e = -7.0 / 6.0;
    const double p0 = (a * u[2] + b * u[3]) + c * u[4];
    const double p1 = (c * u[1] + b * u[2]) + a * u[3];
    const double p2 = (a * u[0] + e * u[1]) + d * u[2];
    um = weno5_side(u[0], u[1], u[2], u[3], u[4], p0, p1, p2, 0.3, 0.6, 0.1, z);
    const double q0 = (d * u[3] + e * u[4]) + a * u[5];
    const double q1 = (a * u[2] + b * u[3]) + c * u[4];
    const double q2 = (c * u[1] + b * u[2]) + a * u[3];
    up = weno5_side(u[1], u[2], u[3], u[4], u[5], q0, q1, q2, 0.1, 0.6, 0.3, z);
}

}  // namespace HTX_NS
