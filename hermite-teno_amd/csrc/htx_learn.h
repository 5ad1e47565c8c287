// htx_learn.h -- the learned troubled-face classifier: face features, the
// TENO cut-off label it is trained on, the MLP's parameter layout and
// activation, and the MFMA inference of one wavefront's 64 faces.
//
// A face sits between cells c and c+1 along direction n (x or y); its
// stencil is the six cells c-2..c+3 that Schemes::Lagrange::teno5 reads
// (Schemes_Lagrange.cpp:481-511).  There is no reference counterpart to the
// classifier itself: it predicts whether plain TENO5 would keep every
// candidate stencil on both sides of the face for all four characteristic
// fields (cutOff, Schemes_Lagrange.cpp:1141-1149), in which case TENO5 is the
// optimal-weight linear scheme (Lagrange::linear, :1219-1233) and the learned
// path uses that instead.
#pragma once
#include "htx_internal.h"

namespace HTX_NS {

constexpr int CLS_IN = 12;
constexpr int CLS_H = 64;
// FP32 parameters, logical layout: W1[CLS_IN][CLS_H] (row = input), b1[CLS_H],
// W2[CLS_H][CLS_H] (row = input unit), b2[CLS_H], w3[CLS_H], b3
constexpr int CLS_W1 = 0;
constexpr int CLS_B1 = CLS_W1 + CLS_IN * CLS_H;
constexpr int CLS_W2 = CLS_B1 + CLS_H;
constexpr int CLS_B2 = CLS_W2 + CLS_H * CLS_H;
constexpr int CLS_W3 = CLS_B2 + CLS_H;
constexpr int CLS_B3 = CLS_W3 + CLS_H;
constexpr int CLS_NPARAM = CLS_B3 + 1;  // 5057

// GELU, x * Phi(x), with the normal CDF Phi replaced by the cubic
// 0.5 + t (A + B t^2) of t = clamp(x, -X0, X0): it has Phi's value and slope
// at 0 (A = 1/sqrt(2 pi)) and reaches 0 and 1 with zero slope at -X0 and X0
// (X0 = 0.75 sqrt(2 pi), B = -1/(4 X0^3)), so the activation is C1; it
// differs from Phi by at most 0.039.  Five vector instructions and no
// transcendental; training and inference use the same function.
constexpr float GELU_A = 0.3989422804014327f;
constexpr float GELU_B = -0.037625858331402064f;
constexpr float GELU_X0 = 1.8799712059732503f;

HTX_HD float gelu_f(float x) {
    float t = fminf(fmaxf(x, -GELU_X0), GELU_X0);
    return x * (0.5f + t * (GELU_A + GELU_B * t * t));
}

// (gelu(x), gelu'(x))
// gelu_f with the polynomial in fused multiply-adds, for the matrix-core
// inference (differs from gelu_f in the last bit at most)
__device__ __forceinline__ float gelu_fma(float x) {
    const float t = fminf(fmaxf(x, -GELU_X0), GELU_X0);
    return x * fmaf(t, fmaf(GELU_B, t * t, GELU_A), 0.5f);
}

HTX_HD void gelu_fd(float x, float& g, float& dg) {
    float t = fminf(fmaxf(x, -GELU_X0), GELU_X0);
    float phi = 0.5f + t * (GELU_A + GELU_B * t * t);
    g = x * phi;
    bool in = x > -GELU_X0 && x < GELU_X0;
    dg = phi + (in ? x * (GELU_A + 3.0f * GELU_B * t * t) : 0.0f);
}

// TENO contrast of one 5-value window given its first differences
// e_k = v_k - v_{k-1} (k = 1..4).  With TENO's smoothness measures beta_k of
// the three 3-cell candidates and tau = |beta5 - (beta0 + beta2 + 4 beta1)/6|
// (Schemes_Lagrange.cpp:1013-1034, beta5 in the difference form of
// teno5_side_mixed), the weights (1 + tau/(beta_k + eps))^6 of the best and the
// worst candidate differ by the ratio R = (1 + tau/(bmin+eps)) /
// (1 + tau/(bmax+eps)) >= 1.  Returns 1 - 1/R in [0, 1): 0 on a polynomial of
// degree <= 2 (tau = 0), near 1 across a jump.  TENO5 cuts a candidate roughly
// when R^6 exceeds 1 / (3 C_T), i.e. for 1 - 1/R above about 0.66.  FP32.
HTX_HD float teno_contrast(float e1, float e2, float e3, float e4, float eps) {
    // no contraction: the features must come out the same in the training
    // extraction and in the stage kernel's fused inference, whatever
    // contraction the including file allows
    HTX_NO_CONTRACT
    const float c1 = 13.0f / 12.0f;
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
    float lo = fminf(fminf(s0, s1), s2) + eps, hi = fmaxf(fmaxf(s0, s1), s2) + eps;
    // 1/R = (1 + tau/hi) / (1 + tau/lo), kept free of overflow: when tau/hi
    // overflows every candidate is flat against tau (R = 1); when only tau/lo
    // does, 1/R = 0
    float x = tau * frcp(hi), y = tau * frcp(lo);
    // x >= 0; not finite (infinite or NaN) unless x <= FLT_MAX.  (A test of
    // x - x would be folded with the product above into one fused
    // multiply-add, which returns the product's rounding error, under
    // contraction.)
    float inv = !(x <= 3.40282347e38f) ? 1.0f : (1.0f + x) * frcp(1.0f + y);
    float r = 1.0f - inv;
    return r > 0.0f ? r : 0.0f;
}

// Contrast of a 6-cell stencil from its five differences: the larger of its
// two 5-cell windows (the windows of u^- and u^+ at the face).
HTX_HD float stencil_contrast(const float d[5], float eps) {
    return fmaxf(teno_contrast(d[0], d[1], d[2], d[3], eps), teno_contrast(d[1], d[2], d[3], d[4], eps));
}

HTX_HD float log_amp(float amp) {
    HTX_NO_CONTRACT
    return flog2(amp + 1e-12f) * (0.30102999566398120f / 12.0f) + 1.0f;
}

// The 12 classifier inputs of a face from the primitive values of its
// stencil cells c-2..c+3 (r = density, un / ut = velocity normal / tangential
// to the face, p = pressure).  Differences are formed in FP64, everything
// after them in FP32.  Scales: rf, pf = face averages of |r|, |p|,
// cf = sqrt(gamma pf / rf).
//   x[0..1]   density differences r_c - r_{c-1}, r_{c+1} - r_c over
//             (sum of the five |differences| + 1e-6 rf)
//   x[2..3]   the same for the pressure
//   x[4..7]   log10(max |difference| / scale + 1e-12) / 12 + 1 for r, p, un,
//             ut (scales rf, pf, cf, cf): 0 for a flat stencil, about 1 for
//             an O(1) jump
//   x[8..11]  TENO contrasts of r, p, un, ut (eps = 1e-12 scale^2)
// floor > 0 (states advanced in FP32 bulk arithmetic, whose rounding leaves
// noise of about 1e-7 of the state between neighbouring cells): a
// difference within floor x its scale counts as zero, so that a stencil that
// is flat up to that noise has the features of a flat one.
HTX_HD void face_features(const double r[6], const double un[6], const double ut[6], const double p[6],
                          double gam, float x[CLS_IN], float floor = 0.0f) {
    HTX_NO_CONTRACT
    float dr[5], dp[5], du[5], dv[5];
    float sr = 0.0f, sp = 0.0f, ar = 0.0f, ap = 0.0f, au = 0.0f, at = 0.0f;
    float zr = 0.0f, zp = 0.0f, zc = 0.0f;
    if (floor > 0.0f) {
        const float fr = (float)(0.5 * (fabs(r[2]) + fabs(r[3]))), fp = (float)(0.5 * (fabs(p[2]) + fabs(p[3])));
        zr = floor * fr;
        zp = floor * fp;
        zc = floor * sqrtf((float)gam * fp * frcp(fr + 1e-30f));
    }
    for (int k = 0; k < 5; ++k) {
        dr[k] = (float)(r[k + 1] - r[k]);
        dp[k] = (float)(p[k + 1] - p[k]);
        du[k] = (float)(un[k + 1] - un[k]);
        dv[k] = (float)(ut[k + 1] - ut[k]);
        if (floor > 0.0f) {
            dr[k] = fabsf(dr[k]) <= zr ? 0.0f : dr[k];
            dp[k] = fabsf(dp[k]) <= zp ? 0.0f : dp[k];
            du[k] = fabsf(du[k]) <= zc ? 0.0f : du[k];
            dv[k] = fabsf(dv[k]) <= zc ? 0.0f : dv[k];
        }
        sr += fabsf(dr[k]);
        sp += fabsf(dp[k]);
        ar = fmaxf(ar, fabsf(dr[k]));
        ap = fmaxf(ap, fabsf(dp[k]));
        au = fmaxf(au, fabsf(du[k]));
        at = fmaxf(at, fabsf(dv[k]));
    }
    float rf = (float)(0.5 * (fabs(r[2]) + fabs(r[3]))), pf = (float)(0.5 * (fabs(p[2]) + fabs(p[3])));
    float irf = frcp(rf + 1e-30f), ipf = frcp(pf + 1e-30f);
    float cf2 = (float)gam * pf * irf;
    float icf = frcp(sqrtf(cf2) + 1e-30f);
    float nr = frcp(sr + 1e-6f * rf + 1e-30f), np = frcp(sp + 1e-6f * pf + 1e-30f);
    x[0] = dr[1] * nr;
    x[1] = dr[2] * nr;
    x[2] = dp[1] * np;
    x[3] = dp[2] * np;
    x[4] = log_amp(ar * irf);
    x[5] = log_amp(ap * ipf);
    x[6] = log_amp(au * icf);
    x[7] = log_amp(at * icf);
    x[8] = stencil_contrast(dr, 1e-12f * rf * rf + 1e-30f);
    x[9] = stencil_contrast(dp, 1e-12f * pf * pf + 1e-30f);
    float ev = 1e-12f * cf2 + 1e-30f;
    x[10] = stencil_contrast(du, ev);
    x[11] = stencil_contrast(dv, ev);
}

// Primitive (r, un, ut, p) of a conserved state (rho, m_n, m_t, E).
HTX_HD void prim_nt(double g1, double q0, double qn, double qt, double q3, double& r, double& un, double& ut,
                    double& p) {
    r = q0;
    un = qn / q0;
    ut = qt / q0;
    p = (q3 - 0.5 * (qn * un + qt * ut)) * g1;
}

// Training label of a face: 1 when plain FP64 TENO5 and the linear scheme give
// the same face values on both sides for all four characteristic fields, to
// within 1e-12 M + 1e-9 sum_k |u_{k+1} - u_k| of the field's own stencil, M =
// the largest |value| of any field on the stencil; else 0.  That holds when
// TENO5 keeps every candidate, and also where it cuts one only because the
// data are flat up to rounding (a smoothness measure left a few ulps above
// zero, or rounding noise on a field that is zero), where all candidates
// coincide to rounding.  c[f][m] = characteristic values of field f on the
// six stencil cells (FV_System2D.cpp:372-400).
HTX_HD int teno_label(const double c[4][6]) {
    double mx = 0.0;
    for (int f = 0; f < 4; ++f)
        for (int m = 0; m < 6; ++m) mx = fmax(mx, fabs(c[f][m]));
    for (int f = 0; f < 4; ++f) {
        double tm, tp, lm, lp, tv = 0.0;
        teno5(c[f], tm, tp);
        linear5(c[f], lm, lp);
        for (int m = 0; m < 5; ++m) tv += fabs(c[f][m + 1] - c[f][m]);
        double tol = 1e-12 * mx + 1e-9 * tv;
        if (!(fabs(tm - lm) <= tol && fabs(tp - lp) <= tol)) return 0;
    }
    return 1;
}

// ---------------------------------------------------------------------------
// MFMA inference.  The bf16 weights are packed on the host into per-lane
// fragments of the two matrix layers, computed in transposed form
// (Z^T = W^T X^T) so that a layer's output fragment is the next layer's B
// operand without any exchange between lanes:
//   layer 1, v_mfma_f32_16x16x16_bf16, per 16-unit block mb = 0..3:
//     A lane (i = lane % 16, g = lane / 16) holds W1[4g + c][16 mb + i], c = 0..3
//     (zero for inputs >= 12); B lane holds features 4g..4g+3 of face i;
//     D lane holds units 16 mb + 4g + c of face i
//   layer 2, v_mfma_f32_16x16x32_bf16, per (mb, kb = 0..1):
//     K slot 8g + c' of block kb is hidden unit 16 (2 kb + c'/4) + 4g + c'%4,
//     which is where layer 1's outputs of blocks 2kb and 2kb+1 sit, so
//     A lane holds W2[that unit][16 mb + i] for c' = 0..7
//   layer 3 (64 -> 1) on the vector ALU, summed over the four lane groups.
// Fragment buffer: frag1 (4 x 64 x 4 bf16) then frag2 (4 x 2 x 64 x 8 bf16);
// cvec = b1[64], b2[64], w3[64], b3 in FP32.
constexpr int CLS_FRAG1 = 4 * 64 * 4;
constexpr int CLS_FRAG2 = 4 * 2 * 64 * 8;
constexpr int CLS_NFRAG = CLS_FRAG1 + CLS_FRAG2;  // bf16 values
constexpr int CLS_NCVEC = 3 * CLS_H + 1;

// logical hidden unit of layer 2's K slot (kb, g, c')
HTX_HD int cls_kslot_unit(int kb, int g, int cp) { return 16 * (2 * kb + cp / 4) + 4 * g + (cp % 4); }

// bf16 round-to-nearest-even of an FP32 value (NaN kept quiet)
HTX_HD unsigned short f32_to_bf16(float f) {
    unsigned int u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

HTX_HD float bf16_to_f32(unsigned short h) {
    unsigned int u = (unsigned int)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// Host-side packing of the FP32 parameters into MFMA fragments.
inline void cls_pack(const float* prm, unsigned short* frag, float* cvec) {
    for (int mb = 0; mb < 4; ++mb)
        for (int lane = 0; lane < 64; ++lane) {
            int i = lane % 16, g = lane / 16;
            for (int c = 0; c < 4; ++c) {
                int k = 4 * g + c;
                float w = k < CLS_IN ? prm[CLS_W1 + k * CLS_H + 16 * mb + i] : 0.0f;
                frag[(mb * 64 + lane) * 4 + c] = f32_to_bf16(w);
            }
        }
    unsigned short* f2 = frag + CLS_FRAG1;
    for (int mb = 0; mb < 4; ++mb)
        for (int kb = 0; kb < 2; ++kb)
            for (int lane = 0; lane < 64; ++lane) {
                int i = lane % 16, g = lane / 16;
                for (int cp = 0; cp < 8; ++cp) {
                    int k = cls_kslot_unit(kb, g, cp);
                    f2[((mb * 2 + kb) * 64 + lane) * 8 + cp] = f32_to_bf16(prm[CLS_W2 + k * CLS_H + 16 * mb + i]);
                }
            }
    for (int u = 0; u < CLS_H; ++u) {
        cvec[u] = prm[CLS_B1 + u];
        cvec[CLS_H + u] = prm[CLS_B2 + u];
        cvec[2 * CLS_H + u] = prm[CLS_W3 + u];
    }
    cvec[3 * CLS_H] = prm[CLS_B3];
}

typedef short cls_s4 __attribute__((ext_vector_type(4)));
typedef __bf16 cls_b8 __attribute__((ext_vector_type(8)));
typedef float cls_f4 __attribute__((ext_vector_type(4)));

// Logit of the calling lane's face for the 64 faces of one wavefront.
// xs: this wave's LDS scratch of 64 x 16 bf16 (2 KB); every lane of the
// wave must call (faces without data pass zeros).  The caller orders the
// scratch writes before the reads (a barrier between cls_stage_features and
// cls_logit).
__device__ __forceinline__ void cls_stage_features(unsigned short* xs, const float x[CLS_IN]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 16; ++k) xs[lane * 16 + k] = k < CLS_IN ? f32_to_bf16(x[k]) : (unsigned short)0;
}

__device__ __forceinline__ float cls_logit(const unsigned short* xs, const unsigned short* frag, const float* cvec) {
#if defined(__HIP_DEVICE_COMPILE__)
    const int lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
    const cls_s4* F1 = reinterpret_cast<const cls_s4*>(frag);
    const cls_b8* F2 = reinterpret_cast<const cls_b8*>(frag + CLS_FRAG1);
    float zmine = 0.0f;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        cls_s4 bx = *reinterpret_cast<const cls_s4*>(xs + (16 * nb + i) * 16 + 4 * g);
        cls_b8 b2[2];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            // the biases start the accumulators
            const float* b1 = cvec + 16 * mb + 4 * g;
            cls_f4 d = {b1[0], b1[1], b1[2], b1[3]};
            d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(F1[mb * 64 + lane], bx, d, 0, 0, 0);
#pragma unroll
            for (int c = 0; c < 4; ++c) b2[mb >> 1][(mb & 1) * 4 + c] = (__bf16)gelu_fma(d[c]);
        }
        float part = 0.0f;
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const float* bb = cvec + CLS_H + 16 * mb + 4 * g;
            cls_f4 d = {bb[0], bb[1], bb[2], bb[3]};
            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F2[(mb * 2 + 0) * 64 + lane], b2[0], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F2[(mb * 2 + 1) * 64 + lane], b2[1], d, 0, 0, 0);
#pragma unroll
            for (int c = 0; c < 4; ++c) part = fmaf(cvec[2 * CLS_H + 16 * mb + 4 * g + c], gelu_fma(d[c]), part);
        }
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        if (nb == g) zmine = part + cvec[3 * CLS_H];
    }
    return zmine;
#else
    return 0.0f;
#endif
}


}  // namespace HTX_NS
