// htx_learn.hip -- training-set extraction, training and standalone
// inference of the learned troubled-face classifier (htx_learn.h).
//
// The labels come from plain FP64 TENO5 cut-off decisions, so this file is
// compiled like the strict drop-in build (no contraction, divisions kept, in
// its own namespace): a label is exactly the decision the drop-in solver
// makes on that face.
//
// Training is FP32, deterministic: each workgroup accumulates the gradient of
// its share of a minibatch in registers and writes it to its own row of a
// partial-gradient buffer; one update kernel sums the rows in a fixed order
// and applies Adam.
#define HTX_NS htx_lrn
#define HTX_EXACT_DIVISION 1
#include "htx_learn.h"
#include "htx_learn_api.h"

namespace htx_lrn {

namespace {

struct Stencil6 {
    double q[6][4];  // conserved state of cells c-2..c+3
};

template <bool YDIR>
__device__ void load_stencil(const double* q, int nx, int ny, int pitch, long long fs, int bc, int i, int j,
                             Stencil6& st) {
    for (int m = 0; m < 6; ++m) {
        int gi = YDIR ? map_index(i, nx, bc) : map_index(i - 2 + m, nx, bc);
        int gj = YDIR ? map_index(j - 2 + m, ny, bc) : map_index(j, ny, bc);
        long long off = (long long)gj * pitch + gi;
        for (int e = 0; e < 4; ++e) st.q[m][e] = q[off + e * fs];
    }
}

template <bool YDIR>
__device__ void stencil_features(const Stencil6& st, double gam, float x[CLS_IN]) {
    double r[6], un[6], ut[6], p[6];
    for (int m = 0; m < 6; ++m)
        prim_nt(gam - 1.0, st.q[m][0], YDIR ? st.q[m][2] : st.q[m][1], YDIR ? st.q[m][1] : st.q[m][2], st.q[m][3],
                r[m], un[m], ut[m], p[m]);
    face_features(r, un, ut, p, gam, x);
}

template <bool YDIR>
__device__ int stencil_label(const Stencil6& st, double g1) {
    const double* a = st.q[2];
    const double* b = st.q[3];
    Eig E = roe_eig(g1, YDIR, cell_prim(g1, a[0], a[1], a[2], a[3]), cell_prim(g1, b[0], b[1], b[2], b[3]));
    double c[4][6];
    for (int m = 0; m < 6; ++m)
        eig_left<YDIR>(E, st.q[m][0], st.q[m][1], st.q[m][2], st.q[m][3], c[0][m], c[1][m], c[2][m], c[3][m]);
    return teno_label(c);
}

// one work item per face (cell (i, j) and its +1 neighbour along the direction)
template <bool YDIR>
__global__ void __launch_bounds__(256) k_extract(LrnExtractArgs a) {
    __shared__ unsigned int red[4];
    long long n = (long long)a.nx * a.ny;
    long long id = blockIdx.x * 256ll + threadIdx.x;
    unsigned int smooth = 0;
    if (id < n) {
        int i = (int)(id % a.nx), j = (int)(id / a.nx);
        Stencil6 st;
        load_stencil<YDIR>(a.q, a.nx, a.ny, a.pitch, a.fstride, a.bc, i, j, st);
        float x[CLS_IN];
        stencil_features<YDIR>(st, a.gam, x);
        long long o = a.base + id;
        // a face of a non-finite state enters the set as troubled with zero
        // features rather than poisoning the training with NaN
        bool finite = true;
        for (int k = 0; k < CLS_IN; ++k) finite = finite && fabsf(x[k]) <= 3.40282347e38f;
        for (int k = 0; k < CLS_IN; ++k) a.feat[o * CLS_IN + k] = finite ? x[k] : 0.0f;
        int lab = finite ? stencil_label<YDIR>(st, a.gam - 1.0) : 0;
        a.label[o] = (unsigned char)lab;
        smooth = (unsigned)lab;
    }
    for (int off = 32; off > 0; off >>= 1) smooth += __shfl_xor(smooth, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = smooth;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(a.nsmooth, (unsigned long long)(red[0] + red[1] + red[2] + red[3]));
}

// ---------------------------------------------------------------------------
// training: chunks of 32 samples per workgroup iteration
constexpr int TS = 32;     // samples per chunk
constexpr int RS = 65;     // padded row stride of the activation buffers

__device__ __forceinline__ float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ long long sample_index(const LrnTrainArgs& a, long long k) {
    return (long long)((a.pa * (unsigned long long)k + a.pb) % (unsigned long long)a.n);
}

__global__ void __launch_bounds__(256) k_train_grad(LrnTrainArgs a) {
    __shared__ float sW1[CLS_IN * CLS_H], sB1[CLS_H], sW2[CLS_H * CLS_H], sB2[CLS_H], sW3[CLS_H], sB3;
    __shared__ float sX[TS * CLS_IN], sZ1[TS * RS], sH1[TS * RS], sZ2[TS * RS], sH2[TS * RS];
    __shared__ float sG3[TS], sY[TS], sWt[TS], sL[TS];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int p = tid; p < CLS_IN * CLS_H; p += 256) sW1[p] = a.prm[CLS_W1 + p];
    for (int p = tid; p < CLS_H * CLS_H; p += 256) sW2[p] = a.prm[CLS_W2 + p];
    if (tid < CLS_H) {
        sB1[tid] = a.prm[CLS_B1 + tid];
        sB2[tid] = a.prm[CLS_B2 + tid];
        sW3[tid] = a.prm[CLS_W3 + tid];
    }
    if (tid == 0) sB3 = a.prm[CLS_B3];
    float gW2[16], gW1[3], gv = 0.0f, lsum = 0.0f;
    for (int r = 0; r < 16; ++r) gW2[r] = 0.0f;
    for (int r = 0; r < 3; ++r) gW1[r] = 0.0f;
    __syncthreads();

    for (int ch = 0; ch < a.chunks; ++ch) {
        long long kbase = a.k0 + ((long long)blockIdx.x * a.chunks + ch) * TS;
        // samples, labels and weights of the chunk
        for (int p = tid; p < TS * CLS_IN; p += 256) {
            int s = p / CLS_IN, k = p % CLS_IN;
            long long kk = kbase + s;
            sX[p] = kk < a.kend ? a.feat[sample_index(a, kk) * CLS_IN + k] : 0.0f;
        }
        if (tid < TS) {
            long long kk = kbase + tid;
            bool ok = kk < a.kend;
            float y = ok ? (float)a.label[sample_index(a, kk)] : 0.0f;
            sY[tid] = y;
            sWt[tid] = ok ? (y > 0.5f ? a.wpos : a.wneg) : 0.0f;
        }
        __syncthreads();
        // layer 1
        for (int r = 0; r < TS * CLS_H / 256; ++r) {
            int o = tid + 256 * r, s = o >> 6, u = o & 63;
            float z = sB1[u];
            for (int k = 0; k < CLS_IN; ++k) z += sX[s * CLS_IN + k] * sW1[k * CLS_H + u];
            sZ1[s * RS + u] = z;
            sH1[s * RS + u] = gelu_f(z);
        }
        __syncthreads();
        // layer 2
        for (int r = 0; r < TS * CLS_H / 256; ++r) {
            int o = tid + 256 * r, s = o >> 6, u = o & 63;
            float z = sB2[u];
            for (int k = 0; k < CLS_H; ++k) z += sH1[s * RS + k] * sW2[k * CLS_H + u];
            sZ2[s * RS + u] = z;
            sH2[s * RS + u] = gelu_f(z);
        }
        __syncthreads();
        // output logit, loss and its gradient: 8 samples per wavefront
        for (int r = 0; r < TS / 4; ++r) {
            int s = wv * (TS / 4) + r;
            float z = wave_sum(sH2[s * RS + lane] * sW3[lane]) + sB3;
            if (lane == 0) {
                float y = sY[s], w = sWt[s];
                float sg = 1.0f / (1.0f + expf(-z));
                sG3[s] = w * (sg - y);
                float sp = fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z)));
                sL[s] = w * (sp - y * z);
            }
        }
        __syncthreads();
        // G2 = g3 w3 gelu'(Z2), in place of Z2
        for (int r = 0; r < TS * CLS_H / 256; ++r) {
            int o = tid + 256 * r, s = o >> 6, u = o & 63;
            float g, dg;
            gelu_fd(sZ2[s * RS + u], g, dg);
            sZ2[s * RS + u] = sG3[s] * sW3[u] * dg;
        }
        __syncthreads();
        // G1 = gelu'(Z1) (G2 W2^T), in place of Z1; lanes run over samples
        for (int r = 0; r < TS * CLS_H / 256; ++r) {
            int o = tid + 256 * r, k = o / TS, s = o % TS;
            float acc = 0.0f;
            for (int j = 0; j < CLS_H; ++j) acc += sZ2[s * RS + j] * sW2[k * CLS_H + j];
            float g, dg;
            gelu_fd(sZ1[s * RS + k], g, dg);
            sZ1[s * RS + k] = acc * dg;
        }
        __syncthreads();
        // gradient accumulation
        for (int r = 0; r < 16; ++r) {
            int o = tid + 256 * r, k = o >> 6, u = o & 63;
            float acc = 0.0f;
            for (int s = 0; s < TS; ++s) acc += sH1[s * RS + k] * sZ2[s * RS + u];
            gW2[r] += acc;
        }
        for (int r = 0; r < 3; ++r) {
            int o = tid + 256 * r, k = o >> 6, u = o & 63;
            float acc = 0.0f;
            for (int s = 0; s < TS; ++s) acc += sX[s * CLS_IN + k] * sZ1[s * RS + u];
            gW1[r] += acc;
        }
        {
            float acc = 0.0f;
            if (tid < 64) {
                for (int s = 0; s < TS; ++s) acc += sZ1[s * RS + tid];
            } else if (tid < 128) {
                for (int s = 0; s < TS; ++s) acc += sZ2[s * RS + tid - 64];
            } else if (tid < 192) {
                for (int s = 0; s < TS; ++s) acc += sH2[s * RS + tid - 128] * sG3[s];
            } else if (tid == 192) {
                for (int s = 0; s < TS; ++s) acc += sG3[s];
            } else if (tid == 193) {
                for (int s = 0; s < TS; ++s) lsum += sL[s];
            }
            gv += acc;
        }
        __syncthreads();
    }
    float* out = a.partial + (long long)blockIdx.x * (CLS_NPARAM + 1);
    for (int r = 0; r < 16; ++r) out[CLS_W2 + tid + 256 * r] = gW2[r];
    for (int r = 0; r < 3; ++r) out[CLS_W1 + tid + 256 * r] = gW1[r];
    if (tid < 64) out[CLS_B1 + tid] = gv;
    else if (tid < 128) out[CLS_B2 + tid - 64] = gv;
    else if (tid < 192) out[CLS_W3 + tid - 128] = gv;
    else if (tid == 192) out[CLS_B3] = gv;
    else if (tid == 193) out[CLS_NPARAM] = lsum;
}

__global__ void __launch_bounds__(256) k_train_update(LrnUpdateArgs a) {
    int p = blockIdx.x * 256 + threadIdx.x;
    if (p > CLS_NPARAM) return;
    float g = 0.0f;
    for (int b = 0; b < a.nblk; ++b) g += a.partial[(long long)b * (CLS_NPARAM + 1) + p];
    if (p == CLS_NPARAM) {
        *a.loss += (double)g;
        return;
    }
    if (a.grad_out) {
        a.grad_out[p] = g;
        return;
    }
    float m = a.beta1 * a.m[p] + (1.0f - a.beta1) * g;
    float v = a.beta2 * a.v[p] + (1.0f - a.beta2) * g * g;
    a.m[p] = m;
    a.v[p] = v;
    a.prm[p] -= a.lr_t * m / (sqrtf(v) + a.eps);
}

// bf16 rounding of an FP32 value, kept in FP32
__device__ __forceinline__ float bf16_round(float x) { return bf16_to_f32(f32_to_bf16(x)); }

// FP32 forward of every sample (one per work item): confusion counts at the
// logit threshold, summed loss, optionally the logits.  With
// inference_rounding the weights of the two matrix layers, the features and
// layer 1's activations are rounded to bf16 where the solver's inference
// rounds them (the sums stay FP32).
__global__ void __launch_bounds__(256) k_eval(LrnEvalArgs a) {
    __shared__ float sW1[CLS_IN * CLS_H], sB1[CLS_H], sW2[CLS_H * CLS_H], sB2[CLS_H], sW3[CLS_H], sB3;
    __shared__ unsigned long long red[4][4];
    const int tid = threadIdx.x;
    const bool rb = a.inference_rounding != 0;
    for (int p = tid; p < CLS_IN * CLS_H; p += 256) sW1[p] = rb ? bf16_round(a.prm[CLS_W1 + p]) : a.prm[CLS_W1 + p];
    for (int p = tid; p < CLS_H * CLS_H; p += 256) sW2[p] = rb ? bf16_round(a.prm[CLS_W2 + p]) : a.prm[CLS_W2 + p];
    if (tid < CLS_H) {
        sB1[tid] = a.prm[CLS_B1 + tid];
        sB2[tid] = a.prm[CLS_B2 + tid];
        sW3[tid] = a.prm[CLS_W3 + tid];
    }
    if (tid == 0) sB3 = a.prm[CLS_B3];
    __syncthreads();
    unsigned long long cnt[4] = {0, 0, 0, 0};  // smooth->smooth, rough->smooth, rough->rough, smooth->rough
    for (long long s = blockIdx.x * 256ll + tid; s < a.n; s += (long long)gridDim.x * 256) {
        float x[CLS_IN], h1[CLS_H];
        for (int k = 0; k < CLS_IN; ++k) x[k] = rb ? bf16_round(a.feat[s * CLS_IN + k]) : a.feat[s * CLS_IN + k];
        for (int u = 0; u < CLS_H; ++u) {
            float z = sB1[u];
            for (int k = 0; k < CLS_IN; ++k) z += x[k] * sW1[k * CLS_H + u];
            h1[u] = rb ? bf16_round(gelu_f(z)) : gelu_f(z);
        }
        float z3 = sB3;
        for (int u = 0; u < CLS_H; ++u) {
            float z = sB2[u];
            for (int k = 0; k < CLS_H; ++k) z += h1[k] * sW2[k * CLS_H + u];
            z3 += sW3[u] * gelu_f(z);
        }
        if (a.logits) a.logits[s] = z3;
        bool smooth = a.label[s] != 0, pred = z3 > a.threshold;
        cnt[smooth ? (pred ? 0 : 3) : (pred ? 1 : 2)]++;
    }
    for (int c = 0; c < 4; ++c) {
        unsigned long long v = cnt[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((tid & 63) == 0) red[c][tid >> 6] = v;
    }
    __syncthreads();
    if (tid < 4) atomicAdd(&a.counts[tid], red[tid][0] + red[tid][1] + red[tid][2] + red[tid][3]);
}

// standalone MFMA inference on every face of one direction (the fused stage
// kernel runs the same cls_logit on its tile's faces)
template <bool YDIR>
__global__ void __launch_bounds__(256) k_classify(LrnClassifyArgs a) {
    __shared__ unsigned short xs[4][64 * 16];
    long long n = (long long)a.nx * a.ny;
    long long id = blockIdx.x * 256ll + threadIdx.x;
    float x[CLS_IN];
    for (int k = 0; k < CLS_IN; ++k) x[k] = 0.0f;
    if (id < n) {
        int i = (int)(id % a.nx), j = (int)(id / a.nx);
        Stencil6 st;
        load_stencil<YDIR>(a.q, a.nx, a.ny, a.pitch, a.fstride, a.bc, i, j, st);
        stencil_features<YDIR>(st, a.gam, x);
        if (a.feat)
            for (int k = 0; k < CLS_IN; ++k) a.feat[id * CLS_IN + k] = x[k];
    }
    cls_stage_features(xs[threadIdx.x >> 6], x);
    __syncthreads();
    float z = cls_logit(xs[threadIdx.x >> 6], a.frag, a.cvec);
    if (id < n) a.logits[id] = z;
}

// Work block of the face-flag kernel: FA faces along the stencil direction
// by FB grid lines across it, with the primitive variables of the
// (FA + 5) x FB cells their stencils read computed once into LDS.  In both
// families 32 lanes run along x, the contiguous direction of the state.
template <bool YDIR>
struct FlagBlock {
    static constexpr int FA = YDIR ? 8 : 32;
    static constexpr int FB = YDIR ? 32 : 8;
    static constexpr int NA = FA + 5;  // cells along: a0 - 2 .. a0 + FA + 2
    static constexpr int NCELL = NA * FB;
    // LDS slot of cell ca (along) of line cb; x-contiguous in both families
    static HTX_HD int slot(int ca, int cb) { return YDIR ? ca * FB + cb : cb * NA + ca; }
};
static_assert(FlagBlock<false>::FA * FlagBlock<false>::FB == 256 && FlagBlock<true>::FA * FlagBlock<true>::FB == 256,
              "one face per thread of a 256-thread workgroup");
constexpr int FLAG_NCELL_MAX = FlagBlock<false>::NCELL > FlagBlock<true>::NCELL ? FlagBlock<false>::NCELL
                                                                                : FlagBlock<true>::NCELL;

// classifier flags of every face of one family, including the faces one
// cell past each boundary (the stage kernel's tiles reach them).  Face a of
// line b lies between cells a and a + 1 along the family's direction,
// a = -1 .. n_along; its flag goes to flags_x[b][a + 1] (x faces, a row of
// nx + 2) or flags_y[a + 1][b] (y faces, a row of nx).  The features are
// those of stencil_features on the same six cells, bit for bit.
// the features' noise floor on FP32-bulk states, relative to each quantity's
// scale: the rounding noise of those states is about 1e-7 to 1e-6 of the
// state, and a real jump of 1e-5 of the state moves a linear face value by
// less than that again
#ifndef HTX_LRN_FP32_FLOOR
#define HTX_LRN_FP32_FLOOR 1e-5f
#endif

template <bool YDIR>
__global__ void __launch_bounds__(256) k_face_flags(LrnFlagsArgs a) {
    using FB_ = FlagBlock<YDIR>;
    __shared__ unsigned short xs[4][64 * 16];
    __shared__ double cell[7][FLAG_NCELL_MAX];  // q0..q3, un, ut, p (r = q0)
    if (a.ctl && !static_cast<const Ctl*>(a.ctl)->active) return;  // uniform over the grid
    const int tid = threadIdx.x;
    const int n_along = YDIR ? a.ny : a.nx, n_across = YDIR ? a.nx : a.ny;
    const int a0 = -1 + FB_::FA * (int)blockIdx.x, b0 = FB_::FB * (int)blockIdx.y;
    const double g1 = a.gam - 1.0;
    for (int w = tid; w < FB_::NCELL; w += 256) {
        const int ca = YDIR ? w / FB_::FB : w % FB_::NA;
        const int cb = YDIR ? w % FB_::FB : w / FB_::NA;
        const int gb = b0 + cb;
        if (gb >= n_across) continue;  // read only by faces that are not written
        const int ga = map_index(a0 - 2 + ca, n_along, a.bc);
        const int gi = YDIR ? gb : ga, gj = YDIR ? ga : gb;
        const long long off = (long long)gj * a.pitch + gi;
        double q[4];
        for (int e = 0; e < 4; ++e) q[e] = a.q[off + e * a.fstride];
        double r, un, ut, p;
        prim_nt(g1, q[0], YDIR ? q[2] : q[1], YDIR ? q[1] : q[2], q[3], r, un, ut, p);
        for (int e = 0; e < 4; ++e) cell[e][w] = q[e];
        cell[4][w] = un;
        cell[5][w] = ut;
        cell[6][w] = p;
    }
    __syncthreads();
    const int fa = YDIR ? tid / FB_::FB : tid % FB_::FA;
    const int fb = YDIR ? tid % FB_::FB : tid / FB_::FA;
    const int fal = a0 + fa, fbl = b0 + fb;
    const bool valid = fal <= n_along && fbl < n_across;
    bool flat = true;  // faces beyond the array count as flat
    int c[6];
    if (valid) {
        for (int m = 0; m < 6; ++m) c[m] = FB_::slot(fa + m, fb);
        for (int m = 1; m < 6; ++m)
            for (int e = 0; e < 4; ++e) flat = flat && cell[e][c[m]] == cell[e][c[0]];
    }
    unsigned char* out = YDIR ? a.flags_y : a.flags_x;
    const long long o = YDIR ? (long long)(fal + 1) * a.nx + fbl : (long long)fbl * (a.nx + 2) + (fal + 1);
    // A stencil of six equal states is smooth for TENO5 (every smoothness
    // measure is zero, no candidate is cut): when all the workgroup's faces
    // are such, the matrix cores are skipped (the classifier agrees there,
    // with a logit far above any threshold).
    if (__syncthreads_and(flat ? 1 : 0)) {
        if (valid) out[o] = 1;
        return;
    }
    float x[CLS_IN];
    for (int k = 0; k < CLS_IN; ++k) x[k] = 0.0f;
    if (valid) {
        double r[6], un[6], ut[6], p[6];
        for (int m = 0; m < 6; ++m) {
            r[m] = cell[0][c[m]];
            un[m] = cell[4][c[m]];
            ut[m] = cell[5][c[m]];
            p[m] = cell[6][c[m]];
        }
        face_features(r, un, ut, p, a.gam, x, a.bulk32 ? HTX_LRN_FP32_FLOOR : 0.0f);
    }
    cls_stage_features(xs[tid >> 6], x);
    __syncthreads();
    float z = cls_logit(xs[tid >> 6], a.frag, a.cvec);
    if (valid) out[o] = (flat || z > a.threshold) ? 1 : 0;
}

}  // namespace

}  // namespace htx_lrn

using namespace htx_lrn;

hipError_t lrn_launch_extract(const LrnExtractArgs& a, int dir, hipStream_t s) {
    long long n = (long long)a.nx * a.ny;
    dim3 g((unsigned)((n + 255) / 256));
    if (dir == 0)
        hipLaunchKernelGGL(k_extract<false>, g, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_extract<true>, g, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t lrn_launch_train(const LrnTrainArgs& a, int nblk, const LrnUpdateArgs& u, hipStream_t s) {
    hipLaunchKernelGGL(k_train_grad, dim3(nblk), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_train_update, dim3((CLS_NPARAM + 1 + 255) / 256), dim3(256), 0, s, u);
    return hipGetLastError();
}

hipError_t lrn_launch_eval(const LrnEvalArgs& a, hipStream_t s) {
    long long nb = (a.n + 255) / 256;
    if (nb > 4096) nb = 4096;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_eval, dim3((unsigned)nb), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t lrn_launch_classify(const LrnClassifyArgs& a, int dir, hipStream_t s) {
    long long n = (long long)a.nx * a.ny;
    dim3 g((unsigned)((n + 255) / 256));
    if (dir == 0)
        hipLaunchKernelGGL(k_classify<false>, g, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_classify<true>, g, dim3(256), 0, s, a);
    return hipGetLastError();
}

size_t lrn_flag_bytes_x(int nx, int ny) { return (size_t)(nx + 2) * ny; }
size_t lrn_flag_bytes_y(int nx, int ny) { return (size_t)nx * (ny + 2); }

hipError_t lrn_launch_flags(const LrnFlagsArgs& a, hipStream_t s) {
    using BX = FlagBlock<false>;
    using BY = FlagBlock<true>;
    // faces a = -1 .. n along each family's direction: n + 2 of them
    dim3 gx((unsigned)((a.nx + 2 + BX::FA - 1) / BX::FA), (unsigned)((a.ny + BX::FB - 1) / BX::FB));
    dim3 gy((unsigned)((a.ny + 2 + BY::FA - 1) / BY::FA), (unsigned)((a.nx + BY::FB - 1) / BY::FB));
    hipLaunchKernelGGL(k_face_flags<false>, gx, dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_face_flags<true>, gy, dim3(256), 0, s, a);
    return hipGetLastError();
}

void lrn_pack(const float* prm, unsigned short* frag, float* cvec) { cls_pack(prm, frag, cvec); }
