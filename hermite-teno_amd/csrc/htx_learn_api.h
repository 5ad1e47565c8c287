// htx_learn_api.h -- argument blocks and launchers of htx_learn.hip, in plain
// types so that the C ABI (htx_api.hip, another build of the numerics) can
// call them.
#pragma once
#include <hip/hip_runtime.h>

struct LrnExtractArgs {
    const double* q;
    int nx, ny, pitch, bc;
    long long fstride;
    double gam;
    float* feat;               // [base + face][12]
    unsigned char* label;      // [base + face]
    long long base;
    unsigned long long* nsmooth;
};

struct LrnTrainArgs {
    const float* feat;
    const unsigned char* label;
    long long n;               // samples in the set
    const float* prm;          // FP32 parameters
    float* partial;            // [nblk][CLS_NPARAM + 1]
    long long k0, kend;        // epoch positions k0 <= k < kend of this minibatch
    unsigned long long pa, pb; // sample(k) = (pa k + pb) mod n, gcd(pa, n) = 1
    float wpos, wneg;          // loss weight of a smooth / troubled sample (per minibatch)
    int chunks;                // chunks of 32 samples per workgroup
};

struct LrnUpdateArgs {
    const float* partial;
    int nblk;
    float* prm;
    float* m;
    float* v;
    double* loss;
    float* grad_out;   // if set: write the summed gradient here and leave the parameters
    float lr_t, beta1, beta2, eps;
};

struct LrnEvalArgs {
    const float* feat;
    const unsigned char* label;
    long long n;
    const float* prm;
    float threshold;
    unsigned long long* counts;  // smooth->smooth, troubled->smooth, troubled->troubled, smooth->troubled
    float* logits;               // optional
    int inference_rounding;      // 1: round W1, W2, the features and layer 1's outputs to bf16 as the
                                 // solver's matrix-core inference does
};

struct LrnClassifyArgs {
    const double* q;
    int nx, ny, pitch, bc;
    long long fstride;
    double gam;
    const unsigned short* frag;
    const float* cvec;
    float* logits;   // [face]
    float* feat;     // optional [face][12]
};

// Classifier flags of every face for the learned precision's stage kernel
// (1 = smooth: logit above the threshold).  x faces (cell i | i+1 of row j,
// i = -1 .. nx) at flags_x[j (nx + 2) + i + 1]; y faces (cell j | j+1 of
// column i, j = -1 .. ny) at flags_y[(j + 1) nx + i].  Cells outside the
// grid follow the boundary condition.  With ctl set, nothing is done when
// the step is inactive.
struct LrnFlagsArgs {
    const double* q;
    int nx, ny, pitch, bc;
    long long fstride;
    double gam;
    const unsigned short* frag;
    const float* cvec;
    float threshold;
    const void* ctl;            // Ctl* of the solver, or null
    unsigned char* flags_x;
    unsigned char* flags_y;
    int bulk32;                 // the state is advanced in FP32 bulk arithmetic (noise floor in the features)
};

hipError_t lrn_launch_extract(const LrnExtractArgs& a, int dir, hipStream_t s);
hipError_t lrn_launch_flags(const LrnFlagsArgs& a, hipStream_t s);
size_t lrn_flag_bytes_x(int nx, int ny);
size_t lrn_flag_bytes_y(int nx, int ny);
hipError_t lrn_launch_train(const LrnTrainArgs& a, int nblk, const LrnUpdateArgs& u, hipStream_t s);
hipError_t lrn_launch_eval(const LrnEvalArgs& a, hipStream_t s);
hipError_t lrn_launch_classify(const LrnClassifyArgs& a, int dir, hipStream_t s);
void lrn_pack(const float* prm, unsigned short* frag, float* cvec);
