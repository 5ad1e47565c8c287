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

hipError_t lrn_launch_extract(const LrnExtractArgs& a, int dir, hipStream_t s);
hipError_t lrn_launch_train(const LrnTrainArgs& a, int nblk, const LrnUpdateArgs& u, hipStream_t s);
hipError_t lrn_launch_eval(const LrnEvalArgs& a, hipStream_t s);
hipError_t lrn_launch_classify(const LrnClassifyArgs& a, int dir, hipStream_t s);
void lrn_pack(const float* prm, unsigned short* frag, float* cvec);
