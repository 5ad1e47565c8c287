// htx_batch.hip -- many independent 2D Euler problems of uneven size
// advanced together (include/htx.h, "Batches").
//
// TENO5 instances run in ONE persistent launch per RK stage: the grid is the
// GPU's resident capacity of workgroups, each workgroup walks a contiguous
// run of the batch's global tile list (all instances' tiles back to back), and
// a tile finds its instance and that stage's argument block through a
// tile -> instance table.  Hermite instances run batched versions of the five
// Hermite launches (limiter, recon x/y, Gauss-point fluxes x/y, update), each
// over the concatenated work items of all Hermite instances, a workgroup never
// straddling two instances.  Every instance keeps its own control block (t,
// dt, CFL speeds, end time), so instances of different sizes and speeds step
// with their own dt; one batched kernel computes all dt and one advances all
// times.  A whole step can be captured once in a hipGraph and replayed.  The
// tile shape of the TENO5 launch comes from htx_batch_tune, which times the
// built shapes on the batch itself.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/htx.h"
#include "htx_batch_tables.h"
#include "htx_learn_api.h"

int htx_internal_fail(int code, const std::string& msg);
// the strict FP64 build's single-problem launchers (htx_strict.hip)
hipError_t htx_strict_launch_fv2d_stage(const void* a, int flux, int tile_x, int tile_y, hipStream_t s);
hipError_t htx_strict_launch_fvh2d_stage(const void* a, int flux, hipStream_t s);
hipError_t htx_strict_launch_max_speed(const double* q, int nx, int ny, int pitch, long long fstride, double gam,
                                       int quirks, void* ctl, hipStream_t s);

namespace htx {

namespace {

struct BatchTenoArgs {
    const Fv2dArgs* args;   // [mode - 1][instance]
    const int* tile_inst;   // global tile -> TENO instance
    const int* tile0;       // first global tile of each instance
    int ninst, ntiles, mode;
};

template <int TX, int TY, int FLUX, bool MIXED>
__global__ void __launch_bounds__(256, 2) k_batch_teno(BatchTenoArgs B) {
    using T = Fv2dTile<TX, TY>;
    __shared__ T s;
    const int per = (B.ntiles + gridDim.x - 1) / gridDim.x;
    const int g0 = blockIdx.x * per;
    const int g1 = min(g0 + per, B.ntiles);
    for (int g = g0; g < g1; ++g) {
        const int b = B.tile_inst[g];
        const Fv2dArgs& a = B.args[(B.mode - 1) * B.ninst + b];
        if (!a.ctl->active) continue;  // uniform over the workgroup
        const double dt = a.ctl->dt;
        stage_tile<TX, TY, FLUX, MIXED, false>(a, s, g - B.tile0[b], dt, threadIdx.x);
        __syncthreads();
    }
}

struct BatchHermArgs {
    const FvhArgs* args;    // [mode - 1][instance]
    const int* inst[6];     // block -> Hermite instance, per launch kind
    const int* first[6];    // first block of each instance, per launch kind
    int ninst, mode;
};

__device__ __forceinline__ const FvhArgs& herm_args(const BatchHermArgs& B, int kind, int& lb) {
    const int b = B.inst[kind][blockIdx.x];
    lb = blockIdx.x - B.first[kind][b];
    return B.args[(B.mode - 1) * B.ninst + b];
}

__global__ void __launch_bounds__(LIM_THREADS) kb_fvh_limit(BatchHermArgs B) {
    int lb;
    const FvhArgs& a = herm_args(B, HK_LIMIT, lb);
    if (!a.ctl->active) return;
    fvh_limit_block(a, lb, (4 * a.ny + 63) / 64);
}

// FT: the face values' storage type (float in the FP32 bulk precision)
template <bool YDIR, bool MIXED, typename FT = double>
__global__ void __launch_bounds__(256) kb_fvh_recon(BatchHermArgs B) {
    int lb;
    const FvhArgs& a = herm_args(B, YDIR ? HK_RECON_Y : HK_RECON_X, lb);
    if (!a.ctl->active) return;
    long long nf = YDIR ? (long long)a.nx * (a.ny + 1) : (long long)(a.nx + 1) * a.ny;
    long long t = lb * 256ll + threadIdx.x;
    if (t >= nf) return;
    if (YDIR) {
        int i = (int)(t % a.nx), fj = (int)(t / a.nx);
        fvh_recon_face<true, MIXED>(a, i, fj - 1, reinterpret_cast<FT*>(a.fy), t, nf);
    } else {
        int fi = (int)(t % (a.nx + 1)), j = (int)(t / (a.nx + 1));
        fvh_recon_face<false, MIXED>(a, fi - 1, j, reinterpret_cast<FT*>(a.fx), t, nf);
    }
}

template <bool YDIR, int FLUX, bool MIXED>
__global__ void __launch_bounds__(192) kb_fvh_flux(BatchHermArgs B) {
    __shared__ double lds[24 * 3 * 64];
    int lb;
    const FvhArgs& a = herm_args(B, YDIR ? HK_FLUX_Y : HK_FLUX_X, lb);
    if (!a.ctl->active) return;
    long long nf = YDIR ? (long long)a.nx * (a.ny + 1) : (long long)(a.nx + 1) * a.ny;
    fvh_flux_block<YDIR, FLUX, MIXED>(a, lb * 64ll, nf, lds);
}

template <bool YDIR, int FLUX>
__global__ void __launch_bounds__(192, 3) kb_fvh_flux_f32(BatchHermArgs B) {
    __shared__ double lds[12 * 3 * 64];
    int lb;
    const FvhArgs& a = herm_args(B, YDIR ? HK_FLUX_Y : HK_FLUX_X, lb);
    if (!a.ctl->active) return;
    long long nf = YDIR ? (long long)a.nx * (a.ny + 1) : (long long)(a.nx + 1) * a.ny;
    fvh_flux_block_f32<YDIR, FLUX>(a, lb * 64ll, nf, lds);
}

__global__ void __launch_bounds__(256) kb_fvh_update(BatchHermArgs B) {
    __shared__ double red[2][4];
    int lb;
    const FvhArgs& a = herm_args(B, HK_UPDATE, lb);
    if (!a.ctl->active) return;
    const double dt = a.ctl->dt;
    long long n = (long long)a.nx * a.ny;
    long long t = lb * 256ll + threadIdx.x;
    double smx = 0.0, smy = 0.0;
    bool bad = false;
    if (t < n) fvh_update_cell(a, (int)(t % a.nx), (int)(t / a.nx), dt, smx, smy, &bad);
    if (a.mode == 3) {
        flag_nonfinite(bad, a.ctl);
        smx = wave_max(smx);
        smy = wave_max(smy);
        if ((threadIdx.x & 63) == 0) {
            red[0][threadIdx.x >> 6] = smx;
            red[1][threadIdx.x >> 6] = smy;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int k = 1; k < 4; ++k) {
                smx = fmax(smx, red[0][k]);
                smy = fmax(smy, red[1][k]);
            }
            atomicMax(&a.ctl->smax_x, (unsigned long long)__double_as_longlong(smx));
            atomicMax(&a.ctl->smax_y, (unsigned long long)__double_as_longlong(smy));
        }
    }
}

__global__ void k_batch_dt(Ctl* c, const double* dxdy, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) ctl_dt(c + i, dxdy[2 * i], dxdy[2 * i + 1]);
}

__global__ void k_batch_tadvance(Ctl* c, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) ctl_tadvance(c + i);
}

// ---------------------------------------------------------------------------
// tile shapes of the TENO5 launch (htx_batch_tune picks one)
struct Shape {
    int tx, ty;
};
constexpr Shape SHAPES[] = {{16, 16}, {32, 8}, {8, 32}};
constexpr int NSHAPES = 3;

template <int TX, int TY>
hipError_t launch_teno_shape(const BatchTenoArgs& B, int flux, int prec, int nblk, hipStream_t s) {
    dim3 g(nblk), b(256);
    if (prec != 0 && prec != 1) return hipErrorInvalidValue;  // batches run fp64_fast or fp32_mixed
    if (flux == FLUX_RUSANOV_CHAR) {
        if (prec == 0) hipLaunchKernelGGL((k_batch_teno<TX, TY, FLUX_RUSANOV_CHAR, false>), g, b, 0, s, B);
        else hipLaunchKernelGGL((k_batch_teno<TX, TY, FLUX_RUSANOV_CHAR, true>), g, b, 0, s, B);
    } else {
        if (prec == 0) hipLaunchKernelGGL((k_batch_teno<TX, TY, FLUX_HLLC, false>), g, b, 0, s, B);
        else hipLaunchKernelGGL((k_batch_teno<TX, TY, FLUX_HLLC, true>), g, b, 0, s, B);
    }
    return hipGetLastError();
}

hipError_t launch_teno(int shape, const BatchTenoArgs& B, int flux, int prec, int nblk, hipStream_t s) {
    switch (shape) {
        case 0: return launch_teno_shape<16, 16>(B, flux, prec, nblk, s);
        case 1: return launch_teno_shape<32, 8>(B, flux, prec, nblk, s);
        default: return launch_teno_shape<8, 32>(B, flux, prec, nblk, s);
    }
}

template <bool MIXED>
void launch_herm_faces(const BatchHermArgs& B, const int nblocks[6], int flux, hipStream_t s) {
    hipLaunchKernelGGL((kb_fvh_recon<false, MIXED>), dim3(nblocks[HK_RECON_X]), dim3(256), 0, s, B);
    hipLaunchKernelGGL((kb_fvh_recon<true, MIXED>), dim3(nblocks[HK_RECON_Y]), dim3(256), 0, s, B);
    dim3 fb(64, 3), gx(nblocks[HK_FLUX_X]), gy(nblocks[HK_FLUX_Y]);
    if (flux == FLUX_RUSANOV_CHAR) {
        hipLaunchKernelGGL((kb_fvh_flux<false, FLUX_RUSANOV_CHAR, MIXED>), gx, fb, 0, s, B);
        hipLaunchKernelGGL((kb_fvh_flux<true, FLUX_RUSANOV_CHAR, MIXED>), gy, fb, 0, s, B);
    } else if (flux == FLUX_HLLC) {
        hipLaunchKernelGGL((kb_fvh_flux<false, FLUX_HLLC, MIXED>), gx, fb, 0, s, B);
        hipLaunchKernelGGL((kb_fvh_flux<true, FLUX_HLLC, MIXED>), gy, fb, 0, s, B);
    } else {
        hipLaunchKernelGGL((kb_fvh_flux<false, FLUX_HLLC_TORO, MIXED>), gx, fb, 0, s, B);
        hipLaunchKernelGGL((kb_fvh_flux<true, FLUX_HLLC_TORO, MIXED>), gy, fb, 0, s, B);
    }
}

// FP32 bulk precision: FP32 face values from the FP32-mixed reconstruction,
// FP32 fluxes
void launch_herm_faces_f32(const BatchHermArgs& B, const int nblocks[6], int flux, hipStream_t s) {
    hipLaunchKernelGGL((kb_fvh_recon<false, true, float>), dim3(nblocks[HK_RECON_X]), dim3(256), 0, s, B);
    hipLaunchKernelGGL((kb_fvh_recon<true, true, float>), dim3(nblocks[HK_RECON_Y]), dim3(256), 0, s, B);
    dim3 fb(64, 3), gx(nblocks[HK_FLUX_X]), gy(nblocks[HK_FLUX_Y]);
    if (flux == FLUX_RUSANOV_CHAR) {
        hipLaunchKernelGGL((kb_fvh_flux_f32<false, FLUX_RUSANOV_CHAR>), gx, fb, 0, s, B);
        hipLaunchKernelGGL((kb_fvh_flux_f32<true, FLUX_RUSANOV_CHAR>), gy, fb, 0, s, B);
    } else if (flux == FLUX_HLLC) {
        hipLaunchKernelGGL((kb_fvh_flux_f32<false, FLUX_HLLC>), gx, fb, 0, s, B);
        hipLaunchKernelGGL((kb_fvh_flux_f32<true, FLUX_HLLC>), gy, fb, 0, s, B);
    } else {
        hipLaunchKernelGGL((kb_fvh_flux_f32<false, FLUX_HLLC_TORO>), gx, fb, 0, s, B);
        hipLaunchKernelGGL((kb_fvh_flux_f32<true, FLUX_HLLC_TORO>), gy, fb, 0, s, B);
    }
}

// prec: the Hermite instances' shared precision, 0 fp64_fast, 1 fp32_mixed,
// 3 fp32 (bulk)
hipError_t launch_herm(const BatchHermArgs& B, const int nblocks[6], int flux, int prec, hipStream_t s) {
    hipLaunchKernelGGL(kb_fvh_limit, dim3(nblocks[HK_LIMIT]), dim3(LIM_THREADS), 0, s, B);
    if (prec == 3)
        launch_herm_faces_f32(B, nblocks, flux, s);
    else if (prec == 1)
        launch_herm_faces<true>(B, nblocks, flux, s);
    else
        launch_herm_faces<false>(B, nblocks, flux, s);
    hipLaunchKernelGGL(kb_fvh_update, dim3(nblocks[HK_UPDATE]), dim3(256), 0, s, B);
    return hipGetLastError();
}

}  // namespace
}  // namespace htx

using namespace htx;

struct htx_batch {
    int n = 0;
    std::vector<htx_params> p;
    std::vector<int> teno, herm;           // instance ids per scheme
    std::vector<long long> off;            // instance -> offset of its state in each state copy
    std::vector<double> dxdy;              // host copy
    long long total = 0;                   // doubles per state copy
    double* state = nullptr;               // 3 copies: A | B | C
    Ctl* ctl = nullptr;
    double* dDxdy = nullptr;
    // TENO5 launch
    int teno_flux = FLUX_HLLC, teno_prec = BP_FP32_MIXED;  // a BatchPrec code
    int shape = 0, ntiles = 0, nblk_teno = 0;
    Fv2dArgs* dTenoArgs = nullptr;
    int *dTileInst = nullptr, *dTile0 = nullptr;
    // direct TENO5 launches (batch_prec_direct): one stage launch (default tile) per
    // instance from these host argument blocks, and for the learned
    // precisions one face-flag launch per instance and step
    bool teno_direct = false;
    std::vector<Fv2dArgs> teno_args;        // [mode - 1][TENO instance]
    std::vector<unsigned char*> dFlagsX, dFlagsY;
    // Hermite launches
    int herm_flux = FLUX_RUSANOV_CHAR;
    int herm_prec = 0;                      // Hermite instances: 0 fp64_fast, 1 fp32_mixed, 3 fp32 (bulk)
    bool herm_strict = false;               // Hermite instances in FP64 (strict build, one launch per instance)
    std::vector<FvhArgs> herm_args;         // [mode - 1][Hermite instance], host copy
    FvhArgs* dHermArgs = nullptr;
    int* dHermMap = nullptr;
    int herm_blocks[6] = {0, 0, 0, 0, 0, 0};
    long long herm_map_off[12] = {0};
    double* herm_scratch = nullptr;
    // classifier
    unsigned short* dFrag = nullptr;
    float* dCvec = nullptr;
    float cls_thr = 0.0f;
    bool have_cls = false;
    // execution
    hipStream_t stream = nullptr;
    bool use_graph = false;
    hipGraphExec_t graph = nullptr;
    int cu = 256;
};

namespace {

#define BHIP(expr)                                                                                      \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return htx_internal_fail(HTX_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)


int flux_code(const htx_params& p) {
    if (p.flux == HTX_FLUX_RUSANOV_CHAR) return FLUX_RUSANOV_CHAR;
    return p.ref_quirks ? FLUX_HLLC : FLUX_HLLC_TORO;
}

double* inst_state(htx_batch* b, int copy, int i) { return b->state + copy * b->total + b->off[i]; }

BatchLayout layout_of(const htx_batch* b) {
    BatchLayout L;
    L.p = b->p;
    L.teno = b->teno;
    L.herm = b->herm;
    L.off = b->off;
    L.dxdy = b->dxdy;
    L.total = b->total;
    return L;
}

// TENO5 argument blocks of the current tile shape, tile tables
int build_teno_tables(htx_batch* b) {
    const int nt = (int)b->teno.size();
    if (nt == 0) return HTX_OK;
    BatchLayout L = layout_of(b);
    const bool lrn = batch_prec_learned(b->teno_prec);
    TenoTables T = make_teno_tables(L, b->state, b->ctl, SHAPES[b->shape].tx, SHAPES[b->shape].ty, b->teno_prec,
                                    lrn ? b->dFlagsX.data() : nullptr, lrn ? b->dFlagsY.data() : nullptr);
    if (b->teno_direct) {
        // the single-problem launchers size their own grids from these blocks
        b->teno_args = T.args;
        b->ntiles = T.ntiles;
        return HTX_OK;
    }
    const int g = T.ntiles;
    b->ntiles = g;
    // the old tables may still be read by launches in flight
    BHIP(hipStreamSynchronize(b->stream));
    if (b->dTenoArgs) (void)hipFree(b->dTenoArgs);
    if (b->dTileInst) (void)hipFree(b->dTileInst);
    if (b->dTile0) (void)hipFree(b->dTile0);
    b->dTenoArgs = nullptr;
    b->dTileInst = b->dTile0 = nullptr;
    BHIP(hipMalloc(&b->dTenoArgs, sizeof(Fv2dArgs) * T.args.size()));
    BHIP(hipMalloc(&b->dTileInst, sizeof(int) * T.tile_inst.size()));
    BHIP(hipMalloc(&b->dTile0, sizeof(int) * T.tile0.size()));
    BHIP(copy_sync(b->dTenoArgs, T.args.data(), sizeof(Fv2dArgs) * T.args.size(), hipMemcpyHostToDevice, b->stream));
    BHIP(copy_sync(b->dTileInst, T.tile_inst.data(), sizeof(int) * T.tile_inst.size(), hipMemcpyHostToDevice,
                   b->stream));
    BHIP(copy_sync(b->dTile0, T.tile0.data(), sizeof(int) * T.tile0.size(), hipMemcpyHostToDevice, b->stream));
    // persistent grid: the resident capacity (2 workgroups per CU), at most
    // one workgroup per tile
    b->nblk_teno = std::min(g, 2 * b->cu);
    return HTX_OK;
}

// the learned precisions' face-flag arrays, one pair per TENO5 instance
int alloc_teno_flags(htx_batch* b) {
    const int nt = (int)b->teno.size();
    b->dFlagsX.assign(nt, nullptr);
    b->dFlagsY.assign(nt, nullptr);
    for (int k = 0; k < nt; ++k) {
        const htx_params& p = b->p[b->teno[k]];
        BHIP(hipMalloc(&b->dFlagsX[k], lrn_flag_bytes_x(p.nx, p.ny)));
        BHIP(hipMalloc(&b->dFlagsY[k], lrn_flag_bytes_y(p.nx, p.ny)));
    }
    return HTX_OK;
}

int build_herm_tables(htx_batch* b) {
    const int nh = (int)b->herm.size();
    if (nh == 0) return HTX_OK;
    BatchLayout L = layout_of(b);
    BHIP(hipMalloc(&b->herm_scratch, sizeof(double) * (size_t)herm_scratch_size(L)));
    HermTables H = make_herm_tables(L, b->state, b->ctl, b->herm_scratch);
    b->herm_args = H.args;
    long long msize = 0;
    for (int k = 0; k < 12; ++k) {
        b->herm_map_off[k] = msize;
        msize += (long long)H.maps[k].size();
    }
    for (int kind = 0; kind < 6; ++kind) b->herm_blocks[kind] = (int)H.maps[kind].size();
    BHIP(hipMalloc(&b->dHermArgs, sizeof(FvhArgs) * H.args.size()));
    BHIP(copy_sync(b->dHermArgs, H.args.data(), sizeof(FvhArgs) * H.args.size(), hipMemcpyHostToDevice, b->stream));
    BHIP(hipMalloc(&b->dHermMap, sizeof(int) * (size_t)msize));
    for (int k = 0; k < 12; ++k)
        if (!H.maps[k].empty())
            BHIP(copy_sync(b->dHermMap + b->herm_map_off[k], H.maps[k].data(), sizeof(int) * H.maps[k].size(),
                           hipMemcpyHostToDevice, b->stream));
    return HTX_OK;
}

int enqueue_batch_step(htx_batch* b, hipStream_t s) {
    const int n = b->n;
    hipLaunchKernelGGL(k_batch_dt, dim3((n + 255) / 256), dim3(256), 0, s, b->ctl, b->dDxdy, n);
    for (int mode = 1; mode <= 3; ++mode) {
        if (!b->teno.empty() && b->teno_direct) {
            const int nt = (int)b->teno.size();
            for (int k = 0; k < nt; ++k) {
                const Fv2dArgs& a = b->teno_args[(size_t)(mode - 1) * nt + k];
                if (a.learned && mode == 1) {
                    // classify the faces of u^n once per step (nothing when the
                    // instance is past its end time)
                    LrnFlagsArgs f{};
                    f.q = a.qin;
                    f.nx = a.nx;
                    f.ny = a.ny;
                    f.pitch = a.pitch;
                    f.bc = a.bc;
                    f.fstride = a.fstride;
                    f.gam = a.gam;
                    f.frag = b->dFrag;
                    f.cvec = b->dCvec;
                    f.threshold = b->cls_thr;
                    f.ctl = a.ctl;
                    f.flags_x = b->dFlagsX[k];
                    f.flags_y = b->dFlagsY[k];
                    f.bulk32 = a.bulk32;
                    BHIP(lrn_launch_flags(f, s));
                }
                if (b->teno_prec == BP_FP64)
                    BHIP(htx_strict_launch_fv2d_stage(&a, b->teno_flux, 0, 0, s));
                else
                    BHIP(launch_fv2d_stage(a, b->teno_flux, 0, 0, s));
            }
        } else if (!b->teno.empty()) {
            BatchTenoArgs T{};
            T.args = b->dTenoArgs;
            T.tile_inst = b->dTileInst;
            T.tile0 = b->dTile0;
            T.ninst = (int)b->teno.size();
            T.ntiles = b->ntiles;
            T.mode = mode;
            BHIP(launch_teno(b->shape, T, b->teno_flux, b->teno_prec, b->nblk_teno, s));
        }
        if (!b->herm.empty() && b->herm_strict) {
            const int nh = (int)b->herm.size();
            for (int k = 0; k < nh; ++k)
                BHIP(htx_strict_launch_fvh2d_stage(&b->herm_args[(size_t)(mode - 1) * nh + k], b->herm_flux, s));
        } else if (!b->herm.empty()) {
            BatchHermArgs H{};
            H.args = b->dHermArgs;
            for (int k = 0; k < 6; ++k) {
                H.inst[k] = b->dHermMap + b->herm_map_off[k];
                H.first[k] = b->dHermMap + b->herm_map_off[6 + k];
            }
            H.ninst = (int)b->herm.size();
            H.mode = mode;
            BHIP(launch_herm(H, b->herm_blocks, b->herm_flux, b->herm_prec, s));
        }
    }
    hipLaunchKernelGGL(k_batch_tadvance, dim3((n + 255) / 256), dim3(256), 0, s, b->ctl, n);
    BHIP(hipGetLastError());
    return HTX_OK;
}

int drop_graph(htx_batch* b) {
    if (b->graph) {
        BHIP(hipGraphExecDestroy(b->graph));
        b->graph = nullptr;
    }
    return HTX_OK;
}

int capture_graph(htx_batch* b) {
    int rc = drop_graph(b);
    if (rc) return rc;
    hipGraph_t g = nullptr;
    BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_batch_step(b, b->stream);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return htx_internal_fail(HTX_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return htx_internal_fail(HTX_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    return HTX_OK;
}

// CFL maxima of a newly set state; FP64 instances take the strict build's
// sound speed, as the single-problem solver does
int refresh_instance_cfl(htx_batch* b, int i) {
    const htx_params& p = b->p[i];
    BHIP(launch_reset_smax(b->ctl + i, b->stream));
    if (p.precision == HTX_PREC_FP64)
        BHIP(htx_strict_launch_max_speed(inst_state(b, 0, i), p.nx, p.ny, p.nx, (long long)p.nx * p.ny, p.gamma,
                                         p.ref_quirks ? 1 : 0, b->ctl + i, b->stream));
    else
        BHIP(launch_max_speed(inst_state(b, 0, i), p.nx, p.ny, p.nx, (long long)p.nx * p.ny, p.gamma,
                              p.ref_quirks ? 1 : 0, b->ctl + i, b->stream));
    return HTX_OK;
}

int check_index(htx_batch* b, int i) {
    if (!b) return htx_internal_fail(HTX_EINVAL, "null batch");
    if (i < 0 || i >= b->n) return htx_internal_fail(HTX_EINVAL, "instance index out of range");
    return HTX_OK;
}

// HTX_ENONFINITE naming the first instance whose control block records a
// failure (non-finite state or dt; that instance stops stepping)
int status_error(const Ctl* c, int n, int first = 0) {
    for (int i = 0; i < n; ++i)
        if (c[i].status)
            return htx_internal_fail(HTX_ENONFINITE,
                                     "instance " + std::to_string(first + i) +
                                         ((c[i].status & CTL_NONFINITE_STATE)
                                              ? ": the state holds a NaN or infinite value"
                                              : ": the CFL time step is not finite and positive") +
                                         " at t = " + std::to_string(c[i].t));
    return HTX_OK;
}

int instance_status(htx_batch* b, int i) {
    Ctl c{};
    BHIP(hipMemcpyAsync(&c, b->ctl + i, sizeof(Ctl), hipMemcpyDeviceToHost, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    return status_error(&c, 1, i);
}

}  // namespace

extern "C" {

int htx_batch_create(const htx_params* ps, int n, htx_batch** out) {
    if (!out) return htx_internal_fail(HTX_EINVAL, "out is null");
    *out = nullptr;
    if (!ps || n < 1) return htx_internal_fail(HTX_EINVAL, "need n >= 1 parameter blocks");
    int teno_flux = -1, teno_prec = -1, herm_flux = -1, herm_pc = -1;
    for (int i = 0; i < n; ++i) {
        const htx_params& p = ps[i];
        if (p.struct_size != (int)sizeof(htx_params)) return htx_internal_fail(HTX_EINVAL, "params.struct_size mismatch");
        if (p.dim != 2 || p.nx <= 0 || p.ny <= 0 || !(p.xr > p.xl) || !(p.yt > p.yb) || !(p.gamma > 1.0) ||
            !(p.cfl > 0.0) || (p.bc != HTX_BC_ZERO_GRAD && p.bc != HTX_BC_PERIODIC) ||
            (p.flux != HTX_FLUX_RUSANOV_CHAR && p.flux != HTX_FLUX_HLLC))
            return htx_internal_fail(HTX_EINVAL, "instance " + std::to_string(i) + ": invalid parameters");
        if (p.scheme == HTX_SCHEME_TENO5) {
            int pc = batch_prec_code(p.precision), fc = flux_code(p);
            if (pc < 0) return htx_internal_fail(HTX_EINVAL, "instance " + std::to_string(i) + ": unknown precision");
            // the persistent launch is built for the reference's HLLC and
            // the characteristic Rusanov flux; the direct launches take all
            if (fc == FLUX_HLLC_TORO && !batch_prec_direct(pc))
                return htx_internal_fail(HTX_ENOTSUP, "batched fp64_fast / fp32_mixed HLLC keeps the reference quirks");
            if ((teno_prec >= 0 && pc != teno_prec) || (teno_flux >= 0 && fc != teno_flux))
                return htx_internal_fail(HTX_ENOTSUP, "the TENO5 instances of a batch share flux and precision");
            teno_prec = pc;
            teno_flux = fc;
        } else if (p.scheme == HTX_SCHEME_HTENO) {
            if (p.precision != HTX_PREC_FP64 && p.precision != HTX_PREC_FP64_FAST &&
                p.precision != HTX_PREC_FP32_MIXED && p.precision != HTX_PREC_FP32)
                return htx_internal_fail(HTX_ENOTSUP,
                                         "batched Hermite instances run fp64, fp64_fast, fp32_mixed or fp32 precision");
            int fc = flux_code(p);
            int pm = p.precision == HTX_PREC_FP32_MIXED ? 1
                   : p.precision == HTX_PREC_FP64       ? 2
                   : p.precision == HTX_PREC_FP32       ? 3
                                                        : 0;
            if ((herm_flux >= 0 && fc != herm_flux) || (herm_pc >= 0 && pm != herm_pc))
                return htx_internal_fail(HTX_ENOTSUP, "the Hermite instances of a batch share flux and precision");
            herm_flux = fc;
            herm_pc = pm;
        } else {
            return htx_internal_fail(HTX_EINVAL, "unknown scheme");
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return htx_internal_fail(HTX_EDEVICE, "no HIP device");
    htx_batch* b = new htx_batch();
    b->n = n;
    b->p.assign(ps, ps + n);
    if (teno_prec >= 0) b->teno_prec = teno_prec;
    if (teno_flux >= 0) b->teno_flux = teno_flux;
    b->teno_direct = batch_prec_direct(b->teno_prec);
    if (herm_flux >= 0) b->herm_flux = herm_flux;
    b->herm_prec = herm_pc == 2 || herm_pc < 0 ? 0 : herm_pc;
    b->herm_strict = herm_pc == 2;
    BatchLayout L = make_layout(ps, n);
    b->off = L.off;
    b->dxdy = L.dxdy;
    b->teno = L.teno;
    b->herm = L.herm;
    const long long total = L.total;
    b->total = total;
    int rc = HTX_OK;
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess && prop.multiProcessorCount > 0) b->cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&b->state, sizeof(double) * 3 * (size_t)total) != hipSuccess ||
        hipMalloc(&b->ctl, sizeof(Ctl) * (size_t)n) != hipSuccess ||
        hipMalloc(&b->dDxdy, sizeof(double) * 2 * (size_t)n) != hipSuccess) {
        htx_batch_destroy(b);
        return htx_internal_fail(HTX_ENOMEM, "device allocation failed");
    }
    std::vector<Ctl> c(n);
    for (int i = 0; i < n; ++i) {
        c[i] = Ctl{};
        c[i].t = ps[i].tinit;
        c[i].tfinal = ps[i].tmax;
        c[i].cfl = ps[i].cfl;
    }
    if (copy_sync(b->ctl, c.data(), sizeof(Ctl) * n, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        copy_sync(b->dDxdy, b->dxdy.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        zero_sync(b->state, sizeof(double) * 3 * (size_t)total, b->stream) != hipSuccess) {
        htx_batch_destroy(b);
        return htx_internal_fail(HTX_EDEVICE, "batch initialisation failed");
    }
    if ((batch_prec_learned(b->teno_prec) && !b->teno.empty() && (rc = alloc_teno_flags(b))) ||
        (rc = build_teno_tables(b)) || (rc = build_herm_tables(b))) {
        htx_batch_destroy(b);
        return rc;
    }
    *out = b;
    return HTX_OK;
}

void htx_batch_destroy(htx_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (void* m : {(void*)b->state, (void*)b->ctl, (void*)b->dDxdy, (void*)b->dTenoArgs, (void*)b->dTileInst,
                    (void*)b->dTile0, (void*)b->dHermArgs, (void*)b->dHermMap,
                    (void*)b->herm_scratch, (void*)b->dFrag, (void*)b->dCvec})
        if (m) (void)hipFree(m);
    for (unsigned char* m : b->dFlagsX)
        if (m) (void)hipFree(m);
    for (unsigned char* m : b->dFlagsY)
        if (m) (void)hipFree(m);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int htx_batch_set_state(htx_batch* b, int i, const double* q, size_t ld) {
    int rc = check_index(b, i);
    if (rc) return rc;
    const htx_params& p = b->p[i];
    if (!q || ld < (size_t)p.nx) return htx_internal_fail(HTX_EINVAL, "null state or ld < nx");
    int nf = p.scheme == HTX_SCHEME_HTENO ? 12 : 4;
    BHIP(hipMemcpy2DAsync(inst_state(b, 0, i), sizeof(double) * p.nx, q, sizeof(double) * ld, sizeof(double) * p.nx,
                          (size_t)p.ny * nf, hipMemcpyHostToDevice, b->stream));
    if ((rc = refresh_instance_cfl(b, i))) return rc;
    return instance_status(b, i);
}

int htx_batch_get_state(htx_batch* b, int i, double* q, size_t ld) {
    int rc = check_index(b, i);
    if (rc) return rc;
    const htx_params& p = b->p[i];
    if (!q || ld < (size_t)p.nx) return htx_internal_fail(HTX_EINVAL, "null state or ld < nx");
    int nf = p.scheme == HTX_SCHEME_HTENO ? 12 : 4;
    BHIP(hipMemcpy2DAsync(q, sizeof(double) * ld, inst_state(b, 0, i), sizeof(double) * p.nx, sizeof(double) * p.nx,
                          (size_t)p.ny * nf, hipMemcpyDeviceToHost, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    return HTX_OK;
}

int htx_batch_init_quadrant(htx_batch* b, int i, double xs, double ys, const double* states) {
    int rc = check_index(b, i);
    if (rc) return rc;
    if (!states) return htx_internal_fail(HTX_EINVAL, "null states");
    const htx_params& p = b->p[i];
    PiecewiseArgs a{};
    a.q = inst_state(b, 0, i);
    a.nx = p.nx;
    a.ny = p.ny;
    a.pitch = p.nx;
    a.nfield = p.scheme == HTX_SCHEME_HTENO ? 12 : 4;
    a.fstride = (long long)p.nx * p.ny;
    a.xl = p.xl;
    a.yb = p.yb;
    a.dx = b->dxdy[2 * i];
    a.dy = b->dxdy[2 * i + 1];
    a.gam = p.gamma;
    a.kind = IC_QUADRANT;
    a.xs = xs;
    a.ys = ys;
    for (int k = 0; k < 4; ++k)
        for (int e = 0; e < 4; ++e) a.st[k][e] = states[4 * k + e];
    BHIP(launch_ic_piecewise(a, b->stream));
    if ((rc = refresh_instance_cfl(b, i))) return rc;
    return instance_status(b, i);
}

int htx_batch_load_classifier(htx_batch* b, const float* params, double threshold) {
    if (!b || !params) return htx_internal_fail(HTX_EINVAL, "null argument");
    for (int k = 0; k < CLS_NPARAM; ++k)
        if (!std::isfinite(params[k])) return htx_internal_fail(HTX_EINVAL, "classifier parameters must be finite");
    std::vector<unsigned short> frag(CLS_NFRAG);
    std::vector<float> cvec(CLS_NCVEC);
    cls_pack(params, frag.data(), cvec.data());
    if (!b->dFrag) {
        BHIP(hipMalloc(&b->dFrag, sizeof(unsigned short) * CLS_NFRAG));
        BHIP(hipMalloc(&b->dCvec, sizeof(float) * CLS_NCVEC));
    }
    BHIP(copy_sync(b->dFrag, frag.data(), sizeof(unsigned short) * CLS_NFRAG, hipMemcpyHostToDevice, b->stream));
    BHIP(copy_sync(b->dCvec, cvec.data(), sizeof(float) * CLS_NCVEC, hipMemcpyHostToDevice, b->stream));
    b->cls_thr = (float)threshold;
    b->have_cls = true;
    int rc = build_teno_tables(b);
    if (rc) return rc;
    return drop_graph(b);
}

int htx_batch_steps(htx_batch* b, int nsteps) {
    if (!b) return htx_internal_fail(HTX_EINVAL, "null batch");
    if (nsteps < 0) return htx_internal_fail(HTX_EINVAL, "nsteps < 0");
    if (!b->teno.empty() && batch_prec_learned(b->teno_prec) && !b->have_cls)
        return htx_internal_fail(HTX_ESTATE, "the learned precision needs htx_batch_load_classifier");
    int rc;
    if (b->use_graph && !b->graph && (rc = capture_graph(b))) return rc;
    for (int k = 0; k < nsteps; ++k) {
        if (b->use_graph)
            BHIP(hipGraphLaunch(b->graph, b->stream));
        else if ((rc = enqueue_batch_step(b, b->stream)))
            return rc;
    }
    return HTX_OK;
}

int htx_batch_advance(htx_batch* b, int max_steps, int* nsteps) {
    if (!b) return htx_internal_fail(HTX_EINVAL, "null batch");
    std::vector<Ctl> c(b->n);
    int done = 0, chunk = 8, rc;
    while (done < max_steps) {
        int k = std::min(chunk, max_steps - done);
        if ((rc = htx_batch_steps(b, k))) return rc;
        done += k;
        BHIP(hipMemcpyAsync(c.data(), b->ctl, sizeof(Ctl) * b->n, hipMemcpyDeviceToHost, b->stream));
        BHIP(hipStreamSynchronize(b->stream));
        if ((rc = status_error(c.data(), b->n))) {
            if (nsteps) *nsteps = done;
            return rc;
        }
        bool all = true;
        for (int i = 0; i < b->n; ++i) all = all && !(c[i].t < c[i].tfinal);
        if (all) break;
        chunk = std::min(chunk * 2, 128);
    }
    if (nsteps) *nsteps = done;
    return HTX_OK;
}

int htx_batch_get_time(htx_batch* b, int i, double* t, int* nsteps) {
    int rc = check_index(b, i);
    if (rc) return rc;
    Ctl c{};
    BHIP(hipMemcpyAsync(&c, b->ctl + i, sizeof(Ctl), hipMemcpyDeviceToHost, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    if (t) *t = c.t;
    if (nsteps) *nsteps = c.nsteps;
    return HTX_OK;
}

int htx_batch_sync(htx_batch* b) {
    if (!b) return htx_internal_fail(HTX_EINVAL, "null batch");
    std::vector<Ctl> c(b->n);
    BHIP(hipMemcpyAsync(c.data(), b->ctl, sizeof(Ctl) * b->n, hipMemcpyDeviceToHost, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    return status_error(c.data(), b->n);
}

int htx_batch_use_graph(htx_batch* b, int enable) {
    if (!b) return htx_internal_fail(HTX_EINVAL, "null batch");
    b->use_graph = enable != 0;
    if (!b->use_graph) return drop_graph(b);
    return HTX_OK;
}

int htx_batch_set_tile(htx_batch* b, int tile_x, int tile_y) {
    if (!b) return htx_internal_fail(HTX_EINVAL, "null batch");
    int k = -1;
    for (int s = 0; s < NSHAPES; ++s)
        if (SHAPES[s].tx == tile_x && SHAPES[s].ty == tile_y) k = s;
    if (k < 0) return htx_internal_fail(HTX_EINVAL, "tile shape not built (16x16, 32x8, 8x32)");
    if (b->teno_direct && k != 0)
        return htx_internal_fail(HTX_ENOTSUP, "the direct TENO5 launches of this precision use the default tile");
    BHIP(hipStreamSynchronize(b->stream));
    b->shape = k;
    int rc = build_teno_tables(b);
    if (rc) return rc;
    return drop_graph(b);
}

int htx_batch_tune(htx_batch* b, int steps, int* tile_x, int* tile_y, double* ms) {
    if (!b) return htx_internal_fail(HTX_EINVAL, "null batch");
    if (steps < 1) return htx_internal_fail(HTX_EINVAL, "steps >= 1");
    if (!b->teno.empty() && batch_prec_learned(b->teno_prec) && !b->have_cls)
        return htx_internal_fail(HTX_ESTATE, "the learned precision needs htx_batch_load_classifier");
    if (ms)
        for (int k = 0; k < NSHAPES; ++k) ms[k] = -1.0;  // shapes not timed stay -1
    // Everything the timing changes is put back on every way out: the state
    // and control blocks (saved on the device first), the graph setting, the
    // tile shape with its launch tables (the original one unless a fastest
    // shape was found), and the buffers and events are released.
    struct Guard {
        htx_batch* b;
        double* saved = nullptr;
        Ctl* saved_ctl = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        bool graph;
        int shape;
        bool restore = false;  // saved holds the batch's state
        explicit Guard(htx_batch* b_) : b(b_), graph(b_->use_graph), shape(b_->shape) {}
        ~Guard() {
            if (restore) {
                (void)hipMemcpyAsync(b->state, saved, sizeof(double) * (size_t)b->total, hipMemcpyDeviceToDevice,
                                     b->stream);
                (void)hipMemcpyAsync(b->ctl, saved_ctl, sizeof(Ctl) * (size_t)b->n, hipMemcpyDeviceToDevice,
                                     b->stream);
            }
            (void)hipStreamSynchronize(b->stream);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (saved) (void)hipFree(saved);
            if (saved_ctl) (void)hipFree(saved_ctl);
            b->use_graph = graph;
            if (b->shape != shape) {
                b->shape = shape;
                (void)build_teno_tables(b);
            }
            (void)drop_graph(b);
        }
    } g(b);
    if (hipMalloc(&g.saved, sizeof(double) * (size_t)b->total) != hipSuccess ||
        hipMalloc(&g.saved_ctl, sizeof(Ctl) * (size_t)b->n) != hipSuccess)
        return htx_internal_fail(HTX_ENOMEM, "device allocation failed");
    BHIP(hipEventCreate(&g.e0));
    BHIP(hipEventCreate(&g.e1));
    BHIP(hipMemcpyAsync(g.saved, b->state, sizeof(double) * (size_t)b->total, hipMemcpyDeviceToDevice, b->stream));
    BHIP(hipMemcpyAsync(g.saved_ctl, b->ctl, sizeof(Ctl) * (size_t)b->n, hipMemcpyDeviceToDevice, b->stream));
    g.restore = true;
    b->use_graph = false;
    int best = b->shape, rc;
    double best_ms = 1e300;
    // the direct TENO5 launches have one shape: time it alone
    const int nshapes = b->teno_direct ? 1 : NSHAPES;
    for (int k = 0; k < nshapes; ++k) {
        b->shape = k;
        if ((rc = build_teno_tables(b)) || (rc = enqueue_batch_step(b, b->stream))) return rc;  // warm-up
        BHIP(hipEventRecord(g.e0, b->stream));
        for (int s = 0; s < steps; ++s)
            if ((rc = enqueue_batch_step(b, b->stream))) return rc;
        BHIP(hipEventRecord(g.e1, b->stream));
        BHIP(hipEventSynchronize(g.e1));
        float t = 0.f;
        BHIP(hipEventElapsedTime(&t, g.e0, g.e1));
        if (ms) ms[k] = t / steps;
        if (t < best_ms) {
            best_ms = t;
            best = k;
        }
        BHIP(hipMemcpyAsync(b->state, g.saved, sizeof(double) * (size_t)b->total, hipMemcpyDeviceToDevice, b->stream));
        BHIP(hipMemcpyAsync(b->ctl, g.saved_ctl, sizeof(Ctl) * (size_t)b->n, hipMemcpyDeviceToDevice, b->stream));
    }
    // keep the fastest shape: the guard rebuilds its tables on the way out;
    // build them here once more so that a failure is reported
    g.shape = best;
    b->shape = best;
    if ((rc = build_teno_tables(b))) return rc;
    if (tile_x) *tile_x = SHAPES[best].tx;
    if (tile_y) *tile_y = SHAPES[best].ty;
    return HTX_OK;
}

}  // extern "C"
