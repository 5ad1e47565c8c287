// htx_api.hip -- the C ABI of libhtx.so (include/htx.h): handle lifetime,
// state transfer, the device-resident time loop and stage timing.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/htx.h"
#include "htx_schemes.h"
#include "htx_fv1d.h"
#include "htx_hermite.h"
#include "htx_learn.h"
#include "htx_learn_api.h"

namespace htx {
hipError_t launch_fvh2d_stage(const FvhArgs& a, int flux, hipStream_t s);
hipError_t launch_ic_vortex_hermite(const VortexArgs& a, hipStream_t s);
}  // namespace htx

// strict IEEE build of the same kernels (htx_strict.hip)
hipError_t htx_strict_launch_fv2d_stage(const void* a, int flux, int tile_x, int tile_y, hipStream_t s);
hipError_t htx_strict_launch_fvh2d_stage(const void* a, int flux, hipStream_t s);
hipError_t htx_strict_launch_max_speed(const double* q, int nx, int ny, int pitch, long long fstride, double gam,
                                       int quirks, void* ctl, hipStream_t s);
hipError_t htx_strict_launch_fv1d_stage(const void* a, hipStream_t s);
hipError_t htx_strict_launch_fv1d_max_speed(const double* q, int nx, double gam, void* ctl, hipStream_t s);
hipError_t htx_strict_launch_fv1d_lmpe(const double* q, int nx, int bc, double dx, unsigned char* flags,
                                       hipStream_t s);
hipError_t htx_strict_launch_fv1d_kxrcf(const double* q, int nx, int bc, double radius, unsigned char* flags,
                                        hipStream_t s);

using namespace htx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(HTX_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

// Gauss-Legendre 9-point rule in the reference's node order
// (Schemes_Lagrange.cpp:22-42).
void gauss9(double x[9], double w[9]) {
    const double a[4] = {0.8360311073266358, 0.9681602395076261, 0.3242534234038089, 0.6133714327005904};
    const double wa[4] = {0.1806481606948574, 0.08127438836157441, 0.31234707704000286, 0.26061069640293544};
    x[0] = 0.0;
    w[0] = 0.3302393550012598;
    for (int k = 0; k < 4; ++k) {
        x[1 + 2 * k] = -a[k];
        x[2 + 2 * k] = a[k];
        w[1 + 2 * k] = wa[k];
        w[2 + 2 * k] = wa[k];
    }
}

}  // namespace

// error reporting for the other translation units of the C ABI
int htx_internal_fail(int code, const std::string& msg) { return fail(code, msg); }

struct htx_solver {
    htx_params p{};
    int nx = 0, ny = 0, nfield = 4, pitch = 0;
    long long fstride = 0;
    double dx = 0, dy = 0;
    double* dA = nullptr;  // u^n / current state
    double* dB = nullptr;  // stage 1 result
    double* dC = nullptr;  // stage 2 result
    double* dR = nullptr;  // residual output
    // Hermite scheme only: limited moments, face states and face fluxes
    double* dLV = nullptr;
    double *dFX = nullptr, *dFY = nullptr, *dGX = nullptr, *dGY = nullptr;
    HermConsts hc{};
    Ctl* ctl = nullptr;
    hipStream_t stream = nullptr;
    QuadConsts qc{};
    bool have_state = false;
    bool timing = false;
    int next_stage = 1;  // htx_stage: the stage that may run next
    // learned classifier (precision HTX_PREC_LEARNED): packed bf16 fragments,
    // FP32 biases and output weights, logit threshold
    unsigned short* dFrag = nullptr;
    float* dCvec = nullptr;
    unsigned char* dFlagsX = nullptr;  // classifier flags of the x and y faces (learned precision)
    unsigned char* dFlagsY = nullptr;
    unsigned char* dTroubled = nullptr;  // 1D KXRCF cell flags (allocated on first use)
    unsigned char* dLmpe = nullptr;      // 1D LMPE flags, 3 fields (allocated on first use)
    // the analytic initial condition last set (htx_init_vortex,
    // htx_init_entropy_wave) for htx_l1_density_error: its exact solution
    // is the initial field advected with velocity (adv_u, adv_v)
    ExactArgs exact{};
    double adv_u = 0.0, adv_v = 0.0;
    double* dPartial = nullptr;  // per-workgroup partial sums + the total (allocated on first use)
    float cls_threshold = 0.0f;
    bool have_classifier = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> event_pool;
};

namespace {

bool is_learned(int precision) { return precision == HTX_PREC_LEARNED || precision == HTX_PREC_LEARNED_FP32; }
bool is_bulk32(int precision) { return precision == HTX_PREC_FP32 || precision == HTX_PREC_LEARNED_FP32; }

int check_params(const htx_params* p) {
    if (!p) return fail(HTX_EINVAL, "params is null");
    if (p->struct_size != (int)sizeof(htx_params)) return fail(HTX_EINVAL, "params.struct_size mismatch");
    if (p->dim != 1 && p->dim != 2) return fail(HTX_EINVAL, "dim must be 1 or 2");
    if (p->nx <= 0 || (p->dim == 2 && p->ny <= 0)) return fail(HTX_EINVAL, "grid size must be positive");
    if (!(p->xr > p->xl) || (p->dim == 2 && !(p->yt > p->yb)))
        return fail(HTX_EINVAL, "domain must have positive extent");
    if (p->scheme < HTX_SCHEME_TENO5 || p->scheme > HTX_SCHEME_HLINEAR) return fail(HTX_EINVAL, "unknown scheme");
    if (p->dim == 1) {
        // FV::System1D: TENO5, the shipped TENO5i or any other 5-point
        // reconstruction of Schemes::Lagrange, FP64
        if (htx_is_hermite(p->scheme))
            return fail(HTX_EINVAL, "the 1D solver takes a 5-point reconstruction (teno5, teno5i, weno5js, ...)");
        if (p->precision != HTX_PREC_FP64) return fail(HTX_ENOTSUP, "the 1D path is built in fp64 precision");
    } else if (p->scheme != HTX_SCHEME_TENO5 && p->scheme != HTX_SCHEME_HTENO &&
               p->precision != HTX_PREC_FP64 && p->precision != HTX_PREC_FP64_FAST) {
        // the 2D path takes them in place of TENO5 (FV_System2D.cpp:379, :394)
        // or of HTENO (FVHermite_System2D.cpp:459, :486)
        return fail(HTX_ENOTSUP, "the alternative 2D reconstructions are built in fp64 and fp64_fast");
    }
    if (!(p->gamma > 1.0)) return fail(HTX_EINVAL, "gamma must exceed 1");
    if (!(p->cfl > 0.0)) return fail(HTX_EINVAL, "cfl must be positive");
    if (p->flux != HTX_FLUX_RUSANOV_CHAR && p->flux != HTX_FLUX_HLLC && p->flux != HTX_FLUX_CU)
        return fail(HTX_EINVAL, "unknown flux");
    if (p->flux == HTX_FLUX_CU && p->dim != 1) return fail(HTX_ENOTSUP, "the central-upwind flux is built for 1D");
    if (p->bc != HTX_BC_ZERO_GRAD && p->bc != HTX_BC_PERIODIC) return fail(HTX_EINVAL, "unknown bc");
    if (p->precision < HTX_PREC_FP64 || p->precision > HTX_PREC_LEARNED_FP32)
        return fail(HTX_EINVAL, "unknown precision");
    if (is_learned(p->precision) && p->scheme != HTX_SCHEME_TENO5)
        return fail(HTX_ENOTSUP, "the learned precision is built for the TENO5 scheme only");
    if (p->precision == HTX_PREC_FP32 && p->scheme != HTX_SCHEME_TENO5 && p->scheme != HTX_SCHEME_HTENO)
        return fail(HTX_ENOTSUP, "the fp32 bulk precision is built for the TENO5 and Hermite-TENO schemes");
    return HTX_OK;
}

Fv2dArgs stage_args(htx_solver* h, const double* qin, const double* q0, double* qout, int mode) {
    Fv2dArgs a{};
    a.qin = qin;
    a.q0 = q0;
    a.qout = qout;
    a.ctl = h->ctl;
    a.nx = h->nx;
    a.ny = h->ny;
    a.pitch = h->pitch;
    a.fstride = h->fstride;
    a.bc = h->p.bc;
    a.mode = mode;
    a.quirks = h->p.ref_quirks ? 1 : 0;
    a.mixed = (h->p.precision == HTX_PREC_FP32_MIXED || is_learned(h->p.precision) || is_bulk32(h->p.precision)) ? 1 : 0;
    a.learned = is_learned(h->p.precision) ? 1 : 0;
    a.bulk32 = is_bulk32(h->p.precision) ? 1 : 0;
    a.recon = htx_recon_code(h->p.scheme);
    a.face_flags_x = h->dFlagsX;
    a.face_flags_y = h->dFlagsY;
    a.gam = h->p.gamma;
    a.g1 = h->p.gamma - 1.0;
    a.dx = h->dx;
    a.dy = h->dy;
    a.area = h->dx * h->dy;
    a.qc = h->qc;
    return a;
}

int flux_kind(const htx_solver* h) {
    if (h->p.flux == HTX_FLUX_RUSANOV_CHAR) return FLUX_RUSANOV_CHAR;
    if (h->p.flux == HTX_FLUX_CU) return FLUX_CU;
    return h->p.ref_quirks ? FLUX_HLLC : FLUX_HLLC_TORO;
}

FvhArgs hermite_args(htx_solver* h, const double* qin, const double* q0, double* qout, int mode) {
    FvhArgs a{};
    a.q = qin;
    a.lv = h->dLV;
    a.lv_out = h->dLV;
    a.fx = h->dFX;
    a.fy = h->dFY;
    a.gx = h->dGX;
    a.gy = h->dGY;
    a.q0 = q0;
    a.qout = qout;
    a.ctl = h->ctl;
    a.nx = h->nx;
    a.ny = h->ny;
    a.pitch = h->pitch;
    a.fstride = h->fstride;
    a.bc = h->p.bc;
    a.mode = mode;
    a.quirks = h->p.ref_quirks ? 1 : 0;
    a.flux = h->p.flux;
    a.mixed = (h->p.precision == HTX_PREC_FP32_MIXED || h->p.precision == HTX_PREC_FP32) ? 1 : 0;
    a.bulk32 = h->p.precision == HTX_PREC_FP32 ? 1 : 0;
    a.hrecon = htx_hrecon_code(h->p.scheme);
    a.gam = h->p.gamma;
    a.g1 = h->p.gamma - 1.0;
    a.dx = h->dx;
    a.dy = h->dy;
    a.area = h->dx * h->dy;
    a.qc = h->qc;
    a.hc = h->hc;
    return a;
}

int launch_stage(htx_solver* h, const double* qin, const double* q0, double* qout, int mode) {
    if (is_learned(h->p.precision) && !h->have_classifier)
        return fail(HTX_ESTATE, "precision 'learned' needs a classifier (htx_load_classifier)");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (h->timing) {
        if (!h->event_pool.empty()) {
            e0 = h->event_pool.back().first;
            e1 = h->event_pool.back().second;
            h->event_pool.pop_back();
        } else {
            HIP_TRY(hipEventCreate(&e0));
            HIP_TRY(hipEventCreate(&e1));
        }
        HIP_TRY(hipEventRecord(e0, h->stream));
    }
    const bool strict = h->p.precision == HTX_PREC_FP64;
    if (is_learned(h->p.precision) && mode <= 1) {
        // classify every face of u^n once per step; stages 2 and 3 reuse the flags
        LrnFlagsArgs f{};
        f.q = qin;
        f.nx = h->nx;
        f.ny = h->ny;
        f.pitch = h->pitch;
        f.bc = h->p.bc;
        f.fstride = h->fstride;
        f.gam = h->p.gamma;
        f.frag = h->dFrag;
        f.cvec = h->dCvec;
        f.threshold = h->cls_threshold;
        f.ctl = mode == 1 ? h->ctl : nullptr;
        f.flags_x = h->dFlagsX;
        f.flags_y = h->dFlagsY;
        f.bulk32 = h->p.precision == HTX_PREC_LEARNED_FP32;
        HIP_TRY(lrn_launch_flags(f, h->stream));
    }
    if (h->p.dim == 1) {
        Fv1dArgs a{};
        a.qin = qin;
        a.q0 = q0;
        a.qout = qout;
        a.ctl = h->ctl;
        a.nx = h->nx;
        a.bc = h->p.bc;
        a.mode = mode;
        a.recon = htx_recon_code(h->p.scheme);
        a.flux = flux_kind(h);
        a.gam = h->p.gamma;
        a.g1 = h->p.gamma - 1.0;
        a.dx = h->dx;
        HIP_TRY(htx_strict_launch_fv1d_stage(&a, h->stream));
    } else if (htx_is_hermite(h->p.scheme)) {
        FvhArgs a = hermite_args(h, qin, q0, qout, mode);
        if (strict)
            HIP_TRY(htx_strict_launch_fvh2d_stage(&a, flux_kind(h), h->stream));
        else
            HIP_TRY(launch_fvh2d_stage(a, flux_kind(h), h->stream));
    } else {
        Fv2dArgs a = stage_args(h, qin, q0, qout, mode);
        if (strict)
            HIP_TRY(htx_strict_launch_fv2d_stage(&a, flux_kind(h), h->p.tile_x, h->p.tile_y, h->stream));
        else
            HIP_TRY(launch_fv2d_stage(a, flux_kind(h), h->p.tile_x, h->p.tile_y, h->stream));
    }
    if (h->timing) {
        HIP_TRY(hipEventRecord(e1, h->stream));
        h->events.emplace_back(e0, e1);
    }
    return HTX_OK;
}

int enqueue_step(htx_solver* h) {
    h->next_stage = 1;
    HIP_TRY(launch_dt(h->ctl, h->dx, h->dy, h->stream));
    int rc;
    if ((rc = launch_stage(h, h->dA, h->dA, h->dB, 1))) return rc;
    if ((rc = launch_stage(h, h->dB, h->dA, h->dC, 2))) return rc;
    if ((rc = launch_stage(h, h->dC, h->dA, h->dA, 3))) return rc;
    HIP_TRY(launch_tadvance(h->ctl, h->stream));
    return HTX_OK;
}

int read_ctl(htx_solver* h, Ctl* c) {
    HIP_TRY(hipMemcpyAsync(c, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return HTX_OK;
}

// HTX_ENONFINITE when the control block records a failure (a non-finite
// conserved value met by a stage-3 epilogue or the CFL kernel, or a
// non-finite / non-positive dt).  The reference has no such check: its CFL
// loop takes fmax over the cells, which skips NaN, and it steps on.
int status_error(const Ctl& c) {
    if (!c.status) return HTX_OK;
    char buf[200];
    std::snprintf(buf, sizeof buf, "%s%s at t = %.17g after %d steps",
                  (c.status & CTL_NONFINITE_STATE) ? "the state holds a NaN or infinite value" : "",
                  (c.status & CTL_NONFINITE_STATE) ? "" : "the CFL time step is not finite and positive", c.t, c.nsteps);
    return fail(HTX_ENONFINITE, buf);
}

int read_ctl_checked(htx_solver* h, Ctl* c) {
    int rc = read_ctl(h, c);
    return rc ? rc : status_error(*c);
}

int refresh_cfl(htx_solver* h) {
    HIP_TRY(launch_reset_smax(h->ctl, h->stream));
    if (h->p.dim == 1) {
        HIP_TRY(htx_strict_launch_fv1d_max_speed(h->dA, h->nx, h->p.gamma, h->ctl, h->stream));
        return HTX_OK;
    }
    // (the 12-field Hermite state's first four fields are the cell averages)
    if (h->p.precision == HTX_PREC_FP64)
        HIP_TRY(htx_strict_launch_max_speed(h->dA, h->nx, h->ny, h->pitch, h->fstride, h->p.gamma,
                                            h->p.ref_quirks ? 1 : 0, h->ctl, h->stream));
    else
        HIP_TRY(launch_max_speed(h->dA, h->nx, h->ny, h->pitch, h->fstride, h->p.gamma, h->p.ref_quirks ? 1 : 0,
                                 h->ctl, h->stream));
    return HTX_OK;
}

}  // namespace

extern "C" {

int htx_abi_version(void) { return HTX_ABI_VERSION; }

const char* htx_last_error(void) { return g_err.c_str(); }

int htx_params_default(htx_params* p) {
    if (!p) return fail(HTX_EINVAL, "params is null");
    std::memset(p, 0, sizeof(*p));
    p->struct_size = (int)sizeof(htx_params);
    p->dim = 2;
    p->xl = 0.0;
    p->xr = 1.0;
    p->nx = 64;
    p->yb = 0.0;
    p->yt = 1.0;
    p->ny = 64;
    p->cfl = 0.5;
    p->tmax = 1.0;
    p->tinit = 0.0;
    p->gamma = 1.4;
    p->scheme = HTX_SCHEME_TENO5;
    p->flux = HTX_FLUX_HLLC;
    p->bc = HTX_BC_ZERO_GRAD;
    p->precision = HTX_PREC_FP64;
    p->ref_quirks = 1;
    return HTX_OK;
}

int htx_create(const htx_params* p, htx_solver** out) {
    if (!out) return fail(HTX_EINVAL, "out is null");
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(HTX_EDEVICE, "no HIP device");
    htx_solver* h = new htx_solver();
    h->p = *p;
    h->nx = p->nx;
    h->ny = p->dim == 1 ? 1 : p->ny;
    h->nfield = p->dim == 1 ? 3 : (htx_is_hermite(p->scheme) ? 12 : 4);
    h->pitch = p->nx;
    h->fstride = (long long)h->pitch * h->ny;
    h->dx = (p->xr - p->xl) / p->nx;
    // 1D: dt = cfl dx / max speed (FV_System1D.cpp:277) through the 2D rule
    // cfl min(dx, dy) / max speed with dy = dx
    h->dy = p->dim == 1 ? h->dx : (p->yt - p->yb) / p->ny;
    fill_quad_consts(h->qc);
    if (!check_quad_consts()) {
        delete h;
        return fail(HTX_EDEVICE, "the kernels' compile-time quadrature table differs from the host's sqrt(15) "
                                 "expressions (host math library or compiler rounding)");
    }
    fill_herm_consts(h->hc, h->dx);  // Hermite::setGridSize(dx), FVHermite_System2D.cpp:28
    size_t bytes = sizeof(double) * (size_t)h->fstride * h->nfield;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(HTX_EDEVICE, "hipStreamCreate failed");
    }
    if (hipMalloc(&h->dA, bytes) != hipSuccess || hipMalloc(&h->dB, bytes) != hipSuccess ||
        hipMalloc(&h->dC, bytes) != hipSuccess || hipMalloc(&h->dR, bytes) != hipSuccess ||
        hipMalloc(&h->ctl, sizeof(Ctl)) != hipSuccess) {
        htx_destroy(h);
        return fail(HTX_ENOMEM, "device allocation failed");
    }
    if (is_learned(p->precision) && (hipMalloc(&h->dFlagsX, lrn_flag_bytes_x(p->nx, p->ny)) != hipSuccess ||
                                             hipMalloc(&h->dFlagsY, lrn_flag_bytes_y(p->nx, p->ny)) != hipSuccess)) {
        htx_destroy(h);
        return fail(HTX_ENOMEM, "device allocation failed");
    }
    if (htx_is_hermite(p->scheme)) {
        size_t nxf = (size_t)(p->nx + 1) * p->ny, nyf = (size_t)p->nx * (p->ny + 1);
        if (hipMalloc(&h->dLV, sizeof(double) * 8 * (size_t)h->fstride) != hipSuccess ||
            hipMalloc(&h->dFX, sizeof(double) * 24 * nxf) != hipSuccess ||
            hipMalloc(&h->dFY, sizeof(double) * 24 * nyf) != hipSuccess ||
            hipMalloc(&h->dGX, sizeof(double) * 12 * nxf) != hipSuccess ||
            hipMalloc(&h->dGY, sizeof(double) * 12 * nyf) != hipSuccess) {
            htx_destroy(h);
            return fail(HTX_ENOMEM, "device allocation failed");
        }
    }
    Ctl c{};
    c.t = p->tinit;
    c.tfinal = p->tmax;
    c.cfl = p->cfl;
    if (copy_sync(h->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, h->stream) != hipSuccess) {
        htx_destroy(h);
        return fail(HTX_EDEVICE, "control block upload failed");
    }
    *out = h;
    return HTX_OK;
}

void htx_destroy(htx_solver* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto& e : h->events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (auto& e : h->event_pool) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (h->dA) (void)hipFree(h->dA);
    if (h->dB) (void)hipFree(h->dB);
    if (h->dC) (void)hipFree(h->dC);
    if (h->dR) (void)hipFree(h->dR);
    if (h->ctl) (void)hipFree(h->ctl);
    for (double* b : {h->dLV, h->dFX, h->dFY, h->dGX, h->dGY})
        if (b) (void)hipFree(b);
    if (h->dFrag) (void)hipFree(h->dFrag);
    if (h->dCvec) (void)hipFree(h->dCvec);
    if (h->dFlagsX) (void)hipFree(h->dFlagsX);
    if (h->dFlagsY) (void)hipFree(h->dFlagsY);
    if (h->dTroubled) (void)hipFree(h->dTroubled);
    if (h->dLmpe) (void)hipFree(h->dLmpe);
    if (h->dPartial) (void)hipFree(h->dPartial);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int htx_set_state(htx_solver* h, const double* q, size_t ld) {
    if (!h || !q) return fail(HTX_EINVAL, "null argument");
    if (ld < (size_t)h->nx) return fail(HTX_EINVAL, "ld < nx");
    HIP_TRY(hipMemcpy2DAsync(h->dA, sizeof(double) * h->pitch, q, sizeof(double) * ld, sizeof(double) * h->nx,
                             (size_t)h->ny * h->nfield, hipMemcpyHostToDevice, h->stream));
    int rc = refresh_cfl(h);
    if (rc) return rc;
    h->have_state = true;
    h->next_stage = 1;
    h->exact.kind = EXACT_NONE;
    Ctl c;
    return read_ctl_checked(h, &c);
}

int htx_get_state(htx_solver* h, double* q, size_t ld) {
    if (!h || !q) return fail(HTX_EINVAL, "null argument");
    if (ld < (size_t)h->nx) return fail(HTX_EINVAL, "ld < nx");
    if (!h->have_state) return fail(HTX_ESTATE, "no state set");
    HIP_TRY(hipMemcpy2DAsync(q, sizeof(double) * ld, h->dA, sizeof(double) * h->pitch, sizeof(double) * h->nx,
                             (size_t)h->ny * h->nfield, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return HTX_OK;
}

int htx_init_vortex(htx_solver* h, double cx, double cy, const double* modes, int nmodes) {
    if (!h) return fail(HTX_EINVAL, "null handle");
    if (h->p.dim != 2) return fail(HTX_ENOTSUP, "the vortex is a 2D initial condition");
    if (nmodes < 0 || nmodes > 16 || (nmodes > 0 && !modes)) return fail(HTX_EINVAL, "nmodes must be 0..16");
    VortexArgs a{};
    a.q = h->dA;
    a.nx = h->nx;
    a.ny = h->ny;
    a.pitch = h->pitch;
    a.fstride = h->fstride;
    a.xl = h->p.xl;
    a.yb = h->p.yb;
    a.dx = h->dx;
    a.dy = h->dy;
    a.gam = h->p.gamma;
    a.cx = cx;
    a.cy = cy;
    a.nmodes = nmodes;
    for (int m = 0; m < nmodes; ++m)
        for (int k = 0; k < 6; ++k) a.modes[m][k] = modes[6 * m + k];
    gauss9(a.gx, a.gw);
    if (htx_is_hermite(h->p.scheme))
        HIP_TRY(launch_ic_vortex_hermite(a, h->stream));
    else
        HIP_TRY(launch_ic_vortex(a, h->stream));
    int rc = refresh_cfl(h);
    if (rc) return rc;
    h->have_state = true;
    // the vortex rides on the background flow u = v = 1
    h->exact = ExactArgs{};
    h->exact.kind = EXACT_VORTEX;
    h->exact.cx = cx;
    h->exact.cy = cy;
    h->exact.nmodes = nmodes;
    for (int m = 0; m < nmodes; ++m)
        for (int k = 0; k < 6; ++k) h->exact.modes[m][k] = a.modes[m][k];
    h->adv_u = h->adv_v = 1.0;
    Ctl c;
    return read_ctl_checked(h, &c);
}

int htx_init_entropy_wave(htx_solver* h, double rho, double u, double v, double p, double amplitude) {
    if (!h) return fail(HTX_EINVAL, "null handle");
    if (h->p.dim != 2) return fail(HTX_ENOTSUP, "the entropy wave is a 2D initial condition");
    if (!(rho > 0.0) || !(p > 0.0) || !std::isfinite(u) || !std::isfinite(v) || !std::isfinite(amplitude) ||
        !(std::fabs(amplitude) < rho))
        return fail(HTX_EINVAL, "entropy wave needs rho > |amplitude|, p > 0 and finite u, v");
    EntropyArgs a{};
    a.q = h->dA;
    a.nx = h->nx;
    a.ny = h->ny;
    a.pitch = h->pitch;
    a.nfield = h->nfield;
    a.fstride = h->fstride;
    a.xl = h->p.xl;
    a.yb = h->p.yb;
    a.dx = h->dx;
    a.dy = h->dy;
    a.gam = h->p.gamma;
    a.rho = rho;
    a.u = u;
    a.v = v;
    a.p = p;
    a.amp = amplitude;
    gauss9(a.gx, a.gw);
    HIP_TRY(launch_ic_entropy(a, h->stream));
    int rc = refresh_cfl(h);
    if (rc) return rc;
    h->have_state = true;
    h->next_stage = 1;
    h->exact = ExactArgs{};
    h->exact.kind = EXACT_ENTROPY;
    h->exact.rho = rho;
    h->exact.amp = amplitude;
    h->adv_u = u;
    h->adv_v = v;
    Ctl c;
    return read_ctl_checked(h, &c);
}

int htx_l1_density_error(htx_solver* h, double t_shift, double* err) {
    if (!h || !err) return fail(HTX_EINVAL, "null argument");
    if (h->p.dim != 2) return fail(HTX_ENOTSUP, "the exact solutions are 2D");
    if (!h->have_state || h->exact.kind == EXACT_NONE)
        return fail(HTX_ESTATE, "no analytic initial condition set (htx_init_vortex, htx_init_entropy_wave)");
    if (!std::isfinite(t_shift)) return fail(HTX_EINVAL, "t_shift must be finite");
    const int nb = exact_l1_blocks(h->nx, h->ny);
    if (!h->dPartial) HIP_TRY(hipMalloc(&h->dPartial, sizeof(double) * ((size_t)nb + 1)));
    ExactArgs a = h->exact;
    a.q = h->dA;
    a.nx = h->nx;
    a.ny = h->ny;
    a.pitch = h->pitch;
    a.xl = h->p.xl;
    a.yb = h->p.yb;
    a.dx = h->dx;
    a.dy = h->dy;
    a.gam = h->p.gamma;
    a.sx = h->adv_u * t_shift;
    a.sy = h->adv_v * t_shift;
    gauss9(a.gx, a.gw);
    a.partial = h->dPartial;
    a.out = h->dPartial + nb;
    HIP_TRY(launch_exact_l1(a, h->stream));
    double sum = 0.0;
    HIP_TRY(copy_sync(&sum, a.out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    *err = sum / ((double)h->nx * h->ny);
    return HTX_OK;
}

namespace {
PiecewiseArgs piecewise_args(const htx_solver* h, int kind) {
    PiecewiseArgs a{};
    a.q = h->dA;
    a.nx = h->nx;
    a.ny = h->ny;
    a.pitch = h->pitch;
    a.nfield = h->nfield;
    a.fstride = h->fstride;
    a.xl = h->p.xl;
    a.yb = h->p.yb;
    a.dx = h->dx;
    a.dy = h->dy;
    a.gam = h->p.gamma;
    a.kind = kind;
    return a;
}

int finish_ic(htx_solver* h) {
    h->have_state = true;
    h->exact.kind = EXACT_NONE;
    int rc = refresh_cfl(h);
    if (rc) return rc;
    Ctl c;
    return read_ctl_checked(h, &c);
}
}  // namespace

int htx_init_quadrant(htx_solver* h, double xs, double ys, const double* states) {
    if (!h || !states) return fail(HTX_EINVAL, "null argument");
    if (h->p.dim != 2) return fail(HTX_ENOTSUP, "quadrant data is a 2D initial condition");
    PiecewiseArgs a = piecewise_args(h, IC_QUADRANT);
    a.xs = xs;
    a.ys = ys;
    for (int k = 0; k < 4; ++k)
        for (int e = 0; e < 4; ++e) a.st[k][e] = states[4 * k + e];
    HIP_TRY(launch_ic_piecewise(a, h->stream));
    return finish_ic(h);
}

int htx_init_shock_blobs(htx_solver* h, double xs, double ys, double nx_, double ny_, const double* post,
                         const double* ambient, const double* blobs, int nblob) {
    if (!h || !post || !ambient || (nblob > 0 && !blobs)) return fail(HTX_EINVAL, "null argument");
    if (h->p.dim != 2) return fail(HTX_ENOTSUP, "the shock and blobs are a 2D initial condition");
    if (nblob < 0 || nblob > IC_MAX_BLOBS) return fail(HTX_EINVAL, "nblob must be in [0, 64]");
    PiecewiseArgs a = piecewise_args(h, IC_SHOCK_BLOBS);
    a.xs = xs;
    a.ys = ys;
    a.nrm[0] = nx_;
    a.nrm[1] = ny_;
    for (int e = 0; e < 4; ++e) {
        a.st[0][e] = post[e];
        a.st[1][e] = ambient[e];
    }
    a.nblob = nblob;
    for (int b = 0; b < nblob; ++b)
        for (int e = 0; e < 4; ++e) a.blob[b][e] = blobs[4 * b + e];
    HIP_TRY(launch_ic_piecewise(a, h->stream));
    return finish_ic(h);
}

int htx_rhs(htx_solver* h, double* rhs, size_t ld) {
    if (!h || !rhs) return fail(HTX_EINVAL, "null argument");
    if (ld < (size_t)h->nx) return fail(HTX_EINVAL, "ld < nx");
    if (!h->have_state) return fail(HTX_ESTATE, "no state set");
    int rc = launch_stage(h, h->dA, h->dA, h->dR, 0);
    if (rc) return rc;
    HIP_TRY(hipMemcpy2DAsync(rhs, sizeof(double) * ld, h->dR, sizeof(double) * h->pitch, sizeof(double) * h->nx,
                             (size_t)h->ny * h->nfield, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return HTX_OK;
}

int htx_stage(htx_solver* h, int stage, double dt) {
    if (!h) return fail(HTX_EINVAL, "null handle");
    if (!h->have_state) return fail(HTX_ESTATE, "no state set");
    if (stage < 1 || stage > 3) return fail(HTX_EINVAL, "stage must be 1, 2 or 3");
    if (stage != h->next_stage)
        return fail(HTX_ESTATE, "stage " + std::to_string(stage) + " called where stage " +
                                    std::to_string(h->next_stage) + " is due");
    if (!std::isfinite(dt)) return fail(HTX_EINVAL, "dt must be finite");
    HIP_TRY(launch_set_stage_dt(h->ctl, dt, stage == 1 ? 1 : 0, h->stream));
    int rc;
    if (stage == 1)
        rc = launch_stage(h, h->dA, h->dA, h->dB, 1);
    else if (stage == 2)
        rc = launch_stage(h, h->dB, h->dA, h->dC, 2);
    else
        rc = launch_stage(h, h->dC, h->dA, h->dA, 3);
    if (rc) return rc;
    h->next_stage = stage == 3 ? 1 : stage + 1;
    return HTX_OK;
}

int htx_get_stage_state(htx_solver* h, int stage, double* q, size_t ld) {
    if (!h || !q) return fail(HTX_EINVAL, "null argument");
    if (stage != 1 && stage != 2) return fail(HTX_EINVAL, "stage must be 1 or 2");
    if (ld < (size_t)h->nx) return fail(HTX_EINVAL, "ld < nx");
    if (!h->have_state) return fail(HTX_ESTATE, "no state set");
    HIP_TRY(hipMemcpy2DAsync(q, sizeof(double) * ld, stage == 1 ? h->dB : h->dC, sizeof(double) * h->pitch,
                             sizeof(double) * h->nx, (size_t)h->ny * h->nfield, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return HTX_OK;
}

int htx_compute_dt(htx_solver* h, double* dt) {
    if (!h || !dt) return fail(HTX_EINVAL, "null argument");
    if (!h->have_state) return fail(HTX_ESTATE, "no state set");
    Ctl c;
    int rc = read_ctl_checked(h, &c);
    if (rc) return rc;
    double sx = 0, sy = 0;
    std::memcpy(&sx, &c.smax_x, sizeof(double));
    std::memcpy(&sy, &c.smax_y, sizeof(double));
    double d = c.cfl * std::fmin(h->dx, h->dy) / std::fmax(sx, sy);
    d = (c.t + d > c.tfinal) ? c.tfinal - c.t : d;
    if (c.t < c.tfinal && !(d > 0.0 && std::isfinite(d)))
        return fail(HTX_ENONFINITE, "the CFL time step is not finite and positive");
    *dt = d;
    return HTX_OK;
}

int htx_step(htx_solver* h, double* dt_taken) {
    if (!h) return fail(HTX_EINVAL, "null handle");
    if (!h->have_state) return fail(HTX_ESTATE, "no state set");
    int rc = enqueue_step(h);
    if (rc) return rc;
    Ctl c;
    if ((rc = read_ctl_checked(h, &c))) return rc;
    if (dt_taken) *dt_taken = c.active ? c.dt : 0.0;
    return HTX_OK;
}

int htx_steps_async(htx_solver* h, int n) {
    if (!h) return fail(HTX_EINVAL, "null handle");
    if (!h->have_state) return fail(HTX_ESTATE, "no state set");
    if (n < 0) return fail(HTX_EINVAL, "n < 0");
    for (int k = 0; k < n; ++k) {
        int rc = enqueue_step(h);
        if (rc) return rc;
    }
    return HTX_OK;
}

int htx_advance(htx_solver* h, double tfinal, int* nsteps) {
    if (!h) return fail(HTX_EINVAL, "null handle");
    if (!h->have_state) return fail(HTX_ESTATE, "no state set");
    Ctl c;
    int rc = read_ctl_checked(h, &c);
    if (rc) return rc;
    int n0 = c.nsteps;
    c.tfinal = tfinal;
    HIP_TRY(hipMemcpyAsync(&h->ctl->tfinal, &c.tfinal, sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->p.tmax = tfinal;
    // Steps past tfinal are skipped on the device, so batches may overshoot.
    int batch = 4;
    while (c.t < tfinal) {
        for (int k = 0; k < batch; ++k)
            if ((rc = enqueue_step(h))) return rc;
        if ((rc = read_ctl_checked(h, &c))) return rc;
        if (batch < 64) batch *= 2;
    }
    if (nsteps) *nsteps = c.nsteps - n0;
    return HTX_OK;
}

int htx_get_time(htx_solver* h, double* t, int* nsteps) {
    if (!h) return fail(HTX_EINVAL, "null handle");
    Ctl c;
    int rc = read_ctl(h, &c);
    if (rc) return rc;
    if (t) *t = c.t;
    if (nsteps) *nsteps = c.nsteps;
    return HTX_OK;
}

// Snapshot file, little-endian:
//   char magic[8] "HTXSNAP1"; int32 version (1);
//   int32 nfield, nx, ny, scheme, flux, bc, precision, ref_quirks;
//   double xl, xr, yb, yt, cfl, tmax, tinit, gamma, t; int64 nsteps;
//   double q[nfield][ny][nx] (field-major, interior cells);
//   uint64 checksum of q (snap_checksum).
namespace {
constexpr char SNAP_MAGIC[8] = {'H', 'T', 'X', 'S', 'N', 'A', 'P', '1'};
constexpr int SNAP_VERSION = 1;

// checksum of the state words w_i: sum_i w_i (2 i + 1) mod 2^64 (an odd
// weight per position: any single changed word or bit changes it)
unsigned long long snap_checksum(const double* q, size_t n) {
    unsigned long long s = 0;
    for (size_t i = 0; i < n; ++i) {
        unsigned long long w;
        std::memcpy(&w, q + i, sizeof(w));
        s += w * (2ull * i + 1ull);
    }
    return s;
}

struct SnapHeader {
    int32_t nfield, nx, ny, scheme, flux, bc, precision, ref_quirks;
    double xl, xr, yb, yt, cfl, tmax, tinit, gamma, t;
    int64_t nsteps;
};

struct FileCloser {
    FILE* f;
    ~FileCloser() {
        if (f) std::fclose(f);
    }
};
}  // namespace

int htx_save(htx_solver* h, const char* path) {
    if (!h || !path) return fail(HTX_EINVAL, "null argument");
    if (!h->have_state) return fail(HTX_ESTATE, "no state set");
    std::vector<double> q((size_t)h->nfield * h->nx * h->ny);
    int rc = htx_get_state(h, q.data(), (size_t)h->nx);
    if (rc) return rc;
    double t = 0.0;
    int nsteps = 0;
    if ((rc = htx_get_time(h, &t, &nsteps))) return rc;
    SnapHeader hd{};
    hd.nfield = h->nfield;
    hd.nx = h->nx;
    hd.ny = h->ny;
    hd.scheme = h->p.scheme;
    hd.flux = h->p.flux;
    hd.bc = h->p.bc;
    hd.precision = h->p.precision;
    hd.ref_quirks = h->p.ref_quirks;
    hd.xl = h->p.xl;
    hd.xr = h->p.xr;
    hd.yb = h->p.yb;
    hd.yt = h->p.yt;
    hd.cfl = h->p.cfl;
    hd.tmax = h->p.tmax;
    hd.tinit = h->p.tinit;
    hd.gamma = h->p.gamma;
    hd.t = t;
    hd.nsteps = nsteps;
    const int32_t ver = SNAP_VERSION;
    const unsigned long long hsh = snap_checksum(q.data(), q.size());
    FileCloser fc{std::fopen(path, "wb")};
    if (!fc.f) return fail(HTX_EIO, std::string("cannot open ") + path + " for writing");
    bool ok = std::fwrite(SNAP_MAGIC, 1, 8, fc.f) == 8 && std::fwrite(&ver, sizeof(ver), 1, fc.f) == 1 &&
              std::fwrite(&hd.nfield, sizeof(int32_t), 8, fc.f) == 8 && std::fwrite(&hd.xl, sizeof(double), 9, fc.f) == 9 &&
              std::fwrite(&hd.nsteps, sizeof(int64_t), 1, fc.f) == 1 &&
              std::fwrite(q.data(), sizeof(double), q.size(), fc.f) == q.size() &&
              std::fwrite(&hsh, sizeof(hsh), 1, fc.f) == 1;
    ok = (std::fclose(fc.f) == 0) && ok;
    fc.f = nullptr;
    if (!ok) return fail(HTX_EIO, std::string("write to ") + path + " failed");
    return HTX_OK;
}

int htx_load(htx_solver* h, const char* path) {
    if (!h || !path) return fail(HTX_EINVAL, "null argument");
    FileCloser fc{std::fopen(path, "rb")};
    if (!fc.f) return fail(HTX_EIO, std::string("cannot open ") + path);
    char magic[8];
    int32_t ver = 0;
    SnapHeader hd{};
    if (std::fread(magic, 1, 8, fc.f) != 8 || std::memcmp(magic, SNAP_MAGIC, 8) != 0)
        return fail(HTX_EIO, std::string(path) + " is not a snapshot");
    if (std::fread(&ver, sizeof(ver), 1, fc.f) != 1 || ver != SNAP_VERSION)
        return fail(HTX_EIO, "unsupported snapshot version");
    if (std::fread(&hd.nfield, sizeof(int32_t), 8, fc.f) != 8 || std::fread(&hd.xl, sizeof(double), 9, fc.f) != 9 ||
        std::fread(&hd.nsteps, sizeof(int64_t), 1, fc.f) != 1)
        return fail(HTX_EIO, "truncated snapshot header");
    if (hd.nfield != h->nfield || hd.nx != h->nx || hd.ny != h->ny || hd.scheme != h->p.scheme)
        return fail(HTX_EINVAL, "snapshot grid or scheme differs from the solver's");
    // the physics of the run must be the solver's; precision, CFL number and
    // end time belong to how it is continued and may differ
    {
        const char* field = nullptr;
        if (hd.bc != h->p.bc) field = "bc";
        else if (hd.flux != h->p.flux) field = "flux";
        else if (hd.ref_quirks != h->p.ref_quirks) field = "ref_quirks";
        else if (hd.gamma != h->p.gamma) field = "gamma";
        else if (hd.xl != h->p.xl || hd.xr != h->p.xr) field = "domain (xl, xr)";
        else if (h->p.dim == 2 && (hd.yb != h->p.yb || hd.yt != h->p.yt)) field = "domain (yb, yt)";
        if (field) return fail(HTX_EINVAL, std::string("snapshot ") + field + " differs from the solver's");
    }
    std::vector<double> q((size_t)h->nfield * h->nx * h->ny);
    unsigned long long hsh = 0;
    if (std::fread(q.data(), sizeof(double), q.size(), fc.f) != q.size() || std::fread(&hsh, sizeof(hsh), 1, fc.f) != 1)
        return fail(HTX_EIO, "truncated snapshot");
    if (hsh != snap_checksum(q.data(), q.size())) return fail(HTX_EIO, "snapshot checksum mismatch");
    int rc = htx_set_state(h, q.data(), (size_t)h->nx);
    if (rc) return rc;
    Ctl c;
    if ((rc = read_ctl(h, &c))) return rc;
    c.t = hd.t;
    c.nsteps = (int)hd.nsteps;
    HIP_TRY(copy_sync(h->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
    return HTX_OK;
}

int htx_sync(htx_solver* h) {
    if (!h) return fail(HTX_EINVAL, "null handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (!h->have_state) return HTX_OK;
    Ctl c;
    return read_ctl_checked(h, &c);
}

int htx_timing_enable(htx_solver* h, int enable) {
    if (!h) return fail(HTX_EINVAL, "null handle");
    h->timing = enable != 0;
    return HTX_OK;
}

int htx_timing_read(htx_solver* h, double* kernel_ms, int* nlaunch) {
    if (!h) return fail(HTX_EINVAL, "null handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    double total = 0.0;
    for (auto& e : h->events) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
        total += ms;
        h->event_pool.push_back(e);
    }
    if (kernel_ms) *kernel_ms = total;
    if (nlaunch) *nlaunch = (int)h->events.size();
    h->events.clear();
    return HTX_OK;
}


// ---------------------------------------------------------------------------
// learned troubled-face classifier: training set, training, inference

}  // extern "C"

struct htx_trainer {
    long long capacity = 0, n = 0;
    float* dFeat = nullptr;
    unsigned char* dLabel = nullptr;
    float *dPrm = nullptr, *dM = nullptr, *dV = nullptr, *dPartial = nullptr;
    unsigned long long* dCount = nullptr;  // [0] smooth labels, [1..4] confusion counts
    double* dLoss = nullptr;
    int partial_rows = 0;
    long long adam_t = 0;
    hipStream_t stream = nullptr;
};

namespace {

unsigned long long splitmix(unsigned long long& s) {
    s += 0x9E3779B97F4A7C15ull;
    unsigned long long z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double splitmix_uniform(unsigned long long& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

unsigned long long gcd_u64(unsigned long long a, unsigned long long b) {
    while (b) {
        unsigned long long t = a % b;
        a = b;
        b = t;
    }
    return a;
}

constexpr int TRAIN_CHUNKS = 4;            // chunks of 32 samples per workgroup and minibatch
constexpr int TRAIN_SAMPLES_PER_BLOCK = 32 * TRAIN_CHUNKS;

}  // namespace

extern "C" {

int htx_trainer_create(long long capacity, htx_trainer** out) {
    if (!out) return fail(HTX_EINVAL, "out is null");
    *out = nullptr;
    if (capacity <= 0 || capacity > (1ll << 31)) return fail(HTX_EINVAL, "capacity must be in [1, 2^31]");
    htx_trainer* t = new htx_trainer();
    t->capacity = capacity;
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess) {
        delete t;
        return fail(HTX_EDEVICE, "hipStreamCreate failed");
    }
    if (hipMalloc(&t->dFeat, sizeof(float) * CLS_IN * (size_t)capacity) != hipSuccess ||
        hipMalloc(&t->dLabel, (size_t)capacity) != hipSuccess ||
        hipMalloc(&t->dPrm, sizeof(float) * CLS_NPARAM) != hipSuccess ||
        hipMalloc(&t->dM, sizeof(float) * CLS_NPARAM) != hipSuccess ||
        hipMalloc(&t->dV, sizeof(float) * CLS_NPARAM) != hipSuccess ||
        hipMalloc(&t->dCount, sizeof(unsigned long long) * 8) != hipSuccess ||
        hipMalloc(&t->dLoss, sizeof(double)) != hipSuccess) {
        htx_trainer_destroy(t);
        return fail(HTX_ENOMEM, "device allocation failed");
    }
    if (zero_sync(t->dCount, sizeof(unsigned long long) * 8, t->stream) != hipSuccess ||
        zero_sync(t->dPrm, sizeof(float) * CLS_NPARAM, t->stream) != hipSuccess) {
        htx_trainer_destroy(t);
        return fail(HTX_EDEVICE, "device initialisation failed");
    }
    *out = t;
    return HTX_OK;
}

void htx_trainer_destroy(htx_trainer* t) {
    if (!t) return;
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    for (void* b : {(void*)t->dFeat, (void*)t->dLabel, (void*)t->dPrm, (void*)t->dM, (void*)t->dV,
                    (void*)t->dPartial, (void*)t->dCount, (void*)t->dLoss})
        if (b) (void)hipFree(b);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
}

int htx_trainer_add_state(htx_trainer* t, htx_solver* h) {
    if (!t || !h) return fail(HTX_EINVAL, "null handle");
    if (!h->have_state) return fail(HTX_ESTATE, "the solver has no state");
    if (h->p.scheme != HTX_SCHEME_TENO5 || h->p.dim != 2)
        return fail(HTX_ENOTSUP, "training faces come from the 2D TENO5 scheme");
    long long nf = 2ll * h->nx * h->ny;
    if (t->n + nf > t->capacity) return fail(HTX_ENOMEM, "training set capacity exceeded");
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int dir = 0; dir < 2; ++dir) {
        LrnExtractArgs a{};
        a.q = h->dA;
        a.nx = h->nx;
        a.ny = h->ny;
        a.pitch = h->pitch;
        a.bc = h->p.bc;
        a.fstride = h->fstride;
        a.gam = h->p.gamma;
        a.feat = t->dFeat;
        a.label = t->dLabel;
        a.base = t->n + (long long)dir * h->nx * h->ny;
        a.nsmooth = t->dCount;
        HIP_TRY(lrn_launch_extract(a, dir, t->stream));
    }
    HIP_TRY(hipStreamSynchronize(t->stream));
    t->n += nf;
    return HTX_OK;
}

int htx_trainer_size(htx_trainer* t, long long* n, long long* n_smooth) {
    if (!t) return fail(HTX_EINVAL, "null handle");
    unsigned long long c = 0;
    HIP_TRY(copy_sync(&c, t->dCount, sizeof c, hipMemcpyDeviceToHost, t->stream));
    if (n) *n = t->n;
    if (n_smooth) *n_smooth = (long long)c;
    return HTX_OK;
}

int htx_trainer_read(htx_trainer* t, long long start, long long n, float* feat, unsigned char* label) {
    if (!t) return fail(HTX_EINVAL, "null handle");
    if (start < 0 || n < 0 || start + n > t->n) return fail(HTX_EINVAL, "range outside the training set");
    if (feat)
        HIP_TRY(copy_sync(feat, t->dFeat + start * CLS_IN, sizeof(float) * CLS_IN * n, hipMemcpyDeviceToHost, t->stream));
    if (label) HIP_TRY(copy_sync(label, t->dLabel + start, (size_t)n, hipMemcpyDeviceToHost, t->stream));
    return HTX_OK;
}

int htx_trainer_init(htx_trainer* t, unsigned long long seed) {
    if (!t) return fail(HTX_EINVAL, "null handle");
    std::vector<float> prm(CLS_NPARAM, 0.0f);
    unsigned long long s = seed;
    auto fill = [&](int off, int fan_in, int fan_out) {
        double lim = std::sqrt(6.0 / (fan_in + fan_out));
        for (int k = 0; k < fan_in * fan_out; ++k) prm[off + k] = (float)((2.0 * splitmix_uniform(s) - 1.0) * lim);
    };
    fill(CLS_W1, CLS_IN, CLS_H);
    fill(CLS_W2, CLS_H, CLS_H);
    fill(CLS_W3, CLS_H, 1);
    HIP_TRY(copy_sync(t->dPrm, prm.data(), sizeof(float) * CLS_NPARAM, hipMemcpyHostToDevice, t->stream));
    HIP_TRY(zero_sync(t->dM, sizeof(float) * CLS_NPARAM, t->stream));
    HIP_TRY(zero_sync(t->dV, sizeof(float) * CLS_NPARAM, t->stream));
    t->adam_t = 0;
    return HTX_OK;
}

int htx_trainer_set_params(htx_trainer* t, const float* params) {
    if (!t || !params) return fail(HTX_EINVAL, "null argument");
    HIP_TRY(copy_sync(t->dPrm, params, sizeof(float) * CLS_NPARAM, hipMemcpyHostToDevice, t->stream));
    return HTX_OK;
}

int htx_trainer_get_params(htx_trainer* t, float* params) {
    if (!t || !params) return fail(HTX_EINVAL, "null argument");
    HIP_TRY(copy_sync(params, t->dPrm, sizeof(float) * CLS_NPARAM, hipMemcpyDeviceToHost, t->stream));
    return HTX_OK;
}

int htx_trainer_gradient(htx_trainer* t, double troubled_weight, float* grad, double* loss) {
    if (!t || !grad) return fail(HTX_EINVAL, "null argument");
    if (t->n < 1) return fail(HTX_ESTATE, "the training set is empty");
    if (!(troubled_weight > 0.0)) return fail(HTX_EINVAL, "troubled_weight must be positive");
    unsigned long long nsm = 0;
    HIP_TRY(copy_sync(&nsm, t->dCount, sizeof nsm, hipMemcpyDeviceToHost, t->stream));
    double fpos = (double)nsm / (double)t->n, fneg = 1.0 - fpos;
    if (fpos <= 0.0 || fneg <= 0.0) return fail(HTX_ESTATE, "the training set holds only one class");
    int nblk = (int)((t->n + TRAIN_SAMPLES_PER_BLOCK - 1) / TRAIN_SAMPLES_PER_BLOCK);
    if (nblk > t->partial_rows) {
        if (t->dPartial) HIP_TRY(hipFree(t->dPartial));
        t->dPartial = nullptr;
        if (hipMalloc(&t->dPartial, sizeof(float) * (CLS_NPARAM + 1) * (size_t)nblk) != hipSuccess)
            return fail(HTX_ENOMEM, "device allocation failed");
        t->partial_rows = nblk;
    }
    float* dG = nullptr;
    if (hipMalloc(&dG, sizeof(float) * CLS_NPARAM) != hipSuccess) return fail(HTX_ENOMEM, "device allocation failed");
    double wpos = 0.5 / fpos, wneg = 0.5 / fneg * troubled_weight;
    double norm = 1.0 / (fpos * wpos + fneg * wneg) / (double)t->n;
    LrnTrainArgs a{};
    a.feat = t->dFeat;
    a.label = t->dLabel;
    a.n = t->n;
    a.prm = t->dPrm;
    a.partial = t->dPartial;
    a.k0 = 0;
    a.kend = t->n;
    a.pa = 1;
    a.pb = 0;
    a.wpos = (float)(wpos * norm);
    a.wneg = (float)(wneg * norm);
    a.chunks = TRAIN_CHUNKS;
    LrnUpdateArgs u{};
    u.partial = t->dPartial;
    u.nblk = nblk;
    u.loss = t->dLoss;
    u.grad_out = dG;
    hipError_t e = hipMemsetAsync(t->dLoss, 0, sizeof(double), t->stream);
    if (e == hipSuccess) e = lrn_launch_train(a, nblk, u, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(grad, dG, sizeof(float) * CLS_NPARAM, hipMemcpyDeviceToHost, t->stream);
    double l = 0.0;
    if (e == hipSuccess) e = hipMemcpyAsync(&l, t->dLoss, sizeof l, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    (void)hipFree(dG);
    if (e != hipSuccess) return fail(HTX_EDEVICE, std::string("gradient: ") + hipGetErrorString(e));
    if (loss) *loss = l;
    return HTX_OK;
}

int htx_trainer_train(htx_trainer* t, int epochs, int batch, double lr, double troubled_weight,
                      unsigned long long seed, double* epoch_loss) {
    if (!t) return fail(HTX_EINVAL, "null handle");
    if (t->n < 1) return fail(HTX_ESTATE, "the training set is empty");
    if (epochs < 0 || batch < TRAIN_SAMPLES_PER_BLOCK || !(lr > 0.0) || !(troubled_weight > 0.0))
        return fail(HTX_EINVAL, "epochs >= 0, batch >= 128, lr > 0 and troubled_weight > 0 required");
    unsigned long long nsm = 0;
    HIP_TRY(copy_sync(&nsm, t->dCount, sizeof nsm, hipMemcpyDeviceToHost, t->stream));
    double fpos = (double)nsm / (double)t->n, fneg = 1.0 - fpos;
    if (fpos <= 0.0 || fneg <= 0.0) return fail(HTX_ESTATE, "the training set holds only one class");
    int nblk = (batch + TRAIN_SAMPLES_PER_BLOCK - 1) / TRAIN_SAMPLES_PER_BLOCK;
    int rows = nblk;
    if (rows > t->partial_rows) {
        if (t->dPartial) HIP_TRY(hipFree(t->dPartial));
        t->dPartial = nullptr;
        if (hipMalloc(&t->dPartial, sizeof(float) * (CLS_NPARAM + 1) * (size_t)rows) != hipSuccess)
            return fail(HTX_ENOMEM, "device allocation failed");
        t->partial_rows = rows;
    }
    long long bsz = (long long)nblk * TRAIN_SAMPLES_PER_BLOCK;
    // balanced classes, troubled faces weighted up; weights sum to about 1 per minibatch
    double wpos = 0.5 / fpos, wneg = 0.5 / fneg * troubled_weight;
    double norm = 1.0 / (fpos * wpos + fneg * wneg) / (double)bsz;
    unsigned long long rs = seed;
    const double b1 = 0.9, b2 = 0.999;
    for (int ep = 0; ep < epochs; ++ep) {
        unsigned long long n = (unsigned long long)t->n, pa;
        do {
            pa = 1 + splitmix(rs) % n;
        } while (n > 1 && gcd_u64(pa, n) != 1);
        unsigned long long pb = splitmix(rs) % n;
        HIP_TRY(hipMemsetAsync(t->dLoss, 0, sizeof(double), t->stream));
        long long nbatch = 0;
        for (long long k0 = 0; k0 < t->n; k0 += bsz) {
            ++t->adam_t;
            LrnTrainArgs a{};
            a.feat = t->dFeat;
            a.label = t->dLabel;
            a.n = t->n;
            a.prm = t->dPrm;
            a.partial = t->dPartial;
            a.k0 = k0;
            a.kend = k0 + bsz < t->n ? k0 + bsz : t->n;
            a.pa = pa;
            a.pb = pb;
            a.wpos = (float)(wpos * norm);
            a.wneg = (float)(wneg * norm);
            a.chunks = TRAIN_CHUNKS;
            LrnUpdateArgs u{};
            u.partial = t->dPartial;
            u.nblk = nblk;
            u.prm = t->dPrm;
            u.m = t->dM;
            u.v = t->dV;
            u.loss = t->dLoss;
            u.beta1 = (float)b1;
            u.beta2 = (float)b2;
            u.eps = 1e-8f;
            u.lr_t = (float)(lr * std::sqrt(1.0 - std::pow(b2, (double)t->adam_t)) /
                             (1.0 - std::pow(b1, (double)t->adam_t)));
            HIP_TRY(lrn_launch_train(a, nblk, u, t->stream));
            ++nbatch;
        }
        if (epoch_loss) {
            double l = 0.0;
            HIP_TRY(hipMemcpyAsync(&l, t->dLoss, sizeof l, hipMemcpyDeviceToHost, t->stream));
            HIP_TRY(hipStreamSynchronize(t->stream));
            epoch_loss[ep] = l / (double)nbatch;
        }
    }
    HIP_TRY(hipStreamSynchronize(t->stream));
    return HTX_OK;
}

namespace {
int trainer_eval(htx_trainer* t, double threshold, long long* counts, float* logits, int rounding);
}

int htx_trainer_eval(htx_trainer* t, double threshold, long long* counts, float* logits) {
    return trainer_eval(t, threshold, counts, logits, 0);
}

int htx_trainer_eval_inference(htx_trainer* t, double threshold, long long* counts, float* logits) {
    return trainer_eval(t, threshold, counts, logits, 1);
}

}  // extern "C"

namespace {
int trainer_eval(htx_trainer* t, double threshold, long long* counts, float* logits, int rounding) {
    if (!t || !counts) return fail(HTX_EINVAL, "null argument");
    if (t->n < 1) return fail(HTX_ESTATE, "the training set is empty");
    float* dLog = nullptr;
    if (logits && hipMalloc(&dLog, sizeof(float) * (size_t)t->n) != hipSuccess)
        return fail(HTX_ENOMEM, "device allocation failed");
    HIP_TRY(hipMemsetAsync(t->dCount + 1, 0, sizeof(unsigned long long) * 4, t->stream));
    LrnEvalArgs a{};
    a.feat = t->dFeat;
    a.label = t->dLabel;
    a.n = t->n;
    a.prm = t->dPrm;
    a.threshold = (float)threshold;
    a.counts = t->dCount + 1;
    a.logits = dLog;
    a.inference_rounding = rounding;
    hipError_t e = lrn_launch_eval(a, t->stream);
    unsigned long long c[4] = {0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(c, t->dCount + 1, sizeof c, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && logits)
        e = hipMemcpyAsync(logits, dLog, sizeof(float) * (size_t)t->n, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    if (dLog) (void)hipFree(dLog);
    if (e != hipSuccess) return fail(HTX_EDEVICE, std::string("classifier evaluation: ") + hipGetErrorString(e));
    for (int k = 0; k < 4; ++k) counts[k] = (long long)c[k];
    return HTX_OK;
}
}  // namespace

extern "C" {

int htx_load_classifier(htx_solver* h, const float* params, double threshold) {
    if (!h || !params) return fail(HTX_EINVAL, "null argument");
    for (int k = 0; k < CLS_NPARAM; ++k)
        if (!std::isfinite(params[k])) return fail(HTX_EINVAL, "classifier parameters must be finite");
    std::vector<unsigned short> frag(CLS_NFRAG);
    std::vector<float> cvec(CLS_NCVEC);
    cls_pack(params, frag.data(), cvec.data());
    if (!h->dFrag && (hipMalloc(&h->dFrag, sizeof(unsigned short) * CLS_NFRAG) != hipSuccess ||
                      hipMalloc(&h->dCvec, sizeof(float) * CLS_NCVEC) != hipSuccess))
        return fail(HTX_ENOMEM, "device allocation failed");
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(copy_sync(h->dFrag, frag.data(), sizeof(unsigned short) * CLS_NFRAG, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(copy_sync(h->dCvec, cvec.data(), sizeof(float) * CLS_NCVEC, hipMemcpyHostToDevice, h->stream));
    h->cls_threshold = (float)threshold;
    h->have_classifier = true;
    return HTX_OK;
}

int htx_kxrcf_flags(htx_solver* h, unsigned char* flags) {
    if (!h || !flags) return fail(HTX_EINVAL, "null argument");
    if (h->p.dim != 1) return fail(HTX_ENOTSUP, "the KXRCF detector belongs to the 1D solver");
    if (!h->have_state) return fail(HTX_ESTATE, "the solver has no state");
    if (!h->dTroubled) HIP_TRY(hipMalloc(&h->dTroubled, (size_t)h->nx));
    // circleRadMod of FV_System1D.cpp:19, with the host's pow as there
    const double radius = std::pow(h->dx / 2.0, 1.5);
    HIP_TRY(htx_strict_launch_fv1d_kxrcf(h->dA, h->nx, h->p.bc, radius, h->dTroubled, h->stream));
    HIP_TRY(copy_sync(flags, h->dTroubled, (size_t)h->nx, hipMemcpyDeviceToHost, h->stream));
    return HTX_OK;
}

int htx_lmpe_flags(htx_solver* h, unsigned char* flags) {
    if (!h || !flags) return fail(HTX_EINVAL, "null argument");
    if (h->p.dim != 1) return fail(HTX_ENOTSUP, "the LMPE detector belongs to the 1D solvers");
    if (!h->have_state) return fail(HTX_ESTATE, "the solver has no state");
    if (!h->dLmpe) HIP_TRY(hipMalloc(&h->dLmpe, 3 * (size_t)h->nx));
    HIP_TRY(htx_strict_launch_fv1d_lmpe(h->dA, h->nx, h->p.bc, h->dx, h->dLmpe, h->stream));
    HIP_TRY(copy_sync(flags, h->dLmpe, 3 * (size_t)h->nx, hipMemcpyDeviceToHost, h->stream));
    return HTX_OK;
}

int htx_classifier_flags(htx_solver* h, unsigned char* fx, unsigned char* fy) {
    if (!h || !fx || !fy) return fail(HTX_EINVAL, "null argument");
    if (!is_learned(h->p.precision) || !h->dFlagsX)
        return fail(HTX_ESTATE, "the solver does not run the learned precision");
    HIP_TRY(copy_sync(fx, h->dFlagsX, lrn_flag_bytes_x(h->nx, h->ny), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(copy_sync(fy, h->dFlagsY, lrn_flag_bytes_y(h->nx, h->ny), hipMemcpyDeviceToHost, h->stream));
    return HTX_OK;
}

int htx_classifier_stats(htx_solver* h, long long* n_faces, long long* n_flagged, long long* n_linear, int reset) {
    if (!h) return fail(HTX_EINVAL, "null handle");
    Ctl c{};
    int rc = read_ctl(h, &c);
    if (rc) return rc;
    if (n_faces) *n_faces = (long long)c.n_faces;
    if (n_flagged) *n_flagged = (long long)c.n_flagged;
    if (n_linear) *n_linear = (long long)c.n_linear;
    if (reset) {
        unsigned long long z[3] = {0, 0, 0};
        HIP_TRY(hipMemcpyAsync(&h->ctl->n_faces, z, sizeof z, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    return HTX_OK;
}

int htx_classify(htx_solver* h, int dir, float* logits, float* feat) {
    if (!h || !logits) return fail(HTX_EINVAL, "null argument");
    if (dir != 0 && dir != 1) return fail(HTX_EINVAL, "dir must be 0 (x faces) or 1 (y faces)");
    if (!h->have_classifier) return fail(HTX_ESTATE, "no classifier loaded");
    if (!h->have_state) return fail(HTX_ESTATE, "the solver has no state");
    if (h->p.scheme != HTX_SCHEME_TENO5 || h->p.dim != 2)
        return fail(HTX_ENOTSUP, "the classifier runs on the 2D TENO5 scheme");
    size_t nf = (size_t)h->nx * h->ny;
    float *dL = nullptr, *dF = nullptr;
    if (hipMalloc(&dL, sizeof(float) * nf) != hipSuccess ||
        (feat && hipMalloc(&dF, sizeof(float) * CLS_IN * nf) != hipSuccess)) {
        if (dL) (void)hipFree(dL);
        return fail(HTX_ENOMEM, "device allocation failed");
    }
    LrnClassifyArgs a{};
    a.q = h->dA;
    a.nx = h->nx;
    a.ny = h->ny;
    a.pitch = h->pitch;
    a.bc = h->p.bc;
    a.fstride = h->fstride;
    a.gam = h->p.gamma;
    a.frag = h->dFrag;
    a.cvec = h->dCvec;
    a.logits = dL;
    a.feat = dF;
    hipError_t e = lrn_launch_classify(a, dir, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(logits, dL, sizeof(float) * nf, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && feat)
        e = hipMemcpyAsync(feat, dF, sizeof(float) * CLS_IN * nf, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dL);
    if (dF) (void)hipFree(dF);
    if (e != hipSuccess) return fail(HTX_EDEVICE, std::string("classify: ") + hipGetErrorString(e));
    return HTX_OK;
}

}  // extern "C"
