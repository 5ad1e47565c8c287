// htx_noalloc.hip -- htx_step_noalloc: one SSP-RK3 step of a 2D TENO5
// problem whose state lives in caller-owned device memory, for callers that
// cannot allocate (code under HIP graph capture, serving loops, harnesses
// that forbid hipMalloc).  It runs the kernels of htx_step
// (FV::System2D::timeStepping + RungeKutta3, FV_System2D.cpp:575-651): the
// CFL maxima of the state, the dt kernel, the three fused stage kernels and
// a kernel that hands the dt taken to the caller.  The two intermediate
// stage states and the control block live in one file-scope device array
// (`scratch`) instead of handle-owned buffers; nothing is allocated or freed.
//
// Every call on a GPU shares that GPU's copy of the array and leaves its
// contents behind (the control block is reinitialised by a kernel at the
// start of each call; the stage states are fully overwritten before they
// are read).  Calls that overlap on one GPU (two host threads, or two
// streams without an ordering between them) race on the array.
#include <mutex>
#include <string>

#include "../../include/htx.h"
#include "htx_internal.h"

hipError_t htx_strict_launch_fv2d_stage(const void* a, int flux, int tile_x, int tile_y, hipStream_t s);
hipError_t htx_strict_launch_max_speed(const double* q, int nx, int ny, int pitch, long long fstride, double gam,
                                       int quirks, void* ctl, hipStream_t s);
int htx_internal_fail(int code, const std::string& msg);

namespace {

using namespace htx;

// The control block takes the first 256 bytes; then u^1 and u^2 of at most
// HTX_NOALLOC_MAX_CELLS cells x 4 fields each (64 MiB in all).
constexpr long long CTL_DOUBLES = 32;
static_assert(sizeof(Ctl) <= CTL_DOUBLES * sizeof(double), "control block exceeds its slot");
constexpr long long SCRATCH_DOUBLES = CTL_DOUBLES + 2ll * 4ll * HTX_NOALLOC_MAX_CELLS;

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

// The array's device address, per device (the name `scratch` is a host-side
// stand-in and must never reach a kernel argument).
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_ptr[MAX_DEVICES] = {};

int scratch_on_current_device(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return htx_internal_fail(HTX_EDEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return htx_internal_fail(HTX_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return htx_internal_fail(HTX_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_ptr[dev];
    return HTX_OK;
}

// every field of the control block set, so nothing of the previous call
// carries over (a kernel, not a memset, so that the call can be captured)
__global__ void k_noalloc_ctl(Ctl* c, double t, double tfinal, double cfl) {
    Ctl z{};
    z.t = t;
    z.tfinal = tfinal;
    z.cfl = cfl;
    *c = z;
}

__global__ void k_noalloc_dt_out(const Ctl* c, double* dt) { *dt = c->active ? c->dt : 0.0; }

}  // namespace

extern "C" int htx_step_noalloc(const htx_params* p, double* q, double t, double* dt_taken, void* stream) {
    if (!p || !q) return htx_internal_fail(HTX_EINVAL, "null argument");
    if (p->struct_size != (int)sizeof(htx_params)) return htx_internal_fail(HTX_EINVAL, "params.struct_size mismatch");
    if (p->dim != 2 || p->scheme != HTX_SCHEME_TENO5)
        return htx_internal_fail(HTX_ENOTSUP, "htx_step_noalloc advances 2D TENO5 problems only");
    if (p->precision != HTX_PREC_FP64 && p->precision != HTX_PREC_FP64_FAST && p->precision != HTX_PREC_FP32_MIXED)
        return htx_internal_fail(HTX_ENOTSUP, "htx_step_noalloc runs the fp64, fp64_fast and fp32_mixed precisions");
    if (p->flux != HTX_FLUX_RUSANOV_CHAR && p->flux != HTX_FLUX_HLLC) return htx_internal_fail(HTX_EINVAL, "unknown flux");
    if (p->bc != HTX_BC_ZERO_GRAD && p->bc != HTX_BC_PERIODIC) return htx_internal_fail(HTX_EINVAL, "unknown bc");
    if (p->nx <= 0 || p->ny <= 0) return htx_internal_fail(HTX_EINVAL, "grid size must be positive");
    if ((long long)p->nx * p->ny > HTX_NOALLOC_MAX_CELLS)
        return htx_internal_fail(HTX_EINVAL, "nx * ny = " + std::to_string((long long)p->nx * p->ny) +
                                                 " exceeds HTX_NOALLOC_MAX_CELLS = " +
                                                 std::to_string((long long)HTX_NOALLOC_MAX_CELLS));
    if (!(p->xr > p->xl) || !(p->yt > p->yb)) return htx_internal_fail(HTX_EINVAL, "domain must have positive extent");
    if (!(p->gamma > 1.0) || !(p->cfl > 0.0)) return htx_internal_fail(HTX_EINVAL, "gamma must exceed 1, cfl be positive");
    double* base = nullptr;
    int rc = scratch_on_current_device(&base);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Ctl* ctl = reinterpret_cast<Ctl*>(base);
    const long long fs = (long long)p->nx * p->ny;
    double* u1 = base + CTL_DOUBLES;
    double* u2 = u1 + 4 * fs;
    const double dx = (p->xr - p->xl) / p->nx, dy = (p->yt - p->yb) / p->ny;
    const bool strict = p->precision == HTX_PREC_FP64;
    const int quirks = p->ref_quirks ? 1 : 0;
    const int flux = p->flux == HTX_FLUX_RUSANOV_CHAR ? FLUX_RUSANOV_CHAR : (quirks ? FLUX_HLLC : FLUX_HLLC_TORO);
#define NA_TRY(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return htx_internal_fail(HTX_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
    hipLaunchKernelGGL(k_noalloc_ctl, dim3(1), dim3(1), 0, s, ctl, t, p->tmax, p->cfl);
    NA_TRY(hipGetLastError());
    if (strict)
        NA_TRY(htx_strict_launch_max_speed(q, p->nx, p->ny, p->nx, fs, p->gamma, quirks, ctl, s));
    else
        NA_TRY(launch_max_speed(q, p->nx, p->ny, p->nx, fs, p->gamma, quirks, ctl, s));
    NA_TRY(launch_dt(ctl, dx, dy, s));
    Fv2dArgs a{};
    a.ctl = ctl;
    a.nx = p->nx;
    a.ny = p->ny;
    a.pitch = p->nx;
    a.fstride = fs;
    a.bc = p->bc;
    a.quirks = quirks;
    a.mixed = p->precision == HTX_PREC_FP32_MIXED ? 1 : 0;
    a.gam = p->gamma;
    a.g1 = p->gamma - 1.0;
    a.dx = dx;
    a.dy = dy;
    a.area = dx * dy;
    fill_quad_consts(a.qc);
    const double* in[3] = {q, u1, u2};
    double* out[3] = {u1, u2, q};
    for (int st = 0; st < 3; ++st) {
        a.qin = in[st];
        a.q0 = q;
        a.qout = out[st];
        a.mode = st + 1;
        if (strict)
            NA_TRY(htx_strict_launch_fv2d_stage(&a, flux, p->tile_x, p->tile_y, s));
        else
            NA_TRY(launch_fv2d_stage(a, flux, p->tile_x, p->tile_y, s));
    }
    if (dt_taken) {
        hipLaunchKernelGGL(k_noalloc_dt_out, dim3(1), dim3(1), 0, s, ctl, dt_taken);
        NA_TRY(hipGetLastError());
    }
#undef NA_TRY
    return HTX_OK;
}
