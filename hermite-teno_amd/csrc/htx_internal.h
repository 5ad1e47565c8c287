// htx_internal.h -- shared definitions between the HIP kernels and the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "htx_numerics.h"
#include "htx_variants.h"

namespace HTX_NS {

// Device-resident time-stepping control block.  The host never needs to read
// it between steps: dt comes from the max wave speed that the previous step's
// last stage accumulated, and inactive steps (t >= tfinal) are skipped by the
// stage kernels themselves.
struct Ctl {
    double t, dt, tfinal, cfl;
    unsigned long long smax_x, smax_y;  // bit patterns of non-negative doubles
    int active;
    int nsteps;
    // failure record (bits): CTL_NONFINITE_STATE when a stage-3 epilogue or
    // the CFL kernel met a NaN or infinite conserved value, CTL_BAD_DT when
    // the CFL dt came out non-finite or non-positive (that step and every
    // later one are skipped); cleared only when a new state is set
    unsigned int status;
    // learned precision, each face once per stage since the last reset: faces
    // reconstructed, faces the classifier flagged smooth, flagged faces that
    // kept the linear scheme (the rest failed the positivity check)
    unsigned long long n_faces, n_flagged, n_linear;
};

enum FluxKind { FLUX_RUSANOV_CHAR = 0, FLUX_HLLC = 1, FLUX_HLLC_TORO = 2, FLUX_CU = 3 };  // FLUX_CU: 1D only

enum CtlStatus : unsigned { CTL_NONFINITE_STATE = 1u, CTL_BAD_DT = 2u };

// every component finite (NaN and infinities fail the comparison)
HTX_HD bool all_finite(const double* u, int n) {
    bool ok = true;
    for (int e = 0; e < n; ++e) ok = ok && fabs(u[e]) <= 1.7976931348623157e308;
    return ok;
}

// One atomic per wavefront that met a non-finite value; called by every lane
// of the wavefront.
__device__ __forceinline__ void flag_nonfinite(bool bad, Ctl* c) {
    if (__any(bad ? 1 : 0) && (threadIdx.x & 63) == 0) atomicOr(&c->status, (unsigned)CTL_NONFINITE_STATE);
}

// dt of the next step from the accumulated max wave speeds
// (FV_System2D.cpp:633,645); a step with t >= tFinal is marked inactive and
// every stage kernel skips it.  The speeds are consumed (zeroed for stage 3
// to accumulate afresh) only by an active step: an inactive step computes
// none, so the speeds of the last state carry over to the first step of a
// later advance.
__device__ inline void ctl_dt(Ctl* c, double dx, double dy) {
    double t = c->t, tf = c->tfinal;
    if (t < tf) {
        double sx = __longlong_as_double((long long)c->smax_x);
        double sy = __longlong_as_double((long long)c->smax_y);
        double dt = c->cfl * fmin(dx, dy) / fmax(sx, sy);
        dt = t + dt > tf ? tf - t : dt;
        if (!(dt > 0.0 && dt <= 1.7976931348623157e308) || (c->status & CTL_NONFINITE_STATE)) {
            // a NaN or infinite wave speed or state: stop stepping
            c->status |= CTL_BAD_DT;
            c->dt = 0.0;
            c->active = 0;
            return;
        }
        c->dt = dt;
        c->active = 1;
        c->smax_x = 0ull;
        c->smax_y = 0ull;
    } else {
        c->dt = 0.0;
        c->active = 0;
    }
}

__device__ inline void ctl_tadvance(Ctl* c) {
    if (c->active) {
        c->t += c->dt;
        c->nsteps += 1;
    }
}

struct Fv2dArgs {
    const double* qin;  // state whose residual is taken
    const double* q0;   // u^n (stages 2 and 3)
    double* qout;       // residual (mode 0) or the stage result
    Ctl* ctl;
    int nx, ny, pitch;  // pitch: elements between rows
    long long fstride;  // elements between fields
    int bc;             // 0 zero-gradient, 1 periodic
    int mode;           // 0 residual only, 1..3 SSP-RK3 stage
    int quirks;         // reference maxSpeed quirk in the CFL epilogue
    int mixed;          // 1: cut-off decisions and nonlinear weights in FP32
    int learned;        // 1: classifier flags pick linear or TENO5 per face (implies mixed)
    int bulk32;         // 1: face values, quadrature, back-projection and flux in FP32 (implies mixed)
    int recon;          // characteristic reconstruction: 0 teno5 (the reference's), 1 teno5i, 2 weno5JS, 3 weno5Z,
                        // RECON_WENO5N .. RECON_LAST the variants of recon5_variant (htx_variants.h)
    int tiles_x, tiles_y;
    double gam, g1, dx, dy, area;
    QuadConsts qc;
    // learned precision: the classifier's flags (1 = smooth) of the step's
    // state, x faces i | i+1 (i = -1 .. nx) of row j at face_flags_x[j (nx + 2)
    // + i + 1], y faces j | j+1 (j = -1 .. ny) of column i at
    // face_flags_y[(j + 1) nx + i]
    const unsigned char* face_flags_x;
    const unsigned char* face_flags_y;
};

struct VortexArgs {
    double* q;
    int nx, ny, pitch;
    long long fstride;
    double xl, yb, dx, dy, gam, cx, cy;
    int nmodes;
    double modes[16][6];  // (k, l, a, phi, Lx, Ly)
    double gx[9], gw[9];
};

// Entropy wave (Equations_Euler2D.cpp:430-511): rho = rho_inf + A sin(pi (x
// + y)), u, v, p constant; 9x9 Gauss cell averages of the conserved state
// and, with nfield 12, of its x and y derivatives (the Hermite moments).
struct EntropyArgs {
    double* q;
    int nx, ny, pitch, nfield;
    long long fstride;
    double xl, yb, dx, dy, gam;
    double rho, u, v, p, amp;
    double gx[9], gw[9];
};

// Exact solutions for the L1 density error of printResult
// (FV_System2D.cpp:653-666, FVHermite_System2D.cpp:821-835): the initial
// field of an analytic condition evaluated at (x - sx, y - sy), cell-averaged
// with the 9x9 Gauss rule; |rho - rho_exact| summed per workgroup into
// partial[blockIdx.x] in a fixed order.
enum ExactKind { EXACT_NONE = 0, EXACT_ENTROPY = 1, EXACT_VORTEX = 2 };
struct ExactArgs {
    const double* q;  // density field of the state
    int nx, ny, pitch;
    double xl, yb, dx, dy, gam;
    int kind;
    double sx, sy;
    double rho, amp;               // entropy wave
    double cx, cy;                 // vortex centre
    int nmodes;
    double modes[16][6];           // vortex density modes (k, l, a, phi, Lx, Ly)
    double gx[9], gw[9];
    double* partial;               // [gridDim.x]
    double* out;                   // the sum of the partials
};

enum IcKind { IC_QUADRANT = 0, IC_SHOCK_BLOBS = 1 };
constexpr int IC_MAX_BLOBS = 64;

// Piecewise initial states, point-sampled at cell centres.
//   IC_QUADRANT:    primitive state st[k], k = (x >= xs) + 2 (y >= ys)
//   IC_SHOCK_BLOBS: st[0] where (x - xs, y - ys) . nrm < 0 (post-shock), else
//                   st[1] with rho + sum_b A_b exp(-|x - c_b|^2 / r_b^2),
//                   blob rows (cx, cy, A, r)
struct PiecewiseArgs {
    double* q;
    int nx, ny, pitch, nfield;
    long long fstride;
    double xl, yb, dx, dy, gam;
    int kind, nblob;
    double xs, ys, nrm[2];
    double st[4][4];
    double blob[IC_MAX_BLOBS][4];
};

// Boundary conditions are applied through the index: zero gradient clamps,
// periodic wraps.  This is what zeroGradBc / periodicBc (FV_System2D.cpp
// :114-152) leave in every ghost cell a stencil can reach.
// Every stencil reaches less than one grid width past an edge, so one add
// or subtract wraps it; the integer remainder (some thirty instructions on
// the device) is left to indices further out (grids narrower than a
// stencil).
HTX_HD int map_index(int g, int n, int bc) {
    if (bc == 1) {
        g = g < 0 ? g + n : (g >= n ? g - n : g);
        if (g < 0 || g >= n) {
            g = g % n;
            g = g < 0 ? g + n : g;
        }
        return g;
    }
    return g < 0 ? 0 : (g >= n ? n - 1 : g);
}

// Host-side copies and fills of the library's objects run on the object's own
// stream and are waited for.  Those streams are non-blocking: a copy issued
// on the null stream is not ordered against their kernels, and a pageable
// host-to-device copy may return before its data has landed, so a kernel
// launched next on the object's stream could read the old contents.
inline hipError_t copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

inline hipError_t zero_sync(void* dst, size_t bytes, hipStream_t s) {
    hipError_t e = hipMemsetAsync(dst, 0, bytes, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

// launchers (htx_fv2d.hip)
hipError_t launch_fv2d_stage(const Fv2dArgs& a, int flux, int tile_x, int tile_y, hipStream_t s);
hipError_t launch_dt(Ctl* c, double dx, double dy, hipStream_t s);
hipError_t launch_tadvance(Ctl* c, hipStream_t s);
hipError_t launch_reset_smax(Ctl* c, hipStream_t s);
hipError_t launch_set_stage_dt(Ctl* c, double dt, int fresh, hipStream_t s);
hipError_t launch_max_speed(const double* q, int nx, int ny, int pitch, long long fstride, double gam,
                            int quirks, Ctl* c, hipStream_t s);
hipError_t launch_ic_vortex(const VortexArgs& a, hipStream_t s);
// htx_ic.hip (fast build only)
hipError_t launch_ic_piecewise(const PiecewiseArgs& a, hipStream_t s);
hipError_t launch_ic_entropy(const EntropyArgs& a, hipStream_t s);
int exact_l1_blocks(int nx, int ny);
hipError_t launch_exact_l1(const ExactArgs& a, hipStream_t s);

}  // namespace HTX_NS
