// htx_fv2d.hip -- fused FP64 residual + SSP-RK3 stage kernel for the 2D
// Euler FV path (FV::System2D::rhs + RungeKutta3, FV_System2D.cpp:363-612),
// plus the CFL control kernels.
//
// One workgroup owns a TX x TY tile of cells.  It stages the (TX+6)x(TY+6)
// halo tile of the conserved state in LDS (boundary conditions applied by
// clamping or wrapping the load index, which is what zeroGradBc/periodicBc
// produce for every ghost the stencil reaches), derives per-cell Roe inputs,
// and then runs, all on chip:
//   x faces: characteristic TENO5 on (TX+1)x(TY+4) faces -> LDS,
//            3-point transverse WENO quadrature + back-projection + Riemann
//            flux on (TX+1)xTY faces -> LDS -> x part of the divergence;
//   y faces: the same with the roles of x and y exchanged;
//   epilogue: residual, SSP-RK3 combine, and in stage 3 the CFL max wave
//            speed (wavefront shuffle reduction, one atomic per workgroup).
// Face states never leave the chip; HBM sees the state read, u^n read
// (stages 2-3) and the stage result write.
#include <type_traits>

#include "htx_fv2d_tile.h"

namespace HTX_NS {

// BULK32: the FP32 bulk precision (face values stored in FP32, quadrature,
// back-projection and flux in FP32): a smaller tile in LDS and fewer
// registers, three workgroups per CU.
// Tiles that form the cells' Roe inputs per face (tile_prim_on_the_fly) also
// fit three per CU.

// RECON: the characteristic reconstruction (0 TENO5, the reference's; 1
// teno5i, 2 weno5JS, 3 weno5Z), built for the FP64 precisions
template <int TX, int TY, int FLUX, bool MIXED, bool LEARNED, bool BULK32 = false, int RECON = 0>
__global__ void __launch_bounds__(256, (BULK32 || tile_prim_on_the_fly<MIXED, LEARNED, BULK32>()) ? 3 : 2)
    fv2d_stage_kernel(Fv2dArgs a) {
    using T = Fv2dTile<TX, TY, typename std::conditional<BULK32, float, double>::type,
                       tile_prim_on_the_fly<MIXED, LEARNED, BULK32>()>;
    __shared__ T s;
    double dt = 0.0;
    if (a.mode > 0) {
        if (!a.ctl->active) return;
        dt = a.ctl->dt;
    }
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    stage_tile<TX, TY, FLUX, MIXED, LEARNED, RECON>(a, s, tile, dt, threadIdx.x);
}

// one-thread control kernels around a step (ctl_dt, ctl_tadvance)
__global__ void k_dt(Ctl* c, double dx, double dy) { ctl_dt(c, dx, dy); }

__global__ void k_tadvance(Ctl* c) { ctl_tadvance(c); }

// an explicit stage (htx_stage): the stage kernels take dt from the control
// block and run; stage 1 starts the CFL maxima afresh, as ctl_dt does
__global__ void k_set_stage_dt(Ctl* c, double dt, int fresh) {
    c->dt = dt;
    c->active = 1;
    if (fresh) {
        c->smax_x = 0ull;
        c->smax_y = 0ull;
    }
}

// a new state: fresh CFL maxima and no failure record
__global__ void k_reset_smax(Ctl* c) {
    c->smax_x = 0ull;
    c->smax_y = 0ull;
    c->status = 0u;
}

__global__ void __launch_bounds__(256) k_max_speed(const double* q, int nx, int ny, int pitch,
                                                   long long fs, double gam, int quirks, Ctl* c) {
    __shared__ double red[2][4];
    double g1 = gam - 1.0;
    double mx = 0.0, my = 0.0;
    bool bad = false;
    long long n = (long long)nx * ny;
    for (long long id = blockIdx.x * 256ll + threadIdx.x; id < n; id += (long long)gridDim.x * 256) {
        int i = (int)(id % nx), j = (int)(id / nx);
        long long off = (long long)j * pitch + i;
        double q0 = q[off], q1 = q[off + fs], q2 = q[off + 2 * fs], q3 = q[off + 3 * fs];
        const double qq[4] = {q0, q1, q2, q3};
        bad = bad || !all_finite(qq, 4);
        double wu = q1 / q0, wv = q2 / q0;
        double p = (q3 - 0.5 * (q1 * wu + q2 * wv)) * g1;
        if (quirks) {
            mx = fmax(mx, max_speed<true>(gam, wu, p, q0));
            my = fmax(my, max_speed<true>(gam, wv, p, q0));
        } else {
            mx = fmax(mx, max_speed<false>(gam, wu, p, q0));
            my = fmax(my, max_speed<false>(gam, wv, p, q0));
        }
    }
    flag_nonfinite(bad, c);
    mx = wave_max(mx);
    my = wave_max(my);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = mx;
        red[1][threadIdx.x >> 6] = my;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; ++k) {
            mx = fmax(mx, red[0][k]);
            my = fmax(my, red[1][k]);
        }
        atomicMax(&c->smax_x, (unsigned long long)__double_as_longlong(mx));
        atomicMax(&c->smax_y, (unsigned long long)__double_as_longlong(my));
    }
}

// Isentropic vortex (Equations_Euler2D.cpp:656-666, b = 5) centred at
// (cx, cy), plus a density perturbation sum_m a_m sin(2 pi (k x/Lx + l y/Ly)
// + phi) at fixed velocity and pressure, averaged over each cell with the
// 9x9 Gauss rule of Euler2D::isentropicVortex (:513-528).
__global__ void __launch_bounds__(256) k_ic_vortex(VortexArgs a) {
    long long n = (long long)a.nx * a.ny;
    long long id = blockIdx.x * 256ll + threadIdx.x;
    if (id >= n) return;
    int i = (int)(id % a.nx), j = (int)(id / a.nx);
    const double pi = 4.0 * atan(1.0);
    double g1 = a.gam - 1.0;
    double x = ((double)i + 0.5) * a.dx + a.xl;
    double y = ((double)j + 0.5) * a.dy + a.yb;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int ii = 0; ii < 9; ++ii)
        for (int jj = 0; jj < 9; ++jj) {
            double px = x + a.dx * a.gx[ii] / 2.0, py = y + a.dy * a.gx[jj] / 2.0;
            double xr = px - a.cx, yr = py - a.cy;
            double r = xr * xr + yr * yr, b = 5.0;
            double w0 = 1.0 - g1 * (b * b) / 8.0 / a.gam / (pi * pi) * exp(1 - r);
            w0 = pow(w0, 1.0 / g1);
            double w1 = 1.0 - b * yr / 2.0 / pi * exp(0.5 * (1 - r));
            double w2 = 1.0 + b * xr / 2.0 / pi * exp(0.5 * (1 - r));
            double w3 = pow(w0, a.gam);
            for (int m = 0; m < a.nmodes; ++m)
                w0 += a.modes[m][2] *
                      sin(2.0 * pi * (a.modes[m][0] * px / a.modes[m][4] + a.modes[m][1] * py / a.modes[m][5]) +
                          a.modes[m][3]);
            double q0 = w0, q1 = w0 * w1, q2 = w0 * w2;
            double q3 = w3 / g1 + 0.5 * (q1 * w1 + q2 * w2);
            double ww = a.gw[ii] * a.gw[jj];
            acc[0] += ww * q0 / 4.0;
            acc[1] += ww * q1 / 4.0;
            acc[2] += ww * q2 / 4.0;
            acc[3] += ww * q3 / 4.0;
        }
    long long off = (long long)j * a.pitch + i;
    for (int e = 0; e < 4; ++e) a.q[off + e * a.fstride] = acc[e];
}

template <int TX, int TY, bool MIXED, bool LEARNED, bool BULK32 = false, int RECON = 0>
static hipError_t launch_tile(const Fv2dArgs& a0, int flux, hipStream_t s) {
    Fv2dArgs a = a0;
    a.tiles_x = (a.nx + TX - 1) / TX;
    a.tiles_y = (a.ny + TY - 1) / TY;
    dim3 grid(a.tiles_x * a.tiles_y), block(256);
    if (flux == FLUX_RUSANOV_CHAR)
        hipLaunchKernelGGL((fv2d_stage_kernel<TX, TY, FLUX_RUSANOV_CHAR, MIXED, LEARNED, BULK32, RECON>), grid, block, 0,
                           s, a);
    else if (flux == FLUX_HLLC)
        hipLaunchKernelGGL((fv2d_stage_kernel<TX, TY, FLUX_HLLC, MIXED, LEARNED, BULK32, RECON>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((fv2d_stage_kernel<TX, TY, FLUX_HLLC_TORO, MIXED, LEARNED, BULK32, RECON>), grid, block, 0, s,
                           a);
    return hipGetLastError();
}

template <int TX, int TY>
static hipError_t launch_prec(const Fv2dArgs& a, int flux, hipStream_t s) {
    if (a.learned && (!a.face_flags_x || !a.face_flags_y)) return hipErrorInvalidValue;
    // the alternative reconstructions: FP64 precisions, the default tile
    if (a.recon != 0) {
        if constexpr (TX == 16 && TY == 15) {
            if (a.mixed || a.learned || a.bulk32) return hipErrorInvalidValue;
            if (a.recon == 1) return launch_tile<TX, TY, false, false, false, 1>(a, flux, s);
            if (a.recon == 2) return launch_tile<TX, TY, false, false, false, 2>(a, flux, s);
            if (a.recon == 3) return launch_tile<TX, TY, false, false, false, 3>(a, flux, s);
            if (a.recon >= RECON_WENO5N && a.recon <= RECON_LAST)
                return launch_tile<TX, TY, false, false, false, RECON_WENO5N>(a, flux, s);
        }
        return hipErrorInvalidValue;
    }
#if !HTX_EXACT_DIVISION
    // the FP32 bulk precision is built for the 16-wide tiles
    if constexpr (TX == 16) {
        if (a.bulk32) {
            if (a.learned) return launch_tile<TX, TY, true, true, true>(a, flux, s);
            return launch_tile<TX, TY, true, false, true>(a, flux, s);
        }
    }
#endif
    if (a.learned) return launch_tile<TX, TY, true, true>(a, flux, s);
    if (a.mixed) return launch_tile<TX, TY, true, false>(a, flux, s);
    return launch_tile<TX, TY, false, false>(a, flux, s);
}

// Tile shapes: 16 x 15 (the default: its four face loops, 323 + 320
// reconstructed and 255 + 256 flux faces, fill 19 wavefront passes with 94 %
// of the lanes busy, against 22 passes and 87 % for 16 x 16 cells; the
// learned precisions gain 8-15 % from it too), 16 x 16, and in the fast
// build 16 x 8 and 32 x 8 (smaller LDS footprints, more workgroups per CU for
// kernels with few registers); any other request takes the default.  Every
// shape gives the same results bit for bit, except in the learned
// precisions, whose linear transverse quadrature is decided per wavefront
// and so depends on which faces share one.
hipError_t launch_fv2d_stage(const Fv2dArgs& a, int flux, int tile_x, int tile_y, hipStream_t s) {
    if (a.recon != 0) return launch_prec<16, 15>(a, flux, s);
    if (tile_x == 16 && tile_y == 16) return launch_prec<16, 16>(a, flux, s);
#if !HTX_EXACT_DIVISION
    if (tile_x == 16 && tile_y == 8) return launch_prec<16, 8>(a, flux, s);
    if (tile_x == 32 && tile_y == 8 && !a.bulk32) return launch_prec<32, 8>(a, flux, s);
#endif
    return launch_prec<16, 15>(a, flux, s);
}

hipError_t launch_dt(Ctl* c, double dx, double dy, hipStream_t s) {
    hipLaunchKernelGGL(k_dt, dim3(1), dim3(1), 0, s, c, dx, dy);
    return hipGetLastError();
}

hipError_t launch_tadvance(Ctl* c, hipStream_t s) {
    hipLaunchKernelGGL(k_tadvance, dim3(1), dim3(1), 0, s, c);
    return hipGetLastError();
}

hipError_t launch_set_stage_dt(Ctl* c, double dt, int fresh, hipStream_t s) {
    hipLaunchKernelGGL(k_set_stage_dt, dim3(1), dim3(1), 0, s, c, dt, fresh);
    return hipGetLastError();
}

hipError_t launch_reset_smax(Ctl* c, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_smax, dim3(1), dim3(1), 0, s, c);
    return hipGetLastError();
}

hipError_t launch_max_speed(const double* q, int nx, int ny, int pitch, long long fstride, double gam,
                            int quirks, Ctl* c, hipStream_t s) {
    long long n = (long long)nx * ny;
    long long nb = (n + 255) / 256;
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(k_max_speed, dim3((unsigned)nb), dim3(256), 0, s, q, nx, ny, pitch, fstride, gam,
                       quirks, c);
    return hipGetLastError();
}

hipError_t launch_ic_vortex(const VortexArgs& a, hipStream_t s) {
    long long n = (long long)a.nx * a.ny;
    hipLaunchKernelGGL(k_ic_vortex, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace HTX_NS
