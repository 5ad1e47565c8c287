// htx_fvh2d.hip -- FP64 Hermite-TENO stage (FVHermite::System2D::rhs +
// RungeKutta3, FVHermite_System2D.cpp:393-782) as five launches per stage:
//   k_fvh_limit     the Gauss-Seidel derivative limiter, one sequential chain
//                   per (grid line, field): rows for v, columns for w
//   k_fvh_recon_x/y one work item per face: Roe eigensystem, characteristic
//                   HTENO on (u, normal moment) and TENO5 on the tangential
//                   moment -> 24 face values
//   k_fvh_flux_x/y  64 faces x 3 Gauss points per workgroup: transverse
//                   3-point quadrature of the 24 series (each series' three
//                   points on one work item), back-projection, HLLC or
//                   characteristic Rusanov for u and dRusanov for the moments
//                   per (face, Gauss point); the points are summed in order
//                   through LDS
//   k_fvh_update    divergence of the 12 fluxes, SSP-RK3 stage, CFL maxima
// The work-item bodies live in htx_hermite.h and are shared with the CPU
// build used by the tests.
#include "htx_hermite.h"

namespace HTX_NS {

__device__ __forceinline__ double wave_max_fvh(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// The limiter: fvh_limit_block (htx_hermite.h), 64 chains per workgroup.
__global__ void __launch_bounds__(LIM_THREADS) k_fvh_limit(FvhArgs a, int g0) { fvh_limit_block(a, blockIdx.x, g0); }

// FT: the face values' storage type (float in the FP32 bulk precision)
template <bool MIXED, bool VARIANT = false, typename FT = double>
__global__ void __launch_bounds__(256) k_fvh_recon_x(FvhArgs a) {
    long long nf = (long long)(a.nx + 1) * a.ny;
    long long t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= nf) return;
    int fi = (int)(t % (a.nx + 1)), j = (int)(t / (a.nx + 1));
    fvh_recon_face<false, MIXED, VARIANT>(a, fi - 1, j, reinterpret_cast<FT*>(a.fx), t, nf);
}

template <bool MIXED, bool VARIANT = false, typename FT = double>
__global__ void __launch_bounds__(256) k_fvh_recon_y(FvhArgs a) {
    long long nf = (long long)a.nx * (a.ny + 1);
    long long t = blockIdx.x * 256ll + threadIdx.x;
    if (t >= nf) return;
    int i = (int)(t % a.nx), fj = (int)(t / a.nx);
    fvh_recon_face<true, MIXED, VARIANT>(a, i, fj - 1, reinterpret_cast<FT*>(a.fy), t, nf);
}

// block (64 faces) x (3 Gauss points)
template <bool YDIR, int FLUX, bool MIXED>
__global__ void __launch_bounds__(192) k_fvh_flux(FvhArgs a) {
    __shared__ double lds[24 * 3 * 64];
    long long nf = YDIR ? (long long)a.nx * (a.ny + 1) : (long long)(a.nx + 1) * a.ny;
    fvh_flux_block<YDIR, FLUX, MIXED>(a, blockIdx.x * 64ll, nf, lds);
}

#if !HTX_EXACT_DIVISION
// FP32 bulk precision: 18 KB of LDS per 64 faces x 3 Gauss points
template <bool YDIR, int FLUX>
__global__ void __launch_bounds__(192, 3) k_fvh_flux_f32(FvhArgs a) {
    __shared__ double lds[12 * 3 * 64];
    long long nf = YDIR ? (long long)a.nx * (a.ny + 1) : (long long)(a.nx + 1) * a.ny;
    fvh_flux_block_f32<YDIR, FLUX>(a, blockIdx.x * 64ll, nf, lds);
}
#endif

__global__ void __launch_bounds__(256) k_fvh_update(FvhArgs a) {
    __shared__ double red[2][4];
    double dt = 0.0;
    if (a.mode > 0) {
        if (!a.ctl->active) return;
        dt = a.ctl->dt;
    }
    long long n = (long long)a.nx * a.ny;
    long long t = blockIdx.x * 256ll + threadIdx.x;
    double smx = 0.0, smy = 0.0;
    bool bad = false;
    if (t < n) fvh_update_cell(a, (int)(t % a.nx), (int)(t / a.nx), dt, smx, smy, &bad);
    if (a.mode == 3) {
        flag_nonfinite(bad, a.ctl);
        smx = wave_max_fvh(smx);
        smy = wave_max_fvh(smy);
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

// Isentropic vortex with derivative moments (FVHermite_System2D.cpp:885-900,
// Equations_Euler2D.cpp:513-528, 656-745) centred at (cx, cy), including the
// reference's /pi where /pi^2 is meant in the x-derivative of the density
// (:678); the density modes add a sin(theta) to rho at fixed u, v, p and their
// analytic derivatives to the moments.  9x9 Gauss cell averages of all 12
// fields.
__global__ void __launch_bounds__(256) k_ic_vortex_hermite(VortexArgs a) {
    long long n = (long long)a.nx * a.ny;
    long long id = blockIdx.x * 256ll + threadIdx.x;
    if (id >= n) return;
    int i = (int)(id % a.nx), j = (int)(id / a.nx);
    const double pi = 4.0 * atan(1.0);
    const double gam = a.gam, g1 = gam - 1.0;
    double x0 = ((double)i + 0.5) * a.dx + a.xl;
    double y0 = ((double)j + 0.5) * a.dy + a.yb;
    double acc[12];
    for (int k = 0; k < 12; ++k) acc[k] = 0.0;
    for (int ii = 0; ii < 9; ++ii)
        for (int jj = 0; jj < 9; ++jj) {
            double x = x0 + a.dx * a.gx[ii] / 2.0, y = y0 + a.dy * a.gx[jj] / 2.0;
            double X = x - a.cx, Y = y - a.cy;
            double r = X * X + Y * Y;
            double w0 = 1.0 - g1 * 25.0 / 8.0 / gam / (pi * pi) * exp(1 - r);
            w0 = pow(w0, 1.0 / g1);
            double w1 = 1.0 - 5.0 * Y / 2.0 / pi * exp(0.5 * (1 - r));
            double w2 = 1.0 + 5.0 * X / 2.0 / pi * exp(0.5 * (1 - r));
            double w3 = pow(w0, gam);
            double a0 = 1.0 - g1 * 25.0 / 8.0 / gam / pi * exp(1 - r);
            double rx = 25.0 * X * exp(1 - r) / 4.0 / (pi * pi) / gam * pow(a0, (2.0 - gam) / g1);
            double ux = 5.0 * X * Y / 2.0 / pi * exp(0.5 * (1 - r));
            double vx = 5.0 * (1 - X * X) / 2.0 / pi * exp(0.5 * (1 - r));
            double b0 = 1.0 - g1 * 25.0 / 8.0 / gam / (pi * pi) * exp(1 - r);
            double ry = 25.0 * Y * exp(1 - r) / 4.0 / (pi * pi) / gam * pow(b0, (2.0 - gam) / g1);
            double uy = 5.0 * (Y * Y - 1.0) / 2.0 / pi * exp(0.5 * (1 - r));
            double vy = -5.0 * X * Y / 2.0 / pi * exp(0.5 * (1 - r));
            double pxv = gam * pow(w0, g1) * rx, pyv = gam * pow(w0, g1) * ry;
            double rho = w0, ddx = 0.0, ddy = 0.0;
            for (int m = 0; m < a.nmodes; ++m) {
                double th = 2.0 * pi * (a.modes[m][0] * x / a.modes[m][4] + a.modes[m][1] * y / a.modes[m][5]) +
                            a.modes[m][3];
                rho += a.modes[m][2] * sin(th);
                ddx += a.modes[m][2] * cos(th) * (2.0 * pi * a.modes[m][0] / a.modes[m][4]);
                ddy += a.modes[m][2] * cos(th) * (2.0 * pi * a.modes[m][1] / a.modes[m][5]);
            }
            double q[12];
            q[0] = rho;
            q[1] = rho * w1;
            q[2] = rho * w2;
            q[3] = w3 / g1 + 0.5 * (q[1] * w1 + q[2] * w2);
            double Rx = rx + ddx, Ry = ry + ddy;
            q[4] = Rx;
            q[5] = rho * ux + Rx * w1;
            q[6] = rho * vx + Rx * w2;
            q[7] = pxv / g1 + 0.5 * Rx * (w1 * w1 + w2 * w2) + 0.5 * rho * (2.0 * w1 * ux + 2.0 * w2 * vx);
            q[8] = Ry;
            q[9] = rho * uy + Ry * w1;
            q[10] = rho * vy + Ry * w2;
            q[11] = pyv / g1 + 0.5 * Ry * (w1 * w1 + w2 * w2) + 0.5 * rho * (2.0 * w1 * uy + 2.0 * w2 * vy);
            double ww = a.gw[ii] * a.gw[jj];
            for (int k = 0; k < 12; ++k) acc[k] += ww * q[k] / 4.0;
        }
    long long off = (long long)j * a.pitch + i;
    for (int k = 0; k < 12; ++k) a.q[off + k * a.fstride] = acc[k];
}

template <bool MIXED>
static void launch_recon_flux(const FvhArgs& a, int flux, hipStream_t s) {
    long long nxf = (long long)(a.nx + 1) * a.ny, nyf = (long long)a.nx * (a.ny + 1);
    if (!MIXED && a.hrecon != 0) {
        hipLaunchKernelGGL((k_fvh_recon_x<false, true>), dim3((unsigned)((nxf + 255) / 256)), dim3(256), 0, s, a);
        hipLaunchKernelGGL((k_fvh_recon_y<false, true>), dim3((unsigned)((nyf + 255) / 256)), dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL(k_fvh_recon_x<MIXED>, dim3((unsigned)((nxf + 255) / 256)), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_fvh_recon_y<MIXED>, dim3((unsigned)((nyf + 255) / 256)), dim3(256), 0, s, a);
    }
    dim3 fb(64, 3), gx((unsigned)((nxf + 63) / 64)), gy((unsigned)((nyf + 63) / 64));
    if (flux == FLUX_RUSANOV_CHAR) {
        hipLaunchKernelGGL((k_fvh_flux<false, FLUX_RUSANOV_CHAR, MIXED>), gx, fb, 0, s, a);
        hipLaunchKernelGGL((k_fvh_flux<true, FLUX_RUSANOV_CHAR, MIXED>), gy, fb, 0, s, a);
    } else if (flux == FLUX_HLLC) {
        hipLaunchKernelGGL((k_fvh_flux<false, FLUX_HLLC, MIXED>), gx, fb, 0, s, a);
        hipLaunchKernelGGL((k_fvh_flux<true, FLUX_HLLC, MIXED>), gy, fb, 0, s, a);
    } else {
        hipLaunchKernelGGL((k_fvh_flux<false, FLUX_HLLC_TORO, MIXED>), gx, fb, 0, s, a);
        hipLaunchKernelGGL((k_fvh_flux<true, FLUX_HLLC_TORO, MIXED>), gy, fb, 0, s, a);
    }
}

#if !HTX_EXACT_DIVISION
// FP32 bulk precision: the FP32-mixed reconstruction writing FP32 face values,
// then the FP32 flux kernels
static void launch_recon_flux_f32(const FvhArgs& a, int flux, hipStream_t s) {
    long long nxf = (long long)(a.nx + 1) * a.ny, nyf = (long long)a.nx * (a.ny + 1);
    hipLaunchKernelGGL((k_fvh_recon_x<true, false, float>), dim3((unsigned)((nxf + 255) / 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL((k_fvh_recon_y<true, false, float>), dim3((unsigned)((nyf + 255) / 256)), dim3(256), 0, s, a);
    dim3 fb(64, 3), gx((unsigned)((nxf + 63) / 64)), gy((unsigned)((nyf + 63) / 64));
    if (flux == FLUX_RUSANOV_CHAR) {
        hipLaunchKernelGGL((k_fvh_flux_f32<false, FLUX_RUSANOV_CHAR>), gx, fb, 0, s, a);
        hipLaunchKernelGGL((k_fvh_flux_f32<true, FLUX_RUSANOV_CHAR>), gy, fb, 0, s, a);
    } else if (flux == FLUX_HLLC) {
        hipLaunchKernelGGL((k_fvh_flux_f32<false, FLUX_HLLC>), gx, fb, 0, s, a);
        hipLaunchKernelGGL((k_fvh_flux_f32<true, FLUX_HLLC>), gy, fb, 0, s, a);
    } else {
        hipLaunchKernelGGL((k_fvh_flux_f32<false, FLUX_HLLC_TORO>), gx, fb, 0, s, a);
        hipLaunchKernelGGL((k_fvh_flux_f32<true, FLUX_HLLC_TORO>), gy, fb, 0, s, a);
    }
}
#endif

hipError_t launch_fvh2d_stage(const FvhArgs& a, int flux, hipStream_t s) {
    hipLaunchKernelGGL(k_fvh_limit, dim3(lim_blocks(a.nx, a.ny)), dim3(LIM_THREADS), 0, s, a, (4 * a.ny + 63) / 64);
#if HTX_EXACT_DIVISION
    if (a.mixed || a.bulk32) return hipErrorInvalidValue;  // the strict build is FP64 only
    launch_recon_flux<false>(a, flux, s);
#else
    if (a.bulk32)
        launch_recon_flux_f32(a, flux, s);
    else if (a.mixed)
        launch_recon_flux<true>(a, flux, s);
    else
        launch_recon_flux<false>(a, flux, s);
#endif
    long long n = (long long)a.nx * a.ny;
    hipLaunchKernelGGL(k_fvh_update, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ic_vortex_hermite(const VortexArgs& a, hipStream_t s) {
    long long n = (long long)a.nx * a.ny;
    hipLaunchKernelGGL(k_ic_vortex_hermite, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace HTX_NS
