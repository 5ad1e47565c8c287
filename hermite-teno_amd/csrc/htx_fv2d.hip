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
#include "htx_internal.h"

namespace htx {

__device__ __forceinline__ int map_index(int g, int n, int bc) {
    if (bc == 1) {
        g = g % n;
        return g < 0 ? g + n : g;
    }
    return g < 0 ? 0 : (g >= n ? n - 1 : g);
}

// Workgroups are dealt to the 8 XCDs round-robin; give each XCD one
// contiguous run of tiles so neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    int per = nb / 8, rem = nb % 8;
    int x = b % 8, k = b / 8;
    int start = x * per + (x < rem ? x : rem);
    return start + k;
}

__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

template <int TX, int TY, int FLUX>
__global__ void __launch_bounds__(256) fv2d_stage_kernel(Fv2dArgs a) {
    constexpr int NT = 256;
    constexpr int HX = TX + 6, HY = TY + 6, HC = HX * HY;
    constexpr int NXF = (TX + 1) * (TY + 4);
    constexpr int NYF = (TX + 4) * (TY + 1);
    constexpr int NF = NXF > NYF ? NXF : NYF;
    constexpr int NXQ = (TX + 1) * TY;
    constexpr int NYQ = TX * (TY + 1);
    constexpr int NQ = NXQ > NYQ ? NXQ : NYQ;
    constexpr int NC = TX * TY;
    constexpr int KC = (NC + NT - 1) / NT;

    __shared__ double sQ[4][HC];
    __shared__ double sP[4][HC];
    __shared__ double sF[8][NF];
    __shared__ double sFl[4][NQ];
    __shared__ double sRed[2][NT / 64];

    const int mode = a.mode;
    double dt = 0.0;
    if (mode > 0) {
        if (!a.ctl->active) return;
        dt = a.ctl->dt;
    }
    const int tid = threadIdx.x;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int i0 = (tile % a.tiles_x) * TX;
    const int j0 = (tile / a.tiles_x) * TY;
    const long long fs = a.fstride;
    const double g1 = a.g1, gam = a.gam;

    // ---- halo tile load + per-cell Roe inputs (calcPrimitiveVariable) ----
    for (int w = tid; w < HC; w += NT) {
        int lx = w % HX, ly = w / HX;
        int gi = map_index(i0 - 3 + lx, a.nx, a.bc);
        int gj = map_index(j0 - 3 + ly, a.ny, a.bc);
        long long off = (long long)gj * a.pitch + gi;
        double q0 = a.qin[off], q1 = a.qin[off + fs], q2 = a.qin[off + 2 * fs], q3 = a.qin[off + 3 * fs];
        sQ[0][w] = q0; sQ[1][w] = q1; sQ[2][w] = q2; sQ[3][w] = q3;
        Prim P = cell_prim(g1, q0, q1, q2, q3);
        sP[0][w] = P.sr; sP[1][w] = P.u; sP[2][w] = P.v; sP[3][w] = P.h;
    }
    __syncthreads();

    double dxp[KC][4];

    // ---- x direction: characteristic TENO5 on x faces (FV_System2D.cpp:372-385) ----
    for (int w = tid; w < NXF; w += NT) {
        int fi = w % (TX + 1), fj = w / (TX + 1);
        int ca = (1 + fj) * HX + (2 + fi);  // left cell of the face
        Prim A{sP[0][ca], sP[1][ca], sP[2][ca], sP[3][ca]};
        Prim B{sP[0][ca + 1], sP[1][ca + 1], sP[2][ca + 1], sP[3][ca + 1]};
        Eig E = roe_eig(g1, false, A, B);
        double c[4][6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            int id = ca - 2 + m;
            eig_left(E, sQ[0][id], sQ[1][id], sQ[2][id], sQ[3][id], c[0][m], c[1][m], c[2][m], c[3][m]);
        }
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            double um, up;
            teno5(c[f], um, up);
            sF[2 * f][w] = um;
            sF[2 * f + 1][w] = up;
        }
    }
    __syncthreads();

    // ---- x faces: transverse quadrature, back-projection, flux (:402-497) ----
    for (int w = tid; w < NXQ; w += NT) {
        int fi = w % (TX + 1), r = w / (TX + 1);
        double gl[4][3], gr[4][3];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const double* L = sF[2 * f];
            const double* R = sF[2 * f + 1];
            int b = r * (TX + 1) + fi;
            constexpr int S = TX + 1;
            weno3q(a.qc, L[b], L[b + S], L[b + 2 * S], L[b + 3 * S], L[b + 4 * S], gl[f][0], gl[f][1], gl[f][2]);
            weno3q(a.qc, R[b], R[b + S], R[b + 2 * S], R[b + 3 * S], R[b + 4 * S], gr[f][0], gr[f][1], gr[f][2]);
        }
        int ca = (3 + r) * HX + (2 + fi);
        Prim A{sP[0][ca], sP[1][ca], sP[2][ca], sP[3][ca]};
        Prim B{sP[0][ca + 1], sP[1][ca + 1], sP[2][ca + 1], sP[3][ca + 1]};
        Eig E = roe_eig(g1, false, A, B);
        double acc[3][4];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            double ql[4], qr[4], fl[4];
            eig_right(E, gl[0][m], gl[1][m], gl[2][m], gl[3][m], ql[0], ql[1], ql[2], ql[3]);
            eig_right(E, gr[0][m], gr[1][m], gr[2][m], gr[3][m], qr[0], qr[1], qr[2], qr[3]);
            if (FLUX == FLUX_RUSANOV_CHAR)
                rusanov_char(gam, g1, E, ql, qr, fl);
            else if (FLUX == FLUX_HLLC)
                hllc<true>(gam, g1, ql, qr, fl);
            else
                hllc<false>(gam, g1, ql, qr, fl);
            const double gw = (m == 1) ? 8.0 / 9.0 / 2.0 : 5.0 / 9.0 / 2.0;
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[m][e] = gw * fl[e];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) sFl[e][w] = (acc[0][e] + acc[1][e]) + acc[2][e];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        int w = tid + k * NT;
        if (w < NC) {
            int c = w % TX, r = w / TX;
            int b = r * (TX + 1) + c;
#pragma unroll
            for (int e = 0; e < 4; ++e) dxp[k][e] = (sFl[e][b] - sFl[e][b + 1]) * a.dy;
        }
    }
    __syncthreads();

    // ---- y direction: characteristic TENO5 on y faces (:387-400) ----
    for (int w = tid; w < NYF; w += NT) {
        int fi = w % (TX + 4), fj = w / (TX + 4);
        int ca = (2 + fj) * HX + (1 + fi);  // lower cell of the face
        Prim A{sP[0][ca], sP[1][ca], sP[2][ca], sP[3][ca]};
        Prim B{sP[0][ca + HX], sP[1][ca + HX], sP[2][ca + HX], sP[3][ca + HX]};
        Eig E = roe_eig(g1, true, A, B);
        double c[4][6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            int id = ca + (m - 2) * HX;
            eig_left(E, sQ[0][id], sQ[2][id], sQ[1][id], sQ[3][id], c[0][m], c[1][m], c[2][m], c[3][m]);
        }
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            double um, up;
            teno5(c[f], um, up);
            sF[2 * f][w] = um;
            sF[2 * f + 1][w] = up;
        }
    }
    __syncthreads();

    // ---- y faces: transverse quadrature, back-projection, flux ----
    for (int w = tid; w < NYQ; w += NT) {
        int c0 = w % TX, fj = w / TX;
        double gl[4][3], gr[4][3];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const double* L = sF[2 * f];
            const double* R = sF[2 * f + 1];
            int b = fj * (TX + 4) + c0;
            weno3q(a.qc, L[b], L[b + 1], L[b + 2], L[b + 3], L[b + 4], gl[f][0], gl[f][1], gl[f][2]);
            weno3q(a.qc, R[b], R[b + 1], R[b + 2], R[b + 3], R[b + 4], gr[f][0], gr[f][1], gr[f][2]);
        }
        int ca = (2 + fj) * HX + (3 + c0);
        Prim A{sP[0][ca], sP[1][ca], sP[2][ca], sP[3][ca]};
        Prim B{sP[0][ca + HX], sP[1][ca + HX], sP[2][ca + HX], sP[3][ca + HX]};
        Eig E = roe_eig(g1, true, A, B);
        double acc[3][4];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            double ql[4], qr[4], fl[4];
            eig_right(E, gl[0][m], gl[1][m], gl[2][m], gl[3][m], ql[0], ql[1], ql[2], ql[3]);
            eig_right(E, gr[0][m], gr[1][m], gr[2][m], gr[3][m], qr[0], qr[1], qr[2], qr[3]);
            if (FLUX == FLUX_RUSANOV_CHAR)
                rusanov_char(gam, g1, E, ql, qr, fl);
            else if (FLUX == FLUX_HLLC)
                hllc<true>(gam, g1, ql, qr, fl);
            else
                hllc<false>(gam, g1, ql, qr, fl);
            const double gw = (m == 1) ? 8.0 / 9.0 / 2.0 : 5.0 / 9.0 / 2.0;
            // rotated (rho, n=y, t=x, E) -> (rho, x, y, E)
            acc[m][0] = gw * fl[0];
            acc[m][1] = gw * fl[2];
            acc[m][2] = gw * fl[1];
            acc[m][3] = gw * fl[3];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) sFl[e][w] = (acc[0][e] + acc[1][e]) + acc[2][e];
    }
    __syncthreads();

    // ---- epilogue: divergence (:562-572), SSP-RK3 (:575-612), CFL (:623-631) ----
    double smx = 0.0, smy = 0.0;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        int w = tid + k * NT;
        if (w >= NC) continue;
        int c = w % TX, r = w / TX;
        int gi = i0 + c, gj = j0 + r;
        if (gi >= a.nx || gj >= a.ny) continue;
        long long off = (long long)gj * a.pitch + gi;
        int lc = (3 + r) * HX + (3 + c);
        double unew[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            double rhs = (dxp[k][e] + (sFl[e][r * TX + c] - sFl[e][(r + 1) * TX + c]) * a.dx) / a.area;
            double us = sQ[e][lc];
            double v;
            if (mode == 0)
                v = rhs;
            else if (mode == 1)
                v = us + dt * rhs;
            else if (mode == 2)
                v = 0.75 * a.q0[off + e * fs] + 0.25 * (us + rhs * dt);
            else
                v = (a.q0[off + e * fs] + 2.0 * (us + dt * rhs)) / 3.0;
            unew[e] = v;
            a.qout[off + e * fs] = v;
        }
        if (mode == 3) {
            double wu = unew[1] / unew[0], wv = unew[2] / unew[0];
            double p = (unew[3] - 0.5 * (unew[1] * wu + unew[2] * wv)) * g1;
            if (a.quirks) {
                smx = fmax(smx, max_speed<true>(gam, wu, p, unew[0]));
                smy = fmax(smy, max_speed<true>(gam, wv, p, unew[0]));
            } else {
                smx = fmax(smx, max_speed<false>(gam, wu, p, unew[0]));
                smy = fmax(smy, max_speed<false>(gam, wv, p, unew[0]));
            }
        }
    }
    if (mode == 3) {
        smx = wave_max(smx);
        smy = wave_max(smy);
        if ((tid & 63) == 0) {
            sRed[0][tid >> 6] = smx;
            sRed[1][tid >> 6] = smy;
        }
        __syncthreads();
        if (tid == 0) {
            double mx = 0.0, my = 0.0;
#pragma unroll
            for (int k = 0; k < NT / 64; ++k) {
                mx = fmax(mx, sRed[0][k]);
                my = fmax(my, sRed[1][k]);
            }
            atomicMax(&a.ctl->smax_x, (unsigned long long)__double_as_longlong(mx));
            atomicMax(&a.ctl->smax_y, (unsigned long long)__double_as_longlong(my));
        }
    }
}

// dt from the accumulated max wave speeds (FV_System2D.cpp:633,645); a step
// with t >= tFinal is marked inactive and every stage kernel skips it.
__global__ void k_dt(Ctl* c, double dx, double dy) {
    double t = c->t, tf = c->tfinal;
    if (t < tf) {
        double sx = __longlong_as_double((long long)c->smax_x);
        double sy = __longlong_as_double((long long)c->smax_y);
        double dt = c->cfl * fmin(dx, dy) / fmax(sx, sy);
        dt = t + dt > tf ? tf - t : dt;
        c->dt = dt;
        c->active = 1;
    } else {
        c->dt = 0.0;
        c->active = 0;
    }
    c->smax_x = 0ull;
    c->smax_y = 0ull;
}

__global__ void k_tadvance(Ctl* c) {
    if (c->active) {
        c->t += c->dt;
        c->nsteps += 1;
    }
}

__global__ void k_reset_smax(Ctl* c) {
    c->smax_x = 0ull;
    c->smax_y = 0ull;
}

__global__ void __launch_bounds__(256) k_max_speed(const double* q, int nx, int ny, int pitch,
                                                   long long fs, double gam, int quirks, Ctl* c) {
    __shared__ double red[2][4];
    double g1 = gam - 1.0;
    double mx = 0.0, my = 0.0;
    long long n = (long long)nx * ny;
    for (long long id = blockIdx.x * 256ll + threadIdx.x; id < n; id += (long long)gridDim.x * 256) {
        int i = (int)(id % nx), j = (int)(id / nx);
        long long off = (long long)j * pitch + i;
        double q0 = q[off], q1 = q[off + fs], q2 = q[off + 2 * fs], q3 = q[off + 3 * fs];
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
        mx = fmax(mx, red[0][0]);
        my = fmax(my, red[1][0]);
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

void fill_quad_consts(QuadConsts& C) {
    // Schemes_Lagrange.cpp:64-105, same expressions, evaluated on the host.
    double sq15 = sqrt(15.0);
    C.gp[0][0] = 1.0 / 30.0 + sq15 / 20.0;
    C.gp[1][0] = -1.0 / 15.0 - sq15 / 5.0;
    C.gp[2][0] = 31.0 / 30.0 + 3.0 * sq15 / 20.0;
    C.gp[0][1] = 1.0 / 30.0 - sq15 / 20.0;
    C.gp[1][1] = 14.0 / 15.0;
    C.gp[2][1] = 1.0 / 30.0 + sq15 / 20.0;
    C.gp[0][2] = 31.0 / 30.0 - 3.0 * sq15 / 20.0;
    C.gp[1][2] = -1.0 / 15.0 + sq15 / 5.0;
    C.gp[2][2] = 1.0 / 30.0 - sq15 / 20.0;
    C.gm[0][0] = 1.0 / 30.0 - sq15 / 20.0;
    C.gm[1][0] = -1.0 / 15.0 + sq15 / 5.0;
    C.gm[2][0] = 31.0 / 30.0 - 3.0 * sq15 / 20.0;
    C.gm[0][1] = 1.0 / 30.0 + sq15 / 20.0;
    C.gm[1][1] = 14.0 / 15.0;
    C.gm[2][1] = 1.0 / 30.0 - sq15 / 20.0;
    C.gm[0][2] = 31.0 / 30.0 + 3.0 * sq15 / 20.0;
    C.gm[1][2] = -1.0 / 15.0 - sq15 / 5.0;
    C.gm[2][2] = 1.0 / 30.0 + sq15 / 20.0;
    double lin0[3] = {-9.0 / 80.0, 49.0 / 40.0, -9.0 / 80.0};
    double p[3], m[3];
    for (int i = 0; i < 3; ++i) {
        p[i] = 0.5 * (lin0[i] + 3.0 * fabs(lin0[i]));
        m[i] = p[i] - lin0[i];
    }
    C.sumw0p = (p[0] + p[1]) + p[2];
    C.sumw0m = (m[0] + m[1]) + m[2];
    for (int i = 0; i < 3; ++i) {
        C.linw0p[i] = p[i] / C.sumw0p;
        C.linw0m[i] = m[i] / C.sumw0m;
    }
    double den = 15720.0 * sq15 + 10480.0;
    C.linwp[0] = (2882.0 * sq15 - 1179.0) / den;
    C.linwp[1] = (9672.0 * sq15 + 6448.0) / den;
    C.linwp[2] = (3166.0 * sq15 + 5211.0) / den;
    den = 15720.0 * sq15 - 10480;
    C.linwm[0] = (2882.0 * sq15 + 1179.0) / den;
    C.linwm[1] = (9672.0 * sq15 - 6448.0) / den;
    C.linwm[2] = (3166.0 * sq15 - 5211.0) / den;
}

template <int TX, int TY>
static hipError_t launch_tile(const Fv2dArgs& a0, int flux, hipStream_t s) {
    Fv2dArgs a = a0;
    a.tiles_x = (a.nx + TX - 1) / TX;
    a.tiles_y = (a.ny + TY - 1) / TY;
    dim3 grid(a.tiles_x * a.tiles_y), block(256);
    if (flux == FLUX_RUSANOV_CHAR)
        hipLaunchKernelGGL((fv2d_stage_kernel<TX, TY, FLUX_RUSANOV_CHAR>), grid, block, 0, s, a);
    else if (flux == FLUX_HLLC)
        hipLaunchKernelGGL((fv2d_stage_kernel<TX, TY, FLUX_HLLC>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((fv2d_stage_kernel<TX, TY, FLUX_HLLC_TORO>), grid, block, 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fv2d_stage(const Fv2dArgs& a, int flux, int tile_x, int tile_y, hipStream_t s) {
    (void)tile_x;
    (void)tile_y;
    return launch_tile<16, 16>(a, flux, s);
}

hipError_t launch_dt(Ctl* c, double dx, double dy, hipStream_t s) {
    hipLaunchKernelGGL(k_dt, dim3(1), dim3(1), 0, s, c, dx, dy);
    return hipGetLastError();
}

hipError_t launch_tadvance(Ctl* c, hipStream_t s) {
    hipLaunchKernelGGL(k_tadvance, dim3(1), dim3(1), 0, s, c);
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

}  // namespace htx
