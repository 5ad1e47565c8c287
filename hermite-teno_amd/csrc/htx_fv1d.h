// htx_fv1d.h -- the 1D Euler path (FV::System1D, FV_System1D.cpp:121-292;
// Equations::Euler1D, Equations_Euler1D.cpp): per-face bodies shared by the
// GPU kernel and its CPU build.  Every operation is one IEEE operation in the
// reference's order (the strict build includes this file).
#pragma once
#include "htx_internal.h"

namespace HTX_NS {

struct Fv1dArgs {
    const double* qin;  // state whose residual is taken (3 fields x nx)
    const double* q0;   // u^n (stages 2 and 3)
    double* qout;       // residual (mode 0) or the stage result
    Ctl* ctl;
    int nx, bc, mode;   // bc: 0 zero-gradient, 1 periodic
    int recon;          // 0 teno5, 1 teno5i (the reference's shipped 1D scheme), 2 weno5JS, 3 weno5Z,
                        // RECON_WENO5N .. RECON_LAST the variants of recon5_variant (diff = dx)
    int flux;           // FLUX_RUSANOV_CHAR, FLUX_HLLC (quirks), FLUX_HLLC_TORO, FLUX_CU
    double gam, g1, dx;
};

// Euler1D::primitive / flux (Equations_Euler1D.cpp:18-37)
HTX_HD void e1_prim(double g1, const double q[3], double w[3]) {
    HTX_NO_CONTRACT
    w[0] = q[0];
    w[1] = q[1] / q[0];
    w[2] = (q[2] - 0.5 * q[1] * w[1]) * g1;
}

HTX_HD void e1_phys_flux(const double q[3], const double w[3], double f[3]) {
    HTX_NO_CONTRACT
    f[0] = q[1];
    f[1] = q[1] * w[1] + w[2];
    f[2] = w[1] * (q[2] + w[2]);
}

// 3x3 product in the row-by-row, left-to-right order of the reference's
// fixed-size products
HTX_HD void e1_mv(const double M[9], const double x[3], double y[3]) {
    HTX_NO_CONTRACT
#pragma unroll
    for (int r = 0; r < 3; ++r) y[r] = (M[3 * r] * x[0] + M[3 * r + 1] * x[1]) + M[3 * r + 2] * x[2];
}

// FV::System1D::eigen1DEuler (FV_System1D.cpp:121-156): Roe average of
// cells a, b and the left / right eigenvector matrices (row-major)
HTX_HD void e1_eigen(double g1, const double qa[3], const double qb[3], double L[9], double R[9]) {
    HTX_NO_CONTRACT
    double wa[3], wb[3];
    e1_prim(g1, qa, wa);
    e1_prim(g1, qb, wb);
    const double sa = sqrt(fabs(wa[0])), sb = sqrt(fabs(wb[0]));
    const double ha = (qa[2] + wa[2]) / wa[0], hb = (qb[2] + wb[2]) / wb[0];
    const double t0 = sa / (sa + sb), t1 = 1.0 - t0;
    const double vxm = t0 * wa[1] + t1 * wb[1];
    const double hm = t0 * ha + t1 * hb;
    const double qm = 0.5 * vxm * vxm;
    const double cm = sqrt(fabs(g1 * (hm - qm)));
    const double tc = vxm * cm;
    R[0] = 1.0; R[1] = 1.0; R[2] = 1.0;
    R[3] = vxm - cm; R[4] = vxm; R[5] = vxm + cm;
    R[6] = hm - tc; R[7] = qm; R[8] = hm + tc;
    const double rcm = 1.0 / cm, b1 = g1 * rcm * rcm, b2 = qm * b1, tr = vxm * rcm, tb = b1 * vxm, t2 = 0.5 * b1;
    L[0] = 0.5 * (b2 + tr); L[1] = -0.5 * (tb + rcm); L[2] = t2;
    L[3] = 1.0 - b2; L[4] = tb; L[5] = -b1;
    L[6] = 0.5 * (b2 - tr); L[7] = -0.5 * (tb - rcm); L[8] = t2;
}

// Euler1D::HLLC (Equations_Euler1D.cpp:156-210); QUIRKS: the sound speeds
// without gamma (:174-175) and p_L in the right star energy (:204)
template <bool QUIRKS>
HTX_HD void e1_hllc(double gam, double g1, const double ql[3], const double qr[3], double f[3]) {
    HTX_NO_CONTRACT
    double wl[3], wr[3], fl[3], fr[3], qs[3];
    e1_prim(g1, ql, wl);
    e1_prim(g1, qr, wr);
    const double hl = (ql[2] + wl[2]) / wl[0], hr = (qr[2] + wr[2]) / wr[0];
    const double rr = sqrt(fabs(wr[0] / wl[0]));
    const double rrr = 1.0 / (1.0 + rr);
    const double vm = (wl[1] + wr[1] * rr) * rrr;
    const double qm = 0.5 * vm * vm;
    const double hm = (hl + hr * rr) * rrr;
    const double cm = sqrt(fabs(g1 * (hm - qm)));
    const double cl = QUIRKS ? sqrt(fabs(wl[2] / wl[0])) : sqrt(fabs(gam * wl[2] / wl[0]));
    const double cr = QUIRKS ? sqrt(fabs(wr[2] / wr[0])) : sqrt(fabs(gam * wr[2] / wr[0]));
    const double sl = fmin(wl[1] - cl, vm - cm);
    const double sr = fmax(wr[1] + cr, vm + cm);
    const double sm = (qr[1] * (sr - wr[1]) - ql[1] * (sl - wl[1]) + wl[2] - wr[2]) /
                      (wr[0] * (sr - wr[1]) - wl[0] * (sl - wl[1]));
    const double ps = wl[2] + wl[0] * (wl[1] - sl) * (wl[1] - sm);
    e1_phys_flux(ql, wl, fl);
    e1_phys_flux(qr, wr, fr);
    if (sl > 0.0) {
        for (int k = 0; k < 3; ++k) f[k] = fl[k];
    } else if (sl <= 0.0 && sm > 0.0) {
        const double nom = sl - wl[1], den = sl - sm;
        qs[0] = wl[0] * nom / den;
        qs[1] = (nom * ql[1] + ps - wl[2]) / den;
        qs[2] = (nom * ql[2] - wl[2] * wl[1] + ps * sm) / den;
        for (int k = 0; k < 3; ++k) f[k] = fl[k] + sl * (qs[k] - ql[k]);
    } else if (sm <= 0.0 && sr >= 0.0) {
        const double nom = sr - wr[1], den = sr - sm;
        qs[0] = wr[0] * nom / den;
        qs[1] = (nom * qr[1] + ps - wr[2]) / den;
        qs[2] = (nom * qr[2] - (QUIRKS ? wl[2] : wr[2]) * wr[1] + ps * sm) / den;
        for (int k = 0; k < 3; ++k) f[k] = fr[k] + sr * (qs[k] - qr[k]);
    } else {
        for (int k = 0; k < 3; ++k) f[k] = fr[k];
    }
}

// Euler1D::Rusanov with maxLocalWaveSpeed (FV_System1D.cpp:218-222):
// 0.5 (f_L + f_R - R diag(a) L (q_R - q_L)), a per characteristic field
HTX_HD void e1_rusanov_char(double gam, double g1, const double L[9], const double R[9], const double ql[3],
                            const double qr[3], double f[3]) {
    HTX_NO_CONTRACT
    double wl[3], wr[3], fl[3], fr[3], a[3], dd[3], ld[3], rd[3];
    e1_prim(g1, ql, wl);
    e1_prim(g1, qr, wr);
    const double cl = sqrt(fabs(gam * wl[2] / wl[0])), cr = sqrt(fabs(gam * wr[2] / wr[0]));
    a[0] = fmax(fabs(wl[1] - cl), fabs(wr[1] - cr));
    a[1] = fmax(fabs(wl[1]), fabs(wr[1]));
    a[2] = fmax(fabs(wl[1] + cl), fabs(wr[1] + cr));
    e1_phys_flux(ql, wl, fl);
    e1_phys_flux(qr, wr, fr);
    for (int e = 0; e < 3; ++e) dd[e] = qr[e] - ql[e];
    e1_mv(L, dd, ld);
    for (int e = 0; e < 3; ++e) ld[e] = a[e] * ld[e];
    e1_mv(R, ld, rd);
    for (int e = 0; e < 3; ++e) f[e] = 0.5 * ((fl[e] + fr[e]) - rd[e]);
}

// Euler1D::CUFlux (Equations_Euler1D.cpp:137-154), the central-upwind flux
// the reference's 1D solver offers in a comment (FV_System1D.cpp:234): the
// one-sided speeds a+ = max(s_max, s_min, 0), a- = min(s_max, s_min, 0) of the
// componentwise larger and smaller state, s = localWaveSpeed (:108-119:
// |u - c|, |u|, |u + c|), then
//   F = (a+ f(q_L) - a- f(q_R)) / (a+ - a-) + R (a+ a- / (a+ - a-) L (q_R - q_L)),
// each division a multiplication by the inverse as the reference writes it.
// With the absolute speeds a- is zero, so F = (a+ f(q_L)) (1 / a+).
HTX_HD void e1_cuflux(double gam, double g1, const double L[9], const double R[9], const double ql[3],
                      const double qr[3], double f[3]) {
    HTX_NO_CONTRACT
    double qmax[3], qmin[3], wmax[3], wmin[3], wl[3], wr[3], fl[3], fr[3], ap[3], am[3], inv[3], dd[3], ld[3], rd[3];
    for (int e = 0; e < 3; ++e) {
        qmax[e] = ql[e] < qr[e] ? qr[e] : ql[e];  // cwiseMax / cwiseMin
        qmin[e] = qr[e] < ql[e] ? qr[e] : ql[e];
    }
    e1_prim(g1, qmax, wmax);
    e1_prim(g1, qmin, wmin);
    const double cx = sqrt(fabs(gam * wmax[2] / wmax[0])), cn = sqrt(fabs(gam * wmin[2] / wmin[0]));
    const double sx[3] = {fabs(wmax[1] - cx), fabs(wmax[1]), fabs(wmax[1] + cx)};
    const double sn[3] = {fabs(wmin[1] - cn), fabs(wmin[1]), fabs(wmin[1] + cn)};
    for (int e = 0; e < 3; ++e) {
        const double hi = sx[e] < sn[e] ? sn[e] : sx[e], lo = sn[e] < sx[e] ? sn[e] : sx[e];
        ap[e] = hi < 0.0 ? 0.0 : hi;
        am[e] = 0.0 < lo ? 0.0 : lo;
        inv[e] = 1.0 / (ap[e] - am[e]);
    }
    e1_prim(g1, ql, wl);
    e1_prim(g1, qr, wr);
    e1_phys_flux(ql, wl, fl);
    e1_phys_flux(qr, wr, fr);
    for (int e = 0; e < 3; ++e) dd[e] = qr[e] - ql[e];
    e1_mv(L, dd, ld);
    for (int e = 0; e < 3; ++e) ld[e] = ((ap[e] * am[e]) * inv[e]) * ld[e];
    e1_mv(R, ld, rd);
    for (int e = 0; e < 3; ++e) f[e] = (ap[e] * fl[e] - am[e] * fr[e]) * inv[e] + rd[e];
}

// Flux through the face between cells c and c + 1 (c = -1 .. nx - 1):
// characteristic reconstruction (FV_System1D.cpp:194-208), back to
// conservative variables (FV_System1D.h:48-51), and the flux (:210-223).
HTX_HD void fv1d_face_flux(const Fv1dArgs& a, int c, double F[3]) {
    HTX_NO_CONTRACT
    double qa[3], qb[3], L[9], R[9];
    const int ia = map_index(c, a.nx, a.bc), ib = map_index(c + 1, a.nx, a.bc);
    for (int e = 0; e < 3; ++e) {
        qa[e] = a.qin[e * a.nx + ia];
        qb[e] = a.qin[e * a.nx + ib];
    }
    e1_eigen(a.g1, qa, qb, L, R);
    double C[3][6];
    for (int m = 0; m < 6; ++m) {
        const int im = map_index(c - 2 + m, a.nx, a.bc);
        double cc[3], y[3];
        for (int e = 0; e < 3; ++e) cc[e] = a.qin[e * a.nx + im];
        e1_mv(L, cc, y);
        for (int e = 0; e < 3; ++e) C[e][m] = y[e];
    }
    double ul[3], ur[3];
    for (int e = 0; e < 3; ++e) {
        if (a.recon == 1)
            teno5i(C[e], ul[e], ur[e]);
        else if (a.recon == 2 || a.recon == 3)
            weno5(C[e], a.recon == 3, ul[e], ur[e]);
        else if (a.recon >= RECON_WENO5N)
            recon5_variant(a.recon, C[e], a.dx, ul[e], ur[e]);
        else
            teno5(C[e], ul[e], ur[e]);
    }
    double ql[3], qr[3];
    e1_mv(R, ul, ql);
    e1_mv(R, ur, qr);
    if (a.flux == FLUX_HLLC)
        e1_hllc<true>(a.gam, a.g1, ql, qr, F);
    else if (a.flux == FLUX_HLLC_TORO)
        e1_hllc<false>(a.gam, a.g1, ql, qr, F);
    else if (a.flux == FLUX_CU)
        e1_cuflux(a.gam, a.g1, L, R, ql, qr, F);
    else
        e1_rusanov_char(a.gam, a.g1, L, R, ql, qr, F);
}

// residual (FV_System1D.cpp:225-227) and SSP-RK3 stage (:230-262 with
// Schemes_SSPRK.cpp:3-15) of cell c from the fluxes of its faces
HTX_HD double fv1d_stage_value(const Fv1dArgs& a, int c, int e, double fm, double fp, double dt) {
    HTX_NO_CONTRACT
    const double rhs = (fm - fp) / a.dx;
    const double us = a.qin[e * a.nx + c];
    if (a.mode == 0) return rhs;
    if (a.mode == 1) return us + dt * rhs;
    if (a.mode == 2) return 0.75 * a.q0[e * a.nx + c] + 0.25 * (us + rhs * dt);
    return (a.q0[e * a.nx + c] + 2.0 * (us + dt * rhs)) / 3.0;
}

// Euler1D::maxSpeed (Equations_Euler1D.cpp:91-94) of a conservative state
HTX_HD double fv1d_max_speed(double gam, double g1, const double q[3]) {
    HTX_NO_CONTRACT
    double w[3];
    e1_prim(g1, q, w);
    const double c = sqrt(fabs(gam * w[2] / w[0]));
    return fmax(fabs(w[1] - c), fabs(w[1] + c));
}

// KXRCF troubled-cell flag of cell c (FV::System1D::rhs, FV_System1D.cpp:
// 166-192): the density at the face upwind of the momentum's sign,
// reconstructed from the 3-cell stencils of the two cells sharing that face;
// flagged when |h1 - h2| / (radius |rho_c|) > 1, radius = (dx/2)^1.5 (:19).
// rho is the density row of the state, ghosts by the boundary condition.
// The reference's loop (i = s+1 .. nxx-s-2, s = 2) covers every cell; the
// first and last cells read the ghost values.
HTX_HD double fv1d_rho_at(const double* rho, int g, int nx, int bc) { return rho[map_index(g, nx, bc)]; }

HTX_HD bool fv1d_kxrcf(const double* rho, double mom, int c, int nx, int bc, double radius) {
    HTX_NO_CONTRACT
    const double rm2 = fv1d_rho_at(rho, c - 2, nx, bc), rm1 = fv1d_rho_at(rho, c - 1, nx, bc), r0 = rho[c];
    const double rp1 = fv1d_rho_at(rho, c + 1, nx, bc), rp2 = fv1d_rho_at(rho, c + 2, nx, bc);
    double h1, h2;
    if (mom < 0.0) {
        h1 = (2.0 * rm1 + 5.0 * r0 - rp1) / 6.0;
        h2 = (-rm2 + 5.0 * rm1 + 2.0 * r0) / 6.0;
    } else {
        h2 = (2.0 * r0 + 5.0 * rp1 - rp2) / 6.0;
        h1 = (-rm1 + 5.0 * r0 + 2.0 * rp1) / 6.0;
    }
    return fabs(h1 - h2) / (radius * fabs(r0)) > 1.0;
}

// LMPE troubled-cell detector of FVHermite::System1D::rhs
// (FVHermite_System1D.cpp:239-257), from a reset state.  The test at cell g
// of one field u (ghosts by the boundary condition):
//     A = |-u_{g-1} + 2 u_g - u_{g+1}| / 2 > dx / 2  and
//     0 <= xc <= 3 dx,  xc = (-2 u_{g-1} + 3 u_g - u_{g+1}) / (-u_{g-1} + 2 u_g - u_{g+1}) dx.
HTX_HD bool fv1d_lmpe_test(const double* u, int g, int nx, int bc, double dx) {
    HTX_NO_CONTRACT
    const double um = fv1d_rho_at(u, g - 1, nx, bc), u0 = fv1d_rho_at(u, g, nx, bc), up = fv1d_rho_at(u, g + 1, nx, bc);
    const double A = fabs(-um + 2.0 * u0 - up) / 2.0;
    if (!(A > 0.5 * dx)) return false;
    const double xc = (-2. * um + 3. * u0 - up) / (-um + 2.0 * u0 - up) * dx;
    return 0.0 <= xc && xc <= 3. * dx;
}

// The reference sweeps g = -2 .. nx+1 in order (k = 4 ghosts, s = 1) and
// tests a cell only when it is not flagged yet; a successful test flags
// g-1, g, g+1.  From a reset state a cell is flagged before its turn only by
// its left neighbour's test, so along a run of consecutive successful tests
// starting at r the tests at r, r+2, r+4, ... act and the others are
// skipped: active(g) = test(g) and (g - r) even.
HTX_HD bool fv1d_lmpe_active(const double* u, int g, int nx, int bc, double dx) {
    if (g < -2 || g > nx + 1 || !fv1d_lmpe_test(u, g, nx, bc, dx)) return false;
    int r = g;
    while (r - 1 >= -2 && fv1d_lmpe_test(u, r - 1, nx, bc, dx)) --r;
    return ((g - r) & 1) == 0;
}

// the flag of interior cell c after the sweep
HTX_HD bool fv1d_lmpe(const double* u, int c, int nx, int bc, double dx) {
    return fv1d_lmpe_active(u, c - 1, nx, bc, dx) || fv1d_lmpe_active(u, c, nx, bc, dx) ||
           fv1d_lmpe_active(u, c + 1, nx, bc, dx);
}

}  // namespace HTX_NS
