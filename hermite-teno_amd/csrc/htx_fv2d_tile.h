// htx_fv2d_tile.h -- the phases of the fused 2D FV stage kernel, written as
// host+device functions of (thread id, shared tile) so that the kernel and a
// CPU build of the same tiling (used by the tests) run identical code.
//
// Tile geometry (TX x TY cells, NT threads):
//   halo tile   (TX+6) x (TY+6) cells, LDS cell (lx, ly) = global (i0-3+lx, j0-3+ly)
//   x faces     recon (TX+1) x (TY+4): face fi sits between LDS cols 2+fi and 3+fi,
//               row fj at LDS row 1+fj; flux (TX+1) x TY at rows 3..TY+2
//   y faces     recon (TX+4) x (TY+1): column fi at LDS col 1+fi, face fj between
//               LDS rows 2+fj and 3+fj; flux TX x (TY+1) at cols 3..TX+2
#pragma once
#include "htx_learn.h"

namespace HTX_NS {

// FT: storage type of the characteristic face values (double; float in the
// FP32 bulk precision, which halves their LDS footprint)
template <int TX, int TY, typename FT = double, bool OF = false>
struct Fv2dTile {
    static constexpr int NT = 256;
    static constexpr int HX = TX + 6, HY = TY + 6, HC = HX * HY;
    static constexpr int NXF = (TX + 1) * (TY + 4);
    static constexpr int NYF = (TX + 4) * (TY + 1);
    static constexpr int NF = NXF > NYF ? NXF : NYF;
    static constexpr int NXQ = (TX + 1) * TY;
    static constexpr int NYQ = TX * (TY + 1);
    static constexpr int NQ = NXQ > NYQ ? NXQ : NYQ;
    static constexpr int NC = TX * TY;
    static constexpr int KC = (NC + NT - 1) / NT;
    double sQ[4][HC];  // conserved state
    // sqrt|rho|, u, v, H of each cell; with OF the tile keeps none and forms
    // them from sQ where a face needs them (a third less LDS: three
    // workgroups per CU instead of two, for a few more operations per face)
    double sP[OF ? 1 : 4][OF ? 1 : HC];
    FT sF[8][NF];      // characteristic face values (left/right per field)
    double sFl[4][NQ]; // face fluxes
    double sRed[2][NT / 64];
    unsigned char sFlag[2][NF];  // learned precision: 1 = linear scheme on this recon face
    unsigned int sCnt[3][NT / 64];
};

// Which stage kernels keep no per-cell Roe inputs in LDS: the strict build
// and the learned FP64 path, whose registers fit the three-workgroup budget
// (168 VGPRs) that the smaller tile opens.  The fp64_fast and fp32_mixed
// kernels need more registers than that and keep the stored inputs at two
// workgroups per CU.
template <bool MIXED, bool LEARNED, bool BULK32>
constexpr bool tile_prim_on_the_fly() {
    return HTX_EXACT_DIVISION || (LEARNED && !BULK32);
}

// halo load with the boundary condition applied through the index, plus the
// per-cell Roe inputs of calcPrimitiveVariable (FV_System2D.cpp:226-263)
template <int TX, int TY, typename FT, bool OF>
HTX_HD void ph_load(const Fv2dArgs& a, Fv2dTile<TX, TY, FT, OF>& s, int i0, int j0, int tid) {
    using T = Fv2dTile<TX, TY, FT, OF>;
    const long long fs = a.fstride;
    for (int w = tid; w < T::HC; w += T::NT) {
        int lx = w % T::HX, ly = w / T::HX;
        int gi = map_index(i0 - 3 + lx, a.nx, a.bc);
        int gj = map_index(j0 - 3 + ly, a.ny, a.bc);
        long long off = (long long)gj * a.pitch + gi;
        double q0 = a.qin[off], q1 = a.qin[off + fs], q2 = a.qin[off + 2 * fs], q3 = a.qin[off + 3 * fs];
        s.sQ[0][w] = q0;
        s.sQ[1][w] = q1;
        s.sQ[2][w] = q2;
        s.sQ[3][w] = q3;
        if constexpr (!OF) {
            Prim P = cell_prim(a.g1, q0, q1, q2, q3);
            s.sP[0][w] = P.sr;
            s.sP[1][w] = P.u;
            s.sP[2][w] = P.v;
            s.sP[3][w] = P.h;
        }
    }
}

template <int TX, int TY, typename FT, bool OF>
HTX_HD Prim prim_at(const Fv2dTile<TX, TY, FT, OF>& s, int id, double g1) {
    if constexpr (OF)
        return cell_prim(g1, s.sQ[0][id], s.sQ[1][id], s.sQ[2][id], s.sQ[3][id]);
    else
        return Prim{s.sP[0][id], s.sP[1][id], s.sP[2][id], s.sP[3][id]};
}

// Learned precision: the classifier flag of face g (cells g | g+1 along a
// grid line of n cells) from a line of flags stored for g = -1 .. n; faces
// further out follow the boundary condition (periodic: the same face of
// the wrapped grid) or take TENO5 (zero gradient: they feed only halo faces
// whose fluxes no cell uses).
HTX_HD unsigned char line_flag(const unsigned char* line, int g, int n, int bc, int stride) {
    if (g >= -1 && g <= n) return line[(long long)(g + 1) * stride];
    if (bc == 1) {
        g %= n;
        g = g < 0 ? g + n : g;
        return line[(long long)(g + 1) * stride];
    }
    return 0;
}

// Learned precision: the flags of the tile's recon faces, from the face-flag
// arrays of the step (rows of x faces and columns of y faces outside the grid
// map like the cells their stencils read).
template <int TX, int TY, typename FT, bool OF>
HTX_HD void ph_flags_load(const Fv2dArgs& a, Fv2dTile<TX, TY, FT, OF>& s, int i0, int j0, int tid) {
    using T = Fv2dTile<TX, TY, FT, OF>;
    for (int w = tid; w < T::NXF; w += T::NT) {
        int gi = i0 - 1 + w % (TX + 1), gj = map_index(j0 - 2 + w / (TX + 1), a.ny, a.bc);
        s.sFlag[0][w] = line_flag(a.face_flags_x + (long long)gj * (a.nx + 2), gi, a.nx, a.bc, 1);
    }
    for (int w = tid; w < T::NYF; w += T::NT) {
        int gi = map_index(i0 - 2 + w % (TX + 4), a.nx, a.bc), gj = j0 - 1 + w / (TX + 4);
        s.sFlag[1][w] = line_flag(a.face_flags_y + gi, gj, a.ny, a.bc, a.nx);
    }
}

// FP64 positivity of the conservative state that characteristic face values
// c[0..3] back-project to: density > 0 and pressure (gamma - 1)(E - m^2 / 2
// rho) > 0, the latter tested as rho E - m^2 / 2 > 0 (same sign for rho > 0
// and gamma > 1, no division).
HTX_HD bool face_state_positive(const Eig& E, const double c[4]) {
    double q0, qn, qt, q3;
    eig_right(E, c[0], c[1], c[2], c[3], q0, qn, qt, q3);
    return q0 > 0.0 && q3 * q0 - 0.5 * (qn * qn + qt * qt) > 0.0;
}


// Whether recon face (fi, fj) of family YDIR is one the tile owns: every
// global face is owned by exactly one tile (an x face by the tile of its left
// cell, a y face by the tile of its lower cell).
template <int TX, int TY, bool YDIR>
HTX_HD bool recon_face_owned(int fi, int fj, int i0, int j0, int nx, int ny) {
    if (!YDIR) return fi >= 1 && fi <= TX && fj >= 2 && fj <= TY + 1 && i0 + fi - 1 < nx && j0 + fj - 2 < ny;
    return fj >= 1 && fj <= TY && fi >= 2 && fi <= TX + 1 && j0 + fj - 1 < ny && i0 + fi - 2 < nx;
}

// characteristic TENO5 on faces (FV_System2D.cpp:372-400).  YDIR selects the
// face family; the state is read in (rho, normal, tangential, E) order.  With
// LEARNED, faces the classifier flagged take Lagrange::linear
// (Schemes_Lagrange.cpp:1219-1233) instead, provided the linear face states
// have positive density and pressure on both sides (checked in FP64; else
// TENO5); the face's flag is then overwritten with the scheme it took (1 =
// linear), and cnt accumulates (owned faces, flagged, linear) of this thread.
template <int TX, int TY, bool YDIR, bool MIXED = false, bool LEARNED = false, int RECON = 0, typename FT = double,
          bool OF = false>
HTX_HD void ph_recon(const Fv2dArgs& a, Fv2dTile<TX, TY, FT, OF>& s, int tid, int i0 = 0, int j0 = 0,
                     unsigned* cnt = nullptr) {
    using T = Fv2dTile<TX, TY, FT, OF>;
    constexpr int NW = YDIR ? T::NYF : T::NXF;
    constexpr int RW = YDIR ? (TX + 4) : (TX + 1);   // faces per row of the face grid
    constexpr int STEP = YDIR ? T::HX : 1;           // LDS stride along the normal
    for (int w = tid; w < NW; w += T::NT) {
        int fi = w % RW, fj = w / RW;
        int ca = YDIR ? (2 + fj) * T::HX + (1 + fi) : (1 + fj) * T::HX + (2 + fi);
        Eig E = roe_eig(a.g1, YDIR, prim_at(s, ca, a.g1), prim_at(s, ca + STEP, a.g1));
        bool lin = false;
        if (LEARNED) {
            const bool flag = s.sFlag[YDIR ? 1 : 0][w] != 0;
            if (flag) {
                // The linear scheme commutes with the characteristic
                // projection: reconstruct the conserved variables and
                // project the two face states only (instead of six cells).
                double um[4], up[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    double u[6];
#pragma unroll
                    for (int m = 0; m < 6; ++m) u[m] = s.sQ[e][ca + (m - 2) * STEP];
                    linear5(u, um[e], up[e]);
                }
                // density and pressure (gamma - 1)(E - m^2 / 2 rho) positive, FP64
                lin = um[0] > 0.0 && um[3] * um[0] - 0.5 * (um[1] * um[1] + um[2] * um[2]) > 0.0 && up[0] > 0.0 &&
                      up[3] * up[0] - 0.5 * (up[1] * up[1] + up[2] * up[2]) > 0.0;
                if (lin) {
                    double cm[4], cp[4];
                    eig_left<YDIR>(E, um[0], um[1], um[2], um[3], cm[0], cm[1], cm[2], cm[3]);
                    eig_left<YDIR>(E, up[0], up[1], up[2], up[3], cp[0], cp[1], cp[2], cp[3]);
#pragma unroll
                    for (int f = 0; f < 4; ++f) {
                        s.sF[2 * f][w] = cm[f];
                        s.sF[2 * f + 1][w] = cp[f];
                    }
                }
            }
            if (cnt && recon_face_owned<TX, TY, YDIR>(fi, fj, i0, j0, a.nx, a.ny)) {
                cnt[0] += 1;
                cnt[1] += flag ? 1 : 0;
                cnt[2] += lin ? 1 : 0;
            }
            // from here on the flag records what the face took (the flux
            // phase of this direction reads it)
            s.sFlag[YDIR ? 1 : 0][w] = lin ? 1 : 0;
        }
        if (!lin) {
            double c[4][6];
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                int id = ca + (m - 2) * STEP;
                eig_left<YDIR>(E, s.sQ[0][id], s.sQ[1][id], s.sQ[2][id], s.sQ[3][id], c[0][m], c[1][m], c[2][m],
                               c[3][m]);
            }
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                double um, up;
                if (MIXED)
                    teno5_mixed(c[f], um, up);
                else if (RECON == 1)
                    teno5i(c[f], um, up);
                else if (RECON == 2 || RECON == 3)
                    weno5(c[f], RECON == 3, um, up);
                else if (RECON == RECON_WENO5N)  // the other variants, chosen at run time
                    recon5_variant(a.recon, c[f], YDIR ? a.dy : a.dx, um, up);
                else
                    teno5(c[f], um, up);
                s.sF[2 * f][w] = um;
                s.sF[2 * f + 1][w] = up;
            }
        }
    }
}

// Learned path: whether flux face (fi, fj) of family YDIR may take the
// linear transverse quadrature.  Its 5 transverse recon faces must all have
// taken the linear scheme (normal smoothness along the window), and the
// classifier must have flagged smooth the 2 x 2 recon faces of the other
// family that sit on the face's two cells across the face's row (x flux)
// or column (y flux): their 6-cell stencils cover the window's transverse
// extent, so a discontinuity running along the face normal through the
// window is seen there.
template <int TX, int TY, bool YDIR, typename FT, bool OF>
HTX_HD bool flux_window_smooth(const Fv2dTile<TX, TY, FT, OF>& s, int fi, int fj) {
    if (!YDIR) {
        const unsigned char* N = s.sFlag[0];
        const int b = fj * (TX + 1) + fi;
        bool ok = N[b] & N[b + (TX + 1)] & N[b + 2 * (TX + 1)] & N[b + 3 * (TX + 1)] & N[b + 4 * (TX + 1)];
        const unsigned char* C = s.sFlag[1];  // y recon face (column 1 + fi, between rows 2 + fj, 3 + fj)
        const int c = fj * (TX + 4) + 1 + fi;
        return ok && (C[c] & C[c + 1] & C[c + (TX + 4)] & C[c + (TX + 4) + 1]);
    }
    const unsigned char* N = s.sFlag[1];
    const int b = fj * (TX + 4) + fi;
    bool ok = N[b] & N[b + 1] & N[b + 2] & N[b + 3] & N[b + 4];
    const unsigned char* C = s.sFlag[0];  // x recon face (between columns 2 + fi, 3 + fi, row 1 + fj)
    const int c = (1 + fj) * (TX + 1) + fi;
    return ok && (C[c] & C[c + 1] & C[c + (TX + 1)] & C[c + (TX + 1) + 1]);
}

// Whether every flux face of the wavefront that handles flux face w (work
// item w = tid + k NT of ph_flux) has a smooth window.  On the device a
// wavefront vote over its active lanes; on the host, where a phase runs item
// by item, the same vote taken over the items of that wavefront.
template <int TX, int TY, bool YDIR, typename FT, bool OF>
HTX_HD bool window_smooth_wave(const Fv2dTile<TX, TY, FT, OF>& s, int w, int tid) {
    using T = Fv2dTile<TX, TY, FT, OF>;
    constexpr int NW = YDIR ? T::NYQ : T::NXQ;
    constexpr int RW = YDIR ? TX : (TX + 1);
#if defined(__HIP_DEVICE_COMPILE__)
    (void)tid;
    (void)NW;
    return __all(flux_window_smooth<TX, TY, YDIR>(s, w % RW, w / RW) ? 1 : 0) != 0;
#else
    const int w0 = w - (tid & 63);
    bool all = true;
    for (int l = 0; l < 64 && w0 + l < NW; ++l)
        all = all && flux_window_smooth<TX, TY, YDIR>(s, (w0 + l) % RW, (w0 + l) / RW);
    return all;
#endif
}

// FP32 bulk precision of ph_flux: the transverse quadrature (or, on smooth
// windows of the learned path, its linear limit), the back-projection with
// the face's R and the Riemann flux at the 3 Gauss points in FP32, the
// Gauss-weighted sum in FP64.  A Gauss-point state whose density is not
// positive and well above the cancellation of its back-projection
// (rho > 2^-10 (|c0| + |c2| + |c3|)), or whose pressure is not positive and
// at least 2^-10 of its energy (rho E - m^2 / 2 > 2^-10 rho E), sends the
// face to the FP64 path: back-projection and flux recomputed in FP64 from
// the same Gauss-point values (on the learned path's linear windows, first
// the nonlinear quadrature).  The margins are state_fp32_ok's.
template <int TX, int TY, bool YDIR, int FLUX, bool LEARNED, bool OF>
HTX_HD void ph_flux32(const Fv2dArgs& a, Fv2dTile<TX, TY, float, OF>& s, int tid) {
    using T = Fv2dTile<TX, TY, float, OF>;
    constexpr int NW = YDIR ? T::NYQ : T::NXQ;
    constexpr int RW = YDIR ? TX : (TX + 1);
    constexpr int RR = YDIR ? (TX + 4) : (TX + 1);
    constexpr int TS = YDIR ? 1 : (TX + 1);
    constexpr int STEP = YDIR ? T::HX : 1;
    const float gam = (float)a.gam, g1 = (float)a.g1;
    for (int w = tid; w < NW; w += T::NT) {
        const int fi = w % RW, fj = w / RW;
        const int b = fj * RR + fi;
        const int ca = YDIR ? (2 + fj) * T::HX + (3 + fi) : (3 + fj) * T::HX + (2 + fi);
        const Eig E = roe_eig(a.g1, YDIR, prim_at(s, ca, a.g1), prim_at(s, ca + STEP, a.g1));
        const Eig32 E32 = eig32(E);
        float gl[4][3], gr[4][3];
        bool lin = LEARNED && window_smooth_wave<TX, TY, YDIR>(s, w, tid);
        auto quadrature = [&](bool linear) {
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const float* L = s.sF[2 * f];
                const float* R = s.sF[2 * f + 1];
                if (linear) {
                    weno3q_linear_f32(kQuad, L[b], L[b + TS], L[b + 2 * TS], L[b + 3 * TS], L[b + 4 * TS], gl[f][0],
                                      gl[f][1], gl[f][2]);
                    weno3q_linear_f32(kQuad, R[b], R[b + TS], R[b + 2 * TS], R[b + 3 * TS], R[b + 4 * TS], gr[f][0],
                                      gr[f][1], gr[f][2]);
                } else {
                    weno3q_f32(kQuad, L[b], L[b + TS], L[b + 2 * TS], L[b + 3 * TS], L[b + 4 * TS], gl[f][0], gl[f][1],
                               gl[f][2]);
                    weno3q_f32(kQuad, R[b], R[b + TS], R[b + 2 * TS], R[b + 3 * TS], R[b + 4 * TS], gr[f][0], gr[f][1],
                               gr[f][2]);
                }
            }
        };
        quadrature(lin);
        double acc[3][4];
        bool ok = true;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            float ql[4], qr[4], fl[4];
            eig_right_f32(E32, gl[0][m], gl[1][m], gl[2][m], gl[3][m], ql[0], ql[1], ql[2], ql[3]);
            eig_right_f32(E32, gr[0][m], gr[1][m], gr[2][m], gr[3][m], qr[0], qr[1], qr[2], qr[3]);
            ok = ok && state_fp32_ok(ql, gl[0][m], gl[2][m], gl[3][m]) && state_fp32_ok(qr, gr[0][m], gr[2][m], gr[3][m]);
            if (FLUX == FLUX_RUSANOV_CHAR)
                rusanov_char_f32(gam, g1, E32, ql, qr, fl);
            else if (FLUX == FLUX_HLLC)
                hllc_f32<true>(gam, g1, ql, qr, fl);
            else
                hllc_f32<false>(gam, g1, ql, qr, fl);
            const double gw = (m == 1) ? 8.0 / 9.0 / 2.0 : 5.0 / 9.0 / 2.0;
            acc[m][0] = gw * (double)fl[0];
            acc[m][1] = gw * (double)(YDIR ? fl[2] : fl[1]);
            acc[m][2] = gw * (double)(YDIR ? fl[1] : fl[2]);
            acc[m][3] = gw * (double)fl[3];
        }
        if (!ok) {
            if (lin) quadrature(false);
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                double ql[4], qr[4], fl[4];
                eig_right(E, gl[0][m], gl[1][m], gl[2][m], gl[3][m], ql[0], ql[1], ql[2], ql[3]);
                eig_right(E, gr[0][m], gr[1][m], gr[2][m], gr[3][m], qr[0], qr[1], qr[2], qr[3]);
                if (FLUX == FLUX_RUSANOV_CHAR)
                    rusanov_char<YDIR>(a.gam, a.g1, E, ql, qr, fl);
                else if (FLUX == FLUX_HLLC)
                    hllc<true>(a.gam, a.g1, ql, qr, fl);
                else
                    hllc<false>(a.gam, a.g1, ql, qr, fl);
                const double gw = (m == 1) ? 8.0 / 9.0 / 2.0 : 5.0 / 9.0 / 2.0;
                acc[m][0] = gw * fl[0];
                acc[m][1] = gw * (YDIR ? fl[2] : fl[1]);
                acc[m][2] = gw * (YDIR ? fl[1] : fl[2]);
                acc[m][3] = gw * fl[3];
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) s.sFl[e][w] = (acc[0][e] + acc[1][e]) + acc[2][e];
    }
}

// transverse 3-point quadrature, back-projection with the face's own R,
// Riemann flux at the Gauss points weighted by (5/18, 4/9, 5/18)
// (FV_System2D.cpp:402-497); the flux is stored in (rho, x, y, E) order.
// With LEARNED, when every face a wavefront handles has a smooth window
// (flux_window_smooth), those faces take the quadrature's linear limit,
// except a face whose back-projected Gauss-point states then include a
// non-positive density or pressure (checked in FP64); every other face takes
// the nonlinear quadrature.  The wavefront-wide decision keeps the two paths
// from running one after the other in the same wavefront.
template <int TX, int TY, bool YDIR, int FLUX, bool MIXED = false, bool LEARNED = false, typename FT = double, bool OF = false>
HTX_HD void ph_flux(const Fv2dArgs& a, Fv2dTile<TX, TY, FT, OF>& s, int tid) {
    if constexpr (sizeof(FT) == sizeof(float)) {
        ph_flux32<TX, TY, YDIR, FLUX, LEARNED>(a, s, tid);
    } else {
        using T = Fv2dTile<TX, TY, FT, OF>;
        constexpr int NW = YDIR ? T::NYQ : T::NXQ;
        constexpr int RW = YDIR ? TX : (TX + 1);          // flux faces per row
        constexpr int RR = YDIR ? (TX + 4) : (TX + 1);    // recon faces per row
        constexpr int TS = YDIR ? 1 : (TX + 1);           // recon stride along the transverse direction
        constexpr int STEP = YDIR ? T::HX : 1;
        for (int w = tid; w < NW; w += T::NT) {
            int fi = w % RW, fj = w / RW;
            // first of the 5 transverse recon faces, and the face's left/lower cell
            int b = YDIR ? fj * RR + fi : fj * RR + fi;
            int ca = YDIR ? (2 + fj) * T::HX + (3 + fi) : (3 + fj) * T::HX + (2 + fi);
            // the learned path needs the eigensystem for its positivity check;
            // the others form it after the quadrature (fewer live registers)
            Eig E;
            if (LEARNED) E = roe_eig(a.g1, YDIR, prim_at(s, ca, a.g1), prim_at(s, ca + STEP, a.g1));
            double gl[4][3], gr[4][3];
            bool linq = false;
            if (LEARNED && window_smooth_wave<TX, TY, YDIR>(s, w, tid)) {
    #pragma unroll
                for (int f = 0; f < 4; ++f) {
                    const double* L = s.sF[2 * f];
                    const double* R = s.sF[2 * f + 1];
                    weno3q_linear(kQuad, L[b], L[b + TS], L[b + 2 * TS], L[b + 3 * TS], L[b + 4 * TS], gl[f][0], gl[f][1],
                                  gl[f][2]);
                    weno3q_linear(kQuad, R[b], R[b + TS], R[b + 2 * TS], R[b + 3 * TS], R[b + 4 * TS], gr[f][0], gr[f][1],
                                  gr[f][2]);
                }
                linq = true;
    #pragma unroll
                for (int m = 0; m < 3; ++m) {
                    double gm[4] = {gl[0][m], gl[1][m], gl[2][m], gl[3][m]};
                    double gp[4] = {gr[0][m], gr[1][m], gr[2][m], gr[3][m]};
                    linq = linq && face_state_positive(E, gm) && face_state_positive(E, gp);
                }
            }
            if (!linq) {
    #pragma unroll
                for (int f = 0; f < 4; ++f) {
                    const double* L = s.sF[2 * f];
                    const double* R = s.sF[2 * f + 1];
                    if (MIXED) {
                        weno3q_mixed(kQuad, kQuad.lw32, L[b], L[b + TS], L[b + 2 * TS], L[b + 3 * TS], L[b + 4 * TS],
                                     gl[f][0], gl[f][1], gl[f][2]);
                        weno3q_mixed(kQuad, kQuad.lw32, R[b], R[b + TS], R[b + 2 * TS], R[b + 3 * TS], R[b + 4 * TS],
                                     gr[f][0], gr[f][1], gr[f][2]);
                    } else {
                        weno3q(kQuad, L[b], L[b + TS], L[b + 2 * TS], L[b + 3 * TS], L[b + 4 * TS], gl[f][0], gl[f][1],
                               gl[f][2]);
                        weno3q(kQuad, R[b], R[b + TS], R[b + 2 * TS], R[b + 3 * TS], R[b + 4 * TS], gr[f][0], gr[f][1],
                               gr[f][2]);
                    }
                }
            }
            if (!LEARNED) E = roe_eig(a.g1, YDIR, prim_at(s, ca, a.g1), prim_at(s, ca + STEP, a.g1));
            double acc[3][4];
    #pragma unroll
            for (int m = 0; m < 3; ++m) {
                double ql[4], qr[4], fl[4];
                eig_right(E, gl[0][m], gl[1][m], gl[2][m], gl[3][m], ql[0], ql[1], ql[2], ql[3]);
                eig_right(E, gr[0][m], gr[1][m], gr[2][m], gr[3][m], qr[0], qr[1], qr[2], qr[3]);
                if (FLUX == FLUX_RUSANOV_CHAR)
                    rusanov_char<YDIR>(a.gam, a.g1, E, ql, qr, fl);
                else if (FLUX == FLUX_HLLC)
                    hllc<true>(a.gam, a.g1, ql, qr, fl);
                else
                    hllc<false>(a.gam, a.g1, ql, qr, fl);
                const double gw = (m == 1) ? 8.0 / 9.0 / 2.0 : 5.0 / 9.0 / 2.0;
                acc[m][0] = gw * fl[0];
                acc[m][1] = gw * (YDIR ? fl[2] : fl[1]);
                acc[m][2] = gw * (YDIR ? fl[1] : fl[2]);
                acc[m][3] = gw * fl[3];
            }
    #pragma unroll
            for (int e = 0; e < 4; ++e) s.sFl[e][w] = (acc[0][e] + acc[1][e]) + acc[2][e];
        }
    }
}

// x part of the divergence, (fx_{i-1} - fx_i) * dy, kept per thread
template <int TX, int TY, typename FT, bool OF>
HTX_HD void ph_xdiv(const Fv2dArgs& a, const Fv2dTile<TX, TY, FT, OF>& s, int tid, double dxp[][4]) {
    using T = Fv2dTile<TX, TY, FT, OF>;
#pragma unroll
    for (int k = 0; k < T::KC; ++k) {
        int w = tid + k * T::NT;
        if (w < T::NC) {
            int c = w % TX, r = w / TX;
            int b = r * (TX + 1) + c;
#pragma unroll
            for (int e = 0; e < 4; ++e) dxp[k][e] = (s.sFl[e][b] - s.sFl[e][b + 1]) * a.dy;
        }
    }
}

// residual (FV_System2D.cpp:562-572), SSP-RK3 stage (:575-612 with
// Schemes_SSPRK.cpp:3-15) and, in stage 3, the per-thread CFL maxima
// (:623-631 with Equations_Euler2D.cpp:207-219)
template <int TX, int TY, typename FT, bool OF>
HTX_HD void ph_epilogue(const Fv2dArgs& a, const Fv2dTile<TX, TY, FT, OF>& s, int i0, int j0, int tid,
                        const double dxp[][4], double dt, double& smx, double& smy, bool* bad = nullptr) {
    using T = Fv2dTile<TX, TY, FT, OF>;
    const long long fs = a.fstride;
#pragma unroll
    for (int k = 0; k < T::KC; ++k) {
        int w = tid + k * T::NT;
        if (w >= T::NC) continue;
        int c = w % TX, r = w / TX;
        int gi = i0 + c, gj = j0 + r;
        if (gi >= a.nx || gj >= a.ny) continue;
        long long off = (long long)gj * a.pitch + gi;
        int lc = (3 + r) * T::HX + (3 + c);
        double unew[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
#if HTX_EXACT_DIVISION
            double rhs = (dxp[k][e] + (s.sFl[e][r * TX + c] - s.sFl[e][(r + 1) * TX + c]) * a.dx) / a.area;
#else
            double rhs = (dxp[k][e] + (s.sFl[e][r * TX + c] - s.sFl[e][(r + 1) * TX + c]) * a.dx) * drcp(a.area);
#endif
            double us = s.sQ[e][lc];
            double v;
            if (a.mode == 0)
                v = rhs;
            else if (a.mode == 1)
                v = us + dt * rhs;
            else if (a.mode == 2)
                v = 0.75 * a.q0[off + e * fs] + 0.25 * (us + rhs * dt);
            else
                v = (a.q0[off + e * fs] + 2.0 * (us + dt * rhs)) / 3.0;
            unew[e] = v;
            a.qout[off + e * fs] = v;
        }
        if (a.mode == 3) {
            if (bad && !all_finite(unew, 4)) *bad = true;
#if HTX_EXACT_DIVISION
            double wu = unew[1] / unew[0], wv = unew[2] / unew[0];
#else
            const double ir = drcp(unew[0]);
            double wu = unew[1] * ir, wv = unew[2] * ir;
#endif
            double p = (unew[3] - 0.5 * (unew[1] * wu + unew[2] * wv)) * a.g1;
            if (a.quirks) {
                smx = fmax(smx, max_speed<true>(a.gam, wu, p, unew[0]));
                smy = fmax(smy, max_speed<true>(a.gam, wv, p, unew[0]));
            } else {
                smx = fmax(smx, max_speed<false>(a.gam, wu, p, unew[0]));
                smy = fmax(smy, max_speed<false>(a.gam, wv, p, unew[0]));
            }
        }
    }
}

__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// One RK stage (or residual) of tile `tile` of the instance described by a:
// halo load, x and y reconstruction and fluxes, epilogue, and in stage 3
// the tile's CFL maxima (one atomic per workgroup).  dt = the stage's time
// step.  Ends with every thread past its last LDS access only for mode 3;
// callers that reuse the tile storage synchronise before the next tile.
template <int TX, int TY, int FLUX, bool MIXED, bool LEARNED, int RECON = 0, typename FT, bool OF>
__device__ __forceinline__ void stage_tile(const Fv2dArgs& a, Fv2dTile<TX, TY, FT, OF>& s, int tile, double dt, int tid) {
    using T = Fv2dTile<TX, TY, FT, OF>;
    const int i0 = (tile % a.tiles_x) * TX;
    const int j0 = (tile / a.tiles_x) * TY;
    double dxp[T::KC][4];

    ph_load<TX, TY>(a, s, i0, j0, tid);
    if (LEARNED) ph_flags_load<TX, TY>(a, s, i0, j0, tid);
    __syncthreads();
    unsigned cnt[3] = {0, 0, 0};
    ph_recon<TX, TY, false, MIXED, LEARNED, RECON>(a, s, tid, i0, j0, cnt);
    __syncthreads();
    ph_flux<TX, TY, false, FLUX, MIXED, LEARNED>(a, s, tid);
    __syncthreads();
    ph_xdiv<TX, TY>(a, s, tid, dxp);
    __syncthreads();
    ph_recon<TX, TY, true, MIXED, LEARNED, RECON>(a, s, tid, i0, j0, cnt);
    if (LEARNED) {
        for (int k = 0; k < 3; ++k) {
            unsigned v = cnt[k];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if ((tid & 63) == 0) s.sCnt[k][tid >> 6] = v;
        }
        __syncthreads();
        if (tid == 0) {
            unsigned long long c3[3] = {0, 0, 0};
            for (int k = 0; k < 3; ++k)
                for (int w = 0; w < T::NT / 64; ++w) c3[k] += s.sCnt[k][w];
            atomicAdd(&a.ctl->n_faces, c3[0]);
            atomicAdd(&a.ctl->n_flagged, c3[1]);
            atomicAdd(&a.ctl->n_linear, c3[2]);
        }
    }
    __syncthreads();
    ph_flux<TX, TY, true, FLUX, MIXED, LEARNED>(a, s, tid);
    __syncthreads();
    double smx = 0.0, smy = 0.0;
    bool bad = false;
    ph_epilogue<TX, TY>(a, s, i0, j0, tid, dxp, dt, smx, smy, &bad);
    if (a.mode == 3) {
        flag_nonfinite(bad, a.ctl);
        smx = wave_max(smx);
        smy = wave_max(smy);
        if ((tid & 63) == 0) {
            s.sRed[0][tid >> 6] = smx;
            s.sRed[1][tid >> 6] = smy;
        }
        __syncthreads();
        if (tid == 0) {
            double mx = 0.0, my = 0.0;
#pragma unroll
            for (int k = 0; k < T::NT / 64; ++k) {
                mx = fmax(mx, s.sRed[0][k]);
                my = fmax(my, s.sRed[1][k]);
            }
            atomicMax(&a.ctl->smax_x, (unsigned long long)__double_as_longlong(mx));
            atomicMax(&a.ctl->smax_y, (unsigned long long)__double_as_longlong(my));
        }
    }
}

// Workgroups are dealt to the 8 XCDs round-robin; give each XCD one
// contiguous run of tiles so neighbouring tiles share that XCD's L2.
HTX_HD int xcd_remap(int b, int nb) {
    int per = nb / 8, rem = nb % 8;
    int x = b % 8, k = b / 8;
    int start = x * per + (x < rem ? x : rem);
    return start + k;
}

}  // namespace HTX_NS
