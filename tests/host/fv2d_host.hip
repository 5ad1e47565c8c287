// CPU build of the fused 2D FV stage kernel's tiling: runs the kernel's own
// phase functions (htx_fv2d_tile.h) for every thread id of every tile, with
// the barriers between phases honoured by running phases in order.  Test
// infrastructure: it checks the tile orchestration (halo indexing, boundary
// mapping, face bookkeeping) on the CPU, with no GPU involved.
#include <cstdlib>
#include <vector>

#include "../../hermite-teno_amd/csrc/htx_fv2d_tile.h"

using namespace htx;

// LEARNED: the tiles load their recon-face flags with the kernel's own
// ph_flags_load from face-flag arrays in the kernel's layout (fx: x faces
// i = -1 .. nx of each row at [j (nx + 2) + i + 1]; fy: y faces j = -1 .. ny
// of each column at [(j + 1) nx + i]); null arrays flag every face smooth.
template <int TX, int TY, bool MIXED = false, bool LEARNED = false, typename FT = double, int RECON = 0>
static void run(Fv2dArgs a, int flux, const unsigned char* fx = nullptr, const unsigned char* fy = nullptr) {
    // the tile layout the product kernel of this precision uses
    using T = Fv2dTile<TX, TY, FT, tile_prim_on_the_fly<MIXED, LEARNED, sizeof(FT) == sizeof(float)>()>;
    T* s = new T();
    a.tiles_x = (a.nx + TX - 1) / TX;
    a.tiles_y = (a.ny + TY - 1) / TY;
    std::vector<unsigned char> ones_x((size_t)(a.nx + 2) * a.ny, 1), ones_y((size_t)a.nx * (a.ny + 2), 1);
    a.face_flags_x = fx ? fx : ones_x.data();
    a.face_flags_y = fy ? fy : ones_y.data();
    std::vector<double> dxp((size_t)T::NT * T::KC * 4);
    for (int tile = 0; tile < a.tiles_x * a.tiles_y; ++tile) {
        int i0 = (tile % a.tiles_x) * TX, j0 = (tile / a.tiles_x) * TY;
        for (int t = 0; t < T::NT; ++t) ph_load<TX, TY>(a, *s, i0, j0, t);
        if (LEARNED)
            for (int t = 0; t < T::NT; ++t) ph_flags_load<TX, TY>(a, *s, i0, j0, t);
        for (int t = 0; t < T::NT; ++t) ph_recon<TX, TY, false, MIXED, LEARNED, RECON>(a, *s, t);
        for (int t = 0; t < T::NT; ++t) {
            if (flux == 0) ph_flux<TX, TY, false, 0, MIXED, LEARNED>(a, *s, t);
            else if (flux == 1) ph_flux<TX, TY, false, 1, MIXED, LEARNED>(a, *s, t);
            else ph_flux<TX, TY, false, 2, MIXED, LEARNED>(a, *s, t);
        }
        for (int t = 0; t < T::NT; ++t) ph_xdiv<TX, TY>(a, *s, t, (double(*)[4]) & dxp[(size_t)t * T::KC * 4]);
        for (int t = 0; t < T::NT; ++t) ph_recon<TX, TY, true, MIXED, LEARNED, RECON>(a, *s, t);
        for (int t = 0; t < T::NT; ++t) {
            if (flux == 0) ph_flux<TX, TY, true, 0, MIXED, LEARNED>(a, *s, t);
            else if (flux == 1) ph_flux<TX, TY, true, 1, MIXED, LEARNED>(a, *s, t);
            else ph_flux<TX, TY, true, 2, MIXED, LEARNED>(a, *s, t);
        }
        for (int t = 0; t < T::NT; ++t) {
            double smx = 0, smy = 0;
            ph_epilogue<TX, TY>(a, *s, i0, j0, t, (const double(*)[4]) & dxp[(size_t)t * T::KC * 4], 0.0, smx, smy);
        }
    }
    delete s;
}

// mixed: 0 FP64, 1 FP32-mixed, 2 learned with every face flagged smooth,
// 3 FP32 bulk, 4 learned on the FP32 bulk arithmetic, every face flagged;
// 5, 6, 7: FP64 with teno5i, weno5JS, weno5Z in place of TENO5; 8, 9: FP64 on
// 16 x 15 and 15 x 16 tiles; 100 + k: FP64 with reconstruction code k
// (RECON_WENO5N .. RECON_LAST, htx_variants.h) in place of TENO5
extern "C" int fv2dh_rhs_prec(int nx, int ny, double xl, double xr, double yb, double yt, double gamma, int bc,
                              int flux, int quirks, int mixed, const double* q, double* rhs) {
    Fv2dArgs a{};
    a.qin = q;
    a.q0 = q;
    a.qout = rhs;
    a.nx = nx;
    a.ny = ny;
    a.pitch = nx;
    a.fstride = (long long)nx * ny;
    a.bc = bc;
    a.mode = 0;
    a.quirks = quirks;
    a.gam = gamma;
    a.g1 = gamma - 1.0;
    a.dx = (xr - xl) / nx;
    a.dy = (yt - yb) / ny;
    a.area = a.dx * a.dy;
    fill_quad_consts(a.qc);
    int fk = flux == 0 ? 0 : (quirks ? 1 : 2);
    if (mixed == 8)
        run<16, 15>(a, fk);
    else if (mixed == 9)
        run<15, 16>(a, fk);

    else if (mixed >= 100) {  // 100 + k: the run-time variant k (RECON_WENO5N ..) on the product's 16 x 15 tiles
        a.recon = mixed - 100;
        run<16, 15, false, false, double, RECON_WENO5N>(a, fk);
    } else if (mixed == 5)
        run<16, 16, false, false, double, 1>(a, fk);
    else if (mixed == 6)
        run<16, 16, false, false, double, 2>(a, fk);
    else if (mixed == 7)
        run<16, 16, false, false, double, 3>(a, fk);
    else if (mixed == 4)
        run<16, 15, true, true, float>(a, fk);
    else if (mixed == 3)
        run<16, 16, true, false, float>(a, fk);
    else if (mixed == 2)
        run<16, 15, true, true>(a, fk);
    else if (mixed)
        run<16, 16, true>(a, fk);
    else
        run<16, 16, false>(a, fk);
    return 0;
}

extern "C" int fv2dh_rhs(int nx, int ny, double xl, double xr, double yb, double yt, double gamma, int bc, int flux,
                         int quirks, const double* q, double* rhs) {
    return fv2dh_rhs_prec(nx, ny, xl, xr, yb, yt, gamma, bc, flux, quirks, 0, q, rhs);
}

// one TENO5 evaluation (u_{i-2..i+3} -> u^-, u^+) with the FP64 rule or the
// FP32-mixed rule of the kernels
extern "C" void fv2dh_teno5(const double* u, int mixed, double* out) {
    if (mixed)
        teno5_mixed(u, out[0], out[1]);
    else
        teno5(u, out[0], out[1]);
}

// the transverse quadrature's linear limit (5 values -> +xi, centre, -xi)
// with the constants the kernels receive
extern "C" void fv2dh_linquad(const double* u, double* out) {
    QuadConsts C{};
    fill_quad_consts(C);
    weno3q_linear(C, u[0], u[1], u[2], u[3], u[4], out[0], out[1], out[2]);
}

// residual of the learned path with given face flags (the kernel's layout, see run)
extern "C" int fv2dh_rhs_flags(int nx, int ny, double xl, double xr, double yb, double yt, double gamma, int bc,
                               int flux, int quirks, const double* q, double* rhs, const unsigned char* fx,
                               const unsigned char* fy) {
    Fv2dArgs a{};
    a.qin = q;
    a.q0 = q;
    a.qout = rhs;
    a.nx = nx;
    a.ny = ny;
    a.pitch = nx;
    a.fstride = (long long)nx * ny;
    a.bc = bc;
    a.mode = 0;
    a.quirks = quirks;
    a.gam = gamma;
    a.g1 = gamma - 1.0;
    a.dx = (xr - xl) / nx;
    a.dy = (yt - yb) / ny;
    a.area = a.dx * a.dy;
    fill_quad_consts(a.qc);
    run<16, 15, true, true>(a, flux == 0 ? 0 : (quirks ? 1 : 2), fx, fy);
    return 0;
}

// features of every owned recon face as the stage kernel's fused classifier
// forms them from its tile (x faces into featx[j nx + i][12], y faces into
// featy), for comparison with the standalone extraction
extern "C" int fv2dh_tile_features(int nx, int ny, double gamma, int bc, const double* q, float* featx,
                                   float* featy) {
    using T = Fv2dTile<16, 16>;
    Fv2dArgs a{};
    a.qin = q;
    a.nx = nx;
    a.ny = ny;
    a.pitch = nx;
    a.fstride = (long long)nx * ny;
    a.bc = bc;
    a.gam = gamma;
    a.g1 = gamma - 1.0;
    a.tiles_x = (nx + 15) / 16;
    a.tiles_y = (ny + 15) / 16;
    T* s = new T();
    for (int tile = 0; tile < a.tiles_x * a.tiles_y; ++tile) {
        int i0 = (tile % a.tiles_x) * 16, j0 = (tile / a.tiles_x) * 16;
        for (int t = 0; t < T::NT; ++t) ph_load<16, 16>(a, *s, i0, j0, t);
        for (int w = 0; w < T::NXF + T::NYF; ++w) {
            const bool ydir = w >= T::NXF;
            const int wl = ydir ? w - T::NXF : w;
            int ca, step, fi, fj;
            if (!ydir) {
                fi = wl % 17, fj = wl / 17;
                ca = (1 + fj) * T::HX + (2 + fi);
                step = 1;
                if (!recon_face_owned<16, 16, false>(fi, fj, i0, j0, nx, ny)) continue;
            } else {
                fi = wl % 20, fj = wl / 20;
                ca = (2 + fj) * T::HX + (1 + fi);
                step = T::HX;
                if (!recon_face_owned<16, 16, true>(fi, fj, i0, j0, nx, ny)) continue;
            }
            double r[6], un[6], ut[6], p[6];
            for (int m = 0; m < 6; ++m) {
                int id = ca + (m - 2) * step;
                double u = s->sP[1][id], v = s->sP[2][id];
                r[m] = s->sQ[0][id];
                p[m] = (s->sQ[3][id] - 0.5 * (s->sQ[1][id] * u + s->sQ[2][id] * v)) * a.g1;
                un[m] = ydir ? v : u;
                ut[m] = ydir ? u : v;
            }
            float* out = ydir ? featy + ((size_t)(j0 + fj - 1) * nx + (i0 + fi - 2)) * 12
                              : featx + ((size_t)(j0 + fj - 2) * nx + (i0 + fi - 1)) * 12;
            face_features(r, un, ut, p, a.gam, out);
        }
    }
    delete s;
    return 0;
}
