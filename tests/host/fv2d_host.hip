// CPU build of the fused 2D FV stage kernel's tiling: runs the kernel's own
// phase functions (htx_fv2d_tile.h) for every thread id of every tile, with
// the barriers between phases honoured by running phases in order.  Test
// infrastructure: it checks the tile orchestration (halo indexing, boundary
// mapping, face bookkeeping) on the CPU, with no GPU involved.
#include <cstdlib>
#include <vector>

#include "../../hermite-teno_amd/csrc/htx_fv2d_tile.h"

using namespace htx;

template <int TX, int TY, bool MIXED = false>
static void run(Fv2dArgs a, int flux) {
    using T = Fv2dTile<TX, TY>;
    T* s = new T();
    a.tiles_x = (a.nx + TX - 1) / TX;
    a.tiles_y = (a.ny + TY - 1) / TY;
    std::vector<double> dxp((size_t)T::NT * T::KC * 4);
    for (int tile = 0; tile < a.tiles_x * a.tiles_y; ++tile) {
        int i0 = (tile % a.tiles_x) * TX, j0 = (tile / a.tiles_x) * TY;
        for (int t = 0; t < T::NT; ++t) ph_load<TX, TY>(a, *s, i0, j0, t);
        for (int t = 0; t < T::NT; ++t) ph_recon<TX, TY, false, MIXED>(a, *s, t);
        for (int t = 0; t < T::NT; ++t) {
            if (flux == 0) ph_flux<TX, TY, false, 0, MIXED>(a, *s, t);
            else if (flux == 1) ph_flux<TX, TY, false, 1, MIXED>(a, *s, t);
            else ph_flux<TX, TY, false, 2, MIXED>(a, *s, t);
        }
        for (int t = 0; t < T::NT; ++t) ph_xdiv<TX, TY>(a, *s, t, (double(*)[4]) & dxp[(size_t)t * T::KC * 4]);
        for (int t = 0; t < T::NT; ++t) ph_recon<TX, TY, true, MIXED>(a, *s, t);
        for (int t = 0; t < T::NT; ++t) {
            if (flux == 0) ph_flux<TX, TY, true, 0, MIXED>(a, *s, t);
            else if (flux == 1) ph_flux<TX, TY, true, 1, MIXED>(a, *s, t);
            else ph_flux<TX, TY, true, 2, MIXED>(a, *s, t);
        }
        for (int t = 0; t < T::NT; ++t) {
            double smx = 0, smy = 0;
            ph_epilogue<TX, TY>(a, *s, i0, j0, t, (const double(*)[4]) & dxp[(size_t)t * T::KC * 4], 0.0, smx, smy);
        }
    }
    delete s;
}

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
    if (mixed)
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
