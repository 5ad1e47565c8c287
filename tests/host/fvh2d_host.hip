// CPU build of the Hermite-TENO kernels' per-line / per-face / per-cell
// bodies (htx_hermite.h): the limiter chains, face reconstructions, Gauss-point
// fluxes and the cell update run here in loops in the order the GPU launches
// them.  Test infrastructure, compared with the FP64 oracle.
#include <type_traits>
#include <vector>

#include "../../hermite-teno_amd/csrc/htx_hermite.h"

using namespace htx;

// BULK32: the FP32 bulk precision (with MIXED): FP32 face values and fluxes
template <bool MIXED, bool VARIANT = false, bool BULK32 = false>
static int rhs_impl(int nx, int ny, double xl, double xr, double yb, double yt, double gamma, int bc, int flux,
                    int quirks, const double* q, double* rhs, double* qlim, int hrecon = 0) {
    FvhArgs a{};
    size_t n = (size_t)nx * ny;
    std::vector<double> lv(8 * n), fx(24 * (size_t)ny * (nx + 1)), fy(24 * (size_t)(ny + 1) * nx),
        gx(12 * (size_t)ny * (nx + 1)), gy(12 * (size_t)(ny + 1) * nx);
    a.q = q;
    a.lv = lv.data();
    a.lv_out = lv.data();
    a.fx = fx.data();
    a.fy = fy.data();
    a.gx = gx.data();
    a.gy = gy.data();
    a.q0 = q;
    a.qout = rhs;
    a.nx = nx;
    a.ny = ny;
    a.pitch = nx;
    a.fstride = (long long)n;
    a.bc = bc;
    a.mode = 0;
    a.quirks = quirks;
    a.mixed = MIXED;
    a.bulk32 = BULK32;
    a.hrecon = hrecon;
    a.gam = gamma;
    a.g1 = gamma - 1.0;
    a.dx = (xr - xl) / nx;
    a.dy = (yt - yb) / ny;
    a.area = a.dx * a.dy;
    fill_quad_consts(a.qc);
    fill_herm_consts(a.hc, a.dx);
    int fk = flux == 0 ? 0 : (quirks ? 1 : 2);
    for (int e = 0; e < 4; ++e) {
        for (int j = 0; j < ny; ++j) fvh_limit_line(a, 0, j, e);
        for (int i = 0; i < nx; ++i) fvh_limit_line(a, 1, i, e);
    }
    long long fxs = (long long)ny * (nx + 1), fys = (long long)(ny + 1) * nx;
    using FT = typename std::conditional<BULK32, float, double>::type;
    FT* Fx = reinterpret_cast<FT*>(a.fx);
    FT* Fy = reinterpret_cast<FT*>(a.fy);
    for (int j = 0; j < ny; ++j)
        for (int fi = 0; fi <= nx; ++fi)
            fvh_recon_face<false, MIXED, VARIANT>(a, fi - 1, j, Fx, (long long)j * (nx + 1) + fi, fxs);
    for (int fj = 0; fj <= ny; ++fj)
        for (int i = 0; i < nx; ++i)
            fvh_recon_face<true, MIXED, VARIANT>(a, i, fj - 1, Fy, (long long)fj * nx + i, fys);
    double o[3][12];
    auto face = [&](auto ydir, int i, int j, const FT* F, long long fstr) {
        constexpr bool Y = decltype(ydir)::value;
        if constexpr (BULK32) {
            if (fk == 0) fvh_flux_face_f32<Y, 0>(a, i, j, F, fstr, o);
            else if (fk == 1) fvh_flux_face_f32<Y, 1>(a, i, j, F, fstr, o);
            else fvh_flux_face_f32<Y, 2>(a, i, j, F, fstr, o);
        } else {
            if (fk == 0) fvh_flux_face<Y, 0, MIXED>(a, i, j, F, fstr, o);
            else if (fk == 1) fvh_flux_face<Y, 1, MIXED>(a, i, j, F, fstr, o);
            else fvh_flux_face<Y, 2, MIXED>(a, i, j, F, fstr, o);
        }
    };
    for (int j = 0; j < ny; ++j)
        for (int fi = 0; fi <= nx; ++fi) {
            face(std::false_type{}, fi - 1, j, Fx, fxs);
            for (int c = 0; c < 12; ++c) a.gx[c * fxs + (long long)j * (nx + 1) + fi] = (o[0][c] + o[1][c]) + o[2][c];
        }
    for (int fj = 0; fj <= ny; ++fj)
        for (int i = 0; i < nx; ++i) {
            face(std::true_type{}, i, fj - 1, Fy, fys);
            for (int c = 0; c < 12; ++c) a.gy[c * fys + (long long)fj * nx + i] = (o[0][c] + o[1][c]) + o[2][c];
        }
    double smx = 0, smy = 0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) fvh_update_cell(a, i, j, 0.0, smx, smy);
    if (qlim) {
        for (size_t c = 0; c < 4 * n; ++c) qlim[c] = q[c];
        for (size_t c = 0; c < 8 * n; ++c) qlim[4 * n + c] = lv[c];
    }
    return 0;
}

extern "C" int fvhh_rhs(int nx, int ny, double xl, double xr, double yb, double yt, double gamma, int bc, int flux,
                        int quirks, const double* q, double* rhs, double* qlim) {
    return rhs_impl<false>(nx, ny, xl, xr, yb, yt, gamma, bc, flux, quirks, q, rhs, qlim);
}

// hweno (1), htenoi (2) or linear (3) in hteno's place (the kernels' run-time
// variant instantiation)
extern "C" int fvhh_rhs_variant(int hrecon, int nx, int ny, double xl, double xr, double yb, double yt, double gamma,
                                int bc, int flux, int quirks, const double* q, double* rhs, double* qlim) {
    return rhs_impl<false, true>(nx, ny, xl, xr, yb, yt, gamma, bc, flux, quirks, q, rhs, qlim, hrecon);
}

// the FP32-mixed precision's bodies (HTENO / TENO5 cut-offs and transverse
// weights in FP32)
extern "C" int fvhh_rhs_mixed(int nx, int ny, double xl, double xr, double yb, double yt, double gamma, int bc,
                              int flux, int quirks, const double* q, double* rhs, double* qlim) {
    return rhs_impl<true>(nx, ny, xl, xr, yb, yt, gamma, bc, flux, quirks, q, rhs, qlim);
}

// the FP32 bulk precision's bodies (FP32 face values, quadrature, projections
// and fluxes; FP64 Roe eigensystem, Gauss sums and update)
extern "C" int fvhh_rhs_bulk32(int nx, int ny, double xl, double xr, double yb, double yt, double gamma, int bc,
                               int flux, int quirks, const double* q, double* rhs, double* qlim) {
    return rhs_impl<true, false, true>(nx, ny, xl, xr, yb, yt, gamma, bc, flux, quirks, q, rhs, qlim);
}

// The limiter in one piece (herm_limiter) and split at its chain dependence
// (herm_limiter_pre + herm_limiter_step), over n input rows
// (u0, u1, u2, d0, d2): the two results of each row.
extern "C" void fvhh_limiter_split(int n, double dif, const double* in, double* whole, double* split) {
    for (int r = 0; r < n; ++r) {
        const double* x = in + 5 * r;
        whole[r] = herm_limiter(dif, x[0], x[1], x[2], x[3], x[4]);
        double o[LIM_NV] = {};
        herm_limiter_pre(dif, x[0], x[1], x[2], x[4], o);
        split[r] = herm_limiter_step(dif, o, x[3]);
    }
}
