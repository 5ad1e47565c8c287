// CPU build of the 1D path's per-face and per-cell bodies (htx_fv1d.h): the
// residual and the SSP-RK3 stages of FV::System1D run here in the order the
// GPU launch runs them.  Test infrastructure, compared with the FP64 oracle.
#include <vector>

#include "../../hermite-teno_amd/csrc/htx_fv1d.h"

using namespace htx;

// recon: the kernels' codes (0 teno5, 1 teno5i, 2 weno5JS, 3 weno5Z, RECON_WENO5N ..); flux: 0 Rusanov,
// 1 HLLC, 2 central-upwind; mode 0 residual,
// 1..3 stage with dt (q0 = u^n)
extern "C" void fv1dh_stage(int nx, double xl, double xr, double gamma, int recon, int flux, int quirks, int mode,
                            double dt, const double* q, const double* q0, double* out) {
    Fv1dArgs a{};
    a.qin = q;
    a.q0 = q0;
    a.qout = out;
    a.nx = nx;
    a.bc = 0;
    a.mode = mode;
    a.recon = recon;
    a.flux = flux == 0 ? FLUX_RUSANOV_CHAR : flux == 2 ? FLUX_CU : (quirks ? FLUX_HLLC : FLUX_HLLC_TORO);
    a.gam = gamma;
    a.g1 = gamma - 1.0;
    a.dx = (xr - xl) / nx;
    std::vector<double> F(3 * (size_t)(nx + 1));
    for (int c = -1; c < nx; ++c) {
        double f[3];
        fv1d_face_flux(a, c, f);
        for (int e = 0; e < 3; ++e) F[e * (size_t)(nx + 1) + c + 1] = f[e];
    }
    std::vector<double> res(3 * (size_t)nx);
    for (int c = 0; c < nx; ++c)
        for (int e = 0; e < 3; ++e)
            res[e * (size_t)nx + c] =
                fv1d_stage_value(a, c, e, F[e * (size_t)(nx + 1) + c], F[e * (size_t)(nx + 1) + c + 1], dt);
    for (size_t k = 0; k < res.size(); ++k) out[k] = res[k];
}

// LMPE flags (per field, flags[eq nx + c]) by the device's run-parity form
// of the reference's sequential sweep
extern "C" void fv1dh_lmpe(int nx, double xl, double xr, int bc, const double* q, unsigned char* flags) {
    const double dx = (xr - xl) / nx;
    for (int eq = 0; eq < 3; ++eq)
        for (int c = 0; c < nx; ++c) flags[eq * nx + c] = fv1d_lmpe(q + (size_t)eq * nx, c, nx, bc, dx) ? 1 : 0;
}
