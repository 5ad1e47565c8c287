// FVHermite_System2D.h -- drop-in for the reference's FVHermite::System2D
// (FVHermite_System2D.h:11-24) over the C ABI of libhtx.so (htx.h).
// Header-only; link with -lhtx.
//
// The public interface is the reference's:
//     System2D(xl, xr, nx, yb, yt, ny, cflnumber, tmax, tInit, specificHeatRatio)
//     void entropyWave(); void isentropicVortex(); void Riemann2D();
//     void timeStepping(); void printResult();
// The state carries the cell averages u and the derivative moments v = <dq/dx>,
// w = <dq/dy> (12 fields).  The Hermite-TENO residual (derivative limiter,
// HTENO / TENO5 reconstruction, Gauss-point fluxes), the SSP-RK3 stages and
// the CFL reduction run on the GPU.  The optional trailing htx::Options
// carries what the reference selects by editing code; its defaults are the
// reference's shipped choices (characteristic Rusanov, zero-gradient
// boundaries in rhs(), quirks on).
#pragma once
#include "FV_System2D.h"

namespace FVHermite {

class System2D : public htx::SolverBase {
public:
    System2D(const double& xl, const double& xr, const int& nx, const double& yb, const double& yt, const int& ny,
             const double& cflnumber, const double& tmax, const double& tInit, const double& specificHeatRatio,
             const htx::Options& opt = htx::Options())
        : SolverBase(xl, xr, nx, yb, yt, ny, cflnumber, tmax, tInit, specificHeatRatio, HTX_SCHEME_HTENO, opt) {}

    // FVHermite::System2D::entropyWave (FVHermite_System2D.cpp:52-86):
    // rho = 1 + 0.5 sin(pi (x + y)), u = v = p = 1; u and both moments are
    // 9x9 Gauss cell averages (the moments of dconservativeOfEntropyWave,
    // equal in x and y), generated on the device; exact = advected by
    // (1, 1) tFinal.
    void entropyWave() {
        htx::check(htx_init_entropy_wave(h_, 1.0, 1.0, 1.0, 1.0, 0.5));
        have_exact_ = true;
        exact_shift_ = tFinal_;
    }

    // FVHermite::System2D::isentropicVortex (FVHermite_System2D.cpp:885-912):
    // the vortex centred at the origin (Equations_Euler2D.cpp:659) with its
    // derivative moments, generated on the device; exact = the vortex
    // advected by (tFinal, tFinal) (:906).
    void isentropicVortex() { isentropicVortex(0.0, 0.0); }

    // Same vortex centred at (cx, cy), with optional density modes; exact =
    // the same field advected by (tFinal, tFinal).
    void isentropicVortex(double cx, double cy, const double* modes = nullptr, int nmodes = 0) {
        htx::check(htx_init_vortex(h_, cx, cy, modes, nmodes));
        have_exact_ = true;
        exact_shift_ = tFinal_;
    }

    // FVHermite::System2D::Riemann2D (FVHermite_System2D.cpp:89-107): the
    // configuration-5 quadrants (conservativeOfRiemann2D) point-sampled; the
    // moments are differences of the point values across the cell.
    void Riemann2D() {
        htx::ic::Euler2DIC ic(gamma_);
        const size_t n = (size_t)nx_ * ny_;
        std::vector<double> q(12 * n);
        auto cons = [&](double x, double y, double c[4]) {
            double w[4];
            ic.riemann2d(x, y, w);
            ic.conservative(w, c);
        };
        for (int j = 0; j < ny_; ++j)
            for (int i = 0; i < nx_; ++i) {
                const double x = xc(i), y = yc(j);
                double c[4], xp[4], xm[4], yp[4], ym[4];
                cons(x, y, c);
                cons(x + dx_ / 2.0, y, xp);
                cons(x - dx_ / 2.0, y, xm);
                cons(x, y + dy_ / 2.0, yp);
                cons(x, y - dy_ / 2.0, ym);
                const size_t o = (size_t)j * nx_ + i;
                for (int e = 0; e < 4; ++e) {
                    q[e * n + o] = c[e];
                    q[(4 + e) * n + o] = (xp[e] - xm[e]) / dx_;
                    q[(8 + e) * n + o] = (yp[e] - ym[e]) / dy_;
                }
            }
        upload(q);
        have_exact_ = false;
    }

    // FVHermite::System2D::timeStepping (FVHermite_System2D.cpp:784-819)
    void timeStepping() {
        int n = 0;
        htx::check(htx_advance(h_, tFinal_, &n));
        steps_ = n;
    }

    // FVHermite::System2D::printResult (FVHermite_System2D.cpp:821-835): L1
    // density error against the exact solution over nx ny cells, then
    // write() (:837-883) to Options::output_path when one is given.
    void printResult() {
        if (have_exact_) std::cout << "Numerical error:" << l1_density_error() << std::endl;
        if (!opt_.output_path.empty()) write_dat(opt_.output_path);
    }

    int steps() const { return steps_; }

private:
    int steps_ = 0;
};

}  // namespace FVHermite
