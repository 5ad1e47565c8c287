// FV_System2D.h -- drop-in for the reference's FV::System2D (FV_System2D.h:8-108)
// over the C ABI of libhtx.so (htx.h).  Header-only; link with -lhtx.
//
// The public interface is the reference's:
//     System2D(xl, xr, nx, yb, yt, ny, cflnumber, tmax, tInit, specificHeatRatio)
//     void entropyWave(); void isentropicVortex(); void Riemann2D();
//     void timeStepping(); void printResult();
// The residual, SSP-RK3 stages and CFL reduction run on the GPU; the initial
// conditions are evaluated here on the host with the reference's formulas and
// uploaded once.  Choices that the reference makes by commenting code in or
// out (flux, boundary condition, quirks) are carried by an optional trailing
// htx::Options argument whose defaults are the reference's shipped code:
// characteristic Rusanov (FV_System2D.cpp:462-464), zero-gradient boundaries
// in rhs() (:366), reference quirks on.
#pragma once
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "htx.h"
#include "htx_ic.h"

namespace htx {

struct Options {
    int flux = HTX_FLUX_RUSANOV_CHAR;  // FV_System2D.cpp:462-466
    int bc = HTX_BC_ZERO_GRAD;         // FV_System2D.cpp:365-366
    int precision = HTX_PREC_FP64;
    int ref_quirks = 1;
    std::string output_path;           // write(): empty -> no file
    // the reconstruction a user of the reference swaps in for teno5
    // (FV_System2D.cpp:379, :394: an htx_scheme of the 5-point family, e.g.
    // HTX_SCHEME_WENO5Z) or for hteno (FVHermite_System2D.cpp:459, :486:
    // HTX_SCHEME_HWENO, HTX_SCHEME_HTENOI, HTX_SCHEME_HLINEAR); -1 keeps the
    // class's own (teno5 / hteno)
    int reconstruction = -1;
};

inline bool is_hermite_scheme(int s) {
    return s == HTX_SCHEME_HTENO || s == HTX_SCHEME_HWENO || s == HTX_SCHEME_HTENOI || s == HTX_SCHEME_HLINEAR;
}

inline void check(int rc) {
    if (rc != HTX_OK) throw std::runtime_error(std::string("htx: ") + htx_last_error());
}

// Shared state handling of the 2D façades.
class SolverBase {
public:
    SolverBase(const SolverBase&) = delete;
    SolverBase& operator=(const SolverBase&) = delete;
    ~SolverBase() { htx_destroy(h_); }

    // Device-side handle, for callers that want the C ABI directly.
    htx_solver* handle() const { return h_; }
    double time() const {
        double t = 0.0;
        check(htx_get_time(h_, &t, nullptr));
        return t;
    }
    std::vector<double> state() const {
        std::vector<double> q((size_t)nf_ * nx_ * ny_);
        check(htx_get_state(h_, q.data(), (size_t)nx_));
        return q;
    }
    // Snapshot of the state, time, step count and configuration (htx_save),
    // and its restoration into a solver of the same grid and scheme.
    void saveSnapshot(const std::string& path) { check(htx_save(h_, path.c_str())); }
    // One SSP-RK3 stage with a given dt (the stage loop of RungeKutta3,
    // FV_System2D.cpp:575-612, private in the reference): stages 1, 2, 3 in order.
    void stage(int k, double dt) { check(htx_stage(h_, k, dt)); }
    void loadSnapshot(const std::string& path) { check(htx_load(h_, path.c_str())); }

protected:
    SolverBase(double xl, double xr, int nx, double yb, double yt, int ny, double cfl, double tmax, double tinit,
               double gamma, int scheme, const Options& o)
        : nx_(nx), ny_(ny), nf_(is_hermite_scheme(scheme) ? 12 : 4), xl_(xl), yb_(yb), gamma_(gamma),
          tFinal_(tmax), opt_(o) {
        if (o.reconstruction >= 0) {
            if (is_hermite_scheme(o.reconstruction) != is_hermite_scheme(scheme))
                throw std::invalid_argument("htx: Options::reconstruction belongs to the other solver class");
            scheme = o.reconstruction;
        }
        htx_params p;
        check(htx_params_default(&p));
        p.xl = xl; p.xr = xr; p.nx = nx;
        p.yb = yb; p.yt = yt; p.ny = ny;
        p.cfl = cfl; p.tmax = tmax; p.tinit = tinit; p.gamma = gamma;
        p.scheme = scheme;
        p.flux = o.flux;
        p.bc = o.bc;
        p.precision = o.precision;
        p.ref_quirks = o.ref_quirks;
        check(htx_create(&p, &h_));
        dx_ = (xr - xl) / nx;
        dy_ = (yt - yb) / ny;
    }

    // cell centre, FV_System2D.cpp:37-38 (interior index)
    double xc(int i) const { return ((double)i + 0.5) * dx_ + xl_; }
    double yc(int j) const { return ((double)j + 0.5) * dy_ + yb_; }

    void upload(const std::vector<double>& q) { check(htx_set_state(h_, q.data(), (size_t)nx_)); }

    // printResult's numerical error, reduced on the device against the exact
    // solution of the last analytic initial condition
    double l1_density_error() const {
        double err = 0.0;
        check(htx_l1_density_error(h_, exact_shift_, &err));
        return err;
    }

    void write_dat(const std::string& path) const {
        std::vector<double> q = state();
        std::ofstream out(path);
        if (!out.good()) throw std::runtime_error("htx: cannot open " + path);
        const int w = 20;
        out << "t: " << time() << "\n";
        out << std::right << std::setw(w) << "x" << std::setw(w) << "y" << std::setw(w) << "Density"
            << std::setw(w) << "Ux" << std::setw(w) << "Uy" << std::setw(w) << "Pressure" << "\n";
        size_t n = (size_t)nx_ * ny_;
        for (int j = 0; j < ny_; ++j)
            for (int i = 0; i < nx_; ++i) {
                size_t c = (size_t)j * nx_ + i;
                out << std::setprecision(8) << std::setw(w) << xc(i) << std::setw(w) << yc(j);
                for (int e = 0; e < 4; ++e) out << std::setw(w) << q[e * n + c];
                out << "\n";
            }
    }

    htx_solver* h_ = nullptr;
    int nx_, ny_, nf_;
    double xl_, yb_, dx_ = 0, dy_ = 0, gamma_, tFinal_;
    Options opt_;
    bool have_exact_ = false;    // an analytic initial condition is set (printResult reports its error)
    double exact_shift_ = 0.0;   // time its exact solution is advected by
};

}  // namespace htx

namespace FV {

class System2D : public htx::SolverBase {
public:
    System2D(const double& xl, const double& xr, const int& nx, const double& yb, const double& yt, const int& ny,
             const double& cflnumber, const double& tmax, const double& tInit, const double& specificHeatRatio,
             const htx::Options& opt = htx::Options())
        : SolverBase(xl, xr, nx, yb, yt, ny, cflnumber, tmax, tInit, specificHeatRatio, HTX_SCHEME_TENO5, opt) {}

    // FV::System2D::entropyWave (FV_System2D.cpp:49-74): rho = 1 + 0.5 sin(pi(x+y)),
    // u = v = p = 1, 9x9 Gauss cell averages, generated on the device; exact =
    // advected by (1, 1) tFinal.
    void entropyWave() {
        htx::check(htx_init_entropy_wave(h_, 1.0, 1.0, 1.0, 1.0, 0.5));
        have_exact_ = true;
        exact_shift_ = tFinal_;
    }

    // FV::System2D::isentropicVortex (FV_System2D.cpp:76-99): vortex centred at
    // the origin (Equations_Euler2D.cpp:659); exact = initial field (:93).
    void isentropicVortex() { isentropicVortex(0.0, 0.0); }

    // Same vortex centred at (cx, cy), generated on the device; exact = the
    // initial field, as for the reference's vortex.
    void isentropicVortex(double cx, double cy, const double* modes = nullptr, int nmodes = 0) {
        htx::check(htx_init_vortex(h_, cx, cy, modes, nmodes));
        have_exact_ = true;
        exact_shift_ = 0.0;
    }

    // FV::System2D::Riemann2D (FV_System2D.cpp:101-112) loads the Rayleigh-Taylor
    // state conservativeOfTaylorIns2D (Equations_Euler2D.cpp:619-635), point-sampled.
    void Riemann2D() {
        htx::ic::Euler2DIC ic(gamma_);
        std::vector<double> q((size_t)4 * nx_ * ny_);
        size_t n = (size_t)nx_ * ny_;
        for (int j = 0; j < ny_; ++j)
            for (int i = 0; i < nx_; ++i) {
                double w[4], c[4];
                ic.taylor_instability(xc(i), yc(j), w);
                ic.conservative(w, c);
                for (int e = 0; e < 4; ++e) q[e * n + (size_t)j * nx_ + i] = c[e];
            }
        upload(q);
        have_exact_ = false;
    }

    // FV::System2D::timeStepping (FV_System2D.cpp:614-651)
    void timeStepping() {
        int n = 0;
        htx::check(htx_advance(h_, tFinal_, &n));
        steps_ = n;
    }

    // FV::System2D::printResult (FV_System2D.cpp:653-666)
    void printResult() {
        if (have_exact_) std::cout << "Numerical error:" << l1_density_error() << std::endl;
        if (!opt_.output_path.empty()) write_dat(opt_.output_path);
    }

    int steps() const { return steps_; }

private:
    int steps_ = 0;
};

}  // namespace FV
