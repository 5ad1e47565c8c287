// FV_System1D.h -- drop-in for the reference's FV::System1D
// (FV_System1D.h:10-26) over the C ABI of libhtx.so (htx.h, dim = 1).
// Header-only; link with -lhtx.
//
// The public interface is the reference's:
//     System1D(xl, xr, nx, cflnumber, tmax, tInit, specificHeatRatio)
//     void entropyWave(); void blastWave(); void timeStepping(); void printResult();
// plus sodShock() (Euler1D::sodShock, the 1D configuration's initial state)
// and troubledCells() (the KXRCF flags of FV::System1D::rhs).
// The residual (characteristic TENO5i or TENO5, characteristic Rusanov or
// HLLC), the SSP-RK3 stages and the CFL reduction run on the GPU.  The
// optional trailing htx::Options carries the reference's code-toggled
// choices; its defaults are the shipped ones (TENO5i, characteristic
// Rusanov, zero-gradient boundaries, FV_System1D.cpp:161, 199-200, 220).
#pragma once
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "htx.h"
#include "htx_ic.h"

namespace htx {

struct Options1D {
    int scheme = HTX_SCHEME_TENO5I;    // FV_System1D.cpp:199-200 (teno5i)
    int flux = HTX_FLUX_RUSANOV_CHAR;  // :220-222
    int bc = HTX_BC_ZERO_GRAD;         // :160-162
    int ref_quirks = 1;                // HLLC sound speed and star energy (Equations_Euler1D.cpp:174-175, 204)
};

}  // namespace htx

namespace FV {

class System1D {
public:
    System1D(const double& xl, const double& xr, const int& nx, const double& cflnumber, const double& tmax,
             const double& tInit, const double& specificHeatRatio, const htx::Options1D& opt = htx::Options1D())
        : nx_(nx), xl_(xl), gamma_(specificHeatRatio), tFinal_(tmax) {
        htx_params p;
        check(htx_params_default(&p));
        p.dim = 1;
        p.xl = xl; p.xr = xr; p.nx = nx;
        p.cfl = cflnumber; p.tmax = tmax; p.tinit = tInit; p.gamma = specificHeatRatio;
        p.scheme = opt.scheme;
        p.flux = opt.flux;
        p.bc = opt.bc;
        p.precision = HTX_PREC_FP64;
        p.ref_quirks = opt.ref_quirks;
        check(htx_create(&p, &h_));
        dx_ = (xr - xl) / nx;
    }
    System1D(const System1D&) = delete;
    System1D& operator=(const System1D&) = delete;
    ~System1D() { htx_destroy(h_); }

    // FV::System1D::entropyWave (FV_System1D.cpp:35-55): rho = 1 + 0.2 sin(pi x),
    // u = p = 1, 9-point Gauss cell averages; exact = advected by u t.
    void entropyWave() {
        std::vector<double> q = average(0.0);
        upload(q);
        std::vector<double> e = average(tFinal_);
        exact_.assign(e.begin(), e.begin() + nx_);
    }

    // FV::System1D::blastWave (FV_System1D.cpp:57-65): Euler1D::shuOsher
    // (Equations_Euler1D.cpp:334-345), point-sampled.
    void blastWave() {
        std::vector<double> q(3 * (size_t)nx_);
        for (int i = 0; i < nx_; ++i) {
            const double x = xc(i);
            double w[3];
            if (x < -4.0) {
                w[0] = 3.857143; w[1] = 2.629369; w[2] = 10.333333;
            } else {
                w[0] = 1.0 + 0.2 * std::sin(5.0 * x); w[1] = 0.0; w[2] = 1.0;
            }
            put(q, i, w);
        }
        upload(q);
        exact_.clear();
    }

    // Euler1D::sodShock (Equations_Euler1D.cpp:317-332), point-sampled.
    void sodShock() {
        std::vector<double> q(3 * (size_t)nx_);
        for (int i = 0; i < nx_; ++i) {
            double w[3];
            if (xc(i) < 0.5) {
                w[0] = 1.0; w[1] = 0.0; w[2] = 1.0;
            } else {
                w[0] = 0.125; w[1] = 0.0; w[2] = 0.1;
            }
            put(q, i, w);
        }
        upload(q);
        exact_.clear();
    }

    // FV::System1D::timeStepping (FV_System1D.cpp:264-283)
    void timeStepping() {
        int n = 0;
        check(htx_advance(h_, tFinal_, &n));
        steps_ = n;
    }

    // FV::System1D::printResult (FV_System1D.cpp:285-300): L1 and Linf of the
    // density against the exact solution.
    void printResult() {
        if (exact_.empty()) return;
        std::vector<double> q = state();
        double err = 0.0, merr = -1.0;
        for (int i = 0; i < nx_; ++i) {
            err += std::fabs(q[i] - exact_[i]);
            merr = std::fmax(merr, std::fabs(q[i] - exact_[i]));
        }
        std::cout << "Numerical error:" << err / nx_ << std::endl;
        std::cout << "L inf: " << merr << std::endl;
    }

    // The per-cell discontinue flags FV::System1D::rhs sets with its KXRCF
    // detector (FV_System1D.cpp:166-192), for the current state.
    std::vector<unsigned char> troubledCells() const {
        std::vector<unsigned char> f((size_t)nx_);
        check(htx_kxrcf_flags(h_, f.data()));
        return f;
    }

    std::vector<double> state() const {
        std::vector<double> q(3 * (size_t)nx_);
        check(htx_get_state(h_, q.data(), (size_t)nx_));
        return q;
    }
    double time() const {
        double t = 0.0;
        check(htx_get_time(h_, &t, nullptr));
        return t;
    }
    int steps() const { return steps_; }
    htx_solver* handle() const { return h_; }

private:
    static void check(int rc) {
        if (rc != HTX_OK) throw std::runtime_error(std::string("htx: ") + htx_last_error());
    }
    double xc(int i) const { return ((double)i + 0.5) * dx_ + xl_; }  // FV_System1D.cpp:25
    void put(std::vector<double>& q, int i, const double w[3]) const {  // Euler1D::conservative (:25-30)
        const double g1 = gamma_ - 1.0;
        q[i] = w[0];
        q[nx_ + i] = w[0] * w[1];
        q[2 * (size_t)nx_ + i] = w[2] / g1 + 0.5 * q[nx_ + i] * w[1];
    }
    void upload(const std::vector<double>& q) { check(htx_set_state(h_, q.data(), (size_t)nx_)); }
    // Euler1D::entropyWaves (Equations_Euler1D.cpp:263-277) shifted by -shift
    std::vector<double> average(double shift) const {
        htx::ic::Euler2DIC ic(gamma_);  // the 9 Gauss nodes and weights
        std::vector<double> q(3 * (size_t)nx_);
        for (int i = 0; i < nx_; ++i) {
            double acc[3] = {0, 0, 0};
            for (int a = 0; a < 9; ++a) {
                const double x = xc(i) - shift + dx_ * ic.gx[a] / 2.0;
                double w[3] = {1.0 + 0.2 * std::sin(ic.pi * x), 1.0, 1.0}, c[3];
                c[0] = w[0];
                c[1] = w[0] * w[1];
                c[2] = w[2] / (gamma_ - 1.0) + 0.5 * c[1] * w[1];
                for (int e = 0; e < 3; ++e) acc[e] += ic.gw[a] * c[e] / 2.0;
            }
            for (int e = 0; e < 3; ++e) q[e * (size_t)nx_ + i] = acc[e];
        }
        return q;
    }

    htx_solver* h_ = nullptr;
    int nx_;
    double xl_, dx_ = 0, gamma_, tFinal_;
    int steps_ = 0;
    std::vector<double> exact_;
};

}  // namespace FV
