// htx_ic.h -- host-side initial conditions of the 2D Euler façades, restating
// the reference's point formulas (Equations_Euler2D.cpp) and its 9x9
// Gauss-Legendre cell averaging (Equations_Euler2D.cpp:454-528 with the rule of
// Schemes_Lagrange.cpp:22-42).  Header-only, no GPU code.
#pragma once
#include <cmath>
#include <cstddef>
#include <vector>

namespace htx {
namespace ic {

struct Euler2DIC {
    double gamma, g1, pi;
    double gx[9], gw[9];

    explicit Euler2DIC(double gam) : gamma(gam), g1(gam - 1.0), pi(4.0 * std::atan(1.0)) {
        const double a[4] = {0.8360311073266358, 0.9681602395076261, 0.3242534234038089, 0.6133714327005904};
        const double w[4] = {0.1806481606948574, 0.08127438836157441, 0.31234707704000286, 0.26061069640293544};
        gx[0] = 0.0;
        gw[0] = 0.3302393550012598;
        for (int k = 0; k < 4; ++k) {
            gx[1 + 2 * k] = -a[k];
            gx[2 + 2 * k] = a[k];
            gw[1 + 2 * k] = w[k];
            gw[2 + 2 * k] = w[k];
        }
    }

    // Euler2D::conservative, Equations_Euler2D.cpp:28-34
    void conservative(const double w[4], double q[4]) const {
        q[0] = w[0];
        q[1] = w[0] * w[1];
        q[2] = w[0] * w[2];
        q[3] = w[3] / g1 + 0.5 * (q[1] * w[1] + q[2] * w[2]);
    }

    // conservativeOfEntropyWave (primitive part), :430-440
    void entropy_wave(double rho, double u, double v, double p, double amp, double x, double y, double w[4]) const {
        w[0] = rho + amp * std::sin(pi * (x + y));
        w[1] = u;
        w[2] = v;
        w[3] = p;
    }

    // dconservativeOfEntropyWave, :442-452: the x (equally y) derivative of
    // the conservative entropy-wave state (density depends on x + y only)
    void d_entropy_wave(double u, double v, double amp, double x, double y, double c[4]) const {
        c[0] = amp * std::cos(pi * (x + y)) * pi;
        c[1] = u * c[0];
        c[2] = v * c[0];
        c[3] = (0.5 * (u * u + v * v)) * c[0];
    }

    // conservativeOfIsentropicVortex (primitive part), :656-666, b = 5, at the origin
    void vortex(double x, double y, double w[4]) const {
        double r = std::pow(x, 2) + std::pow(y, 2), b = 5.0;
        w[0] = 1.0 - g1 * std::pow(b, 2) / 8.0 / gamma / std::pow(pi, 2) * std::exp(1 - r);
        w[0] = std::pow(w[0], 1.0 / g1);
        w[1] = 1.0 - b * y / 2.0 / pi * std::exp(0.5 * (1 - r));
        w[2] = 1.0 + b * x / 2.0 / pi * std::exp(0.5 * (1 - r));
        w[3] = std::pow(w[0], gamma);
    }

    // conservativeOfRiemann2D (primitive part), :530-617 (four-quadrant data, split at 0.5)
    void riemann2d(double x, double y, double w[4]) const {
        if ((x > 0.5) && (y > 0.5)) {
            w[0] = 1.0; w[1] = 0.75; w[2] = -0.5; w[3] = 1.0;
        } else if ((x < 0.5) && (y > 0.5)) {
            w[0] = 2.0; w[1] = 0.75; w[2] = 0.5; w[3] = 1.0;
        } else if ((x < 0.5) && (y < 0.5)) {
            w[0] = 1.0; w[1] = -0.75; w[2] = 0.5; w[3] = 1.0;
        } else {
            w[0] = 3.0; w[1] = -0.75; w[2] = -0.5; w[3] = 1.0;
        }
    }

    // conservativeOfTaylorIns2D (primitive part), :619-635
    void taylor_instability(double x, double y, double w[4]) const {
        if (y < 0.5) {
            w[0] = 2.0; w[1] = 0.0; w[3] = 1.0 + 2.0 * y;
        } else {
            w[0] = 1.0; w[1] = 0.0; w[3] = y + 1.5;
        }
        w[2] = -0.025 * std::sqrt(gamma * w[3] / w[0]) * std::cos(8.0 * pi * x);
    }

    // 9x9 Gauss cell averages of conservative(prim(x, y)) over every interior
    // cell; q is field-major (4, ny, nx).  Summation order as
    // Euler2D::entropyWaves / isentropicVortex (:462-468, :519-524).
    template <class F>
    void average(int nx, int ny, double dx, double dy, double xl, double yb, std::vector<double>& q, F prim) const {
        size_t n = (size_t)nx * ny;
        q.assign(4 * n, 0.0);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double x = ((double)i + 0.5) * dx + xl, y = ((double)j + 0.5) * dy + yb;
                double acc[4] = {0, 0, 0, 0};
                for (int a = 0; a < 9; ++a)
                    for (int b = 0; b < 9; ++b) {
                        double w[4], c[4];
                        prim(x + dx * gx[a] / 2.0, y + dy * gx[b] / 2.0, w);
                        conservative(w, c);
                        for (int e = 0; e < 4; ++e) acc[e] += gw[a] * gw[b] * c[e] / 4.0;
                    }
                for (int e = 0; e < 4; ++e) q[e * n + (size_t)j * nx + i] = acc[e];
            }
    }

    // The same 9x9 Gauss cell averages of a conservative-valued cons(x, y, c)
    // (derivative moments, Euler2D::entropyWavesdx/dy :473-511).
    template <class F>
    void average_conservative(int nx, int ny, double dx, double dy, double xl, double yb, std::vector<double>& q,
                              F cons) const {
        size_t n = (size_t)nx * ny;
        q.assign(4 * n, 0.0);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double x = ((double)i + 0.5) * dx + xl, y = ((double)j + 0.5) * dy + yb;
                double acc[4] = {0, 0, 0, 0};
                for (int a = 0; a < 9; ++a)
                    for (int b = 0; b < 9; ++b) {
                        double c[4];
                        cons(x + dx * gx[a] / 2.0, y + dy * gx[b] / 2.0, c);
                        for (int e = 0; e < 4; ++e) acc[e] += gw[a] * gw[b] * c[e] / 4.0;
                    }
                for (int e = 0; e < 4; ++e) q[e * n + (size_t)j * nx + i] = acc[e];
            }
    }
};

}  // namespace ic
}  // namespace htx
