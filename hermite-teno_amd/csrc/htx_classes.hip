// htx_classes.hip -- host implementations of the drop-in physics and scheme
// classes (include/Equations_Euler2D.h, Equations_Euler1D.h,
// Schemes_Lagrange.h, Schemes_Hermite.h): the reference's public interfaces
// over the host build of the GPU kernels' numerics, strict IEEE arithmetic
// (no contraction, divisions kept), so that a class call computes what the
// drop-in kernels compute for the same data.  Host code only.
#define HTX_NS htx_host
#define HTX_EXACT_DIVISION 1
#include <cmath>
#include <stdexcept>
#include <string>

#include "../../include/Equations_Euler1D.h"
#include "../../include/Equations_Euler2D.h"
#include "../../include/Schemes_Hermite.h"
#include "../../include/Schemes_Lagrange.h"
#include "htx_fv1d.h"
#include "htx_hermite.h"

#pragma clang fp contract(off)

namespace {

void check_dir(int dir) {
    if (dir != 1 && dir != 2) throw std::invalid_argument("Euler2D: dir must be 1 (x) or 2 (y)");
}

// Gauss-Legendre 9-point rule, reference node order (Schemes_Lagrange.cpp:22-42)
void gauss9(double x[9], double w[9]) {
    const double a[4] = {0.8360311073266358, 0.9681602395076261, 0.3242534234038089, 0.6133714327005904};
    const double wa[4] = {0.1806481606948574, 0.08127438836157441, 0.31234707704000286, 0.26061069640293544};
    x[0] = 0.0;
    w[0] = 0.3302393550012598;
    for (int k = 0; k < 4; ++k) {
        x[1 + 2 * k] = -a[k];
        x[2 + 2 * k] = a[k];
        w[1 + 2 * k] = wa[k];
        w[2 + 2 * k] = wa[k];
    }
}

// 9x9 Gauss average of a conservative-valued point function, in the order of
// Euler2D::entropyWaves / isentropicVortex (Equations_Euler2D.cpp:462-468,
// :519-524): acc += w_i w_j f / 4
template <class F>
Vector4d gauss_avg2(double x, double y, double dx, double dy, const Matrix<double, 9, 1>& wt,
                    const Matrix<double, 9, 1>& ab, F f) {
    Vector4d acc;
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) {
            const Vector4d c = f(x + dx * ab(i) / 2.0, y + dy * ab(j) / 2.0);
            for (int e = 0; e < 4; ++e) acc(e) += wt(i) * wt(j) * c(e) / 4.0;
        }
    return acc;
}

template <class F>
Vector3d gauss_avg1(double x, double dx, const Matrix<double, 9, 1>& wt, const Matrix<double, 9, 1>& ab, F f) {
    Vector3d acc;
    for (int i = 0; i < 9; ++i) {
        const Vector3d c = f(x + dx * ab(i) / 2.0);
        for (int e = 0; e < 3; ++e) acc(e) += wt(i) * c(e) / 2.0;
    }
    return acc;
}

}  // namespace

// ---------------------------------------------------------------------------
// Equations::Euler2D (Equations_Euler2D.cpp)
// ---------------------------------------------------------------------------
namespace Equations {

Euler2D::Euler2D(const double& g) : gamma(g), gamma_1(g - 1.0), gamma_3(g - 3.0), pi(4.0 * std::atan(1.0)) {}

// :20-26
Vector4d Euler2D::primitive(const Vector4d& q) const {
    Vector4d w;
    w(0) = q(0);
    w(1) = q(1) / q(0);
    w(2) = q(2) / q(0);
    w(3) = (q(3) - 0.5 * (q(1) * w(1) + q(2) * w(2))) * gamma_1;
    return w;
}

// :28-34
Vector4d Euler2D::conservative(const Vector4d& w) const {
    Vector4d q;
    q(0) = w(0);
    q(1) = w(0) * w(1);
    q(2) = w(0) * w(2);
    q(3) = w(3) / gamma_1 + 0.5 * (q(1) * w(1) + q(2) * w(2));
    return q;
}

// :36-50
Vector4d Euler2D::fluxx(const Vector4d& q, const Vector4d& w) const {
    Vector4d f;
    f << q(1), q(1) * w(1) + w(3), q(1) * w(2), w(1) * (q(3) + w(3));
    return f;
}

Vector4d Euler2D::fluxy(const Vector4d& q, const Vector4d& w) const {
    Vector4d f;
    f << q(2), q(2) * w(1), q(2) * w(2) + w(3), w(2) * (q(3) + w(3));
    return f;
}

// :52-98; the x and y Jacobians are one matrix with the momenta read as
// (normal n, tangential t)
static Matrix4d flux_jacobian(double gamma, double g1, double g3, const Vector4d& q, const Vector4d& w, int dir) {
    const int n = dir, t = 3 - dir;  // velocity components of w
    const double un = w(n), ut = w(t);
    const double sq = un * un;
    const double dyn = 0.5 * g1 * (sq + ut * ut);
    const double h = (q(3) + w(3)) / w(0);
    Matrix4d J;
    J(0, n) = 1.0;
    J(n, 0) = dyn - sq;
    J(n, n) = -g3 * un;
    J(n, t) = -g1 * ut;
    J(n, 3) = g1;
    // the mixed products in the reference's factor order, w(1) before w(2)
    J(t, 0) = -w(1) * w(2);
    J(t, n) = ut;
    J(t, t) = un;
    J(3, 0) = (dyn - h) * un;
    J(3, n) = h - g1 * sq;
    J(3, t) = -g1 * w(1) * w(2);
    J(3, 3) = gamma * un;
    return J;
}

Matrix4d Euler2D::dfluxx(const Vector4d& q, const Vector4d& w) const {
    return flux_jacobian(gamma, gamma_1, gamma_3, q, w, 1);
}

Matrix4d Euler2D::dfluxy(const Vector4d& q, const Vector4d& w) const {
    return flux_jacobian(gamma, gamma_1, gamma_3, q, w, 2);
}

// :100-205.  t0 is the product of the normal Roe velocity and the Roe sound
// speed in rightEigenVector*, the normal velocity over the sound speed in
// leftEigenVector*; t1 = b1 times the normal velocity, t2 = b1 / 2.
static Matrix4d right_eig(int dir, double vxm, double vym, double hm, double qm, double cm, double t0) {
    const int n = dir, t = 3 - dir;
    const double vn = dir == 1 ? vxm : vym, vt = dir == 1 ? vym : vxm;
    Matrix4d R;
    R(0, 0) = 1.0;
    R(0, 2) = 1.0;
    R(0, 3) = 1.0;
    R(n, 0) = vn - cm;
    R(n, 2) = vn;
    R(n, 3) = vn + cm;
    R(t, 0) = vt;
    R(t, 1) = 1.0;
    R(t, 2) = vt;
    R(t, 3) = vt;
    R(3, 0) = hm - t0;
    R(3, 1) = vt;
    R(3, 2) = qm;
    R(3, 3) = hm + t0;
    return R;
}

static Matrix4d left_eig(int dir, double rcm, double vt, double b1, double b2, double t0, double t1, double t2) {
    const int n = dir, t = 3 - dir;
    Matrix4d L;
    L(0, 0) = 0.5 * (b2 + t0);
    L(0, n) = -0.5 * (t1 + rcm);
    L(0, t) = -t2 * vt;
    L(0, 3) = t2;
    L(1, 0) = -vt;
    L(1, t) = 1.0;
    L(2, 0) = 1.0 - b2;
    L(2, n) = t1;
    L(2, t) = b1 * vt;
    L(2, 3) = -b1;
    L(3, 0) = 0.5 * (b2 - t0);
    L(3, n) = -0.5 * (t1 - rcm);
    L(3, t) = -t2 * vt;
    L(3, 3) = t2;
    return L;
}

Matrix4d Euler2D::rightEigenVectorx(const double& vxm, const double& vym, const double& hm, const double& qm,
                                    const double& cm, const double& t0) const {
    return right_eig(1, vxm, vym, hm, qm, cm, t0);
}

Matrix4d Euler2D::leftEigenVectorx(const double& rcm, const double& vym, const double& b1, const double& b2,
                                   const double& t0, const double& t1, const double& t2) const {
    return left_eig(1, rcm, vym, b1, b2, t0, t1, t2);
}

Matrix4d Euler2D::rightEigenVectory(const double& vxm, const double& vym, const double& hm, const double& qm,
                                    const double& cm, const double& t0) const {
    return right_eig(2, vxm, vym, hm, qm, cm, t0);
}

Matrix4d Euler2D::leftEigenVectory(const double& rcm, const double& vxm, const double& b1, const double& b2,
                                   const double& t0, const double& t1, const double& t2) const {
    return left_eig(2, rcm, vxm, b1, b2, t0, t1, t2);
}

// :207-219, without fabs on u - c
double Euler2D::maxSpeed(const Vector4d& w, int dir) const {
    check_dir(dir);
    const double c = std::sqrt(std::fabs(gamma * w(3) / w(0)));
    return std::fmax(w(dir) - c, std::fabs(w(dir) + c));
}

// :274-308
Vector4d Euler2D::localWaveSpeed(const Vector4d& w, int dir) const {
    check_dir(dir);
    const double c = std::sqrt(std::fabs(gamma * w(3) / w(0)));
    Vector4d s;
    s << std::fabs(w(dir) - c), std::fabs(w(dir)), std::fabs(w(dir)), std::fabs(w(dir) + c);
    return s;
}

// :221-246
Vector4d Euler2D::maxLocalWaveSpeed(const Vector4d& wl, const Vector4d& wr, int dir) const {
    return localWaveSpeed(wl, dir).cwiseMax(localWaveSpeed(wr, dir));
}

// :248-282: max over the two states of R diag(|lambda|) L, entry by entry
Matrix4d Euler2D::maxLocalWaveSpeed(const Vector4d& wl, const Vector4d& wr, const Matrix4d& L, const Matrix4d& R,
                                    int dir) const {
    auto dissipation = [&](const Vector4d& a) {
        Matrix4d D;
        for (int i = 0; i < 4; ++i) D(i, i) = a(i);
        return R * D * L;
    };
    return dissipation(localWaveSpeed(wl, dir)).cwiseMax(dissipation(localWaveSpeed(wr, dir)));
}

// :310-330: 1/2 (f_L + f_R - R (a .* (L (q_R - q_L))))
Vector4d Euler2D::Rusanov(const Vector4d& ql, const Vector4d& qr, const Vector4d& wl, const Vector4d& wr,
                          const Matrix4d& L, const Matrix4d& R, const Vector4d& a, int dir) const {
    check_dir(dir);
    const Vector4d fl = dir == 1 ? fluxx(ql, wl) : fluxy(ql, wl);
    const Vector4d fr = dir == 1 ? fluxx(qr, wr) : fluxy(qr, wr);
    return 0.5 * (fl + fr - R * a.cwiseProduct(L * (qr - ql)));
}

// :332-350
Vector4d Euler2D::Rusanov(const Vector4d& ql, const Vector4d& qr, const Vector4d& wl, const Vector4d& wr,
                          const Matrix4d& A, int dir) const {
    check_dir(dir);
    const Vector4d fl = dir == 1 ? fluxx(ql, wl) : fluxy(ql, wl);
    const Vector4d fr = dir == 1 ? fluxx(qr, wr) : fluxy(qr, wr);
    return 0.5 * (fl + fr - A * (qr - ql));
}

// :352-360
Vector4d Euler2D::dRusanov(const Vector4d& dql, const Vector4d& dqr, const Vector4d&, const Vector4d&,
                           const Matrix4d& L, const Matrix4d& R, const Matrix4d& dfl, const Matrix4d& dfr,
                           const Vector4d& a) const {
    return 0.5 * (dfl * dql + dfr * dqr - R * a.cwiseProduct(L * (dqr - dql)));
}

Vector4d Euler2D::dRusanov(const Vector4d& dql, const Vector4d& dqr, const Vector4d&, const Vector4d&,
                           const Matrix4d& dfl, const Matrix4d& dfr, const Matrix4d& A) const {
    return 0.5 * (dfl * dql + dfr * dqr - A * (dqr - dql));
}

// :372-428 (the kernels' HLLC with the reference's quirks), evaluated in
// (rho, normal, tangential, E) order
Vector4d Euler2D::HLLC(const Vector4d& ql, const Vector4d& qr, int dir) const {
    check_dir(dir);
    const int n = dir, t = 3 - dir;
    const double a[4] = {ql(0), ql(n), ql(t), ql(3)}, b[4] = {qr(0), qr(n), qr(t), qr(3)};
    double f[4];
    htx_host::hllc<true>(gamma, gamma_1, a, b, f);
    Vector4d F;
    F(0) = f[0];
    F(n) = f[1];
    F(t) = f[2];
    F(3) = f[3];
    return F;
}

// conservativeOfEntropyWave / dconservativeOfEntropyWave, :430-452
static Vector4d entropy_wave(const Euler2D& e, double rho, double u, double v, double p, double amp, double pi,
                             double x, double y) {
    Vector4d w;
    w << rho + amp * std::sin(pi * (x + y)), u, v, p;
    return e.conservative(w);
}

static Vector4d d_entropy_wave(double u, double v, double amp, double pi, double x, double y) {
    Vector4d q;
    q(0) = amp * std::cos(pi * (x + y)) * pi;
    q(1) = u * q(0);
    q(2) = v * q(0);
    q(3) = (0.5 * (u * u + v * v)) * q(0);
    return q;
}

// :454-511
Vector4d Euler2D::entropyWaves(const double& rhoInf, const double& uInf, const double& vInf, const double& pInf,
                               const double& amp, const double& x, const double& y, const double& dx, const double& dy,
                               const Matrix<double, 9, 1>& wt, const Matrix<double, 9, 1>& ab) const {
    return gauss_avg2(x, y, dx, dy, wt, ab, [&](double px, double py) {
        return entropy_wave(*this, rhoInf, uInf, vInf, pInf, amp, pi, px, py);
    });
}

Vector4d Euler2D::entropyWavesdx(const double&, const double& uInf, const double& vInf, const double&,
                                 const double& amp, const double& x, const double& y, const double& dx, const double& dy,
                                 const Matrix<double, 9, 1>& wt, const Matrix<double, 9, 1>& ab) const {
    return gauss_avg2(x, y, dx, dy, wt, ab,
                      [&](double px, double py) { return d_entropy_wave(uInf, vInf, amp, pi, px, py); });
}

Vector4d Euler2D::entropyWavesdy(const double& rhoInf, const double& uInf, const double& vInf, const double& pInf,
                                 const double& amp, const double& x, const double& y, const double& dx, const double& dy,
                                 const Matrix<double, 9, 1>& wt, const Matrix<double, 9, 1>& ab) const {
    // the density depends on x + y only: the y derivative is the x derivative
    return entropyWavesdx(rhoInf, uInf, vInf, pInf, amp, x, y, dx, dy, wt, ab);
}

// conservativeOfIsentropicVortex (:656-666, strength 5, centred at the origin)
// and its x / y derivatives (:668-712, including the /pi of :678 where the
// y derivative's :702 has /pi^2)
static void vortex_prim(double gamma, double g1, double pi, double x, double y, double w[4]) {
    const double r = std::pow(x, 2) + std::pow(y, 2), b = 5.0;
    w[0] = 1.0 - g1 * std::pow(b, 2) / 8.0 / gamma / std::pow(pi, 2) * std::exp(1 - r);
    w[0] = std::pow(w[0], 1.0 / g1);
    w[1] = 1.0 - b * y / 2.0 / pi * std::exp(0.5 * (1 - r));
    w[2] = 1.0 + b * x / 2.0 / pi * std::exp(0.5 * (1 - r));
    w[3] = std::pow(w[0], gamma);
}

static Vector4d vortex_deriv(double gamma, double g1, double pi, double x, double y, bool ydir) {
    double w[4];
    vortex_prim(gamma, g1, pi, x, y, w);
    const double r = std::pow(x, 2) + std::pow(y, 2);
    Vector4d q;
    if (!ydir) {
        q(0) = 1.0 - g1 * 25.0 / 8.0 / gamma / pi * std::exp(1 - r);
        q(0) = 25.0 * x * std::exp(1 - r) / 4.0 / std::pow(pi, 2) / gamma * std::pow(q(0), (2.0 - gamma) / g1);
        q(1) = 5.0 * x * y / 2.0 / pi * std::exp(0.5 * (1 - r));
        q(2) = 5.0 * (1 - std::pow(x, 2)) / 2.0 / pi * std::exp(0.5 * (1 - r));
    } else {
        q(0) = 1.0 - g1 * 25.0 / 8.0 / gamma / std::pow(pi, 2) * std::exp(1 - r);
        q(0) = 25.0 * y * std::exp(1 - r) / 4.0 / std::pow(pi, 2) / gamma * std::pow(q(0), (2.0 - gamma) / g1);
        q(1) = 5.0 * (std::pow(y, 2) - 1.0) / 2.0 / pi * std::exp(0.5 * (1 - r));
        q(2) = -5.0 * x * y / 2.0 / pi * std::exp(0.5 * (1 - r));
    }
    q(3) = gamma * std::pow(w[0], g1) * q(0) / g1 + 0.5 * q(0) * (std::pow(w[1], 2) + std::pow(w[2], 2)) +
           0.5 * w[0] * (2.0 * w[1] * q(1) + 2.0 * w[2] * q(2));
    q(1) = w[0] * q(1) + q(0) * w[1];
    q(2) = w[0] * q(2) + q(0) * w[2];
    return q;
}

Vector4d Euler2D::isentropicVortex(const double& x, const double& y, const double& dx, const double& dy,
                                   const Matrix<double, 9, 1>& wt, const Matrix<double, 9, 1>& ab) const {
    return gauss_avg2(x, y, dx, dy, wt, ab, [&](double px, double py) {
        double w[4];
        vortex_prim(gamma, gamma_1, pi, px, py, w);
        Vector4d v;
        v << w[0], w[1], w[2], w[3];
        return conservative(v);
    });
}

Vector4d Euler2D::dxIsentropicVortex(const double& x, const double& y, const double& dx, const double& dy,
                                     const Matrix<double, 9, 1>& wt, const Matrix<double, 9, 1>& ab) const {
    return gauss_avg2(x, y, dx, dy, wt, ab,
                      [&](double px, double py) { return vortex_deriv(gamma, gamma_1, pi, px, py, false); });
}

Vector4d Euler2D::dyIsentropicVortex(const double& x, const double& y, const double& dx, const double& dy,
                                     const Matrix<double, 9, 1>& wt, const Matrix<double, 9, 1>& ab) const {
    return gauss_avg2(x, y, dx, dy, wt, ab,
                      [&](double px, double py) { return vortex_deriv(gamma, gamma_1, pi, px, py, true); });
}

// :530-617 (the four quadrants split at (0.5, 0.5))
Vector4d Euler2D::conservativeOfRiemann2D(const double& x, const double& y) const {
    Vector4d w;
    if (x > 0.5 && y > 0.5)
        w << 1.0, 0.75, -0.5, 1.0;
    else if (x < 0.5 && y > 0.5)
        w << 2.0, 0.75, 0.5, 1.0;
    else if (x < 0.5 && y < 0.5)
        w << 1.0, -0.75, 0.5, 1.0;
    else
        w << 3.0, -0.75, -0.5, 1.0;
    return conservative(w);
}

// :619-635
Vector4d Euler2D::conservativeOfTaylorIns2D(const double& x, const double& y) const {
    Vector4d w;
    if (y < 0.5) {
        w(0) = 2.0;
        w(3) = 1.0 + 2.0 * y;
    } else {
        w(0) = 1.0;
        w(3) = y + 1.5;
    }
    w(1) = 0.0;
    w(2) = -0.025 * std::sqrt(gamma * w(3) / w(0)) * std::cos(8.0 * pi * x);
    return conservative(w);
}

// :637-654
Vector4d Euler2D::conservativeOfDoubleMach2D(const double& x, const double& y) const {
    Vector4d w;
    if (x < 1.0 / 6.0 + y / std::sqrt(3.0))
        w << 8.0, 8.25 * 0.5 * std::sqrt(3.0), -8.25 * 0.5, 116.5;
    else
        w << 1.4, 0.0, 0.0, 1.0;
    return conservative(w);
}

// ---------------------------------------------------------------------------
// Equations::Euler1D (Equations_Euler1D.cpp)
// ---------------------------------------------------------------------------
Euler1D::Euler1D(const double& g) : gamma(g), gamma_1(g - 1.0), gamma_3(g - 3.0), pi(4.0 * std::atan(1.0)) {}

// :18-37
Vector3d Euler1D::primitive(const Vector3d& q) const {
    Vector3d w;
    htx_host::e1_prim(gamma_1, q.data(), w.data());
    return w;
}

Vector3d Euler1D::conservative(const Vector3d& w) const {
    Vector3d q;
    q(0) = w(0);
    q(1) = w(0) * w(1);
    q(2) = w(2) / gamma_1 + 0.5 * q(1) * w(1);
    return q;
}

Vector3d Euler1D::flux(const Vector3d& q, const Vector3d& w) const {
    Vector3d f;
    htx_host::e1_phys_flux(q.data(), w.data(), f.data());
    return f;
}

// :39-53
Matrix3d Euler1D::dflux(const Vector3d& q, const Vector3d& w) const {
    Matrix3d J;
    J(0, 1) = 1.0;
    J(1, 0) = 0.5 * gamma_3 * w(1) * w(1);
    J(1, 1) = -gamma_3 * w(1);
    J(1, 2) = gamma_1;
    J(2, 0) = (gamma_1 * std::pow(w(1), 2) - gamma * q(2) / w(0)) * w(1);
    J(2, 1) = gamma * q(2) / w(0) - 1.5 * gamma_1 * w(1) * w(1);
    J(2, 2) = gamma * w(1);
    return J;
}

// :55-89
Matrix3d Euler1D::rightEigenVector(const double& vxm, const double& hm, const double& qm, const double& cm,
                                   const double& t0) const {
    Matrix3d R;
    R << 1.0, 1.0, 1.0, vxm - cm, vxm, vxm + cm, hm - t0, qm, hm + t0;
    return R;
}

Matrix3d Euler1D::leftEigenVector(const double& rcm, const double& b1, const double& b2, const double& t0,
                                  const double& t1, const double& t2) const {
    Matrix3d L;
    L << 0.5 * (b2 + t0), -0.5 * (t1 + rcm), t2, 1.0 - b2, t1, -b1, 0.5 * (b2 - t0), -0.5 * (t1 - rcm), t2;
    return L;
}

// :91-94 (with fabs on both terms)
double Euler1D::maxSpeed(const Vector3d& w) const {
    const double c = std::sqrt(std::fabs(gamma * w(2) / w(0)));
    return std::fmax(std::fabs(w(1) - c), std::fabs(w(1) + c));
}

// :109-127
Vector3d Euler1D::localWaveSpeed(const Vector3d& w) const {
    const double c = std::sqrt(std::fabs(gamma * w(2) / w(0)));
    Vector3d s;
    s << std::fabs(w(1) - c), std::fabs(w(1)), std::fabs(w(1) + c);
    return s;
}

// :96-107
Vector3d Euler1D::maxLocalWaveSpeed(const Vector3d& wl, const Vector3d& wr) const {
    return localWaveSpeed(wl).cwiseMax(localWaveSpeed(wr));
}

// :129-135
Vector3d Euler1D::Rusanov(const Vector3d& ql, const Vector3d& qr, const Vector3d& wl, const Vector3d& wr,
                          const Matrix3d& L, const Matrix3d& R, const Vector3d& a) const {
    return 0.5 * (flux(ql, wl) + flux(qr, wr) - R * a.cwiseProduct(L * (qr - ql)));
}

// Kurganov's central-upwind flux (:137-154): the 1D kernels' e1_cuflux
// (one-sided local speeds a+ >= 0, a- <= 0 from the componentwise max / min
// of the two states, the HLL-type average of the fluxes and the
// characteristic anti-diffusion term)
static void cu_speeds(const Euler1D& e, const Vector3d& ql, const Vector3d& qr, Vector3d& ap, Vector3d& am) {
    const Vector3d smax = e.localWaveSpeed(e.primitive(ql.cwiseMax(qr)));
    const Vector3d smin = e.localWaveSpeed(e.primitive(ql.cwiseMin(qr)));
    ap = smax.cwiseMax(smin).cwiseMax(0.0);
    am = smax.cwiseMin(smin).cwiseMin(0.0);
}

Vector3d Euler1D::CUFlux(const Vector3d& ql, const Vector3d& qr, const Matrix3d& L, const Matrix3d& R) const {
    double Lr[9], Rr[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            Lr[3 * r + c] = L(r, c);
            Rr[3 * r + c] = R(r, c);
        }
    Vector3d f;
    htx_host::e1_cuflux(gamma, gamma_1, Lr, Rr, ql.data(), qr.data(), f.data());
    return f;
}

// :156-210 (the kernels' 1D HLLC with the reference's quirks)
Vector3d Euler1D::HLLC(const Vector3d& ql, const Vector3d& qr) const {
    Vector3d f;
    htx_host::e1_hllc<true>(gamma, gamma_1, ql.data(), qr.data(), f.data());
    return f;
}

// :212-220
Vector3d Euler1D::dRusanov(const Vector3d& dql, const Vector3d& dqr, const Vector3d&, const Vector3d&,
                           const Matrix3d& L, const Matrix3d& R, const Matrix3d& dfl, const Matrix3d& dfr,
                           const Vector3d& a) const {
    return 0.5 * (dfl * dql + dfr * dqr - R * a.cwiseProduct(L * (dqr - dql)));
}

// :222-241: the derivative moments' central-upwind flux as the reference
// writes it (its anti-diffusion term multiplies R by the moment jump without
// a left projection)
Vector3d Euler1D::dCUFlux(const Vector3d& ql, const Vector3d& qr, const Matrix3d&, const Matrix3d& R,
                          const Vector3d& dql, const Vector3d& dqr, const Matrix3d& dfl, const Matrix3d& dfr) const {
    Vector3d ap, am;
    cu_speeds(*this, ql, qr, ap, am);
    Vector3d f = ap.cwiseProduct(dfl * dql) - am.cwiseProduct(dfr * dqr) + R * ap.cwiseProduct(am).cwiseProduct(dqr - dql);
    return f.cwiseProduct((ap - am).cwiseInverse());
}

// :243-293
Vector3d Euler1D::entropyWaves(const double& rhoInf, const double& uInf, const double& pInf, const double& amp,
                               const double& x, const double& dx, const Matrix<double, 9, 1>& wt,
                               const Matrix<double, 9, 1>& ab) const {
    return gauss_avg1(x, dx, wt, ab, [&](double px) {
        Vector3d w;
        w << rhoInf + amp * std::sin(pi * (px)), uInf, pInf;
        return conservative(w);
    });
}

Vector3d Euler1D::entropyWavesdx(const double&, const double& uInf, const double&, const double& amp,
                                 const double& x, const double& dx, const Matrix<double, 9, 1>& wt,
                                 const Matrix<double, 9, 1>& ab) const {
    return gauss_avg1(x, dx, wt, ab, [&](double px) {
        Vector3d q;
        q(0) = amp * std::cos(pi * (px)) * pi;
        q(1) = uInf * q(0);
        q(2) = (0.5 * (uInf * uInf)) * q(0);
        return q;
    });
}

// :295-344 (primitive states)
Vector3d Euler1D::interactingBlast(const double& x) const {
    Vector3d w;
    w << 1.0, 0.0, x < 0.1 ? 1000.0 : (x < 0.9 ? 0.01 : 100.0);
    return w;
}

Vector3d Euler1D::sodShock(const double& x) const {
    Vector3d w;
    if (x < 0.5)
        w << 1.0, 0.0, 1.0;
    else
        w << 0.125, 0.0, 0.1;
    return w;
}

Vector3d Euler1D::shuOsher(const double& x) const {
    Vector3d w;
    if (x < -4.0)
        w << 3.857143, 2.629369, 10.333333;
    else
        w << 1.0 + 0.2 * std::sin(5.0 * x), 0.0, 1.0;
    return w;
}

}  // namespace Equations

// ---------------------------------------------------------------------------
// Schemes::Lagrange (Schemes_Lagrange.cpp)
// ---------------------------------------------------------------------------
namespace Schemes {

Lagrange::Lagrange() {
    double x[9], w[9];
    gauss9(x, w);
    for (int k = 0; k < 9; ++k) {
        gaussianQuadrature9Points(k) = x[k];
        gaussianQuadrature9Weights(k) = w[k];
    }
    // :56-66
    gaussianQuadrature2Weights << 1.0, 1.0;
    gaussLobattoQuadrature4Weights << 1.0 / 6.0, 1.0 / 6.0, 5.0 / 6.0, 5.0 / 6.0;
    gaussLegendreQuadrature4Weights << (18.0 + std::sqrt(30.0)) / 36.0, (18.0 + std::sqrt(30.0)) / 36.0,
        (18.0 - std::sqrt(30.0)) / 36.0, (18.0 - std::sqrt(30.0)) / 36.0;
    gaussLegendreQuadrature3Weights << 5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0;
}

Vector9d Lagrange::getQuadrature9Weights() const { return gaussianQuadrature9Weights; }
Vector9d Lagrange::getQuadrature9Points() const { return gaussianQuadrature9Points; }

static Vector2d pair(double a, double b) {
    Vector2d v;
    v << a, b;
    return v;
}

Vector2d Lagrange::teno5(const Vector6d& u) const {
    double um, up;
    htx_host::teno5(u.data(), um, up);
    return pair(um, up);
}

Vector2d Lagrange::linear(const Vector6d& u) const {
    double um, up;
    htx_host::linear5(u.data(), um, up);
    return pair(um, up);
}

Vector2d Lagrange::teno5i(const Vector6d& u) const {
    double um, up;
    htx_host::teno5i(u.data(), um, up);
    return pair(um, up);
}

Vector2d Lagrange::weno5JS(const Vector6d& u) const {
    double um, up;
    htx_host::weno5(u.data(), false, um, up);
    return pair(um, up);
}

Vector2d Lagrange::weno5Z(const Vector6d& u) const {
    double um, up;
    htx_host::weno5(u.data(), true, um, up);
    return pair(um, up);
}

Vector3d Lagrange::weno5JS3PointsQuadrature(const Vector5d& u) const {
    htx_host::QuadConsts C;
    htx_host::fill_quad_consts(C);
    Vector3d o;
    htx_host::weno3q(C, u(0), u(1), u(2), u(3), u(4), o(0), o(1), o(2));
    return o;
}

// The remaining reconstructions and quadratures (htx_variants.h, each citing
// its reference routine)
Vector2d Lagrange::weno3JS(const Vector4d& u) const {
    double a, b;
    htx_host::weno3js(u.data(), a, b);
    return pair(a, b);
}

Vector2d Lagrange::weno3JS2PointsQuadrature(const Vector3d& u) const {
    double a, b;
    htx_host::weno3js_2q(u.data(), a, b);
    return pair(a, b);
}

static Vector2d recon_variant(int kind, const Vector6d& u, double diff) {
    double um, up;
    htx_host::recon5_variant(kind, u.data(), diff, um, up);
    return pair(um, up);
}

Vector2d Lagrange::weno5Z(const Vector6d& u, const double& diff) const {
    return recon_variant(htx_host::RECON_WENO5Z_DIFF, u, diff);
}
Vector2d Lagrange::weno5N(const Vector6d& u) const { return recon_variant(htx_host::RECON_WENO5N, u, 0.0); }
Vector2d Lagrange::teno5a(const Vector6d& u) const { return recon_variant(htx_host::RECON_TENO5A, u, 0.0); }
Vector2d Lagrange::teno5(const Vector6d& u, const double& diff) const {
    return recon_variant(htx_host::RECON_TENO5_DIFF, u, diff);
}
Vector2d Lagrange::teno5N(const Vector6d& u) const { return recon_variant(htx_host::RECON_TENO5N, u, 0.0); }
Vector2d Lagrange::weno5IS(const Vector6d& u) const { return recon_variant(htx_host::RECON_WENO5IS, u, 0.0); }
Vector2d Lagrange::weno5ISmod(const Vector6d& u, const double& diff) const {
    return recon_variant(htx_host::RECON_WENO5ISMOD, u, diff);
}
Vector2d Lagrange::weno5IS_Extended(const Vector6d& u) const {
    return recon_variant(htx_host::RECON_WENO5IS_EXT, u, 0.0);
}

static const htx_host::QuadVar& quad_var() {
    static const htx_host::QuadVar Q = [] {
        htx_host::QuadVar q;
        htx_host::fill_quad_var(q);
        return q;
    }();
    return Q;
}

static Vector2d quad2(int kind, const Vector5d& u) {
    double a, b;
    htx_host::weno5_2q(quad_var(), kind, u.data(), a, b);
    return pair(a, b);
}

Vector2d Lagrange::weno5JS2PointsQuadrature(const Vector5d& u) const { return quad2(0, u); }
Vector2d Lagrange::weno5Z2PointsQuadrature(const Vector5d& u) const { return quad2(1, u); }
Vector2d Lagrange::teno5th2PointsQuadrature(const Vector5d& u) const { return quad2(2, u); }

static Vector4d quad4(bool lobatto, int kind, const Vector5d& u) {
    Vector4d o;
    if (lobatto)
        htx_host::weno5_4lobatto(quad_var(), kind, u.data(), o.data());
    else
        htx_host::weno5_4legendre(quad_var(), kind, u.data(), o.data());
    return o;
}

Vector4d Lagrange::weno5JS4PointsGaussLobatto(const Vector5d& u) const { return quad4(true, 0, u); }
Vector4d Lagrange::weno5Z4PointsGaussLobatto(const Vector5d& u) const { return quad4(true, 1, u); }
Vector4d Lagrange::teno5th4PointsGaussLobatto(const Vector5d& u) const { return quad4(true, 2, u); }
Vector4d Lagrange::weno5JS4PointsGaussLegendre(const Vector5d& u) const { return quad4(false, 0, u); }
Vector4d Lagrange::weno5Z4PointsGaussLegendre(const Vector5d& u) const { return quad4(false, 1, u); }

// ---------------------------------------------------------------------------
// Schemes::Hermite (Schemes_Hermite.cpp)
// ---------------------------------------------------------------------------
Hermite::Hermite() : dif(0.0) {}

// :15-232: every coefficient follows from the grid size; a zero size is an
// error (the reference prints and exits)
void Hermite::setGridSize(const double& gridSize) {
    if (gridSize == 0.0) throw std::invalid_argument("Hermite::setGridSize: the grid size must be nonzero");
    dif = gridSize;
}

static htx_host::HermConsts herm_consts(double dif) {
    if (dif == 0.0) throw std::logic_error("Hermite: setGridSize has not been called");
    htx_host::HermConsts H;
    htx_host::fill_herm_consts(H, dif);
    return H;
}

Vector4d Hermite::hteno(const Vector4d& u, const Vector4d& d) const {
    const htx_host::HermConsts H = herm_consts(dif);
    Vector4d o;
    htx_host::hteno<false>(H, u.data(), d.data(), o.data());
    return o;
}

double Hermite::hermiteLimiter(const Vector3d& u, const Vector3d& d) const {
    (void)herm_consts(dif);
    return htx_host::herm_limiter(dif, u(0), u(1), u(2), d(0), d(2));
}

// :664-668
bool Hermite::KXRCF(const double& uj, const double& unbj, const double& unorm) const {
    const double ik = std::fabs(uj - unbj) / (std::fabs(unorm) * std::pow(dif, 1.5));
    return ik > 1.0;
}

Vector4d Hermite::hweno(const Vector4d& u, const Vector4d& d) const {
    const htx_host::HermConsts H = herm_consts(dif);
    Vector4d o;
    htx_host::hweno(H, u.data(), d.data(), o.data());
    return o;
}

Vector4d Hermite::linear(const Vector4d& u, const Vector4d& d) const {
    const htx_host::HermConsts H = herm_consts(dif);
    Vector4d o;
    htx_host::herm_linear(H, u.data(), d.data(), o.data());
    return o;
}

Vector4d Hermite::htenoi(const Vector4d& u, const Vector4d& d) const {
    const htx_host::HermConsts H = herm_consts(dif);
    Vector4d o;
    htx_host::htenoi(H, u.data(), d.data(), o.data());
    return o;
}

Vector4d Hermite::hweno2PointsGLQ(const Vector3d& u, const Vector3d& d) const {
    (void)herm_consts(dif);
    htx_host::HermGLQ G;
    htx_host::fill_herm_glq(G, dif);
    Vector4d o;
    htx_host::hweno_2q(G, dif, u.data(), d.data(), o.data());
    return o;
}

Matrix<double, 8, 1> Hermite::hweno4PointsGLQ(const Vector3d& u, const Vector3d& d) const {
    const htx_host::HermConsts H = herm_consts(dif);
    htx_host::HermGLQ G;
    htx_host::fill_herm_glq(G, dif);
    Matrix<double, 8, 1> o;
    htx_host::hweno_4q(H, G, u.data(), d.data(), o.data());
    return o;
}

}  // namespace Schemes
