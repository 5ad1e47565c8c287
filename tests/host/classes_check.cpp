// classes_check.cpp -- test driver (CPU): calls the drop-in physics and
// scheme classes (include/Equations_*.h, Schemes_*.h, built into libhtx.so)
// the way code written against the reference's classes calls them, on fixed
// inputs, and prints every result as "name v0 v1 ..." (%.17g) for
// tests/test_classes_cpu.py to compare with the oracle.  Test infrastructure.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "Equations_Euler1D.h"
#include "Equations_Euler2D.h"
#include "Schemes_Hermite.h"
#include "Schemes_Lagrange.h"
#include "Schemes_SSPRK.h"

template <int R, int C>
static void put(const char* name, const Matrix<double, R, C>& m) {
    std::printf("%s", name);
    for (int j = 0; j < C; ++j)
        for (int i = 0; i < R; ++i) std::printf(" %.17g", m(i, j));  // column-major
    std::printf("\n");
}

static void put(const char* name, double v) { std::printf("%s %.17g\n", name, v); }

template <class F>
static void expect_throw(const char* name, F f) {
    try {
        f();
        std::printf("%s no-throw\n", name);
    } catch (const std::logic_error&) {
        std::printf("%s throws\n", name);
    }
}

int main() {
    const double gamma = 1.4;
    Equations::Euler2D eul(gamma);
    Vector4d wl, wr;
    wl << 1.0, 0.3, -0.2, 1.0;
    wr << 0.5, -0.1, 0.4, 0.6;
    const Vector4d ql = eul.conservative(wl), qr = eul.conservative(wr);
    put("e2.ql", ql);
    put("e2.qr", qr);
    put("e2.prim", eul.primitive(ql));
    for (int dir = 1; dir <= 2; ++dir) {
        std::string d = std::to_string(dir);
        put(("e2.hllc" + d).c_str(), eul.HLLC(ql, qr, dir));
        put(("e2.maxspeed" + d).c_str(), eul.maxSpeed(wr, dir));
        put(("e2.dflux" + d).c_str(), dir == 1 ? eul.dfluxx(ql, wl) : eul.dfluxy(ql, wl));
        // the Roe average and eigenvectors as FV::System2D::eigen2DEuler forms
        // them (FV_System2D.cpp:271-298), then the characteristic Rusanov flux
        const double sa = std::sqrt(std::fabs(wl(0))), sb = std::sqrt(std::fabs(wr(0)));
        const double ha = (ql(3) + wl(3)) / wl(0), hb = (qr(3) + wr(3)) / wr(0);
        const double t0 = sa / (sa + sb), t1 = 1.0 - t0;
        const double vxm = t0 * wl(1) + t1 * wr(1), vym = t0 * wl(2) + t1 * wr(2);
        const double hm = t0 * ha + t1 * hb, qm = 0.5 * (vxm * vxm + vym * vym);
        const double cm = std::sqrt(std::fabs((gamma - 1.0) * (hm - qm))), rcm = 1.0 / cm;
        const double b1 = (gamma - 1.0) * rcm * rcm, b2 = qm * b1, tt2 = 0.5 * b1;
        const double vn = dir == 1 ? vxm : vym;
        Matrix4d R = dir == 1 ? eul.rightEigenVectorx(vxm, vym, hm, qm, cm, vn * cm)
                              : eul.rightEigenVectory(vxm, vym, hm, qm, cm, vn * cm);
        Matrix4d L = dir == 1 ? eul.leftEigenVectorx(rcm, vym, b1, b2, vn * rcm, b1 * vn, tt2)
                              : eul.leftEigenVectory(rcm, vxm, b1, b2, vn * rcm, b1 * vn, tt2);
        put(("e2.R" + d).c_str(), R);
        put(("e2.L" + d).c_str(), L);
        put(("e2.rusanov" + d).c_str(), eul.Rusanov(ql, qr, wl, wr, L, R, eul.maxLocalWaveSpeed(wl, wr, dir), dir));
        put(("e2.rusanovM" + d).c_str(), eul.Rusanov(ql, qr, wl, wr, eul.maxLocalWaveSpeed(wl, wr, L, R, dir), dir));
    }
    Schemes::Lagrange lag;
    const Vector9d gw = lag.getQuadrature9Weights(), gx = lag.getQuadrature9Points();
    put("lag.gw", gw);
    put("lag.gx", gx);
    put("e2.vortex", eul.isentropicVortex(0.3, -0.7, 0.2, 0.25, gw, gx));
    put("e2.vortexdx", eul.dxIsentropicVortex(0.3, -0.7, 0.2, 0.25, gw, gx));
    put("e2.vortexdy", eul.dyIsentropicVortex(0.3, -0.7, 0.2, 0.25, gw, gx));
    put("e2.entropy", eul.entropyWaves(1.0, 1.0, 1.0, 1.0, 0.5, 0.3, 0.1, 0.1, 0.1, gw, gx));
    put("e2.riemann", eul.conservativeOfRiemann2D(0.7, 0.2));

    // reconstructions on two stencils: a smooth one and one with a jump
    Vector6d s1, s2;
    s1 << 1.0, 1.1, 1.25, 1.3, 1.32, 1.31;
    s2 << 1.0, 1.0, 1.0, 0.2, 0.21, 0.19;
    const Vector6d st[2] = {s1, s2};
    for (int k = 0; k < 2; ++k) {
        std::string n = std::to_string(k);
        put(("lag.teno5." + n).c_str(), lag.teno5(st[k]));
        put(("lag.linear." + n).c_str(), lag.linear(st[k]));
        put(("lag.teno5i." + n).c_str(), lag.teno5i(st[k]));
        put(("lag.weno5JS." + n).c_str(), lag.weno5JS(st[k]));
        put(("lag.weno5Z." + n).c_str(), lag.weno5Z(st[k]));
        Vector5d v;
        for (int m = 0; m < 5; ++m) v(m) = st[k](m);
        put(("lag.quad3." + n).c_str(), lag.weno5JS3PointsQuadrature(v));
    }
    // the remaining reconstructions and quadratures on eight stencils: the
    // two above, a flat one, a linear one, and four drawn from a fixed LCG
    Vector6d sv[8];
    sv[0] = s1;
    sv[1] = s2;
    sv[2] << 2.0, 2.0, 2.0, 2.0, 2.0, 2.0;
    sv[3] << 0.5, 0.75, 1.0, 1.25, 1.5, 1.75;
    unsigned long long lcg = 0x48544E4F00ull;
    for (int k = 4; k < 8; ++k)
        for (int m = 0; m < 6; ++m) {
            lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
            sv[k](m) = 0.5 + (double)(lcg >> 11) / 9007199254740992.0 * (k == 7 ? 1e-3 : 1.0);
        }
    for (int k = 0; k < 8; ++k) {
        std::string n = "." + std::to_string(k);
        const Vector6d& u = sv[k];
        put(("lag.in" + n).c_str(), u);
        put(("lag.weno5N" + n).c_str(), lag.weno5N(u));
        put(("lag.teno5N" + n).c_str(), lag.teno5N(u));
        put(("lag.teno5a" + n).c_str(), lag.teno5a(u));
        put(("lag.weno5IS" + n).c_str(), lag.weno5IS(u));
        put(("lag.weno5ISmod" + n).c_str(), lag.weno5ISmod(u, 0.05));
        put(("lag.weno5IS_Extended" + n).c_str(), lag.weno5IS_Extended(u));
        put(("lag.weno5Zd" + n).c_str(), lag.weno5Z(u, 0.05));
        put(("lag.teno5d" + n).c_str(), lag.teno5(u, 0.05));
        Vector4d u4;
        Vector3d u3;
        Vector5d u5;
        for (int m = 0; m < 5; ++m) u5(m) = u(m);
        for (int m = 0; m < 4; ++m) u4(m) = u(m);
        for (int m = 0; m < 3; ++m) u3(m) = u(m);
        put(("lag.weno3JS" + n).c_str(), lag.weno3JS(u4));
        put(("lag.weno3JS2q" + n).c_str(), lag.weno3JS2PointsQuadrature(u3));
        put(("lag.q2js" + n).c_str(), lag.weno5JS2PointsQuadrature(u5));
        put(("lag.q2z" + n).c_str(), lag.weno5Z2PointsQuadrature(u5));
        put(("lag.q2teno" + n).c_str(), lag.teno5th2PointsQuadrature(u5));
        put(("lag.q4lojs" + n).c_str(), lag.weno5JS4PointsGaussLobatto(u5));
        put(("lag.q4loz" + n).c_str(), lag.weno5Z4PointsGaussLobatto(u5));
        put(("lag.q4loteno" + n).c_str(), lag.teno5th4PointsGaussLobatto(u5));
        put(("lag.q4lejs" + n).c_str(), lag.weno5JS4PointsGaussLegendre(u5));
        put(("lag.q4lez" + n).c_str(), lag.weno5Z4PointsGaussLegendre(u5));
    }

    Schemes::Hermite her;
    expect_throw("her.unset", [&] { her.hteno(Vector4d(), Vector4d()); });
    her.setGridSize(0.1);
    Vector4d hu, hd;
    hu << 1.0, 1.2, 1.1, 0.4;
    hd << 0.5, 1.0, -2.0, -0.3;
    put("her.hteno", her.hteno(hu, hd));
    Vector3d lu, ld;
    lu << 1.0, 1.2, 1.1;
    ld << 0.5, 1.0, -2.0;
    put("her.limiter", her.hermiteLimiter(lu, ld));
    put("her.kxrcf", her.KXRCF(1.0, 1.001, 1.0) ? 1.0 : 0.0);
    // the Hermite alternatives on the stencils above (cell averages sv[k](0..3),
    // moments from sv[(k + 3) % 8] scaled)
    for (int k = 0; k < 8; ++k) {
        std::string n = "." + std::to_string(k);
        Vector4d u4, d4;
        Vector3d u3, d3;
        for (int m = 0; m < 4; ++m) {
            u4(m) = sv[k](m);
            d4(m) = 10.0 * (sv[(k + 3) % 8](m + 1) - 1.0);
        }
        for (int m = 0; m < 3; ++m) {
            u3(m) = u4(m);
            d3(m) = d4(m);
        }
        put(("her.hweno" + n).c_str(), her.hweno(u4, d4));
        put(("her.htenoi" + n).c_str(), her.htenoi(u4, d4));
        put(("her.linear" + n).c_str(), her.linear(u4, d4));
        put(("her.glq2" + n).c_str(), her.hweno2PointsGLQ(u3, d3));
        put(("her.glq4" + n).c_str(), her.hweno4PointsGLQ(u3, d3));
        put(("her.in" + n).c_str(), u4);
        put(("her.din" + n).c_str(), d4);
    }
    try {
        her.setGridSize(0.0);
        std::printf("her.zero no-throw\n");
    } catch (const std::invalid_argument&) {
        std::printf("her.zero throws\n");
    }

    Schemes::SSPRK rk;
    const double lam = -0.7, dt = 0.3, q0 = 1.3;
    const double q1 = rk.RungeKuttaIII(q0, q0, lam * q0, dt, 1);
    const double q2 = rk.RungeKuttaIII(q0, q1, lam * q1, dt, 2);
    put("rk.scalar", rk.RungeKuttaIII(q0, q2, lam * q2, dt, 3));
    put("rk.vec", rk.RungeKuttaIII(ql, qr, ql - qr, 0.1, 3));

    Equations::Euler1D e1(gamma);
    Vector3d w1l, w1r;
    w1l << 1.0, 0.2, 1.0;
    w1r << 0.125, -0.1, 0.1;
    const Vector3d q1l = e1.conservative(w1l), q1r = e1.conservative(w1r);
    put("e1.ql", q1l);
    put("e1.qr", q1r);
    put("e1.hllc", e1.HLLC(q1l, q1r));
    put("e1.maxspeed", e1.maxSpeed(w1r));
    Matrix3d I3 = Matrix3d::Identity();
    put("e1.cu_same", e1.CUFlux(q1l, q1l, I3, I3));
    put("e1.flux_l", e1.flux(q1l, e1.primitive(q1l)));
    put("e1.sod", e1.sodShock(0.7));
    return 0;
}
