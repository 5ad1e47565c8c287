// Equations_Euler2D.h -- drop-in for the reference's Equations::Euler2D
// (Equations_Euler2D.h:17-79): the 2D Euler physics as a host FP64 class with
// the reference's public interface.  Conserved state q = (rho, rho u, rho v,
// E), primitive w = (rho, u, v, p); dir 1 = x, 2 = y.  The methods evaluate
// the reference's formulas (Equations_Euler2D.cpp) with the same operations
// the GPU kernels use (libhtx.so, host build of the kernel numerics); the
// HLLC flux keeps the reference's c = sqrt(|p / rho|) and p_L in the right
// star energy (:393-394, :423), maxSpeed its missing fabs (:218).
// Unlike the reference the object keeps no scratch members: the methods are
// const and safe to call from several threads.  Link with -lhtx.
#pragma once
#include "htx_linalg.h"

namespace Equations {

class Euler2D {
public:
    Euler2D(const double& specificHeatRatio);

    Vector4d primitive(const Vector4d& q) const;
    Vector4d conservative(const Vector4d& w) const;
    Vector4d fluxx(const Vector4d& q, const Vector4d& w) const;
    Vector4d fluxy(const Vector4d& q, const Vector4d& w) const;
    Matrix4d dfluxx(const Vector4d& q, const Vector4d& w) const;
    Matrix4d dfluxy(const Vector4d& q, const Vector4d& w) const;
    Matrix4d rightEigenVectorx(const double& vxm, const double& vym, const double& hm, const double& qm,
                               const double& cm, const double& t0) const;
    Matrix4d leftEigenVectorx(const double& rcm, const double& vym, const double& b1, const double& b2,
                              const double& t0, const double& t1, const double& t2) const;
    Matrix4d rightEigenVectory(const double& vxm, const double& vym, const double& hm, const double& qm,
                               const double& cm, const double& t0) const;
    Matrix4d leftEigenVectory(const double& rcm, const double& vxm, const double& b1, const double& b2,
                              const double& t0, const double& t1, const double& t2) const;
    double maxSpeed(const Vector4d& w, int dir) const;

    Matrix4d maxLocalWaveSpeed(const Vector4d& wl, const Vector4d& wr, const Matrix4d& leftEigenVec,
                               const Matrix4d& rightEigenVec, int dir) const;
    Vector4d maxLocalWaveSpeed(const Vector4d& wl, const Vector4d& wr, int dir) const;
    Vector4d localWaveSpeed(const Vector4d& w, int dir) const;

    Vector4d Rusanov(const Vector4d& ql, const Vector4d& qr, const Vector4d& wl, const Vector4d& wr,
                     const Matrix4d& leftEigenVec, const Matrix4d& rightEigenVec, const Vector4d& maxSpeed,
                     int dir) const;
    Vector4d Rusanov(const Vector4d& ql, const Vector4d& qr, const Vector4d& wl, const Vector4d& wr,
                     const Matrix4d& maxSpeed, int dir) const;
    Vector4d dRusanov(const Vector4d& dql, const Vector4d& dqr, const Vector4d& wl, const Vector4d& wr,
                      const Matrix4d& leftEigenVec, const Matrix4d& rightEigenVec, const Matrix4d& dfluxl,
                      const Matrix4d& dfluxr, const Vector4d& maxSpeed) const;
    Vector4d dRusanov(const Vector4d& dql, const Vector4d& dqr, const Vector4d& wl, const Vector4d& wr,
                      const Matrix4d& dfluxl, const Matrix4d& dfluxr, const Matrix4d& maxSpeed) const;

    Vector4d HLLC(const Vector4d& ql, const Vector4d& qr, int dir) const;

    Vector4d entropyWaves(const double& rhoInf, const double& uInf, const double& vInf, const double& pInf,
                          const double& amplitude, const double& x, const double& y, const double& dx,
                          const double& dy, const Matrix<double, 9, 1>& weights,
                          const Matrix<double, 9, 1>& absis) const;
    Vector4d entropyWavesdx(const double& rhoInf, const double& uInf, const double& vInf, const double& pInf,
                            const double& amplitude, const double& x, const double& y, const double& dx,
                            const double& dy, const Matrix<double, 9, 1>& weights,
                            const Matrix<double, 9, 1>& absis) const;
    Vector4d entropyWavesdy(const double& rhoInf, const double& uInf, const double& vInf, const double& pInf,
                            const double& amplitude, const double& x, const double& y, const double& dx,
                            const double& dy, const Matrix<double, 9, 1>& weights,
                            const Matrix<double, 9, 1>& absis) const;

    Vector4d isentropicVortex(const double& x, const double& y, const double& dx, const double& dy,
                              const Matrix<double, 9, 1>& weights, const Matrix<double, 9, 1>& absis) const;
    Vector4d dxIsentropicVortex(const double& x, const double& y, const double& dx, const double& dy,
                                const Matrix<double, 9, 1>& weights, const Matrix<double, 9, 1>& absis) const;
    Vector4d dyIsentropicVortex(const double& x, const double& y, const double& dx, const double& dy,
                                const Matrix<double, 9, 1>& weights, const Matrix<double, 9, 1>& absis) const;

    Vector4d conservativeOfRiemann2D(const double& x, const double& y) const;
    Vector4d conservativeOfTaylorIns2D(const double& x, const double& y) const;
    Vector4d conservativeOfDoubleMach2D(const double& x, const double& y) const;

private:
    double gamma, gamma_1, gamma_3, pi;
};

}  // namespace Equations
