// Equations_Euler1D.h -- drop-in for the reference's Equations::Euler1D
// (Equations_Euler1D.h:16-51): the 1D Euler physics as a host FP64 class with
// the reference's public interface.  q = (rho, rho u, E), w = (rho, u, p).
// The methods evaluate the reference's formulas (Equations_Euler1D.cpp) with
// the operations of the GPU kernels' host build; HLLC keeps the reference's
// c = sqrt(|p / rho|) and p_L in the right star energy (:174-175, :204).
// The methods are const (no scratch members).  Link with -lhtx.
#pragma once
#include "htx_linalg.h"

namespace Equations {

class Euler1D {
public:
    Euler1D(const double& specificHeatRatio);

    Vector3d primitive(const Vector3d& q) const;
    Vector3d conservative(const Vector3d& w) const;
    Vector3d flux(const Vector3d& q, const Vector3d& w) const;
    Matrix3d dflux(const Vector3d& q, const Vector3d& w) const;
    Matrix3d rightEigenVector(const double& vxm, const double& hm, const double& qm, const double& cm,
                              const double& t0) const;
    Matrix3d leftEigenVector(const double& rcm, const double& b1, const double& b2, const double& t0,
                             const double& t1, const double& t2) const;
    double maxSpeed(const Vector3d& w) const;

    Vector3d maxLocalWaveSpeed(const Vector3d& wl, const Vector3d& wr) const;
    Vector3d localWaveSpeed(const Vector3d& w) const;
    Vector3d Rusanov(const Vector3d& ql, const Vector3d& qr, const Vector3d& wl, const Vector3d& wr,
                     const Matrix3d& leftEigenVec, const Matrix3d& rightEigenVec, const Vector3d& maxSpeed) const;
    Vector3d CUFlux(const Vector3d& ql, const Vector3d& qr, const Matrix3d& leftEigenVec,
                    const Matrix3d& rightEigenVec) const;
    Vector3d HLLC(const Vector3d& ql, const Vector3d& qr) const;
    Vector3d dRusanov(const Vector3d& dql, const Vector3d& dqr, const Vector3d& wl, const Vector3d& wr,
                      const Matrix3d& leftEigenVec, const Matrix3d& rightEigenVec, const Matrix3d& dfluxl,
                      const Matrix3d& dfluxr, const Vector3d& maxSpeed) const;
    Vector3d dCUFlux(const Vector3d& ql, const Vector3d& qr, const Matrix3d& leftEigenVec,
                     const Matrix3d& rightEigenVec, const Vector3d& dql, const Vector3d& dqr,
                     const Matrix3d& dfluxl, const Matrix3d& dfluxr) const;

    Vector3d entropyWaves(const double& rhoInf, const double& uInf, const double& pInf, const double& amplitude,
                          const double& x, const double& dx, const Matrix<double, 9, 1>& weights,
                          const Matrix<double, 9, 1>& absis) const;
    Vector3d entropyWavesdx(const double& rhoInf, const double& uInf, const double& pInf, const double& amplitude,
                            const double& x, const double& dx, const Matrix<double, 9, 1>& weights,
                            const Matrix<double, 9, 1>& absis) const;

    Vector3d interactingBlast(const double& x) const;
    Vector3d sodShock(const double& x) const;
    Vector3d shuOsher(const double& x) const;

private:
    double gamma, gamma_1, gamma_3, pi;
};

}  // namespace Equations
