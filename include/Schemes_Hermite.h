// Schemes_Hermite.h -- drop-in for the reference's Schemes::Hermite
// (Schemes_Hermite.h:21-36): the Hermite (HWENO / HTENO) reconstruction of
// cell averages u and derivative moments, the derivative limiter and the
// KXRCF indicator, as a host FP64 class with the reference's public interface.
// hteno takes four cell averages u_{i-1..i+2} and their derivative moments
// and returns (u^-, u^+, d^-, d^+) at x_{i+1/2}.  The methods run the numerics
// of the GPU kernels' host build (libhtx.so).  No scratch members: the
// methods are const.  Link with -lhtx.
//
// setGridSize(0) throws std::invalid_argument where the reference prints and
// exits (Schemes_Hermite.cpp:16-25); calling a method before setGridSize
// throws std::logic_error.  hweno, htenoi and linear are also selectable as the
// face reconstruction of the 2D Hermite GPU solver (htx_scheme in htx.h).
#pragma once
#include "htx_linalg.h"

namespace Schemes {

class Hermite {
public:
    Hermite();
    void setGridSize(const double& gridSize);
    Vector4d hweno(const Vector4d& u, const Vector4d& derivatives) const;    // :234-293
    Vector4d hteno(const Vector4d& u, const Vector4d& derivatives) const;    // :295-379
    Vector4d linear(const Vector4d& u, const Vector4d& derivatives) const;   // :757-767

    Vector4d htenoi(const Vector4d& u, const Vector4d& derivatives) const;   // :670-755

    double hermiteLimiter(const Vector3d& u, const Vector3d& derivatives) const;  // :634-662
    Vector4d hweno2PointsGLQ(const Vector3d& u, const Vector3d& derivatives) const;           // :381-425
    Matrix<double, 8, 1> hweno4PointsGLQ(const Vector3d& u, const Vector3d& derivatives) const;  // :427-513
    bool KXRCF(const double& u0, const double& ul, const double& unorm) const;  // :664-668

    double gridSize() const { return dif; }

private:
    double dif;
};

}  // namespace Schemes
