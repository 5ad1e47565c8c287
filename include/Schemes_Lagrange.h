// Schemes_Lagrange.h -- drop-in for the reference's Schemes::Lagrange
// (Schemes_Lagrange.h:20-79): the 5th-order WENO / TENO reconstructions and the
// transverse Gauss-point quadrature as a host FP64 class with the reference's
// public interface.  u holds six consecutive cell values u_{i-2..i+3}; a
// Vector2d result is (u^-_{i+1/2}, u^+_{i+1/2}).  The methods run the
// numerics of the GPU kernels' host build (libhtx.so), each one restating the
// reference routine cited at it in the library source.  The object keeps no
// scratch members (the reference's do, Schemes_Lagrange.h:106-145), so its
// methods are const and thread-safe.  Link with -lhtx.
//
// The remaining reconstructions and the 2- and 4-point quadratures (used by
// the reference's scalar solvers) are the same host+device routines the GPU
// solvers select (htx_variants.h).  The 5-point ones are also selectable as
// the characteristic reconstruction of the 1D and 2D GPU solvers
// (htx_scheme in htx.h).
#pragma once
#include "htx_linalg.h"

namespace Schemes {

class Lagrange {
public:
    Lagrange();

    Vector2d weno3JS(const Vector4d& u) const;                    // :315-335 (both outputs u^-)
    Vector2d weno3JS2PointsQuadrature(const Vector3d& u) const;   // :337-356

    Vector2d weno5JS(const Vector6d& u) const;   // :358-381
    Vector2d linear(const Vector6d& u) const;    // :1219-1233

    Vector2d weno5Z(const Vector6d& u) const;    // :383-408
    Vector2d weno5Z(const Vector6d& u, const double& diff) const;  // :410-438

    Vector2d weno5N(const Vector6d& u) const;    // :440-478

    Vector2d teno5(const Vector6d& u) const;     // :481-511
    Vector2d teno5a(const Vector6d& u) const;    // :1265-1294
    Vector2d teno5i(const Vector6d& u) const;    // :1235-1263
    Vector2d teno5(const Vector6d& u, const double& diff) const;  // :514-551

    Vector2d teno5N(const Vector6d& u) const;    // :553-593

    Vector2d weno5IS(const Vector6d& u) const;   // :596-620
    Vector2d weno5ISmod(const Vector6d& u, const double& diff) const;  // :622-647
    Vector2d weno5IS_Extended(const Vector6d& u) const;  // :649-717

    Vector2d weno5JS2PointsQuadrature(const Vector5d& u) const;   // :719-740
    Vector2d weno5Z2PointsQuadrature(const Vector5d& u) const;    // :742-765
    Vector2d teno5th2PointsQuadrature(const Vector5d& u) const;   // :767-791

    Vector3d weno5JS3PointsQuadrature(const Vector5d& u) const;   // :1171-1217

    Vector4d weno5JS4PointsGaussLobatto(const Vector5d& u) const;    // :793-831
    Vector4d weno5Z4PointsGaussLobatto(const Vector5d& u) const;     // :914-955
    Vector4d teno5th4PointsGaussLobatto(const Vector5d& u) const;    // :957-1004
    Vector4d weno5JS4PointsGaussLegendre(const Vector5d& u) const;   // :833-869
    Vector4d weno5Z4PointsGaussLegendre(const Vector5d& u) const;    // :871-912

    // Gauss-Legendre 9-point rule on [-1, 1] in the reference's node order
    // (0, -a1, +a1, ..., Schemes_Lagrange.cpp:22-42)
    Vector9d getQuadrature9Weights() const;
    Vector9d getQuadrature9Points() const;

    Vector2d getQuadrature2Weights() const { return gaussianQuadrature2Weights; }
    Vector3d getGaussLegendre3Weights() const { return gaussLegendreQuadrature3Weights; }
    Vector4d getGaussLobatto4Weights() const { return gaussLobattoQuadrature4Weights; }
    Vector4d getGaussLegendre4Weights() const { return gaussLegendreQuadrature4Weights; }

    ~Lagrange() {}

protected:
    Vector2d gaussianQuadrature2Weights;
    Vector3d gaussLegendreQuadrature3Weights;
    Vector4d gaussLobattoQuadrature4Weights;
    Vector4d gaussLegendreQuadrature4Weights;
    Vector9d gaussianQuadrature9Weights, gaussianQuadrature9Points;
};

}  // namespace Schemes
