// FVHermite_Scalar1D.h -- declaration of the reference's FVHermite::Scalar1D (FVHermite_Scalar1D.h:10-50), the
// scalar (linear advection / Burgers) test solver.  Outside the rebuilt path:
// declared so that a driver written against the reference's headers (its
// main.cpp includes every solver header) compiles against these; no
// definitions are built, so a program that constructs one does not link.
#pragma once

namespace FVHermite {

class Scalar1D {
public:
    Scalar1D(const double& xl, const double& xr, const int& nx, const double& cflnumber, const double& tmax,
             const double& tInit);
    void initialize();
    void periodicBc();
    void limitDerivative();
    void rhs();
    void RungeKutta3();
    void timeStepping();
    void Exact();
    void printResult();
};

}  // namespace FVHermite
