// FVHermite_System1D.h -- declaration of the reference's FVHermite::System1D
// (FVHermite_System1D.h:9-24), the 1D Hermite solver with its LMPE/KXRCF
// hybrid.  Outside the rebuilt path: declared so that a driver written
// against the reference's headers (its main.cpp includes every solver header
// and, as shipped, runs this solver's entropy-wave case) compiles against
// these; no definitions are built, so a program that constructs one does not
// link.
#pragma once

namespace FVHermite {

class System1D {
public:
    System1D(const double& xl, const double& xr, const int& nx, const double& cflnumber, const double& tmax,
             const double& tInit, const double& specificHeatRatio);
    void entropyWave();
    void blastWave();
    void timeStepping();
    void printResult();

    void calcPrimitiveVariable();
    void periodicBc();
    void zeroGradBc();
    void reflectiveBc();
    void eigen1DEuler();
    void rhs();
    void RungeKutta3();
};

}  // namespace FVHermite
