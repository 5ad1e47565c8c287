// Schemes_SSPRK.h -- drop-in for the reference's Schemes::SSPRK
// (Schemes_SSPRK.h:8-9, Schemes_SSPRK.cpp:3-15): the Shu-Osher SSP-RK3 stage
// combination.  Header-only; T is double or a fixed-size vector of
// htx_linalg.h (the reference instantiates double, Vector3d and Vector4d).
// The GPU stage kernels evaluate the same expressions in the same order.
#pragma once
#include "htx_linalg.h"

namespace Schemes {

class SSPRK {
public:
    // stage 1: q0 + dt rhs; stage 2: 3/4 q0 + 1/4 (qn + rhs dt);
    // stage 3: (q0 + 2 (qn + dt rhs)) / 3 (a division by 3, as written)
    template <typename T>
    T RungeKuttaIII(const T& q0, const T& qn, const T& rhs, const double& dt, const int& stage) const {
        if (stage == 1) return q0 + dt * rhs;
        if (stage == 2) return 0.75 * q0 + 0.25 * (qn + rhs * dt);
        return (q0 + 2.0 * (qn + dt * rhs)) / 3.0;
    }
};

}  // namespace Schemes
