// htx_ic.hip -- device generators of the piecewise initial states: the seeded
// four-quadrant Riemann problem and the planar shock running into a gas that
// carries Gaussian density blobs.  Both are point-sampled at cell centres, as
// the reference does for discontinuous data (FV_System2D.cpp:104-108).  For
// the Hermite scheme the derivative moments are the face differences of the
// point data, (f(x + dx/2) - f(x - dx/2)) / dx, as FVHermite::System2D::
// Riemann2D builds them (FVHermite_System2D.cpp:89-107).
//
// Compiled without FMA contraction so that the host generator (htx/ic.py)
// can restate every value with the same IEEE operations.
#include "htx_internal.h"

namespace htx {

namespace {

struct PwPoint {
    double r, u, v, p;
};

__device__ PwPoint pw_eval(const PiecewiseArgs& a, double x, double y) {
    PwPoint w;
    if (a.kind == IC_QUADRANT) {
        int k = (x >= a.xs ? 1 : 0) + (y >= a.ys ? 2 : 0);
        w.r = a.st[k][0];
        w.u = a.st[k][1];
        w.v = a.st[k][2];
        w.p = a.st[k][3];
    } else {
        double s = (x - a.xs) * a.nrm[0] + (y - a.ys) * a.nrm[1];
        if (s < 0.0) {
            w.r = a.st[0][0];
            w.u = a.st[0][1];
            w.v = a.st[0][2];
            w.p = a.st[0][3];
        } else {
            double r = a.st[1][0];
            for (int b = 0; b < a.nblob; ++b) {
                double dx = x - a.blob[b][0], dy = y - a.blob[b][1];
                r = r + a.blob[b][2] * exp(-(dx * dx + dy * dy) / (a.blob[b][3] * a.blob[b][3]));
            }
            w.r = r;
            w.u = a.st[1][1];
            w.v = a.st[1][2];
            w.p = a.st[1][3];
        }
    }
    return w;
}

// Euler2D::conservative (Equations_Euler2D.cpp:28-34)
__device__ void pw_cons(const PwPoint& w, double g1, double q[4]) {
    q[0] = w.r;
    q[1] = w.r * w.u;
    q[2] = w.r * w.v;
    q[3] = w.p / g1 + 0.5 * (q[1] * w.u + q[2] * w.v);
}

__global__ void __launch_bounds__(256) k_ic_piecewise(PiecewiseArgs a) {
    long long n = (long long)a.nx * a.ny;
    long long id = blockIdx.x * 256ll + threadIdx.x;
    if (id >= n) return;
    int i = (int)(id % a.nx), j = (int)(id / a.nx);
    const double g1 = a.gam - 1.0;
    double x = ((double)i + 0.5) * a.dx + a.xl;
    double y = ((double)j + 0.5) * a.dy + a.yb;
    long long off = (long long)j * a.pitch + i;
    double q[4];
    pw_cons(pw_eval(a, x, y), g1, q);
    for (int e = 0; e < 4; ++e) a.q[off + e * a.fstride] = q[e];
    if (a.nfield == 12) {
        double qp[4], qm[4];
        pw_cons(pw_eval(a, x + a.dx / 2.0, y), g1, qp);
        pw_cons(pw_eval(a, x - a.dx / 2.0, y), g1, qm);
        for (int e = 0; e < 4; ++e) a.q[off + (4 + e) * a.fstride] = (qp[e] - qm[e]) / a.dx;
        pw_cons(pw_eval(a, x, y + a.dy / 2.0), g1, qp);
        pw_cons(pw_eval(a, x, y - a.dy / 2.0), g1, qm);
        for (int e = 0; e < 4; ++e) a.q[off + (8 + e) * a.fstride] = (qp[e] - qm[e]) / a.dy;
    }
}

}  // namespace

hipError_t launch_ic_piecewise(const PiecewiseArgs& a, hipStream_t s) {
    long long n = (long long)a.nx * a.ny;
    hipLaunchKernelGGL(k_ic_piecewise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace htx
