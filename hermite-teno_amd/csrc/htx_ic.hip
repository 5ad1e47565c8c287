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

// Euler2D::entropyWaves / entropyWavesdx / entropyWavesdy
// (Equations_Euler2D.cpp:430-511): the conserved state of the wave and its
// derivative (equal in x and y: the density depends on x + y), averaged over
// the cell with the 9x9 Gauss rule in the reference's order (i outer, j
// inner, each term (w_i w_j) f / 4).
__global__ void __launch_bounds__(256) k_ic_entropy(EntropyArgs a) {
    long long n = (long long)a.nx * a.ny;
    long long id = blockIdx.x * 256ll + threadIdx.x;
    if (id >= n) return;
    int i = (int)(id % a.nx), j = (int)(id / a.nx);
    const double pi = 4.0 * atan(1.0), g1 = a.gam - 1.0;
    double x = ((double)i + 0.5) * a.dx + a.xl;
    double y = ((double)j + 0.5) * a.dy + a.yb;
    double acc[4] = {0.0, 0.0, 0.0, 0.0}, dacc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int ii = 0; ii < 9; ++ii)
        for (int jj = 0; jj < 9; ++jj) {
            const double px = x + a.dx * a.gx[ii] / 2.0, py = y + a.dy * a.gx[jj] / 2.0;
            const double ww = a.gw[ii] * a.gw[jj];
            double q[4];
            q[0] = a.rho + a.amp * sin(pi * (px + py));
            q[1] = q[0] * a.u;
            q[2] = q[0] * a.v;
            q[3] = a.p / g1 + 0.5 * (q[1] * a.u + q[2] * a.v);
            for (int e = 0; e < 4; ++e) acc[e] += ww * q[e] / 4.0;
            if (a.nfield == 12) {
                double d[4];
                d[0] = a.amp * cos(pi * (px + py)) * pi;
                d[1] = a.u * d[0];
                d[2] = a.v * d[0];
                d[3] = (0.5 * (a.u * a.u + a.v * a.v)) * d[0];
                for (int e = 0; e < 4; ++e) dacc[e] += ww * d[e] / 4.0;
            }
        }
    long long off = (long long)j * a.pitch + i;
    for (int e = 0; e < 4; ++e) a.q[off + e * a.fstride] = acc[e];
    if (a.nfield == 12)
        for (int e = 0; e < 4; ++e) {
            a.q[off + (4 + e) * a.fstride] = dacc[e];
            a.q[off + (8 + e) * a.fstride] = dacc[e];
        }
}

// the exact density of the condition at a point
__device__ double exact_density(const ExactArgs& a, double px, double py, double pi) {
    if (a.kind == EXACT_ENTROPY) return a.rho + a.amp * sin(pi * (px + py));
    // conservativeOfIsentropicVortex's density (Equations_Euler2D.cpp:656-666),
    // centred at (cx, cy), plus the density modes
    const double g1 = a.gam - 1.0;
    const double xr = px - a.cx, yr = py - a.cy, b = 5.0;
    const double r = xr * xr + yr * yr;
    double w0 = 1.0 - g1 * (b * b) / 8.0 / a.gam / (pi * pi) * exp(1 - r);
    w0 = pow(w0, 1.0 / g1);
    for (int m = 0; m < a.nmodes; ++m)
        w0 += a.modes[m][2] *
              sin(2.0 * pi * (a.modes[m][0] * px / a.modes[m][4] + a.modes[m][1] * py / a.modes[m][5]) + a.modes[m][3]);
    return w0;
}

// per cell |rho - rho_exact| with rho_exact the 9x9 Gauss average at the
// shifted cell centre (the reference shifts the centre, then spreads the
// Gauss points), summed per workgroup: wavefront shuffles, then the four
// wavefronts in order
__global__ void __launch_bounds__(256) k_exact_l1(ExactArgs a) {
    __shared__ double red[4];
    long long n = (long long)a.nx * a.ny;
    long long id = blockIdx.x * 256ll + threadIdx.x;
    double v = 0.0;
    if (id < n) {
        int i = (int)(id % a.nx), j = (int)(id / a.nx);
        const double pi = 4.0 * atan(1.0);
        const double x = (((double)i + 0.5) * a.dx + a.xl) - a.sx;
        const double y = (((double)j + 0.5) * a.dy + a.yb) - a.sy;
        double e = 0.0;
        for (int ii = 0; ii < 9; ++ii)
            for (int jj = 0; jj < 9; ++jj)
                e += a.gw[ii] * a.gw[jj] *
                     exact_density(a, x + a.dx * a.gx[ii] / 2.0, y + a.dy * a.gx[jj] / 2.0, pi) / 4.0;
        v = fabs(a.q[(long long)j * a.pitch + i] - e);
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) a.partial[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// one workgroup: the partial sums in a fixed order (strided per thread,
// then a fixed tree)
__global__ void __launch_bounds__(256) k_sum_partials(const double* partial, int n, double* out) {
    __shared__ double red[256];
    double v = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) v += partial[k];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace

hipError_t launch_ic_entropy(const EntropyArgs& a, hipStream_t s) {
    long long n = (long long)a.nx * a.ny;
    hipLaunchKernelGGL(k_ic_entropy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// partial must hold exact_l1_blocks(nx, ny) doubles
int exact_l1_blocks(int nx, int ny) { return (int)(((long long)nx * ny + 255) / 256); }

hipError_t launch_exact_l1(const ExactArgs& a, hipStream_t s) {
    const int nb = exact_l1_blocks(a.nx, a.ny);
    hipLaunchKernelGGL(k_exact_l1, dim3((unsigned)nb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, a.partial, nb, a.out);
    return hipGetLastError();
}

hipError_t launch_ic_piecewise(const PiecewiseArgs& a, hipStream_t s) {
    long long n = (long long)a.nx * a.ny;
    hipLaunchKernelGGL(k_ic_piecewise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace htx
