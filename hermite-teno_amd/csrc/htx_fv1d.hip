// htx_fv1d.hip -- the 1D Euler stage (FV::System1D::rhs + RungeKutta3,
// FV_System1D.cpp:158-292) as one launch per RK stage.  A workgroup owns
// FV1_T consecutive cells: its threads compute the FV1_T + 1 face fluxes
// into LDS (characteristic TENO5 / TENO5i, HLLC or characteristic Rusanov),
// then each cell's residual and RK stage, and in stage 3 the workgroup's
// CFL maximum (one 64-bit atomic max).  Built in the strict (FP64 drop-in)
// build only: the 1D path is small and runs its reference arithmetic.
#include "htx_fv1d.h"

namespace HTX_NS {

constexpr int FV1_T = 256;

__device__ __forceinline__ double wave_max_1d(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ void block_max_to_ctl(double v, Ctl* ctl) {
    __shared__ double red[FV1_T / 64];
    v = wave_max_1d(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < FV1_T / 64; ++k) v = fmax(v, red[k]);
        atomicMax(&ctl->smax_x, (unsigned long long)__double_as_longlong(v));
    }
}

__global__ void __launch_bounds__(FV1_T) k_fv1d_stage(Fv1dArgs a) {
    __shared__ double fl[3][FV1_T + 1];
    double dt = 0.0;
    if (a.mode > 0) {
        if (!a.ctl->active) return;  // uniform over the grid
        dt = a.ctl->dt;
    }
    const int c0 = blockIdx.x * FV1_T;
    for (int f = threadIdx.x; f <= FV1_T; f += FV1_T) {
        const int c = c0 - 1 + f;  // face between cells c and c + 1
        if (c <= a.nx - 1) {
            double F[3];
            fv1d_face_flux(a, c, F);
            for (int e = 0; e < 3; ++e) fl[e][f] = F[e];
        }
    }
    __syncthreads();
    const int c = c0 + threadIdx.x;
    double sm = 0.0;
    bool bad = false;
    if (c < a.nx) {
        double u[3];
        for (int e = 0; e < 3; ++e) {
            u[e] = fv1d_stage_value(a, c, e, fl[e][threadIdx.x], fl[e][threadIdx.x + 1], dt);
            a.qout[e * a.nx + c] = u[e];
        }
        if (a.mode == 3) {
            sm = fv1d_max_speed(a.gam, a.g1, u);
            bad = !all_finite(u, 3);
        }
    }
    if (a.mode == 3) {
        flag_nonfinite(bad, a.ctl);
        block_max_to_ctl(sm, a.ctl);
    }
}

// wave-speed maximum of a state (after a state is set)
__global__ void __launch_bounds__(FV1_T) k_fv1d_max_speed(const double* q, int nx, double gam, Ctl* ctl) {
    const int c = blockIdx.x * FV1_T + threadIdx.x;
    double sm = 0.0;
    bool bad = false;
    if (c < nx) {
        const double u[3] = {q[c], q[nx + c], q[2 * nx + c]};
        sm = fv1d_max_speed(gam, gam - 1.0, u);
        bad = !all_finite(u, 3);
    }
    flag_nonfinite(bad, ctl);
    block_max_to_ctl(sm, ctl);
}

// KXRCF flags of every cell of a state (q = 3 fields x nx)
__global__ void __launch_bounds__(FV1_T) k_fv1d_kxrcf(const double* q, int nx, int bc, double radius,
                                                      unsigned char* flags) {
    const int c = blockIdx.x * FV1_T + threadIdx.x;
    if (c < nx) flags[c] = fv1d_kxrcf(q, q[nx + c], c, nx, bc, radius) ? 1 : 0;
}

// LMPE flags of the three fields: flags[eq nx + c]
__global__ void __launch_bounds__(FV1_T) k_fv1d_lmpe(const double* q, int nx, int bc, double dx,
                                                     unsigned char* flags) {
    const int c = blockIdx.x * FV1_T + threadIdx.x, eq = blockIdx.y;
    if (c < nx) flags[eq * nx + c] = fv1d_lmpe(q + (long long)eq * nx, c, nx, bc, dx) ? 1 : 0;
}

hipError_t launch_fv1d_lmpe(const double* q, int nx, int bc, double dx, unsigned char* flags, hipStream_t s) {
    hipLaunchKernelGGL(k_fv1d_lmpe, dim3((nx + FV1_T - 1) / FV1_T, 3), dim3(FV1_T), 0, s, q, nx, bc, dx, flags);
    return hipGetLastError();
}

hipError_t launch_fv1d_stage(const Fv1dArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_fv1d_stage, dim3((a.nx + FV1_T - 1) / FV1_T), dim3(FV1_T), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fv1d_kxrcf(const double* q, int nx, int bc, double radius, unsigned char* flags, hipStream_t s) {
    hipLaunchKernelGGL(k_fv1d_kxrcf, dim3((nx + FV1_T - 1) / FV1_T), dim3(FV1_T), 0, s, q, nx, bc, radius, flags);
    return hipGetLastError();
}

hipError_t launch_fv1d_max_speed(const double* q, int nx, double gam, Ctl* ctl, hipStream_t s) {
    hipLaunchKernelGGL(k_fv1d_max_speed, dim3((nx + FV1_T - 1) / FV1_T), dim3(FV1_T), 0, s, q, nx, gam, ctl);
    return hipGetLastError();
}

}  // namespace HTX_NS
