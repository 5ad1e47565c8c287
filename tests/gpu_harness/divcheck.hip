// divcheck.hip -- test harness (GPU): the strict build's shared-reciprocal
// quotients (quot, quot_const in htx_numerics.h) against the hardware's
// correctly rounded division, bit for bit, on random operands over the
// ranges where the kernels use them.  Test infrastructure only.
#define HTX_NS htx_divcheck
#define HTX_EXACT_DIVISION 1
#include "../../hermite-teno_amd/csrc/htx_numerics.h"

namespace {

__device__ inline unsigned long long splitmix(unsigned long long& s) {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// a double with a random 52-bit mantissa (every eighth one all ones or all
// zeros) and an exponent uniform in [emin, emax]
__device__ inline double rand_double(unsigned long long& s, int emin, int emax, bool neg_ok) {
    unsigned long long r = splitmix(s);
    unsigned long long m = r & ((1ull << 52) - 1);
    const int sel = (int)((r >> 52) & 7);
    if (sel == 0) m = (1ull << 52) - 1;
    if (sel == 1) m = 0;
    unsigned long long r2 = splitmix(s);
    int e = emin + (int)(r2 % (unsigned long long)(emax - emin + 1));
    unsigned long long bits = ((unsigned long long)(e + 1023) << 52) | m;
    if (neg_ok && ((r2 >> 40) & 1)) bits |= 1ull << 63;
    return __longlong_as_double((long long)bits);
}

__global__ void k_divcheck(unsigned long long seed, long long n, unsigned long long* bad, double* ex) {
    const double consts[5] = {5040.0, 6.0, 24.0, 6720.0, 1680.0};
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        unsigned long long s = seed ^ (0x5851F42D4C957F2Dull * (unsigned long long)(i + 1));
        // shared reciprocal: b in [2^-300, 2^300], x = +-0 or |x| in [2^-400, 2^400]
        const double b = rand_double(s, -300, 299, false);
        const htx_divcheck::DivBy D = htx_divcheck::div_by(b);
        for (int k = 0; k < 4; ++k) {
            double x = rand_double(s, -400, 399, true);
            if (k == 3) x = (i & 1) ? -0.0 : 0.0;
            const double q = htx_divcheck::quot(D, x), r = x / b;
            if (__double_as_longlong(q) != __double_as_longlong(r)) {
                if (atomicAdd(bad, 1ull) == 0ull) {
                    ex[0] = x;
                    ex[1] = b;
                }
            }
        }
        // constant divisors, numerators over the whole finite range and zero
        const double c = consts[i % 5];
        double x = rand_double(s, -1020, 1020, true);
        if ((i & 63) == 0) x = (i & 64) ? -0.0 : 0.0;
        const double q = htx_divcheck::quot_const(x, c, 1.0 / c), r = x / c;
        if (__double_as_longlong(q) != __double_as_longlong(r)) {
            if (atomicAdd(bad + 1, 1ull) == 0ull) {
                ex[2] = x;
                ex[3] = c;
            }
        }
    }
}

}  // namespace

// Runs n random cases from seed; bad[0] = mismatches of the shared
// reciprocal (4 numerators per case), bad[1] = of the constant divisors;
// ex = the first mismatching (x, b) of each.  Returns 0 or a HIP error.
extern "C" int divcheck_run(unsigned long long seed, long long n, unsigned long long* bad_host, double* ex_host) {
    unsigned long long* bad = nullptr;
    double* ex = nullptr;
    hipError_t e = hipMalloc(&bad, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&ex, 4 * sizeof(double));
    if (e == hipSuccess) e = hipMemset(bad, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(ex, 0, 4 * sizeof(double));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_divcheck, dim3(4096), dim3(256), 0, 0, seed, n, bad, ex);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(bad_host, bad, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(ex_host, ex, 4 * sizeof(double), hipMemcpyDeviceToHost);
    if (bad) (void)hipFree(bad);
    if (ex) (void)hipFree(ex);
    return e == hipSuccess ? 0 : (int)e;
}
