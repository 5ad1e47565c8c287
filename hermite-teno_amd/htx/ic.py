"""Seeded workload definitions for the 2D Euler configurations.

The random draws use splitmix64 with 53-bit uniforms, so every consumer (GPU
generators, CPU checker, tests) sees bit-identical parameters.
Seeds are 0x48544E4F00 + k for workload k.
"""
import math

MASK = (1 << 64) - 1
SEED_BASE = 0x48544E4F00


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self):
        """U[0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def uniform_in(self, a, b):
        return a + (b - a) * self.uniform()


def vortex_modes(seed=SEED_BASE + 2, nmodes=8, amplitude=1e-3, length=10.0):
    """Density perturbation of the periodic vortex workload: nmodes rows of
    (k, l, a, phi, Lx, Ly) with integer wavenumbers (k, l) in [-4, 4]^2 \\ {0},
    a = amplitude * U[0.5, 1] / nmodes and phi = 2 pi U[0, 1)."""
    rng = SplitMix64(seed)
    rows = []
    while len(rows) < nmodes:
        k = int(rng.uniform() * 9.0) - 4
        l = int(rng.uniform() * 9.0) - 4
        if k == 0 and l == 0:
            continue
        a = amplitude * rng.uniform_in(0.5, 1.0) / nmodes
        phi = 2.0 * math.pi * rng.uniform()
        rows.append([float(k), float(l), a, phi, length, length])
    return rows


def vortex_workload(n=1024):
    """Periodic isentropic vortex on [0,10]^2 (strength 5, background
    rho = u = v = p = 1, centred at (5, 5)) plus 8 seeded Fourier density
    modes of total amplitude 1e-3; gamma 1.4, CFL 0.5, t_end = 2."""
    return dict(xl=0.0, xr=10.0, nx=n, yb=0.0, yt=10.0, ny=n, cfl=0.5, tmax=2.0, gamma=1.4, bc="periodic",
                centre=(5.0, 5.0), modes=vortex_modes())
