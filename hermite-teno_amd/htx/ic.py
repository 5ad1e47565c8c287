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


def quadrant_params(rng):
    """Split point (xs, ys) ~ U[0.3, 0.7]^2, then per quadrant k = 0..3
    (k = (x >= xs) + 2 (y >= ys)) rho ~ U[0.5, 2], u, v ~ U[-0.5, 0.5],
    p ~ U[0.2, 2], drawn in that order."""
    xs, ys = rng.uniform_in(0.3, 0.7), rng.uniform_in(0.3, 0.7)
    states = [[rng.uniform_in(0.5, 2.0), rng.uniform_in(-0.5, 0.5), rng.uniform_in(-0.5, 0.5),
               rng.uniform_in(0.2, 2.0)] for _ in range(4)]
    return xs, ys, states


def quadrant_workload(n=2048, seed=SEED_BASE + 3):
    """Seeded four-quadrant Riemann problem on [0,1]^2, transmissive
    (zero-gradient) boundaries, gamma 1.4, CFL 0.5, t_end = 0.25."""
    xs, ys, states = quadrant_params(SplitMix64(seed))
    return dict(kind="quadrant", xl=0.0, xr=1.0, nx=n, yb=0.0, yt=1.0, ny=n, cfl=0.5, tmax=0.25, gamma=1.4,
                bc="zero_grad", xs=xs, ys=ys, states=states)


def shock_blob_workload(n=4096, seed=SEED_BASE + 4, nblob=32, gamma=1.4):
    """Planar shock of Mach M ~ U[2, 10] at angle theta ~ U[-30, 30] degrees,
    through (0.5, 0.5) with normal (cos theta, sin theta), running into gas
    at rest (rho = 1.4, p = 1, so c = 1) that carries nblob Gaussian density
    blobs: for each blob y_b ~ U[0.05, 0.95], x_b ~ U[x_shock(y_b) + 0.2, 3.9],
    amplitude ~ U[0.1, 0.5], radius ~ U[0.02, 0.1], drawn in that order.  The
    post-shock state follows from the Rankine-Hugoniot relations.  Domain
    [0,4] x [0,1], n x n cells, transmissive boundaries, CFL 0.5,
    t_end = 1.5 / M."""
    rng = SplitMix64(seed)
    mach = rng.uniform_in(2.0, 10.0)
    theta = rng.uniform_in(-30.0, 30.0) * math.pi / 180.0
    rho1, p1, c1 = 1.4, 1.0, 1.0
    m2 = mach * mach
    rho2 = rho1 * (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    p2 = p1 * (2.0 * gamma * m2 - (gamma - 1.0)) / (gamma + 1.0)
    u2 = 2.0 / (gamma + 1.0) * (mach - 1.0 / mach) * c1
    nrm = (math.cos(theta), math.sin(theta))
    blobs = []
    for _ in range(nblob):
        yb = rng.uniform_in(0.05, 0.95)
        xsh = 0.5 - (yb - 0.5) * math.tan(theta)
        xb = rng.uniform_in(xsh + 0.2, 3.9)
        amp = rng.uniform_in(0.1, 0.5)
        rad = rng.uniform_in(0.02, 0.1)
        blobs.append([xb, yb, amp, rad])
    return dict(kind="shock_blobs", xl=0.0, xr=4.0, nx=n, yb=0.0, yt=1.0, ny=n, cfl=0.5, tmax=1.5 / mach,
                gamma=gamma, bc="zero_grad", mach=mach, theta=theta, xs=0.5, ys=0.5, nrm=nrm,
                post=[rho2, u2 * nrm[0], u2 * nrm[1], p2], ambient=[rho1, 0.0, 0.0, p1], blobs=blobs)


def _piecewise_prim(w, x, y):
    """Point values (rho, u, v, p) of a piecewise workload at arrays x, y."""
    import numpy as np

    if w["kind"] == "quadrant":
        k = (x >= w["xs"]).astype(int) + 2 * (y >= w["ys"]).astype(int)
        st = np.array(w["states"])
        return [st[k, e] for e in range(4)]
    s = (x - w["xs"]) * w["nrm"][0] + (y - w["ys"]) * w["nrm"][1]
    r = np.full(np.broadcast(x, y).shape, w["ambient"][0])
    for xb, yb, amp, rad in w["blobs"]:
        dx, dy = x - xb, y - yb
        r = r + amp * np.exp(-(dx * dx + dy * dy) / (rad * rad))
    post = s < 0.0
    out = []
    for e in range(4):
        amb = r if e == 0 else np.full(r.shape, w["ambient"][e])
        out.append(np.where(post, w["post"][e], amb))
    return out


def _cons(prim, g1):
    r, u, v, p = prim
    q1, q2 = r * u, r * v
    return [r, q1, q2, p / g1 + 0.5 * (q1 * u + q2 * v)]


def piecewise_state(w, nx, ny, hermite=False):
    """Host restatement of the device generator (htx_ic.hip): conservative
    point values at cell centres, field-major; with hermite=True the 8
    derivative moments (face differences / dx, / dy) follow."""
    import numpy as np

    dx = (w["xr"] - w["xl"]) / nx
    dy = (w["yt"] - w["yb"]) / ny
    x = (np.arange(nx, dtype=np.float64) + 0.5) * dx + w["xl"]
    y = (np.arange(ny, dtype=np.float64) + 0.5) * dy + w["yb"]
    X, Y = np.meshgrid(x, y)
    g1 = w["gamma"] - 1.0
    q = _cons(_piecewise_prim(w, X, Y), g1)
    if hermite:
        qp, qm = _cons(_piecewise_prim(w, X + dx / 2.0, Y), g1), _cons(_piecewise_prim(w, X - dx / 2.0, Y), g1)
        q += [(qp[e] - qm[e]) / dx for e in range(4)]
        qp, qm = _cons(_piecewise_prim(w, X, Y + dy / 2.0), g1), _cons(_piecewise_prim(w, X, Y - dy / 2.0), g1)
        q += [(qp[e] - qm[e]) / dy for e in range(4)]
    return np.stack(q)


def batch_workload(count=256, seed=SEED_BASE + 5, lo=64, hi=1024, tend=0.2):
    """Serving-style sweep: `count` instances, each with nx, ny =
    round(exp(U[ln lo, ln hi])) drawn independently, the domain
    [0, nx/m] x [0, ny/m] with m = max(nx, ny) (square cells), a seeded
    four-quadrant Riemann problem (quadrant_params) and a scheme drawn with
    probability 1/2 each from TENO5 + HLLC and Hermite-TENO + Rusanov;
    transmissive boundaries, gamma 1.4, CFL 0.5, t_end = tend.  Draw order per
    instance: nx, ny, scheme, then the quadrant parameters."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        nx = int(round(math.exp(rng.uniform_in(math.log(lo), math.log(hi)))))
        ny = int(round(math.exp(rng.uniform_in(math.log(lo), math.log(hi)))))
        scheme = "teno5" if rng.uniform() < 0.5 else "hteno"
        xs, ys, states = quadrant_params(rng)
        m = max(nx, ny)
        # the quadrant split point is drawn on the unit square; scale it into the domain
        out.append(dict(kind="quadrant", scheme=scheme, flux="hllc" if scheme == "teno5" else "rusanov",
                        xl=0.0, xr=nx / m, nx=nx, yb=0.0, yt=ny / m, ny=ny, cfl=0.5, tmax=tend, gamma=1.4,
                        bc="zero_grad", xs=xs * nx / m, ys=ys * ny / m, states=states))
    return out
