"""Pins the Hermite-TENO (FVHermite::System2D) and 1D (FV::System1D) parts of
the FP64 CPU oracle, and the smoothness indicators shared with TENO5.

The three big indicator polynomials of the reference (TENO5 beta5,
Schemes_Lagrange.cpp:1029-1034; HTENO brk and brkd, Schemes_Hermite.cpp
:597-612) are pinned to their definition as smoothness integrals of the
full-stencil polynomial, computed here by exact polynomial algebra:
    beta = sum_l  int_{centre cell} (d^l p / dx^l)^2 dx   (h = 1)
"""
import math

import numpy as np
import pytest
from numpy.polynomial import polynomial as P

from conftest import rel_l1
from oracle import params

GAMMA = 1.4


def cavg(k, c):
    """Average of x^k over the unit cell centred at c."""
    return ((c + 0.5) ** (k + 1) - (c - 0.5) ** (k + 1)) / (k + 1)


def davg(k, c):
    """Average of (x^k)' over the unit cell centred at c."""
    return 0.0 if k == 0 else (c + 0.5) ** k - (c - 0.5) ** k


def smoothness_integral(coef, l0=1, lo=-0.5, hi=0.5):
    b = 0.0
    for l in range(l0, len(coef)):
        d = P.polyder(coef, l)
        integ = P.polyint(P.polymul(d, d))
        b += P.polyval(hi, integ) - P.polyval(lo, integ)
    return b


# --------------------------------------------------------- indicators ----
def test_teno5_indicators_are_smoothness_integrals(orc):
    """beta_k = integral indicator of the quadratic candidate k; beta5 = that of
    the quartic through all 5 averages (both evaluated on the centre cell)."""
    rng = np.random.default_rng(2)
    for _ in range(10):
        u = rng.normal(size=6)
        SI, b35 = orc.teno5_indicators(u)
        for side in range(2):
            v = u[side:side + 5]
            A = np.array([[cavg(k, c) for k in range(5)] for c in range(-2, 3)])
            coef = np.linalg.solve(A, v)
            assert b35[side] == pytest.approx(smoothness_integral(coef), rel=1e-11)
            # candidate stencils {0,1,2}, {-1,0,1}, {-2,-1,0} relative to the centre cell (index 2 of v)
            for k, cells in enumerate([(0, 1, 2), (-1, 0, 1), (-2, -1, 0)]):
                Ak = np.array([[cavg(m, c) for m in range(3)] for c in cells])
                ck = np.linalg.solve(Ak, v[[c + 2 for c in cells]])
                assert SI[3 * side + k] == pytest.approx(smoothness_integral(ck), rel=1e-11)


def test_hermite_indicators_are_smoothness_integrals(orc):
    """brk: quartic from the averages of cells -1, 0, 1 and derivative averages
    of cells -1, 1; brkd: derivative indicator (l >= 2) of the quintic that also
    uses the centre derivative average.  Schemes_Hermite.cpp:597-612."""
    rng = np.random.default_rng(3)
    for _ in range(10):
        u, ud = rng.normal(size=3), rng.normal(size=3)
        SI, SId, brk = orc.hteno_indicators(u, ud)
        A = np.array([[cavg(k, c) for k in range(5)] for c in (-1, 0, 1)] +
                     [[davg(k, c) for k in range(5)] for c in (-1, 1)])
        coef = np.linalg.solve(A, np.r_[u, ud[0], ud[2]])
        assert brk[0] == pytest.approx(smoothness_integral(coef), rel=1e-11)
        A5 = np.array([[cavg(k, c) for k in range(6)] for c in (-1, 0, 1)] +
                      [[davg(k, c) for k in range(6)] for c in (-1, 0, 1)])
        coef5 = np.linalg.solve(A5, np.r_[u, ud])
        assert brk[1] == pytest.approx(smoothness_integral(coef5, l0=2), rel=1e-11)


# ------------------------------------------------------------ hteno ----
@pytest.mark.parametrize("deg", [0, 1, 2])
def test_hteno_exact_for_quadratics(orc, deg):
    """Every Hermite candidate is exact for quadratics, so the face values and
    face derivatives are exact whatever the weights (Schemes_Hermite.cpp:295-379)."""
    rng = np.random.default_rng(10 + deg)
    coef = rng.uniform(-1, 1, deg + 1)
    h = 0.2
    # cells i-1..i+2 with cell i = [-h, 0]; face at x = 0
    u = [cell_avg(coef, (m - 2) * h, (m - 1) * h) for m in range(4)]
    d = [cell_avg(P.polyder(coef) if deg else [0.0], (m - 2) * h, (m - 1) * h) for m in range(4)]
    out = orc.hteno(h, u, d)
    dp = P.polyder(coef) if deg else [0.0]
    np.testing.assert_allclose(out, [P.polyval(0, coef), P.polyval(0, coef), P.polyval(0, dp), P.polyval(0, dp)],
                               rtol=0, atol=1e-12)


def cell_avg(coef, a, b):
    integ = P.polyint(coef)
    return (P.polyval(b, integ) - P.polyval(a, integ)) / (b - a)


def test_hteno_high_order_on_smooth_data(orc):
    """On smooth data the face value converges at >= 4th order."""
    errs = []
    for h in (0.1, 0.05):
        u = [(math.cos((m - 2) * h) - math.cos((m - 1) * h)) / h for m in range(4)]   # avg of sin
        d = [(math.sin((m - 1) * h) - math.sin((m - 2) * h)) / h for m in range(4)]   # avg of cos
        out = orc.hteno(h, u, d)
        errs.append(max(abs(out[0]), abs(out[1])))  # sin(0) = 0
    assert errs[0] / errs[1] > 15.0


def test_hteno_mirror_symmetry(orc):
    """Reversing the stencil (and negating derivatives) swaps the two sides."""
    rng = np.random.default_rng(4)
    for _ in range(20):
        u, d = rng.normal(size=4), rng.normal(size=4)
        a = orc.hteno(0.3, u, d)
        b = orc.hteno(0.3, u[::-1].copy(), -d[::-1].copy())
        assert a[1] == pytest.approx(b[0], rel=1e-12, abs=1e-13)
        assert a[3] == pytest.approx(-b[2], rel=1e-12, abs=1e-12)


# ---------------------------------------------------------- limiter ----
def test_hermite_limiter_linear_and_cubic(orc):
    """Linear data: the limiter returns the slope whatever its weights.  Cubic
    data: the central candidate 0.75 (u2-u0)/h - 0.25 (d0+d2) is the exact
    centre-cell derivative average, and smooth weights select it."""
    h = 0.01
    for s in (0.7, -2.0):
        u = [s * (m - 1) * h for m in range(3)]
        assert orc.hermite_limiter(h, u, [s, s, s]) == pytest.approx(s, rel=1e-12)
    coef = [0.1, 0.2, -0.3, 0.5]
    u = [cell_avg(coef, (m - 1.5) * h, (m - 0.5) * h) for m in range(3)]
    d = [cell_avg(P.polyder(coef), (m - 1.5) * h, (m - 0.5) * h) for m in range(3)]
    assert orc.hermite_limiter(h, u, d) == pytest.approx(d[1], rel=1e-6)


def test_hermite_limiter_at_a_jump_is_one_sided(orc):
    """Across a jump the result stays within the one-sided differences."""
    h = 0.1
    u = [0.0, 0.0, 1.0]
    v = orc.hermite_limiter(h, u, [0.0, 0.0, 0.0])
    assert 0.0 <= v <= 1.0 / h + 1e-12


# ------------------------------------------------------------ dflux ----
@pytest.mark.parametrize("direction", [1, 2])
def test_dflux_is_flux_jacobian(orc, direction):
    rng = np.random.default_rng(direction + 5)
    for _ in range(5):
        w = np.array([rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.3, 2)])
        q = orc.conservative(GAMMA, w)
        M = orc.dflux(GAMMA, direction, q, w)
        eps = 1e-6
        J = np.zeros((4, 4))
        for k in range(4):
            dq = np.zeros(4)
            dq[k] = eps
            fp = flux(orc, direction, q + dq)
            fm = flux(orc, direction, q - dq)
            J[:, k] = (fp - fm) / (2 * eps)
        np.testing.assert_allclose(M, J, rtol=1e-7, atol=1e-7)


def flux(orc, direction, q):
    w = orc.primitive(GAMMA, q)
    if direction == 1:
        return np.array([q[1], q[1] * w[1] + w[3], q[1] * w[2], w[1] * (q[3] + w[3])])
    return np.array([q[2], q[2] * w[1], q[2] * w[2] + w[3], w[2] * (q[3] + w[3])])


# --------------------------------------------------- Hermite 2D solver ----
def hermite_vortex(orc, n, bc="periodic", flux="rusanov", modes=None):
    prm = params(n, n, 0.0, 10.0, 0.0, 10.0, GAMMA, bc, flux, True)
    return prm, orc.ic_vortex_hermite(prm, (5.0, 5.0), modes)


def test_hermite_ic_moments(orc):
    """u equals the FV vortex bit for bit; w (y moments) matches 4th-order
    differences of the cell averages; the x moments carry the reference's /pi
    for /pi^2 in the density derivative (Equations_Euler2D.cpp:678), so inside
    the vortex core they do NOT match the same differences."""
    from htx.ic import vortex_modes

    n = 80
    prm, q = hermite_vortex(orc, n, modes=vortex_modes())
    np.testing.assert_array_equal(q[:4], orc.ic_vortex(prm, (5.0, 5.0), vortex_modes()))
    dx = 10.0 / n

    def d4(a, ax):
        return (np.roll(a, 2, ax) - 8 * np.roll(a, 1, ax) + 8 * np.roll(a, -1, ax) - np.roll(a, -2, ax)) / (12 * dx)

    assert rel_l1(q[8:12], d4(q[:4], 1)) < 1e-3
    x = (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(x, x)
    core = (X - 5) ** 2 + (Y - 5) ** 2 < 4
    assert rel_l1(q[8][core], d4(q[:4], 1)[0][core]) < 1e-3
    assert rel_l1(q[4][core], d4(q[:4], 2)[0][core]) > 0.1  # the quirk


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_hermite_free_stream(orc, flux):
    prm = params(12, 12, 0, 1, 0, 1, GAMMA, "zero_grad", flux, True)
    q = np.zeros((12, 12, 12))
    q[:4] = orc.conservative(GAMMA, [1.2, 0.3, -0.4, 0.8])[:, None, None]
    assert np.abs(orc.fvh2d_rhs(prm, q)).max() < 1e-12


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_hermite_periodic_conservation(orc, flux):
    """All 12 fields are in conservation form: periodic sums of the residual vanish."""
    prm, q = hermite_vortex(orc, 20, "periodic", flux)
    r = orc.fvh2d_rhs(prm, q)
    for g in range(12):
        assert abs(r[g].sum()) <= 1e-10 * max(np.abs(r[g]).sum(), 1e-30)


def test_hermite_transpose_symmetry(orc):
    """x <-> y with momenta swapped and v <-> w swapped transposes the residual
    (square cells; the reference uses dx for both directions)."""
    prm, q = hermite_vortex(orc, 16, "zero_grad", "rusanov")

    def T(a):
        u, v, w = a[:4], a[4:8], a[8:]
        sw = lambda b: np.stack([b[0].T, b[2].T, b[1].T, b[3].T])  # noqa: E731
        return np.concatenate([sw(u), sw(w), sw(v)])

    r = orc.fvh2d_rhs(prm, q)
    rt = orc.fvh2d_rhs(prm, T(q))
    # the y projections sum their terms in a different order than the x ones,
    # and on this coarse grid TENO cut-offs in near-uniform fields follow that
    # rounding; an index or v/w mix-up would show as O(1) differences
    assert rel_l1(T(rt), r) < 1e-8


def test_hermite_vortex_accuracy(orc):
    """Short advection of the (unperturbed) vortex: the Hermite solver stays
    close to the translated initial condition (sanity, not parity)."""
    prm, q = hermite_vortex(orc, 24, "periodic", "rusanov")
    q1, t, n = orc.fvh2d_advance(prm, q, 0.4, 0.0, 0.2)
    assert n > 0 and abs(t - 0.2) < 1e-14
    prm2 = params(24, 24, 0.0, 10.0, 0.0, 10.0, GAMMA, "periodic", "rusanov", True)
    ex = orc.ic_vortex(prm2, (5.2, 5.2), None)
    assert rel_l1(q1[0], ex[0]) < 5e-3


# -------------------------------------------------------------- 1D -----
def sod_exact(x, t, gamma=GAMMA):
    """Exact Sod solution (Toro ch. 4) via Newton on the pressure function."""
    rl, ul, pl, rr, ur, pr = 1.0, 0.0, 1.0, 0.125, 0.0, 0.1
    cl, cr = math.sqrt(gamma * pl / rl), math.sqrt(gamma * pr / rr)

    def f(p, rk, pk, ck):
        if p > pk:
            A, B = 2 / ((gamma + 1) * rk), (gamma - 1) / (gamma + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B))
        return 2 * ck / (gamma - 1) * ((p / pk) ** ((gamma - 1) / (2 * gamma)) - 1)

    p = 0.5 * (pl + pr)
    for _ in range(100):
        g = f(p, rl, pl, cl) + f(p, rr, pr, cr) + ur - ul
        e = 1e-8
        dg = (f(p + e, rl, pl, cl) + f(p + e, rr, pr, cr) - f(p - e, rl, pl, cl) - f(p - e, rr, pr, cr)) / (2 * e)
        p -= g / dg
    us = 0.5 * (ul + ur) + 0.5 * (f(p, rr, pr, cr) - f(p, rl, pl, cl))
    rsl = rl * (p / pl) ** (1 / gamma)
    rsr = rr * ((p / pr + (gamma - 1) / (gamma + 1)) / ((gamma - 1) / (gamma + 1) * p / pr + 1))
    sh = ur + cr * math.sqrt((gamma + 1) / (2 * gamma) * p / pr + (gamma - 1) / (2 * gamma))
    csl = cl * (p / pl) ** ((gamma - 1) / (2 * gamma))
    out = np.empty_like(x)
    for i, xi in enumerate(x):
        s = (xi - 0.5) / t
        if s < ul - cl:
            out[i] = rl
        elif s < us - csl:
            out[i] = rl * (2 / (gamma + 1) + (gamma - 1) / ((gamma + 1) * cl) * (ul - s)) ** (2 / (gamma - 1))
        elif s < us:
            out[i] = rsl
        elif s < sh:
            out[i] = rsr
        else:
            out[i] = rr
    return out


@pytest.mark.parametrize("recon,flux", [("teno5", "hllc"), ("teno5i", "rusanov")])
def test_sod_config1(orc, recon, flux):
    """Config 1: N = 400, CFL 0.5, t = 0.2 (teno5 + HLLC), plus the shipped
    teno5i + Rusanov variant.  Sanity against Toro's exact solution; mass is
    conserved exactly while the waves stay inside the domain."""
    n = 400
    q = orc.ic_sod(n)
    q1, t, steps = orc.fv1d_advance(n, 0.0, 1.0, GAMMA, q, 0.5, 0.0, 0.2, recon=recon, flux=flux)
    assert t == pytest.approx(0.2, abs=1e-15)
    x = (np.arange(n) + 0.5) / n
    err = np.abs(q1[0] - sod_exact(x, 0.2)).mean()
    assert err < 4e-3
    assert abs(q1[0].sum() - q[0].sum()) < 1e-12 * q[0].sum()
    assert 150 < steps < 400


def test_teno5i_tau_is_max_beta(orc):
    """teno5i (Schemes_Lagrange.cpp:1235-1263) takes tau = max beta: on smooth
    data it keeps every stencil and equals the linear scheme to rounding."""
    h = 0.02
    u = [(math.cos((m - 2) * h) - math.cos((m - 3) * h)) / h for m in range(6)]
    np.testing.assert_allclose(orc.teno5i(u), orc.linear5(u), atol=1e-15)
    jump = [1.0, 1.0, 1.0, 1.0, 0.0, 0.0]
    assert orc.teno5i(jump)[0] == pytest.approx(1.0, abs=1e-14)


def test_hllc_1d_quirk_energy_term(orc):
    """1D HLLC (Equations_Euler1D.cpp:156-210) in the right star region with
    Roe-dominated wave speeds (S_L = u~ - c~, S_R = u~ + c~ in both modes, so
    the c-without-gamma quirk of :174-175 does not act): the modes agree in
    mass and momentum and differ in energy flux by exactly
    S_R u_R (p_R - p_L) / (S_R - S*), the p_L substitution of :204.  S_R is
    the Roe speed and S* is recovered from the mass flux."""
    wl, wr = [1.0, 1.0, 1.0], [1.0, -1.0, 1.3]

    def cons(w):
        return np.array([w[0], w[0] * w[1], w[2] / (GAMMA - 1) + 0.5 * w[0] * w[1] * w[1]])

    ql, qr = cons(wl), cons(wr)
    a, b = orc.e1_hllc(GAMMA, ql, qr, True), orc.e1_hllc(GAMMA, ql, qr, False)
    np.testing.assert_array_equal(a[:2], b[:2])
    rr = math.sqrt(wr[0] / wl[0])
    vm = (wl[1] + wr[1] * rr) / (1 + rr)
    hl, hr = (ql[2] + wl[2]) / wl[0], (qr[2] + wr[2]) / wr[0]
    hm = (hl + hr * rr) / (1 + rr)
    cm = math.sqrt((GAMMA - 1) * (hm - 0.5 * vm * vm))
    sr = vm + cm
    assert sr > wr[1] + math.sqrt(GAMMA * wr[2] / wr[0]) and vm - cm < wl[1] - math.sqrt(GAMMA * wl[2] / wl[0])
    k = (b[0] - wr[0] * wr[1]) / sr + wr[0]
    sm = sr - wr[0] * (sr - wr[1]) / k
    assert sm < 0.0
    assert a[2] - b[2] == pytest.approx(sr * wr[1] * (wr[2] - wl[2]) / (sr - sm), rel=1e-12)
