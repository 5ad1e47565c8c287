"""Pins for the oracle parts that closed forms of the whole scheme leave open:
the HLLC wave-speed estimate with the reference's missing gamma (quirk Q1),
the TENO5 and HTENO cut-off constants (C_T, the power 6, epsilon), the
Hermite limiter's nonlinear weights, and the all-linear residual the learned
path is compared with.

Each check is built so that a plausible transcription mistake fails it: a
gamma put back into the HLLC sound speed, a cut-off of 4.99e-4 instead of
5e-4, a power of 5, an epsilon of 1e-30, limiter weights 1 - 2e-4, ...

The smoothness indicators that enter the cut-off ratios come from the
oracle's indicator hooks, which tests/test_oracle_hermite_1d.py pins to their
definition as smoothness integrals; the ratios themselves are then evaluated
in exact rational arithmetic from the reference's cut-off rule
(Schemes_Lagrange.cpp:481-494, 1141-1149; Schemes_Hermite.cpp:295-339,
614-628), so the only freedom left to the oracle is the constants.
"""
import math
from fractions import Fraction as Fr

import numpy as np
import pytest

from conftest import rel_l1
from oracle import params

GAMMA = 1.4


# ------------------------------------------------------------ HLLC, Q1 ----
def _prim(q):
    r = q[0]
    u, v = q[1] / r, q[2] / r
    return np.array([r, u, v, (GAMMA - 1) * (q[3] - 0.5 * r * (u * u + v * v))])


def _cons(w):
    r, u, v, p = w
    return np.array([r, r * u, r * v, p / (GAMMA - 1) + 0.5 * r * (u * u + v * v)])


def _toro_hllc(ql, qr, sl, sr):
    """Toro's HLLC (Riemann Solvers, 3rd ed., eqs. 10.37-10.39 and 10.71-10.73)
    for the x direction with given wave-speed estimates."""
    wl, wr = _prim(ql), _prim(qr)

    def f(q, w):
        return np.array([q[1], q[1] * w[1] + w[3], q[1] * w[2], w[1] * (q[3] + w[3])])

    rl, ul, _, pl = wl
    rr, ur, _, pr = wr
    sm = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur))

    def star(q, w, s):
        r, u, v, p = w
        k = r * (s - u) / (s - sm)
        return k * np.array([1.0, sm, v, q[3] / r + (sm - u) * (sm + p / (r * (s - u)))])

    if sl > 0:
        return f(ql, wl), sm
    if sl <= 0 < sm:
        return f(ql, wl) + sl * (star(ql, wl, sl) - ql), sm
    if sm <= 0 <= sr:
        return f(qr, wr) + sr * (star(qr, wr, sr) - qr), sm
    return f(qr, wr), sm


def _einfeldt_speeds(wl, wr, with_gamma):
    """S_L = min(u_L - c_L, u~ - c~), S_R = max(u_R + c_R, u~ + c~) with the
    Roe averages u~, c~; c_K = sqrt(gamma p / rho), or sqrt(p / rho)."""
    ql, qr = _cons(wl), _cons(wr)
    sa, sb = math.sqrt(wl[0]), math.sqrt(wr[0])
    um = (sa * wl[1] + sb * wr[1]) / (sa + sb)
    vm = (sa * wl[2] + sb * wr[2]) / (sa + sb)
    hl, hr = (ql[3] + wl[3]) / wl[0], (qr[3] + wr[3]) / wr[0]
    hm = (sa * hl + sb * hr) / (sa + sb)
    cm = math.sqrt((GAMMA - 1) * (hm - 0.5 * (um * um + vm * vm)))
    g = GAMMA if with_gamma else 1.0
    cl, cr = math.sqrt(g * wl[3] / wl[0]), math.sqrt(g * wr[3] / wr[0])
    return min(wl[1] - cl, um - cm), max(wr[1] + cr, um + cm)


@pytest.mark.parametrize("wl, wr", [
    ([1.0, 0.0, 0.3, 1.0], [1.0, 2.0, 0.1, 1.0]),
    ([2.0, -0.2, 0.0, 3.0], [0.5, 1.5, -0.4, 0.8]),
])
def test_hllc_q1_sound_speed_without_gamma(orc, wl, wr):
    """In the left star region (S_L <= 0 < S*, where the p_L quirk of :423
    does not act) the left wave speed comes from u_L - c_L.  With quirks the
    oracle equals Toro's HLLC whose c_L, c_R omit gamma
    (Equations_Euler2D.cpp:393-394); without quirks, Toro's with gamma.  The
    two differ at the percent level, so putting gamma back (or dropping it
    from the corrected mode) fails."""
    ql, qr = _cons(np.array(wl)), _cons(np.array(wr))
    sl_q, sr_q = _einfeldt_speeds(wl, wr, with_gamma=False)
    sl_g, sr_g = _einfeldt_speeds(wl, wr, with_gamma=True)
    # Q1 sets S_L: u_L - sqrt(p_L / rho_L) is below the Roe speed
    assert sl_q == pytest.approx(wl[1] - math.sqrt(wl[3] / wl[0]), rel=1e-15)
    fq, smq = _toro_hllc(ql, qr, sl_q, sr_q)
    fg, smg = _toro_hllc(ql, qr, sl_g, sr_g)
    assert sl_q <= 0 < smq and sl_g <= 0 < smg
    assert np.abs(fq - fg).max() > 1e-3
    np.testing.assert_allclose(orc.hllc(GAMMA, 1, ql, qr, True), fq, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(orc.hllc(GAMMA, 1, ql, qr, False), fg, rtol=1e-13, atol=1e-14)
    # the y direction reads the momenta the other way round
    swap = lambda q: np.array([q[0], q[2], q[1], q[3]])  # noqa: E731
    fy = orc.hllc(GAMMA, 2, swap(ql), swap(qr), True)
    np.testing.assert_allclose(swap(fy), fq, rtol=1e-13, atol=1e-14)


# ------------------------------------------------- TENO5 cut-off, 5e-4 ----
EPS_Z = Fr(1e-40)  # Schemes_Lagrange.cpp:7 (the double nearest 1e-40)


def _teno_ratios(orc, u):
    """Exact ratios a_k^6 / sum_j a_j^6 of teno5's left side
    (Schemes_Lagrange.cpp:486-494): a_k = 1 + tau / (beta_k + 1e-40),
    tau = |beta5 - (beta_0 + beta_2 + 4 beta_1) / 6|."""
    SI, b35 = orc.teno5_indicators(u)
    S = [Fr(float(x)) for x in SI[:3]]
    tau = abs(Fr(float(b35[0])) - (S[0] + S[2] + 4 * S[1]) / 6)
    a = [(1 + tau / (s + EPS_Z)) ** 6 for s in S]
    tot = sum(a)
    return [x / tot for x in a]


def _hteno_ratios(orc, u3, ud3, deriv):
    """Exact ratios of one HTENO side (Schemes_Hermite.cpp:314-317, 333-336):
    tau = |brk - (beta_0 + beta_1 + 2 beta_2) / 4|, power 6, epsilon 1e-40."""
    SI, SId, b = orc.hteno_indicators(u3, ud3)
    S = [Fr(float(x)) for x in (SId if deriv else SI)]
    tau = abs(Fr(float(b[1] if deriv else b[0])) - (S[0] + S[1] + 2 * S[2]) / 4)
    a = [(1 + tau / (s + EPS_Z)) ** 6 for s in S]
    tot = sum(a)
    return [x / tot for x in a]


def _straddle(ratio_of, lo, hi, ct, log=False):
    """Two parameters t_keep, t_cut in [lo, hi] whose exact ratio lies just
    above / just below ct (relative distance in (1e-10, 1e-6)), by bisection
    on a bracket where the ratio crosses ct."""
    ct = Fr(ct)
    f = lambda t: ratio_of(t) - ct  # noqa: E731
    a, b = lo, hi
    fa, fb = f(a), f(b)
    assert (fa >= 0) != (fb >= 0), "bracket does not straddle the threshold"
    keep = cut = None
    for _ in range(200):
        m = math.sqrt(a * b) if log else 0.5 * (a + b)
        fm = f(m)
        rel = abs(fm / ct)
        if fm >= 0:
            if 1e-10 < rel < 1e-6:
                keep = m
        elif 1e-10 < rel < 1e-6:
            cut = m
        if keep is not None and cut is not None:
            return keep, cut
        if (fm >= 0) == (fa >= 0):
            a, fa = m, fm
        else:
            b, fb = m, fm
    raise AssertionError("no straddling pair found")


@pytest.mark.parametrize("k", [0, 1])
def test_teno5_cutoff_flips_at_5e4(orc, k):
    """u(s) = (0,0,0,1,2,3) + s (1..6): candidate k's normalised weight
    crosses C_T = 5e-4 as s grows.  Just above C_T the oracle keeps the
    candidate, just below it cuts it; the two stencils differ in the ratio by
    less than 1e-6 relative, so any other threshold, power or epsilon flips
    one of the two decisions."""
    fam = lambda s: np.array([0, 0, 0, 1, 2, 3.0]) + s * np.arange(1, 7)  # noqa: E731
    keep, cut = _straddle(lambda s: _teno_ratios(orc, fam(s))[k], 0.3, 1.0, 5e-4)
    assert orc.teno5_flags(fam(keep))[k] == 1
    assert orc.teno5_flags(fam(cut))[k] == 0


def test_teno5_epsilon_is_1e40(orc):
    """With two flat candidate stencils (beta_1 = beta_2 = 0) their weights are
    (1 + tau / 1e-40)^6, so the third candidate's cut depends on epsilon:
    u = (0, 0, 0, 0, d, 0) straddles C_T near d ~ 1e-20 exactly where
    epsilon = 1e-40 puts it."""
    fam = lambda d: np.array([0, 0, 0, 0, d, 0.0])  # noqa: E731
    SI, _ = orc.teno5_indicators(fam(1e-20))
    assert SI[1] == 0.0 and SI[2] == 0.0 and SI[0] > 0.0
    keep, cut = _straddle(lambda d: _teno_ratios(orc, fam(d))[0], 1e-21, 1e-18, 5e-4, log=True)
    assert orc.teno5_flags(fam(keep))[0] == 1
    assert orc.teno5_flags(fam(cut))[0] == 0
    assert 1e-21 < keep < 1e-18


# ------------------------------------------------- HTENO cut-off, 1e-3 ----
def test_hteno_value_cutoff_flips_at_1e3(orc):
    """u^- of hteno (C_T = 1e-3, Schemes_Hermite.cpp:6, 614-620): the family
    u = (0,0,1,1) + t (0,1,2,3), d = t moves candidate 2's ratio across 1e-3
    between t = 1 and 2."""
    dif = 1.0
    fam = lambda t: (np.array([0, 0, 1, 1.0]) + t * np.arange(4), np.full(4, t))  # noqa: E731

    def r(t):
        u, d = fam(t)
        return _hteno_ratios(orc, u[:3], d[:3] * dif, False)[2]

    keep, cut = _straddle(r, 1.0, 2.0, 1e-3)
    assert orc.hteno_flags(dif, *fam(keep))[2] == 1
    assert orc.hteno_flags(dif, *fam(cut))[2] == 0


def test_hteno_derivative_cutoff_flips_at_1e3(orc):
    """d^- of hteno (ctd = 1e-3, Schemes_Hermite.cpp:6, 622-628): cell
    averages of x^2 with consistent derivative moments plus a kink t (0,0,1,1)
    move the derivative candidate 1's ratio across 1e-3 between t = 0.1 and
    0.3."""
    dif = 1.0
    xs = np.arange(4) - 1.0
    us = ((xs + 0.5) ** 3 - (xs - 0.5) ** 3) / 3.0
    ds = 2.0 * xs
    fam = lambda t: (us + t * np.array([0, 0, 1, 1.0]), ds.copy())  # noqa: E731

    def r(t):
        u, d = fam(t)
        return _hteno_ratios(orc, u[:3], d[:3] * dif, True)[1]

    keep, cut = _straddle(r, 0.1, 0.3, 1e-3)
    assert orc.hteno_flags(dif, *fam(keep))[7] == 1
    assert orc.hteno_flags(dif, *fam(cut))[7] == 0


# --------------------------------------------------- Hermite limiter ----
def _limiter_exact(orc, dif, u, d, eps=Fr(1e-6), lin0=Fr(1 - 2e-5), lin12=Fr(1e-5)):
    """hermiteLimiter (Schemes_Hermite.cpp:634-662) in exact arithmetic from
    the oracle's pinned full-stencil indicator brk."""
    _, _, b = orc.hteno_indicators(np.array(u, float), np.array(d, float) * dif)
    U = [Fr(float(x)) for x in u]
    D = [Fr(float(x)) for x in d]
    h = Fr(float(dif))
    p = [Fr(3, 4) * (U[2] - U[0]) / h - Fr(1, 4) * (D[0] + D[2]), (U[1] - U[0]) / h, (U[2] - U[1]) / h]
    lin = [lin0, lin12, lin12]
    sm = [Fr(float(b[0])), (U[0] - U[1]) ** 2, (U[1] - U[2]) ** 2]
    tau = (abs(sm[0] - sm[1]) + abs(sm[0] - sm[2])) ** 2 / 4
    w = [lin[k] * (1 + tau / (sm[k] + eps)) for k in range(3)]
    tot = sum(w)
    w = [x / tot for x in w]
    return w[0] * (p[0] / lin[0] - (lin[1] / lin[0] * p[1] + lin[2] / lin[0] * p[2])) + w[1] * p[1] + w[2] * p[2]


@pytest.mark.parametrize("u, d", [([0, 0, 1], [0, 0, 0]), ([2, 1, 1], [0, 5, 1]), ([1, 1, 3], [0, 0, 1]),
                                  ([4, 2, 2], [1, 0, -1])])
def test_hermite_limiter_weights_exact(orc, u, d):
    """Integer cell data make the two one-sided indicators integers and the
    weights rationals: the oracle matches the exact value to rounding.  In
    each case one one-sided indicator is 0 or small, so an epsilon of 1e-5 or
    linear weights (1 - 2e-4, 1e-4, 1e-4) change the result by more than a
    factor 2."""
    dif = 1.0
    exact = _limiter_exact(orc, dif, u, d)
    got = orc.hermite_limiter(dif, np.array(u, float), np.array(d, float))
    assert got == pytest.approx(float(exact), rel=1e-13, abs=1e-15)
    for alt in (_limiter_exact(orc, dif, u, d, eps=Fr(1e-5)),
                _limiter_exact(orc, dif, u, d, lin0=Fr(1 - 2e-4), lin12=Fr(1e-4))):
        assert abs(float(alt) - got) > 0.5 * abs(got)


# ------------------------------------------------- all-linear residual ----
def cell_avg_poly(coef, a, b):
    return sum(c * (b ** (k + 1) - a ** (k + 1)) / (k + 1) for k, c in enumerate(coef)) / (b - a)


@pytest.mark.parametrize("seed", range(4))
def test_linear_quadrature_exact_for_quartics(orc, seed):
    """The linear limit of the 3-point transverse quadrature interpolates the
    5 cell averages of a quartic exactly at the Gauss points +-sqrt(15)/10
    and 0 (every candidate is exact only to degree 2, so this fails if a
    linear weight is wrong)."""
    rng = np.random.default_rng(seed)
    coef = rng.uniform(-1, 1, 5)
    u = [cell_avg_poly(coef, m - 2.5, m - 1.5) for m in range(5)]
    xi = math.sqrt(15.0) / 10.0
    val = lambda x: sum(c * x ** k for k, c in enumerate(coef))  # noqa: E731
    np.testing.assert_allclose(orc.weno3q_linear(u), [val(xi), val(0.0), val(-xi)], rtol=0, atol=1e-13)
    # on quadratics the nonlinear quadrature gives the same values
    u2 = [cell_avg_poly(coef[:3], m - 2.5, m - 1.5) for m in range(5)]
    np.testing.assert_allclose(orc.weno3q_linear(u2), orc.weno3q(u2), rtol=0, atol=1e-13)


def test_linear_residual_free_stream_and_conservation(orc):
    q = np.empty((4, 10, 12))
    q[:] = orc.conservative(GAMMA, [1.3, 0.4, -0.2, 0.9])[:, None, None]
    prm = params(12, 10, 0, 1, 0, 1, GAMMA, "zero_grad", "hllc", True)
    assert np.abs(orc.fv2d_rhs_linear(prm, q)).max() < 1e-13
    prm = params(24, 24, 0.0, 10.0, 0.0, 10.0, GAMMA, "periodic", "hllc", True)
    q = orc.ic_vortex(prm, (5.0, 5.0), None)
    r = orc.fv2d_rhs_linear(prm, q)
    for e in range(4):
        assert abs(r[e].sum()) < 1e-11 * np.abs(r[e]).sum()


def test_linear_residual_converges_and_approaches_teno(orc):
    """On the smooth vortex the all-linear residual approximates the exact
    time derivative of the cell averages (translation along (1, 1)) and
    converges like the TENO5 residual (the reference's mixing of the rows'
    characteristic bases in the transverse quadrature, Q12, limits both to
    second order once resolved), and the TENO5 residual approaches it as
    the grid is refined (fewer cut candidates)."""
    errs, errs_t, diffs = [], [], []
    for n in (20, 40, 80):
        prm = params(n, n, 0.0, 10.0, 0.0, 10.0, GAMMA, "periodic", "hllc", True)
        q = orc.ic_vortex(prm, (5.0, 5.0), None)
        eps = 1e-6
        dq = (orc.ic_vortex(prm, (5.0 + eps, 5.0 + eps), None) - orc.ic_vortex(prm, (5.0 - eps, 5.0 - eps), None))
        dq /= 2 * eps
        r, rt = orc.fv2d_rhs_linear(prm, q), orc.fv2d_rhs(prm, q)
        errs.append(np.abs(r - dq).mean())
        errs_t.append(np.abs(rt - dq).mean())
        diffs.append(rel_l1(r, rt))
    assert errs[0] / errs[1] > 8.0 and errs[1] / errs[2] > 4.0
    assert all(e <= et * 1.01 for e, et in zip(errs, errs_t))
    assert diffs[0] > diffs[1] > diffs[2]


def test_entropy_wave_averages_closed_form(orc):
    """The oracle's 9x9 Gauss entropy-wave averages against the closed form:
    the average of sin(pi (x + y)) over a dx x dy cell is
    sin(pi (xc + yc)) S(dx) S(dy) with S(h) = sin(pi h / 2) / (pi h / 2), the
    derivative moments the same with pi cos, momentum and energy linear in
    the density; the exact-solution shift moves the centres by -shift."""
    import math

    nx, ny = 7, 5
    prm = params(nx, ny, 0.0, 1.4, 0.0, 1.5, 1.4, "periodic", "hllc", True)
    u, v, p, amp, rho, g1 = 0.7, -0.4, 0.9, 0.3, 1.0, 0.4
    dx, dy = 1.4 / nx, 1.5 / ny

    def S(h):
        return math.sin(math.pi * h / 2) / (math.pi * h / 2)

    for shift in ((0.0, 0.0), (0.37, -0.11)):
        q = orc.ic_entropy(prm, rho, u, v, p, amp, shift=shift, hermite=True)
        xc = (np.arange(nx) + 0.5) * dx - shift[0]
        yc = (np.arange(ny) + 0.5) * dy - shift[1]
        arg = math.pi * (xc[None, :] + yc[:, None])
        r = rho + amp * np.sin(arg) * S(dx) * S(dy)
        d = amp * math.pi * np.cos(arg) * S(dx) * S(dy)
        np.testing.assert_allclose(q[0], r, rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(q[1], u * r, rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(q[2], v * r, rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(q[3], p / g1 + 0.5 * (u * u + v * v) * r, rtol=1e-14)
        np.testing.assert_allclose(q[4], d, rtol=1e-13, atol=1e-14)
        np.testing.assert_allclose(q[8:12], q[4:8], rtol=0, atol=0)
        np.testing.assert_allclose(q[7], 0.5 * (u * u + v * v) * d, rtol=1e-13, atol=1e-14)
