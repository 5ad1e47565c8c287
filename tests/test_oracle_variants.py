"""Pins of the oracle's remaining Lagrange and Hermite routines
(Schemes_Lagrange.cpp:315-1004, :1265-1294; Schemes_Hermite.cpp:234-293,
:381-513, :670-767) by what the mathematics fixes, independently of their
formulas:

* exactness on cell averages of polynomials: every 3-cell candidate is exact
  for quadratics, so every nonlinear combination of them is (degree 2); the
  2-cell candidates of the WENO-IS family for linear data; the optimal linear
  schemes for degree 4 (Lagrange::linear, Hermite::linear);
* fifth-order convergence on smooth data, which holds only if the nonlinear
  weights tend to the optimal linear weights and those weights and the
  candidate coefficients are right (a wrong weight or a dropped term falls to
  third order or below);
* at a jump, the TENO-type cut-off keeps only the smooth candidates, so the
  value is exactly theirs; the WENO-type weights make it so to O(eps^2);
* mirror symmetry (u -> reversed u swaps u^- and u^+) for the routines whose
  definition is symmetric;
* the quadrature points of the 2- and 4-point rules: exact interpolation of
  quadratics at +-sqrt(3)/6, +-sqrt(5)/10 and the Gauss-Legendre abscissae of
  the cell, and fifth-order convergence there, except where the reference's
  weight tables pair the optimal weights with the mirrored stencils (the
  Gauss-Lobatto end points and the Gauss-Legendre points): there the value
  tracks that mirrored linear combination, which the test forms from first
  principles.
"""
import math

import numpy as np
import pytest

QUAD = ["weno5n", "teno5n", "teno5a", "weno5z_diff", "teno5_diff", "linear"]
LIN = ["weno5is", "weno5ismod", "weno5is_ext"]
ALL5 = QUAD + LIN
TENO_TYPE = ["teno5n", "teno5a", "teno5_diff"]


def cell_avgs(coef, xs, h):
    """Exact averages over [x - h/2, x + h/2] of the polynomial sum c_k x^k."""
    out = []
    for x in xs:
        a, b = x - h / 2, x + h / 2
        out.append(sum(c * (b ** (k + 1) - a ** (k + 1)) / (k + 1) for k, c in enumerate(coef)) / h)
    return np.array(out)


def poly(coef, x):
    return sum(c * x ** k for k, c in enumerate(coef))


def dpoly(coef, x):
    return sum(k * c * x ** (k - 1) for k, c in enumerate(coef) if k > 0)


@pytest.mark.parametrize("kind", ALL5)
def test_polynomial_exactness(orc, kind):
    rng = np.random.default_rng(7)
    deg = 4 if kind == "linear" else (2 if kind in QUAD else 1)
    h = 0.1
    xs = (np.arange(6) - 2) * h  # cells i-2 .. i+3, face at x = h/2
    for _ in range(20):
        coef = rng.uniform(-1, 1, deg + 1)
        coef[0] += 2.0
        u = cell_avgs(coef, xs, h)
        r = orc.recon(kind, u, h)
        ex = poly(coef, h / 2)
        np.testing.assert_allclose(r, [ex, ex], rtol=2e-13, atol=2e-13)


@pytest.mark.parametrize("kind", ALL5 + ["teno5"])
def test_fifth_order_on_smooth_data(orc, kind):
    x0 = 0.1
    errs = []
    for n in (20, 40, 80, 160):
        h = 1.0 / n
        a = x0 + (np.arange(6) - 3) * h
        u = (np.cos(2 * np.pi * a) - np.cos(2 * np.pi * (a + h))) / (2 * np.pi * h)
        errs.append(np.max(np.abs(orc.recon(kind, u, h) - math.sin(2 * np.pi * x0))))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(3)]
    assert min(rates[1:]) > 4.8, (kind, errs, rates)


@pytest.mark.parametrize("kind", ALL5)
def test_jump_takes_the_smooth_candidates(orc, kind):
    # a unit step between u3 and u4: on the minus side (window u0..u4) the
    # stencils u0..u2 and u1..u3 are flat, u2..u4 crosses the jump
    u = np.array([1.0, 1.0, 1.0, 1.0, 0.0, 0.0])
    um = orc.recon(kind, u, 0.01)[0]
    if kind in TENO_TYPE:
        assert um == 1.0
    elif kind == "linear":
        assert um == pytest.approx(0.3 * (1 / 3 + 5 / 6 - 1 / 6 * 0) + 0.6 + 0.1, rel=1e-14)
    else:
        assert abs(um - 1.0) < 1e-6, um


@pytest.mark.parametrize("kind", ["teno5a", "teno5_diff", "weno5is", "weno5ismod", "weno5is_ext", "linear"])
def test_mirror_symmetry(orc, kind):
    rng = np.random.default_rng(11)
    for _ in range(30):
        u = 1.0 + rng.uniform(-0.3, 0.3, 6)
        a = orc.recon(kind, u, 0.05)
        b = orc.recon(kind, u[::-1].copy(), 0.05)
        np.testing.assert_allclose([a[0], a[1]], [b[1], b[0]], rtol=1e-13, atol=1e-14)


def test_weno3js_linear_exact_and_quirk(orc):
    h = 0.1
    xs = (np.arange(4) - 1) * h  # cells i-1..i+2, face at h/2
    rng = np.random.default_rng(3)
    for _ in range(10):
        coef = rng.uniform(-1, 1, 2) + [2.0, 0.0]
        r = orc.weno3js(cell_avgs(coef, xs, h))
        ex = poly(coef, h / 2)
        # the second output is the first (Schemes_Lagrange.cpp:333 stores catchnan(out(0)))
        assert r[0] == r[1]
        assert r[0] == pytest.approx(ex, rel=1e-13)


def test_weno3js_2q_points(orc):
    h = 0.1
    xs = (np.arange(3) - 1) * h
    g = math.sqrt(3.0) / 6.0 * h
    rng = np.random.default_rng(4)
    for _ in range(10):
        coef = rng.uniform(-1, 1, 2) + [2.0, 0.0]
        r = orc.weno3js_2q(cell_avgs(coef, xs, h))
        np.testing.assert_allclose(r, [poly(coef, g), poly(coef, -g)], rtol=1e-13)


# quadrature points of the 5-cell Gauss rules, as offsets from the cell centre in cells
LOBATTO = [0.5, -0.5, math.sqrt(5.0) / 10.0, -math.sqrt(5.0) / 10.0]
_gl_in = 0.5 * math.sqrt((3.0 - 2.0 * math.sqrt(6.0 / 5.0)) / 7.0)
_gl_out = 0.5 * math.sqrt((3.0 + 2.0 * math.sqrt(6.0 / 5.0)) / 7.0)
LEGENDRE = [_gl_in, -_gl_in, _gl_out, -_gl_out]
TWO = [math.sqrt(3.0) / 6.0, -math.sqrt(3.0) / 6.0]


def _quad(orc, rule, kind, u):
    if rule == "two":
        return orc.quad2(kind, u)
    return orc.quad4(kind, u, lobatto=(rule == "lobatto"))


RULES = [("two", 0), ("two", 1), ("two", 2), ("lobatto", 0), ("lobatto", 1), ("lobatto", 2), ("legendre", 0),
         ("legendre", 1)]


@pytest.mark.parametrize("rule,kind", RULES)
def test_quadrature_points_exact_for_quadratics(orc, rule, kind):
    pts = {"two": TWO, "lobatto": LOBATTO, "legendre": LEGENDRE}[rule]
    h = 0.1
    xs = (np.arange(5) - 2) * h
    rng = np.random.default_rng(5)
    for _ in range(10):
        coef = rng.uniform(-1, 1, 3) + [2.0, 0.0, 0.0]
        r = _quad(orc, rule, kind, cell_avgs(coef, xs, h))
        np.testing.assert_allclose(r, [poly(coef, p * h) for p in pts], rtol=2e-13, atol=2e-13)


def cand_coef(xi, cells):
    """Coefficients c with sum_j c_j avg_j = p(xi) for every polynomial p of
    degree len(cells) - 1 (unit cells centred at the given offsets)."""
    n = len(cells)
    A = np.array([[((c + 0.5) ** (m + 1) - (c - 0.5) ** (m + 1)) / (m + 1) for c in cells] for m in range(n)])
    return np.linalg.solve(A, np.array([xi ** m for m in range(n)]))


def optimal_weights(xi):
    """The linear weights that combine the three 3-cell interpolants (left
    stencil first) into the 5-cell one at xi."""
    E = np.zeros((5, 3))
    for k in range(3):
        E[k:k + 3, k] = cand_coef(xi, [k - 2, k - 1, k])
    d, *_ = np.linalg.lstsq(E, cand_coef(xi, [-2, -1, 0, 1, 2]), rcond=None)
    return d


@pytest.mark.parametrize("rule,kind", RULES)
def test_quadrature_points_order(orc, rule, kind):
    """The 2-point rule and the inner Gauss-Lobatto points converge at fifth
    order.  The Gauss-Lobatto end points (dkm / dkp paired with the candidates
    ordered left stencil first, Schemes_Lagrange.cpp:798-811) and every
    Gauss-Legendre point (dkpLegendre_12 / _34, :245-286) combine the
    candidates with the optimal weights mirrored, as the reference's tables
    give them: there the value tends to that mirrored linear combination
    and is third order against the exact value."""
    pts = {"two": TWO, "lobatto": LOBATTO, "legendre": LEGENDRE}[rule]
    mirrored = [False, False] if rule == "two" else ([True, True, False, False] if rule == "lobatto" else [True] * 4)
    x0 = 0.1
    err, err_lin = [], []
    for n in (20, 40, 80, 160):
        h = 1.0 / n
        a = x0 + (np.arange(5) - 2.5) * h
        u = (np.cos(2 * np.pi * a) - np.cos(2 * np.pi * (a + h))) / (2 * np.pi * h)
        r = _quad(orc, rule, kind, u)
        err.append(np.abs(r - np.sin(2 * np.pi * (x0 + np.array(pts) * h))))
        lin = []
        for p, mir in zip(pts, mirrored):
            d = optimal_weights(p)[::-1] if mir else optimal_weights(p)
            lin.append(sum(d[k] * (cand_coef(p, [k - 2, k - 1, k]) @ u[k:k + 3]) for k in range(3)))
        err_lin.append(np.abs(r - np.array(lin)))
    err, err_lin = np.array(err), np.array(err_lin)
    rate = np.log2(err[2] / err[3])
    for m, mir in enumerate(mirrored):
        if mir:
            assert 2.7 < rate[m] < 3.3, (rule, kind, m, rate)
            # the value sits at the mirrored combination, far closer than to the exact value
            assert err_lin[3][m] <= max(0.1 * err[3][m], 1e-14), (rule, kind, m, err_lin, err)
            assert err_lin[3][m] <= max(err_lin[2][m] / 8.0, 1e-14), (rule, kind, m, err_lin)
        else:
            assert rate[m] > 4.8, (rule, kind, m, rate)


# ----------------------------------------------------------------------------
# Hermite routines: cell averages u_k and derivative moments d_k = <p'> of a
# polynomial p over cells i-1..i+2; face x_{i+1/2} at h/2


def herm_data(coef, xs, h):
    u = cell_avgs(coef, xs, h)
    d = np.array([(poly(coef, x + h / 2) - poly(coef, x - h / 2)) / h for x in xs])
    return u, d


@pytest.mark.parametrize("kind,deg", [("hweno", 2), ("htenoi", 2), ("hlinear", 4), ("hteno", 2)])
def test_hermite_polynomial_exactness(orc, kind, deg):
    h = 0.1
    xs = (np.arange(4) - 1) * h
    rng = np.random.default_rng(9)
    for _ in range(20):
        coef = rng.uniform(-1, 1, deg + 1)
        coef[0] += 2.0
        u, d = herm_data(coef, xs, h)
        r = orc.herm_recon(kind, h, u, d)
        ex, dex = poly(coef, h / 2), dpoly(coef, h / 2)
        np.testing.assert_allclose(r, [ex, ex, dex, dex], rtol=2e-12, atol=2e-12)


@pytest.mark.parametrize("kind", ["hweno", "htenoi", "hlinear"])
def test_hermite_convergence(orc, kind):
    """u^+- at fifth order or better on smooth data (the linear Hermite
    interpolant and, through the weights' linear limit, the nonlinear ones)."""
    x0 = 0.1
    errs = []
    for n in (20, 40, 80, 160):
        h = 1.0 / n
        a = x0 + (np.arange(4) - 2) * h
        u = (np.cos(2 * np.pi * a) - np.cos(2 * np.pi * (a + h))) / (2 * np.pi * h)
        d = (np.sin(2 * np.pi * (a + h)) - np.sin(2 * np.pi * a)) / h
        r = orc.herm_recon(kind, h, u, d)
        errs.append(max(abs(r[0] - math.sin(2 * np.pi * x0)), abs(r[1] - math.sin(2 * np.pi * x0))))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(3)]
    assert min(rates[1:]) > 4.5, (kind, errs, rates)


def test_htenoi_jump_keeps_smooth_candidates(orc):
    # flat cells i-1, i, i+1 with zero moments, a jump into cell i+2: u^-
    # (cells i-1..i+1) sees no jump at all, u^+ (cells i..i+2) cuts the
    # candidates that cross it and keeps the flat one
    u = np.array([1.0, 1.0, 1.0, 0.0])
    d = np.zeros(4)
    r = orc.herm_recon("htenoi", 0.1, u, d)
    assert r[0] == 1.0


@pytest.mark.parametrize("npts", [2, 4])
def test_hweno_glq_points_exact(orc, npts):
    h = 0.1
    xs = (np.arange(3) - 1) * h
    pts = TWO if npts == 2 else LOBATTO
    rng = np.random.default_rng(12)
    for _ in range(10):
        coef = rng.uniform(-1, 1, 3) + [2.0, 0.0, 0.0]
        u, d = herm_data(coef, xs, h)
        r = orc.hweno_glq(npts, h, u, d)
        ex = [poly(coef, p * h) for p in pts] + [dpoly(coef, p * h) for p in pts]
        np.testing.assert_allclose(r, ex, rtol=2e-12, atol=2e-12)
