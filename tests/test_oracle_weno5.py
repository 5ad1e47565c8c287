"""The oracle's WENO5-JS and WENO5-Z (Schemes_Lagrange.cpp:358-408) pinned
by what the scheme fixes: exact face values on cell averages of a quadratic
(every candidate is exact and the weights sum to one), fifth-order
convergence on a smooth profile (the linear weights 1/10, 6/10, 3/10 are the
ones that cancel the candidates' third-order errors), mirror symmetry of the
two one-sided values, and on a step the value from the one smooth candidate
(the weights of the candidates crossing the step vanish)."""
import numpy as np
import pytest


def _avg(f_prim, x0, h):
    return (f_prim(x0 + h) - f_prim(x0)) / h


@pytest.mark.parametrize("z", [False, True])
def test_quadratic_exact(orc, z):
    h = 0.1
    prim = lambda x: 2.0 * x + 0.5 * x * x - x ** 3 / 3.0  # noqa: E731  (f = 2 + x - x^2)
    u = np.array([_avg(prim, (k - 2) * h, h) for k in range(6)])  # cells i-2 .. i+3, face x = h
    f_face = 2.0 + h - h * h
    out = orc.weno5(u, z)
    assert abs(out[0] - f_face) < 1e-13 and abs(out[1] - f_face) < 1e-13


@pytest.mark.parametrize("z", [False, True])
def test_fifth_order(orc, z):
    errs = []
    for n in (40, 80, 160, 320):
        h = 1.0 / n
        prim = lambda x: -np.cos(2 * np.pi * x) / (2 * np.pi)  # noqa: E731
        x0 = 0.3
        u = np.array([_avg(prim, x0 + (k - 2) * h, h) for k in range(6)])
        exact = np.sin(2 * np.pi * (x0 + h))
        errs.append(abs(orc.weno5(u, z)[0] - exact))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    # (WENO-JS reaches its fifth order pre-asymptotically: 4.7 .. 5)
    assert rates[-2:].min() > 4.6, rates


@pytest.mark.parametrize("z", [False, True])
def test_mirror(orc, z):
    rng = np.random.default_rng(3)
    for _ in range(20):
        u = rng.normal(size=6)
        a = orc.weno5(u, z)
        b = orc.weno5(u[::-1].copy(), z)
        assert abs(a[0] - b[1]) < 1e-13 * (1 + abs(a[0])) and abs(a[1] - b[0]) < 1e-13 * (1 + abs(a[1]))


@pytest.mark.parametrize("z", [False, True])
def test_step_takes_the_smooth_candidate(orc, z):
    # u- at the face i+1/2 with the step between cells i+1 and i+2: only the
    # stencil (i-2, i-1, i) is smooth; its candidate on constant data is 1
    u = np.array([1.0, 1.0, 1.0, 1.0, 0.0, 0.0])
    out = orc.weno5(u, z)
    assert abs(out[0] - 1.0) < (1e-9 if not z else 1e-6)
