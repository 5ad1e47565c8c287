"""Seeded workload definitions (htx.ic): the draws are reproducible and the
shock-blob workload's post-shock state satisfies the Rankine-Hugoniot jump
conditions of a shock moving at speed M c1 into gas at rest."""
import math

import numpy as np
import pytest


def test_quadrant_workload_ranges_and_determinism():
    from htx.ic import quadrant_workload

    w = quadrant_workload(64)
    assert w == quadrant_workload(64)
    assert 0.3 <= w["xs"] <= 0.7 and 0.3 <= w["ys"] <= 0.7
    for r, u, v, p in w["states"]:
        assert 0.5 <= r <= 2.0 and -0.5 <= u <= 0.5 and -0.5 <= v <= 0.5 and 0.2 <= p <= 2.0
    assert quadrant_workload(64, seed=1)["xs"] != w["xs"]


def test_shock_blob_rankine_hugoniot():
    from htx.ic import shock_blob_workload

    w = shock_blob_workload(256)
    g, M = w["gamma"], w["mach"]
    assert 2.0 <= M <= 10.0 and abs(w["theta"]) <= math.pi / 6
    r1, _, _, p1 = w["ambient"]
    r2, ux, uy, p2 = w["post"]
    S = M * math.sqrt(g * p1 / r1)                      # shock speed, c1 = 1
    un = ux * w["nrm"][0] + uy * w["nrm"][1]            # post-shock normal velocity
    ut = -ux * w["nrm"][1] + uy * w["nrm"][0]
    assert abs(ut) < 1e-12
    # mass, momentum and energy fluxes through the moving shock
    m1, m2 = r1 * (0.0 - S), r2 * (un - S)
    assert m2 == pytest.approx(m1, rel=1e-13)
    assert r2 * (un - S) ** 2 + p2 == pytest.approx(r1 * S ** 2 + p1, rel=1e-13)
    e1 = p1 / (g - 1) / r1 + 0.5 * S ** 2 + p1 / r1
    e2 = p2 / (g - 1) / r2 + 0.5 * (un - S) ** 2 + p2 / r2
    assert e2 == pytest.approx(e1, rel=1e-13)
    assert w["tmax"] == pytest.approx(1.5 / M)
    assert len(w["blobs"]) == 32
    for xb, yb, a, rad in w["blobs"]:
        xsh = 0.5 - (yb - 0.5) * math.tan(w["theta"])
        assert xsh + 0.2 <= xb <= 3.9 and 0.05 <= yb <= 0.95 and 0.1 <= a <= 0.5 and 0.02 <= rad <= 0.1


def test_piecewise_host_generator_hermite_moments():
    """The Hermite moments of the host generator vanish away from the jumps
    and equal the face difference across them."""
    from htx.ic import piecewise_state, quadrant_workload

    w = quadrant_workload(32)
    q = piecewise_state(w, 32, 32, hermite=True)
    assert q.shape == (12, 32, 32)
    dx = 1.0 / 32
    x = (np.arange(32) + 0.5) * dx
    jump_cols = [i for i in range(32) if (x[i] - dx / 2 < w["xs"]) != (x[i] + dx / 2 < w["xs"])]
    smooth_cols = [i for i in range(32) if i not in jump_cols]
    assert np.all(q[4:8][:, :, smooth_cols] == 0.0)
    assert len(jump_cols) == 1
