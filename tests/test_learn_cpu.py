"""CPU checks of the learned classifier's definitions: the oracle's face
labels against closed forms, feature invariances, and the numpy MLP gradient
against finite differences."""
import numpy as np
import pytest

import learn_ref as L
from oracle import params


def cons(r, u, v, p, g=1.4):
    return np.stack([r, r * u, r * v, p / (g - 1) + 0.5 * r * (u * u + v * v)])


def test_labels_of_a_density_step(orc):
    """Contact discontinuity in density (u = v = 0, constant p) between cells
    J-1 and J of every row: TENO5 departs from the linear scheme exactly on
    the faces whose 6-cell stencil straddles the jump, c = J-3 .. J+1 (the u-
    window c-2..c+2 or the u+ window c-1..c+3 contains it).  On the flat
    stencils elsewhere rounding can leave one smoothness measure a few ulps
    above zero, so TENO5 cuts a candidate there too (as the reference does),
    but every candidate has the same value and the label stays 1."""
    nx, ny, J = 24, 4, 11
    r = np.where(np.arange(nx) < J, 1.0, 0.3)[None, :] * np.ones((ny, 1))
    q = cons(r, 0 * r, 0 * r, 1.0 + 0 * r)
    prm = params(nx, ny, 0.0, 1.0, 0.0, 1.0, 1.4, "zero_grad", "hllc", True)
    lab = orc.face_labels(prm, q, 1)
    expect = np.ones(nx, dtype=np.uint8)
    expect[J - 3:J + 2] = 0
    for j in range(ny):
        np.testing.assert_array_equal(lab[j], expect)
    # y faces see a constant column: all smooth
    assert orc.face_labels(prm, q, 2).all()


def test_labels_of_smooth_data(orc):
    """A linear density profile and a constant state keep every candidate."""
    nx, ny = 16, 12
    x = (np.arange(nx) + 0.5) / nx
    r = (1.0 + 0.2 * x)[None, :] * np.ones((ny, 1))
    prm = params(nx, ny, 0.0, 1.0, 0.0, 1.0, 1.4, "zero_grad", "hllc", True)
    q = cons(r, 0.1 + 0 * r, -0.2 + 0 * r, 1.0 + 0 * r)
    inner = orc.face_labels(prm, q, 1)[:, 2:-3]   # away from the clamped ends
    assert inner.all()
    assert orc.face_labels(prm, cons(1 + 0 * r, 0 * r, 0 * r, 1 + 0 * r), 1).all()


def test_labels_transpose(orc):
    """x <-> y mirror of a random state swaps the two face families."""
    rng = np.random.default_rng(3)
    n = 20
    r = 1.0 + 0.5 * (rng.random((n, n)) > 0.7)
    u, v, p = rng.normal(0, 0.1, (3, n, n))
    q = cons(r, u, v, 1.0 + 0.1 * np.abs(p))
    qt = np.stack([q[0].T, q[2].T, q[1].T, q[3].T])
    prm = params(n, n, 0.0, 1.0, 0.0, 1.0, 1.4, "periodic", "hllc", True)
    np.testing.assert_array_equal(orc.face_labels(prm, q, 1), orc.face_labels(prm, qt, 2).T)


def test_feature_invariances():
    """Scaling density and pressure by a common constant leaves every
    feature unchanged (up to FP32 rounding: the differences and amplitudes
    are relative to the face's own scales, the contrasts scale-free); a
    uniform state gives all-zero features; the contrasts lie in [0, 1)."""
    rng = np.random.default_rng(5)
    n = 16
    r = 1.0 + 0.3 * rng.random((n, n))
    u, v = rng.normal(0, 0.2, (2, n, n))
    p = 1.0 + 0.3 * rng.random((n, n))
    x0 = L.features(cons(r, u, v, p), 1.4, "periodic", 0)
    x1 = L.features(cons(4.0 * r, u, v, 4.0 * p), 1.4, "periodic", 0)
    np.testing.assert_allclose(x0, x1, rtol=2e-5, atol=2e-6)
    assert (x0[..., 8:] >= 0).all() and (x0[..., 8:] < 1).all()
    assert (x0[..., 4:8] > 0.3).all()
    xz = L.features(cons(1 + 0 * r, 0.3 + 0 * r, 0 * r, 2 + 0 * r), 1.4, "periodic", 1)
    assert np.abs(xz).max() < 1e-6


def test_contrast_of_polynomials_and_jumps():
    """TENO contrast is 0 on quadratics (tau = 0 up to rounding) and close to
    1 across a jump."""
    t = np.arange(5, dtype=np.float64)
    eps = np.float32(1e-30)
    quad = 1.0 + 0.3 * t - 0.05 * t * t
    assert L.teno_contrast(*L.F32(np.diff(quad)), eps) < 1e-3
    jump = np.where(t < 2.5, 1.0, 2.0)
    assert L.teno_contrast(*L.F32(np.diff(jump)), eps) > 0.99


def test_difference_form_smoothness_measures_match_the_reference_polynomials():
    """The contrast's beta_k in first differences equal TENO5's smoothness
    measures written on the values (Schemes_Lagrange.cpp:1013-1026), and its
    beta5 the 5-point quadratic form /5040 (:1029-1034): on random windows
    the contrast agrees with one computed from the oracle's indicators."""
    from oracle import Oracle

    orc = Oracle()
    rng = np.random.default_rng(2)
    for _ in range(200):
        u = rng.normal(0, 1, 6)
        SI, b35 = orc.teno5_indicators(u)
        s = np.array(SI[:3])
        tau = abs(b35[0] - (s[0] + s[2] + 4 * s[1]) / 6)
        lo, hi = s.min() + 1e-30, s.max() + 1e-30
        ref = 1 - (1 + tau / hi) / (1 + tau / lo)
        got = L.teno_contrast(*L.F32(np.diff(u[:5])), np.float32(1e-30))
        assert abs(got - ref) < 2e-4 * (1 + abs(ref))


def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(7)
    prm = rng.normal(0, 0.3, 5057)
    X = rng.normal(0, 1, (40, 12))
    y = (rng.random(40) > 0.4).astype(np.float64)
    loss, g = L.loss_grad(prm, X, y, 3.0)
    for idx in [0, 500, 770, 900, 3000, 4900, 4990, 5056]:
        h = 1e-6
        pp, pm = prm.copy(), prm.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (L.loss_grad(pp, X, y, 3.0)[0] - L.loss_grad(pm, X, y, 3.0)[0]) / (2 * h)
        assert fd == pytest.approx(g[idx], rel=1e-5, abs=1e-9)


def test_gelu_derivative():
    x = np.linspace(-5, 5, 101)
    h = 1e-6
    np.testing.assert_allclose(L.gelu_d(x), (L.gelu(x + h) - L.gelu(x - h)) / (2 * h), rtol=1e-6, atol=1e-8)


def test_rounding_cut_on_flat_stencil_is_not_troubled(orc):
    """The characteristic value 2.5/7 on six equal cells: the JS smoothness
    measure of one candidate rounds to 3e-33 instead of 0, which makes TENO5
    cut it; the reconstructed value is unchanged, so the face is smooth."""
    u = np.full(6, 2.5 * 0.14285714285714285)
    assert list(orc.teno5_flags(u)) != [1] * 6
    assert np.allclose(orc.teno5(u), orc.linear5(u), rtol=1e-15, atol=0)


@pytest.mark.parametrize("lib_name", ["libfv2d_host_fast.so", "libfv2d_host_contract.so"])
def test_tile_features_match_extraction(lib_name):
    """The features the stage kernel's fused classifier forms from its LDS
    tile equal the training extraction's (numpy restatement), also when the
    tile code is compiled with the product's contraction setting and FMA
    available (the device-side counterpart is
    test_gpu_learn.py::test_fused_classifier_flags_match_standalone)."""
    import ctypes
    import os

    from conftest import ROOT
    from htx.ic import shock_blob_workload, piecewise_state

    path = os.path.join(ROOT, "tests", "host", lib_name)
    if not os.path.exists(path):
        pytest.skip(f"tests/host/{lib_name} not built")
    lib = ctypes.CDLL(path)
    FP, DP = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
    lib.fv2dh_tile_features.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, DP, FP, FP]
    w = shock_blob_workload(4096)
    nx, ny = 128, 32
    q = np.ascontiguousarray(piecewise_state(w, nx, ny))
    for d in (0, 1):
        out = [np.zeros((ny, nx, 12), np.float32), np.zeros((ny, nx, 12), np.float32)]
        lib.fv2dh_tile_features(nx, ny, w["gamma"], 0, q.ctypes.data_as(DP), out[0].ctypes.data_as(FP),
                                out[1].ctypes.data_as(FP))
        ref = np.asarray(L.features(q, w["gamma"], "zero_grad", d)).reshape(ny, nx, 12)
        assert np.abs(out[d] - ref).max() < 1e-5
        assert (ref[..., 8:] > 0.5).any()  # the data hold jumps, so contrasts near 1 must appear
