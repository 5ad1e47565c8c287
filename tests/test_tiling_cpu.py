"""The stage kernel's own tiling code (htx_fv2d_tile.h), built for the CPU
with contraction off, reproduces the FP64 oracle's residual bit for bit:
halo indexing, boundary mapping (clamp / wrap), face bookkeeping, transverse
quadrature windows and the divergence are all checked without a GPU."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import params

LIB = os.path.join(ROOT, "tests", "host", "libfv2d_host.so")
D = ctypes.c_double
DP = ctypes.POINTER(D)


@pytest.fixture(scope="module")
def harness():
    if not os.path.exists(LIB):
        pytest.skip("tests/host/libfv2d_host.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    lib.fv2dh_rhs.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              DP, DP]
    return lib


def run(lib, nx, ny, dom, bc, flux, quirks, q):
    q = np.ascontiguousarray(q)
    r = np.empty_like(q)
    lib.fv2dh_rhs(nx, ny, *dom, 1.4, {"zero_grad": 0, "periodic": 1}[bc], {"rusanov": 0, "hllc": 1}[flux],
                  int(quirks), q.ctypes.data_as(DP), r.ctypes.data_as(DP))
    return r


def quadrant(nx, ny, seed):
    from test_gpu_fv2d import quadrant_state

    return quadrant_state(nx, ny, seed)


@pytest.mark.parametrize("nx,ny", [(32, 32), (48, 40), (17, 23), (1, 1), (5, 3), (33, 16)])
@pytest.mark.parametrize("bc", ["periodic", "zero_grad"])
@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_tiling_bitwise_vortex(orc, harness, nx, ny, bc, flux):
    from htx.ic import vortex_modes

    if bc == "periodic" and min(nx, ny) < 6:
        pytest.skip("the reference's periodicBc reads unfilled ghosts when a side is shorter than its "
                    "6-cell ghost width (FV_System2D.cpp:117-131); the device wraps exactly instead")
    dom = (0.0, 10.0, 0.0, 10.0)
    prm = params(nx, ny, *dom, 1.4, bc, flux, True)
    q = orc.ic_vortex(prm, (5.0, 5.0), vortex_modes())
    np.testing.assert_array_equal(run(harness, nx, ny, dom, bc, flux, True, q), orc.fv2d_rhs(prm, q))


@pytest.mark.parametrize("quirks", [True, False])
@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_tiling_bitwise_riemann(orc, harness, flux, quirks):
    dom = (0.0, 1.0, 0.0, 1.0)
    for nx, ny, seed in [(40, 36, 3), (64, 20, 8)]:
        prm = params(nx, ny, *dom, 1.4, "zero_grad", flux, quirks)
        q = quadrant(nx, ny, seed)
        np.testing.assert_array_equal(run(harness, nx, ny, dom, "zero_grad", flux, quirks, q),
                                      orc.fv2d_rhs(prm, q))


LIB_FAST = os.path.join(ROOT, "tests", "host", "libfv2d_host_fast.so")


@pytest.fixture(scope="module")
def harness_fast():
    if not os.path.exists(LIB_FAST):
        pytest.skip("tests/host/libfv2d_host_fast.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_FAST)
    lib.fv2dh_rhs.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              DP, DP]
    return lib


@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_strength_reduced_divisions_within_tolerance(orc, harness_fast, flux):
    """The product's strength-reduced divisions (one reciprocal per common
    denominator) move the residual by rounding only: relative L1 <= 1e-13 on
    smooth and shock data, and the stage increment dt * |dL| stays below
    1e-15 of |q|."""
    from htx.ic import vortex_modes

    for nx, ny, dom, bc, q in [
        (64, 64, (0.0, 10.0, 0.0, 10.0), "periodic", None),
        (40, 36, (0.0, 1.0, 0.0, 1.0), "zero_grad", quadrant(40, 36, 3)),
    ]:
        prm = params(nx, ny, *dom, 1.4, bc, flux, True)
        if q is None:
            q = orc.ic_vortex(prm, (5.0, 5.0), vortex_modes())
        r = run(harness_fast, nx, ny, dom, bc, flux, True, q)
        rc = orc.fv2d_rhs(prm, q)
        dt = orc.fv2d_dt(prm, q, 0.5, 0.0, 10.0)
        assert np.abs(r - rc).sum() / np.abs(rc).sum() < 1e-13
        assert dt * np.abs(r - rc).sum() / np.abs(q).sum() < 1e-15


LIB_HERM = os.path.join(ROOT, "tests", "host", "libfvh2d_host.so")


@pytest.fixture(scope="module")
def harness_hermite():
    if not os.path.exists(LIB_HERM):
        pytest.skip("tests/host/libfvh2d_host.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_HERM)
    lib.fvhh_rhs.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             DP, DP, DP]
    return lib


@pytest.mark.parametrize("nx,ny,bc,flux", [(24, 24, "periodic", "rusanov"), (24, 20, "zero_grad", "hllc"),
                                           (17, 13, "zero_grad", "rusanov"), (9, 6, "periodic", "hllc")])
def test_hermite_kernels_cpu_build(orc, harness_hermite, nx, ny, bc, flux):
    """The Hermite kernels' bodies (limiter chains, face reconstructions,
    Gauss-point fluxes, cell update) run on the CPU in launch order with the
    strict build's evaluation (no contraction, divisions kept): the limited
    moments and the residual are bit-identical to the oracle's."""
    from htx.ic import vortex_modes

    dom = (0.0, 10.0, 0.0, 10.0)
    prm = params(nx, ny, *dom, 1.4, bc, flux, True)
    q = orc.ic_vortex_hermite(prm, (5.0, 5.0), vortex_modes())
    r = np.empty_like(q)
    ql = np.empty_like(q)
    harness_hermite.fvhh_rhs(nx, ny, *dom, 1.4, {"zero_grad": 0, "periodic": 1}[bc],
                             {"rusanov": 0, "hllc": 1}[flux], 1, q.ctypes.data_as(DP), r.ctypes.data_as(DP),
                             ql.ctypes.data_as(DP))
    rc, qlc = orc.fvh2d_rhs_lim(prm, q)
    np.testing.assert_array_equal(ql, qlc)
    np.testing.assert_array_equal(r, rc)


@pytest.mark.parametrize("recon,code", [("hweno", 1), ("htenoi", 2), ("hlinear", 3)])
@pytest.mark.parametrize("bc,flux", [("periodic", "rusanov"), ("zero_grad", "hllc")])
def test_hermite_variants_cpu_build(orc, harness_hermite, recon, code, bc, flux):
    """The Hermite solver with hweno, htenoi or the linear Hermite scheme in
    hteno's place: the strict CPU build of the kernels' bodies equals the
    oracle's residual and limited moments bit for bit, on the vortex with
    density modes and on quadrant data."""
    from htx.ic import vortex_modes

    lib = harness_hermite
    lib.fvhh_rhs_variant.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, DP, DP, DP]
    nx, ny = (24, 20) if bc == "periodic" else (19, 16)
    dom = (0.0, 10.0, 0.0, 10.0)
    prm = params(nx, ny, *dom, 1.4, bc, flux, True, recon=recon)
    q = orc.ic_vortex_hermite(prm, (5.0, 5.0), vortex_modes())
    if bc == "zero_grad":
        q[:4] *= np.where(np.arange(nx) < nx // 2, 1.0, 0.6)[None, None, :]
    q = np.ascontiguousarray(q)
    r, ql = np.empty_like(q), np.empty_like(q)
    lib.fvhh_rhs_variant(code, nx, ny, *dom, 1.4, {"zero_grad": 0, "periodic": 1}[bc],
                         {"rusanov": 0, "hllc": 1}[flux], 1, q.ctypes.data_as(DP), r.ctypes.data_as(DP),
                         ql.ctypes.data_as(DP))
    rc, qlc = orc.fvh2d_rhs_lim(prm, q)
    np.testing.assert_array_equal(ql, qlc)
    np.testing.assert_array_equal(r, rc)
    # and it is not the hteno residual
    assert not np.array_equal(r, orc.fvh2d_rhs(params(nx, ny, *dom, 1.4, bc, flux, True), q))


def test_hermite_limiter_split_is_bit_identical(harness_hermite):
    """The limiter split at its chain dependence (the chain-independent
    subexpressions computed ahead, the rest per step) performs the same
    operations as the limiter in one piece: bit-identical results on random
    data of every scale, on flat and linear data, at jumps and at zero."""
    lib = harness_hermite
    lib.fvhh_limiter_split.argtypes = [ctypes.c_int, D, DP, DP, DP]
    rng = np.random.default_rng(21)
    rows = []
    for scale in [1e-8, 1e-3, 1.0, 1e3, 1e8]:
        rows.append(rng.normal(size=(4000, 5)) * scale)
    flat = np.ones((50, 5)) * rng.uniform(0.1, 10, size=(50, 1))
    flat[:, 3:] = 0.0
    lin = np.cumsum(np.ones((50, 5)), axis=1) * rng.uniform(-2, 2, size=(50, 1))
    jump = np.zeros((50, 5))
    jump[:, 2] = rng.uniform(0.5, 5, 50)
    rows += [flat, lin, jump, np.zeros((1, 5))]
    x = np.ascontiguousarray(np.concatenate(rows))
    for dif in [0.01, 0.37, 1.0]:
        whole = np.empty(len(x))
        split = np.empty(len(x))
        lib.fvhh_limiter_split(len(x), dif, x.ctypes.data_as(DP), whole.ctypes.data_as(DP), split.ctypes.data_as(DP))
        np.testing.assert_array_equal(whole.view(np.int64), split.view(np.int64))


LIB_HERM_FAST = os.path.join(ROOT, "tests", "host", "libfvh2d_host_fast.so")


@pytest.mark.parametrize("bc,flux", [("periodic", "rusanov"), ("zero_grad", "hllc")])
def test_hermite_kernels_cpu_build_fast(orc, bc, flux):
    """The fast build's Hermite bodies (reassociated limiter step, shared
    reciprocals) stay within 1e-12 relative L1 of the oracle, limited
    moments and residual, on the vortex."""
    if not os.path.exists(LIB_HERM_FAST):
        pytest.skip("tests/host/libfvh2d_host_fast.so not built (run __graft_entry__.build())")
    from htx.ic import vortex_modes

    lib = ctypes.CDLL(LIB_HERM_FAST)
    lib.fvhh_rhs.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             DP, DP, DP]
    nx, ny, dom = 24, 20, (0.0, 10.0, 0.0, 10.0)
    prm = params(nx, ny, *dom, 1.4, bc, flux, True)
    q = orc.ic_vortex_hermite(prm, (5.0, 5.0), vortex_modes())
    r = np.empty_like(q)
    ql = np.empty_like(q)
    lib.fvhh_rhs(nx, ny, *dom, 1.4, {"zero_grad": 0, "periodic": 1}[bc], {"rusanov": 0, "hllc": 1}[flux], 1,
                 q.ctypes.data_as(DP), r.ctypes.data_as(DP), ql.ctypes.data_as(DP))
    rc, qlc = orc.fvh2d_rhs_lim(prm, q)
    assert np.abs(ql - qlc).sum() / np.abs(qlc).sum() < 1e-12
    assert np.abs(r - rc).sum() / np.abs(rc).sum() < 1e-12


@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_hermite_fp32_mixed_residual_close_to_fp64(orc, flux):
    """The Hermite FP32-mixed bodies (HTENO and TENO5 cut-off rules and the
    transverse WENO weights in FP32, every value, projection and flux in
    FP64): the 12-field residual within 2e-8 relative L1 of the FP64 oracle
    on the vortex (measured 5.2e-9) and on quadrant data with zero moments
    (measured 2.7e-10); the FP64 fast bodies on the same data are within
    1e-13.  An indexing or term mistake shows at O(1e-3) or more."""
    if not os.path.exists(LIB_HERM_FAST):
        pytest.skip("tests/host/libfvh2d_host_fast.so not built (run __graft_entry__.build())")
    from htx.ic import vortex_modes

    lib = ctypes.CDLL(LIB_HERM_FAST)
    lib.fvhh_rhs_mixed.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, DP, DP, DP]
    for case in ("vortex", "quadrant"):
        if case == "vortex":
            nx, ny, dom, bc = 32, 28, (0.0, 10.0, 0.0, 10.0), "periodic"
            prm = params(nx, ny, *dom, 1.4, bc, flux, True)
            q = orc.ic_vortex_hermite(prm, (5.0, 5.0), vortex_modes())
        else:
            nx, ny, dom, bc = 40, 36, (0.0, 1.0, 0.0, 1.0), "zero_grad"
            prm = params(nx, ny, *dom, 1.4, bc, flux, True)
            q = np.concatenate([quadrant(nx, ny, 9), np.zeros((8, ny, nx))])
        q = np.ascontiguousarray(q)
        r = np.empty_like(q)
        ql = np.empty_like(q)
        lib.fvhh_rhs_mixed(nx, ny, *dom, 1.4, {"zero_grad": 0, "periodic": 1}[bc], {"rusanov": 0, "hllc": 1}[flux],
                           1, q.ctypes.data_as(DP), r.ctypes.data_as(DP), ql.ctypes.data_as(DP))
        rc = orc.fvh2d_rhs(prm, q)
        err = np.abs(r - rc).sum() / np.abs(rc).sum()
        assert err < 2e-8, (case, err)


@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_hermite_fp32_bulk_residual_near_fp64(orc, flux):
    """The Hermite FP32 bulk bodies (FP32 face values, transverse quadrature,
    back-projection, u and moment fluxes; FP64 Roe eigensystem, Gauss sums,
    update): each field group of the residual within 1e-4 relative L1 of
    the FP64 oracle on the vortex and on quadrant data with zero moments,
    and a uniform state's residual exactly zero."""
    if not os.path.exists(LIB_HERM_FAST):
        pytest.skip("tests/host/libfvh2d_host_fast.so not built (run __graft_entry__.build())")
    from htx.ic import vortex_modes

    lib = ctypes.CDLL(LIB_HERM_FAST)
    lib.fvhh_rhs_bulk32.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, DP, DP, DP]

    def rhs(nx, ny, dom, bc, q):
        q = np.ascontiguousarray(q)
        r = np.empty_like(q)
        ql = np.empty_like(q)
        lib.fvhh_rhs_bulk32(nx, ny, *dom, 1.4, {"zero_grad": 0, "periodic": 1}[bc], {"rusanov": 0, "hllc": 1}[flux],
                            1, q.ctypes.data_as(DP), r.ctypes.data_as(DP), ql.ctypes.data_as(DP))
        return r

    for case in ("vortex", "quadrant"):
        if case == "vortex":
            nx, ny, dom, bc = 32, 28, (0.0, 10.0, 0.0, 10.0), "periodic"
            prm = params(nx, ny, *dom, 1.4, bc, flux, True)
            q = orc.ic_vortex_hermite(prm, (5.0, 5.0), vortex_modes())
        else:
            nx, ny, dom, bc = 40, 36, (0.0, 1.0, 0.0, 1.0), "zero_grad"
            prm = params(nx, ny, *dom, 1.4, bc, flux, True)
            q = np.concatenate([quadrant(nx, ny, 9), np.zeros((8, ny, nx))])
        r = rhs(nx, ny, dom, bc, q)
        rc = orc.fvh2d_rhs(prm, q)
        for g in range(0, 12, 4):
            if not np.abs(rc[g:g + 4]).any():
                assert np.abs(r[g:g + 4]).max() < 1e-6 * np.abs(rc[:4]).max(), (case, g)
                continue
            err = np.abs(r[g:g + 4] - rc[g:g + 4]).sum() / np.abs(rc[g:g + 4]).sum()
            assert 0.0 < err < 1e-4, (case, g, err)
    u = np.zeros((12, 20, 24))
    u[:4] = np.array([1.2, 0.3, -0.4, 2.5])[:, None, None]
    assert np.abs(rhs(24, 20, (0.0, 1.0, 0.0, 1.0), "periodic", u)).max() == 0.0


@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_fp32_mixed_residual_close_to_fp64(orc, harness_fast, flux):
    """FP32-mixed mode (cut-off decisions and nonlinear weights in FP32,
    values, projections, fluxes in FP64): the residual stays within 1e-8
    relative L1 of the FP64 oracle on smooth and on shock data."""
    from htx.ic import vortex_modes

    harness_fast.fv2dh_rhs_prec.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, DP, DP]
    for nx, ny, dom, bc, q in [
        (64, 64, (0.0, 10.0, 0.0, 10.0), "periodic", None),
        (96, 80, (0.0, 1.0, 0.0, 1.0), "zero_grad", quadrant(96, 80, 9)),
    ]:
        prm = params(nx, ny, *dom, 1.4, bc, flux, True)
        if q is None:
            q = orc.ic_vortex(prm, (5.0, 5.0), vortex_modes())
        q = np.ascontiguousarray(q)
        r = np.empty_like(q)
        harness_fast.fv2dh_rhs_prec(nx, ny, *dom, 1.4, {"zero_grad": 0, "periodic": 1}[bc],
                                    {"rusanov": 0, "hllc": 1}[flux], 1, 1, q.ctypes.data_as(DP), r.ctypes.data_as(DP))
        rc = orc.fv2d_rhs(prm, q)
        assert np.abs(r - rc).sum() / np.abs(rc).sum() < 1e-8


def test_fp32_mixed_cutoff_never_overflows(harness_fast):
    """The FP32 cut-off rule on (nearly) flat stencils: the raw weights
    (1 + tau / (beta + eps))^6 can overflow FP32 there (their sum can, even
    when each one fits), which once cut every candidate and returned
    catchnan's 1e-15.  The rule is evaluated on ratios to the largest
    weight, so the face value stays the flat value."""
    harness_fast.fv2dh_teno5.argtypes = [DP, ctypes.c_int, DP]
    rng = np.random.default_rng(9)
    out = np.empty(2)
    worst = 0.0
    for k in range(20000):
        base = 10.0 ** rng.uniform(-3, 3)
        u = base * (1.0 + rng.integers(-2, 3, 6) * 2.0 ** -52)
        harness_fast.fv2dh_teno5(u.ctypes.data_as(DP), 1, out.ctypes.data_as(DP))
        worst = max(worst, np.abs(out - base).max() / base)
    assert worst < 1e-14
    # stencils from a shock-plus-blobs state ahead of the shock: flat up to a
    # ramp of relative size 1e-12 .. 1e-9, where every candidate was cut
    for u in ([2.1450916869464454, 2.1450916869464436, 2.1450916869464436, 2.145091686946442,
               2.1450916869431627, 2.1450916860808373],
              [2.1450916869464542, 2.1450916869464542, 2.1450916869464454, 2.1450916869464454,
               2.1450916869464542, 2.1450916869313623],
              [2.1450916869464507, 2.145091686946447, 2.1450916869464454, 2.1450916869464525,
               2.1450916869313623, 2.1450916851272535]):
        u = np.array(u)
        ref = np.empty(2)
        harness_fast.fv2dh_teno5(u.ctypes.data_as(DP), 0, ref.ctypes.data_as(DP))
        harness_fast.fv2dh_teno5(u.ctypes.data_as(DP), 1, out.ctypes.data_as(DP))
        assert np.all(np.abs(out - ref) < 1e-9 * u[0]), (out, ref)


def test_teno5_decision_on_the_threshold_uses_pow(orc, harness):
    """A stencil whose third left-candidate weight ratio lies on C_T = 5e-4 to
    the last bit: the reference's (a)^6 by pow gives 0.0004999999999999999
    (cut), a cube squared gives exactly 0.0005 (kept).  The strict build
    forms the powers as cubes squared but recomputes them with pow where a
    ratio lies within 1e-13 of C_T, so it reconstructs what the oracle (and
    the reference) do."""
    import math

    u = np.array([float.fromhex(h) for h in (
        "0x1.46cb19df16295p+2", "0x1.34e5a6f5e6534p-2", "-0x1.7d87893f7dc1cp-1", "-0x1.7f2e1e186e38ep+1",
        "0x1.842efadc39048p+1", "0x1.2d452282a474ep+1")])
    SI, b = orc.teno5_indicators(u)
    s = list(SI[:3])
    tau = abs(b[0] - (s[0] + s[2] + 4.0 * s[1]) / 6.0)
    a = [1.0 + tau / (x + 1e-40) for x in s]
    gp = [math.pow(x, 6.0) for x in a]
    gc = [(x * x * x) * (x * x * x) for x in a]
    assert gp[2] / ((gp[0] + gp[1]) + gp[2]) < 5e-4 <= gc[2] / ((gc[0] + gc[1]) + gc[2])
    assert orc.teno5_flags(u)[2] == 0
    harness.fv2dh_teno5.argtypes = [DP, ctypes.c_int, DP]
    out = np.empty(2)
    harness.fv2dh_teno5(u.ctypes.data_as(DP), 0, out.ctypes.data_as(DP))
    assert out.tobytes() == orc.teno5(u).tobytes()


def test_linear_quadrature_exact_for_quartics(harness_fast):
    """The learned path's transverse quadrature on smooth windows (weno3q
    with linear weights, folded into one 5-point filter per Gauss point):
    cell averages of any quartic give its exact values at +xi, 0, -xi;
    quadratics agree with the oracle's nonlinear quadrature."""
    from test_oracle import SQ15, cell_avg_poly, poly

    harness_fast.fv2dh_linquad.argtypes = [DP, DP]
    out = np.empty(3)
    xi = SQ15 / 10.0
    rng = np.random.default_rng(4)
    for deg in range(5):
        coef = rng.uniform(-2, 2, deg + 1)
        u = np.array([cell_avg_poly(coef, m - 2.5, m - 1.5) for m in range(5)])
        harness_fast.fv2dh_linquad(u.ctypes.data_as(DP), out.ctypes.data_as(DP))
        np.testing.assert_allclose(out, [poly(coef, xi), poly(coef, 0.0), poly(coef, -xi)], rtol=0, atol=5e-14)
    # a sixth-degree monomial is not reproduced (the filter is 5th order, not more)
    u = np.array([cell_avg_poly([0, 0, 0, 0, 0, 0, 1.0], m - 2.5, m - 1.5) for m in range(5)])
    harness_fast.fv2dh_linquad(u.ctypes.data_as(DP), out.ctypes.data_as(DP))
    assert abs(out[0] - xi ** 6) > 1e-3


def _rhs_prec(lib, nx, ny, dom, bc, flux, mixed, q):
    lib.fv2dh_rhs_prec.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, DP, DP]
    q = np.ascontiguousarray(q)
    r = np.empty_like(q)
    lib.fv2dh_rhs_prec(nx, ny, *dom, 1.4, {"zero_grad": 0, "periodic": 1}[bc], {"rusanov": 0, "hllc": 1}[flux], 1,
                       mixed, q.ctypes.data_as(DP), r.ctypes.data_as(DP))
    return r


@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_learned_all_smooth_residual(orc, harness_fast, flux):
    """The learned path with every face flagged smooth (linear scheme normal
    to the face, linear transverse quadrature, positivity fallbacks): a
    uniform state has zero residual, the periodic residual conserves every
    field, and on the smooth vortex its distance to plain TENO5's residual
    is discretisation error, shrinking at high order (> 16x per halving of
    the cell size)."""
    from htx.ic import vortex_modes

    dom = (0.0, 10.0, 0.0, 10.0)
    errs = []
    for n in (48, 96):
        prm = params(n, n, *dom, 1.4, "periodic", flux, True)
        q = orc.ic_vortex(prm, (5.0, 5.0), vortex_modes())
        r = _rhs_prec(harness_fast, n, n, dom, "periodic", flux, 2, q)
        rc = orc.fv2d_rhs(prm, q)
        errs.append(np.abs(r - rc).sum() / np.abs(rc).sum())
        for f in range(4):
            assert abs(r[f].sum()) <= 1e-12 * np.abs(r[f]).sum()
    assert 0.0 < errs[1] < 1e-4 and errs[0] / errs[1] > 16.0, errs
    n = 40
    u = np.ones((4, n, n)) * np.array([1.2, 0.3, -0.4, 2.5])[:, None, None]
    assert np.abs(_rhs_prec(harness_fast, n, n, dom, "periodic", flux, 2, u)).max() < 1e-13


@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_learned_all_smooth_equals_oracle_linear_residual(orc, harness_fast, flux):
    """With every face flagged smooth the tile phases compute the oracle's
    all-linear residual (orc_fv2d_rhs_linear: Lagrange::linear on the
    characteristic values, the transverse quadrature at its linear weights)
    to rounding: the kernel reconstructs the conserved variables and projects
    the two face states, and folds the linear quadrature into one filter per
    Gauss point, which are the same linear maps.  (Periodic boundaries: with
    zero-gradient ones the faces past the boundary keep TENO5, and the
    wavefronts whose transverse windows reach them the nonlinear
    quadrature.)"""
    from htx.ic import vortex_modes

    dom = (0.0, 10.0, 0.0, 10.0)
    n, bc = 40, "periodic"
    prm = params(n, n, *dom, 1.4, bc, flux, True)
    q = orc.ic_vortex(prm, (5.0, 5.0), vortex_modes())
    r = _rhs_prec(harness_fast, n, n, dom, bc, flux, 2, q)
    rl = orc.fv2d_rhs_linear(prm, q)
    assert np.abs(r - rl).sum() / np.abs(rl).sum() < 1e-12


def _rhs_flags(lib, nx, ny, dom, bc, flux, q, fx, fy):
    UP = ctypes.POINTER(ctypes.c_ubyte)
    lib.fv2dh_rhs_flags.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, DP, DP, UP, UP]
    q = np.ascontiguousarray(q)
    fx, fy = np.ascontiguousarray(fx, dtype=np.uint8), np.ascontiguousarray(fy, dtype=np.uint8)
    r = np.empty_like(q)
    lib.fv2dh_rhs_flags(nx, ny, *dom, 1.4, {"zero_grad": 0, "periodic": 1}[bc], {"rusanov": 0, "hllc": 1}[flux], 1,
                        q.ctypes.data_as(DP), r.ctypes.data_as(DP), fx.ctypes.data_as(UP), fy.ctypes.data_as(UP))
    return r


def _periodic_flag_arrays(cx, cy):
    """Extended face-flag arrays of the stage kernel from per-face flags of a
    periodic grid: cx[j, i] for x face i | i+1, cy[j, i] for y face j | j+1."""
    fx = np.concatenate([cx[:, -1:], cx, cx[:, :1]], axis=1)
    fy = np.concatenate([cy[-1:, :], cy, cy[:1, :]], axis=0)
    return fx, fy


@pytest.mark.parametrize("bc", ["periodic", "zero_grad"])
def test_learned_flag_reaches_its_face(harness_fast, bc):
    """One face flagged smooth, every other face TENO5, on rough data (a
    uniform flow with 3 % cell-to-cell noise, where TENO5 and the linear
    scheme differ on every face): the residual changes only in the cells
    whose fluxes use that recon face (x face i | i+1 of row j: cells i and
    i+1 of rows j-2 .. j+2, through the face's own flux and the transverse
    quadratures of its neighbours; a y face: the transpose), and most in the
    two cells of the face itself (the centre of its own transverse window),
    wherever the face sits relative to the 16 x 15 tiles and the boundary."""
    nx, ny, dom = 48, 40, (0.0, 1.0, 0.0, 1.0)
    rng = np.random.default_rng(21)
    w = np.array([1.0, 0.3, -0.2, 1.0])[:, None, None] * (1.0 + 0.03 * rng.standard_normal((4, ny, nx)))
    q = np.array([w[0], w[0] * w[1], w[0] * w[2], w[3] / 0.4 + 0.5 * w[0] * (w[1] ** 2 + w[2] ** 2)])
    zx, zy = np.zeros((ny, nx + 2), np.uint8), np.zeros((ny + 2, nx), np.uint8)
    r0 = _rhs_flags(harness_fast, nx, ny, dom, bc, "hllc", q, zx, zy)
    for d, i, j in [(0, 20, 21), (0, 15, 16), (0, 16, 31), (0, 31, 15), (0, 1, 3), (0, 46, 38), (1, 17, 15),
                    (1, 31, 20), (1, 5, 0), (1, 40, 16), (1, 47, 38)]:
        fx, fy = zx.copy(), zy.copy()
        if d == 0:
            fx[j, i + 1] = 1
        else:
            fy[j + 1, i] = 1
        r = _rhs_flags(harness_fast, nx, ny, dom, bc, "hllc", q, fx, fy)
        delta = np.abs(r - r0).max(axis=0)
        jj, ii = np.nonzero(delta > 0.0)
        jm, im = np.unravel_index(np.argmax(delta), delta.shape)
        near = lambda c, n, span: {(c + k) % n for k in span}  # noqa: E731  (periodic wrap; no-op inside)
        if d == 0:
            assert set(ii) <= near(i, nx, (0, 1)) and set(jj) <= near(j, ny, range(-2, 3)), (d, i, j, set(ii), set(jj))
            assert jm == j and im in near(i, nx, (0, 1)), (d, i, j, jm, im)
        else:
            assert set(jj) <= near(j, ny, (0, 1)) and set(ii) <= near(i, nx, range(-2, 3)), (d, i, j, set(ii), set(jj))
            assert im == i and jm in near(j, ny, (0, 1)), (d, i, j, jm, im)


def test_learned_flag_placement_is_shift_equivariant(orc, harness_fast):
    """The learned path reads each recon face's flag from the face-flag arrays
    through the tile's index mapping: on a periodic grid, shifting the state
    and the per-face flags by an offset that is not a multiple of the tile
    size (so every face lands in another tile, at another position, or in
    another tile's halo) shifts the residual bit for bit."""
    from htx.ic import vortex_modes

    nx, ny, dom = 48, 40, (0.0, 10.0, 0.0, 10.0)
    prm = params(nx, ny, *dom, 1.4, "periodic", "hllc", True)
    q = orc.ic_vortex(prm, (5.0, 5.0), vortex_modes())
    rng = np.random.default_rng(12)
    cx = (rng.random((ny, nx)) < 0.7).astype(np.uint8)
    cy = (rng.random((ny, nx)) < 0.7).astype(np.uint8)
    r = _rhs_flags(harness_fast, nx, ny, dom, "periodic", "hllc", q, *_periodic_flag_arrays(cx, cy))
    sj, si = 7, 5
    roll = lambda a: np.roll(np.roll(a, sj, axis=-2), si, axis=-1)  # noqa: E731
    r2 = _rhs_flags(harness_fast, nx, ny, dom, "periodic", "hllc", roll(q), *_periodic_flag_arrays(roll(cx), roll(cy)))
    np.testing.assert_array_equal(r2, roll(r))
    # and the flags matter: all-TENO flags give another residual
    r0 = _rhs_flags(harness_fast, nx, ny, dom, "periodic", "hllc", q, *_periodic_flag_arrays(0 * cx, 0 * cy))
    assert np.abs(r0 - r).max() > 0.0


@pytest.mark.parametrize("recon,mode", [("teno5i", 5), ("weno5js", 6), ("weno5z", 7)])
@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_tiling_bitwise_reconstruction_variants(orc, harness, recon, mode, flux):
    """The 2D path with teno5i, weno5JS or weno5Z in TENO5's place: the
    strict CPU build of the tile phases equals the oracle's residual with
    the same reconstruction bit for bit, on the vortex (periodic) and on
    quadrant data with seeded noise (zero gradient); on the noisy data the
    residual differs from the TENO5 one (on clean jumps teno5i makes TENO5's
    cut-off decisions)."""
    from htx.ic import vortex_modes

    rng = np.random.default_rng(4)
    for nx, ny, bc, dom in [(48, 40, "periodic", (0.0, 10.0, 0.0, 10.0)), (40, 36, "zero_grad", (0.0, 1.0, 0.0, 1.0))]:
        prm = params(nx, ny, *dom, 1.4, bc, flux, True, recon=recon)
        if bc == "periodic":
            q = orc.ic_vortex(prm, (5.0, 5.0), vortex_modes())
        else:
            q = quadrant(nx, ny, 3)
            q[0] *= 1.0 + 0.05 * rng.random((ny, nx))
        got = _rhs_prec(harness, nx, ny, dom, bc, flux, mode, q)
        np.testing.assert_array_equal(got, orc.fv2d_rhs(prm, q))
        if bc == "zero_grad":
            assert not np.array_equal(got, orc.fv2d_rhs(params(nx, ny, *dom, 1.4, bc, flux, True), q))


VARIANTS = ["weno5n", "teno5n", "teno5a", "weno5is", "weno5is_ext", "weno5ismod", "weno5z_diff", "teno5_diff",
            "linear"]


@pytest.mark.parametrize("recon", VARIANTS)
def test_tiling_bitwise_remaining_reconstructions(orc, harness, recon):
    """The 2D path with each remaining 5-point routine of Schemes::Lagrange in
    TENO5's place (the run-time variant instantiation of the stage kernel, on
    its 16 x 15 tiles): the strict CPU build of the tile phases equals the
    oracle's residual bit for bit, periodic vortex and noisy quadrants."""
    from htx.ic import vortex_modes
    from oracle import RECON

    rng = np.random.default_rng(5)
    for nx, ny, bc, dom, flux in [(48, 40, "periodic", (0.0, 10.0, 0.0, 10.0), "hllc"),
                                  (40, 36, "zero_grad", (0.0, 1.0, 0.0, 1.0), "rusanov")]:
        prm = params(nx, ny, *dom, 1.4, bc, flux, True, recon=recon)
        if bc == "periodic":
            q = orc.ic_vortex(prm, (5.0, 5.0), vortex_modes())
        else:
            q = quadrant(nx, ny, 3)
            q[0] *= 1.0 + 0.05 * rng.random((ny, nx))
        got = _rhs_prec(harness, nx, ny, dom, bc, flux, 100 + RECON[recon], q)
        np.testing.assert_array_equal(got, orc.fv2d_rhs(prm, q))


@pytest.mark.parametrize("mode,shape", [(8, "16x15"), (9, "15x16")])
@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_tiling_bitwise_other_tile_shapes(orc, harness, mode, shape, flux):
    """The default 16 x 15 tile (and its transpose) of the strict build:
    residual bit-identical to the oracle on ragged grids of both boundary
    conditions, including grids smaller than one tile."""
    from htx.ic import vortex_modes

    for nx, ny, bc in [(48, 40, "periodic"), (33, 31, "zero_grad"), (17, 46, "periodic"), (5, 3, "zero_grad"),
                       (40, 70, "zero_grad")]:
        dom = (0.0, 10.0, 0.0, 10.0)
        prm = params(nx, ny, *dom, 1.4, bc, flux, True)
        q = orc.ic_vortex(prm, (5.0, 5.0), vortex_modes())
        np.testing.assert_array_equal(_rhs_prec(harness, nx, ny, dom, bc, flux, mode, q), orc.fv2d_rhs(prm, q))
