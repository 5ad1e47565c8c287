"""GPU checks of the learned troubled-face classifier: training-set
extraction against the oracle's labels and the numpy features, the bf16
matrix-core inference against a bf16-rounded numpy forward pass, the FP32
training gradient against numpy, and that training separates the classes."""
import numpy as np
import pytest

import learn_ref as L
from oracle import params

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def solver_with(q, n, bc, dom=(0.0, 1.0, 0.0, 1.0)):
    import htx

    s = htx.System2D(dom[0], dom[1], n, dom[2], dom[3], n, 0.5, 10.0, 0.0, GAMMA, bc=bc)
    s.set_state(q)
    return s


def quadrant_q(n, seed):
    from htx.ic import SplitMix64, piecewise_state, quadrant_params

    xs, ys, st = quadrant_params(SplitMix64(seed))
    w = dict(kind="quadrant", xl=0.0, xr=1.0, yb=0.0, yt=1.0, gamma=GAMMA, xs=xs, ys=ys, states=st)
    return piecewise_state(w, n, n)


def vortex_q(orc, n):
    from htx.ic import vortex_modes

    prm = params(n, n, 0.0, 10.0, 0.0, 10.0, GAMMA, "periodic", "hllc", True)
    return orc.ic_vortex(prm, (5.0, 5.0), vortex_modes())


@pytest.mark.parametrize("case", ["vortex", "quadrant"])
def test_extraction_matches_oracle_labels_and_features(orc, case):
    from htx.learn import Trainer

    n = 40
    if case == "vortex":
        q, bc, dom = vortex_q(orc, n), "periodic", (0.0, 10.0, 0.0, 10.0)
    else:
        q, bc, dom = quadrant_q(n, 17), "zero_grad", (0.0, 1.0, 0.0, 1.0)
    print(f"[case] extraction {case} n={n}", flush=True)
    s = solver_with(q, n, bc, dom)
    t = Trainer(4 * n * n)
    t.add_state(s)
    nf, ns = t.size()
    assert nf == 2 * n * n
    feat, lab = t.read(0, nf)
    prm = params(n, n, *dom, GAMMA, bc, "hllc", True)
    for d in range(2):
        lab_o = orc.face_labels(prm, q, d + 1).ravel()
        got = lab[d * n * n:(d + 1) * n * n]
        np.testing.assert_array_equal(got, lab_o)
        fr = L.features(q, GAMMA, bc, d).reshape(-1, 12)
        np.testing.assert_allclose(feat[d * n * n:(d + 1) * n * n], fr, rtol=1e-5, atol=2e-6)
    assert ns == int(lab.sum())
    if case == "quadrant":
        assert 0 < ns < nf


def test_mfma_inference_matches_bf16_reference(orc):
    n = 64
    q = quadrant_q(n, 23)
    print(f"[case] MFMA classifier inference n={n}", flush=True)
    s = solver_with(q, n, "zero_grad")
    rng = np.random.default_rng(11)
    prm = (rng.normal(0, 0.4, 5057)).astype(np.float32)
    s.load_classifier(prm, 0.0)
    for d in range(2):
        lg, ft = s.classify(d, features=True)
        fr = L.features(q, GAMMA, "zero_grad", d)
        np.testing.assert_allclose(ft, fr, rtol=1e-5, atol=2e-6)
        ref = L.forward_bf16(prm, ft.reshape(-1, 12)).reshape(n, n)
        err = np.abs(lg - ref)
        print(f"dir {d}: max |logit - bf16 reference| {err.max():.3e}, median {np.median(err):.3e}, "
              f"logit scale {np.abs(ref).max():.2f}", flush=True)
        assert np.median(err) < 1e-3 * (1 + np.abs(ref).max())
        assert err.max() < 3e-2 * (1 + np.abs(ref).max())
        # and the FP64 forward pass without bf16 rounding is close too
        exact = L.forward(prm, ft.reshape(-1, 12)).reshape(n, n)
        assert np.corrcoef(exact.ravel(), lg.ravel())[0, 1] > 0.999


def test_training_gradient_matches_numpy(orc):
    from htx.learn import Trainer

    n = 32
    q = quadrant_q(n, 29)
    print(f"[case] training gradient n={n}", flush=True)
    s = solver_with(q, n, "zero_grad")
    t = Trainer(2 * n * n)
    t.add_state(s)
    rng = np.random.default_rng(13)
    prm = (rng.normal(0, 0.3, 5057)).astype(np.float32)
    t.set_params(prm)
    g, loss = t.gradient(troubled_weight=3.0)
    feat, lab = t.read(0, 2 * n * n)
    loss_r, g_r = L.loss_grad(prm.astype(np.float64), feat, lab.astype(np.float64), 3.0)
    print(f"loss {loss:.6e} (numpy {loss_r:.6e}); gradient rel. L2 error "
          f"{np.linalg.norm(g - g_r) / np.linalg.norm(g_r):.3e}", flush=True)
    assert loss == pytest.approx(loss_r, rel=1e-4)
    assert np.linalg.norm(g - g_r) <= 1e-4 * np.linalg.norm(g_r)


def test_training_separates_classes(orc):
    import htx
    from htx.learn import Trainer

    t = Trainer(1 << 20)
    n = 64
    print(f"[case] training run, snapshots at n={n}", flush=True)
    for seed in (31, 37, 41):
        s = solver_with(quadrant_q(n, seed), n, "zero_grad")
        t.add_state(s)
        s.steps_async(8)
        t.add_state(s)
    sv = htx.System2D(0.0, 10.0, n, 0.0, 10.0, n, 0.5, 10.0, 0.0, GAMMA, bc="periodic")
    from htx.ic import vortex_modes
    sv.isentropicVortex((5.0, 5.0), vortex_modes())
    t.add_state(sv)
    nf, ns = t.size()
    print(f"{nf} faces, {ns} smooth", flush=True)
    t.init(seed=3)
    before = t.evaluate(0.0)
    losses = t.train(epochs=40, batch=4096, lr=3e-3, troubled_weight=2.0, seed=5)
    after = t.evaluate(0.0)
    acc = (after["smooth_as_smooth"] + after["troubled_as_troubled"]) / nf
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}; accuracy {acc:.4f}; {after}", flush=True)
    assert losses[-1] < 0.6 * losses[0]
    assert acc > 0.9
    assert after["troubled_as_smooth"] < 0.2 * (nf - ns)
    assert before != after


def _pair(precision, n, w, threshold=None, params_=None):
    import htx

    s = htx.System2D(w["xl"], w["xr"], n, w["yb"], w["yt"], n, w["cfl"], w["tmax"], 0.0, w["gamma"], flux="hllc",
                     bc=w["bc"], precision=precision)
    if precision == "learned":
        s.load_classifier(params_, threshold)
    s.init_workload(w)
    return s


@pytest.mark.parametrize("case", ["quadrant", "vortex"])
def test_learned_mode_limits(case):
    """With a threshold no logit reaches, precision "learned" sends every face
    but the exactly flat ones (where TENO5 and the linear scheme agree) to
    TENO5: its residual is fp32_mixed's (up to the rounding of another
    compilation of the same code, which on shock data can move an FP32
    cut-off decision on a few faces); with a threshold every logit passes it
    runs the linear scheme everywhere; the face counts cover every face once
    per stage."""
    from conftest import rel_l1
    from htx.ic import quadrant_workload, vortex_workload

    n = 64
    w = quadrant_workload(n) if case == "quadrant" else vortex_workload(n)
    rng = np.random.default_rng(3)
    prm = rng.normal(0, 0.3, 5057).astype(np.float32)
    print(f"[case] learned limits {case} n={n}", flush=True)
    ref = _pair("fp32_mixed", n, w)
    never = _pair("learned", n, w, 1e30, prm)
    r_ref, r_nev = ref.rhs(), never.rhs()
    differ = np.abs(r_nev - r_ref) > 1e-10 * (np.abs(r_ref) + 1.0)
    print(f"residual rel L1 {rel_l1(r_nev, r_ref):.3e}; cells differing {differ.mean():.4f}", flush=True)
    if case == "vortex":
        assert rel_l1(r_nev, r_ref) < 1e-13
    else:
        assert differ.mean() < 0.01
    # faces whose six stencil states are equal are smooth whatever the
    # threshold (TENO5 cuts nothing there); the quadrant data hold many
    q0 = never.get_state()
    flat = sum(int((np.abs(st - st[0:1]).max(axis=(0, 1)) == 0).sum()) for st in
               (L.stencils(q0, w["bc"], 0), L.stencils(q0, w["bc"], 1)))
    assert never.classifier_stats(reset=True) == (2 * n * n, flat, flat)
    if case == "vortex":
        assert flat == 0
        never.steps_async(4)
        assert never.classifier_stats() == (2 * n * n * 3 * 4, 0, 0)
    always = _pair("learned", n, w, -1e30, prm)
    always.steps_async(4)
    f, g, l = always.classifier_stats(reset=True)
    assert f == g == 2 * n * n * 3 * 4
    assert l <= g and (case == "quadrant" or l == g)
    assert always.classifier_stats() == (0, 0, 0)
    q = always.get_state()
    assert np.isfinite(q).all()
    if case == "vortex":
        ref.steps_async(4)
        # smooth data: the linear scheme stays within discretization error of TENO
        assert rel_l1(q, ref.get_state()) < 1e-4


def test_learned_all_smooth_matches_cpu_tiling():
    """Every face flagged smooth (a threshold every logit passes): the GPU
    residual equals the CPU build of the same tile phases with all flags set
    (linear scheme, linear transverse quadrature, FP64 positivity
    fallbacks) to the rounding of FMA contraction."""
    import ctypes
    import os

    from conftest import ROOT, rel_l1
    from htx.ic import vortex_workload

    lib_path = os.path.join(ROOT, "tests", "host", "libfv2d_host_fast.so")
    if not os.path.exists(lib_path):
        pytest.skip("tests/host/libfv2d_host_fast.so not built")
    lib = ctypes.CDLL(lib_path)
    D, DP = ctypes.c_double, ctypes.POINTER(ctypes.c_double)
    lib.fv2dh_rhs_prec.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, DP, DP]
    n = 64
    w = vortex_workload(n)
    print(f"[case] learned all-smooth residual vs CPU tiling, vortex n={n}", flush=True)
    prm = np.random.default_rng(3).normal(0, 0.3, 5057).astype(np.float32)
    s = _pair("learned", n, w, -1e30, prm)
    q = np.ascontiguousarray(s.get_state())
    r_gpu = s.rhs()
    r_cpu = np.empty_like(q)
    lib.fv2dh_rhs_prec(n, n, w["xl"], w["xr"], w["yb"], w["yt"], w["gamma"], 1, 1, 1, 2, q.ctypes.data_as(DP),
                       r_cpu.ctypes.data_as(DP))
    err = rel_l1(r_gpu, r_cpu)
    print(f"rel L1 {err:.3e}", flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_learned_all_smooth_matches_oracle_linear_residual(orc, flux):
    """Every face flagged smooth (threshold -1e30): on the smooth vortex no
    positivity fallback acts, so the learned residual is the reference's
    residual with Lagrange::linear (Schemes_Lagrange.cpp:1219-1233) on every
    face and the transverse quadrature at its linear weights, which the
    oracle computes independently (orc_fv2d_rhs_linear: linear on the six
    characteristic values, the quadrature's candidates with the linear
    weights); they agree to rounding."""
    import htx
    from conftest import rel_l1
    from htx.ic import vortex_modes

    n = 48
    print(f"[case] learned all-smooth vs oracle linear residual, vortex n={n} {flux}", flush=True)
    prm = params(n, n, 0.0, 10.0, 0.0, 10.0, GAMMA, "periodic", flux, True)
    q = orc.ic_vortex(prm, (5.0, 5.0), vortex_modes())
    s = htx.System2D(0.0, 10.0, n, 0.0, 10.0, n, 0.5, 2.0, 0.0, GAMMA, flux=flux, bc="periodic",
                     precision="learned")
    s.load_classifier(np.random.default_rng(3).normal(0, 0.3, 5057).astype(np.float32), -1e30)
    s.set_state(q)
    r = s.rhs()
    f, g, lin = s.classifier_stats(reset=True)
    err = rel_l1(r, orc.fv2d_rhs_linear(prm, q))
    print(f"faces {f} flagged {g} linear {lin}; rel L1 {err:.3e}", flush=True)
    s.close()
    assert f == g == lin == 2 * n * n
    assert err < 1e-12
    # and it is not the TENO5 residual (the comparison can tell them apart)
    assert rel_l1(r, orc.fv2d_rhs(prm, q)) > 1e-6


@pytest.mark.parametrize("case", ["shock_blobs", "quadrant"])
def test_face_flags_match_standalone_and_drive_the_stage(case):
    """The per-step face-flag kernel's flags equal the standalone inference's
    logits against the threshold face by face (only logits within rounding
    of the threshold may differ), and the learned stage kernel's residual
    equals the CPU build of the same tile phases driven by those very flags
    (to the rounding of FMA contraction, which on shock data may move an
    FP32 cut-off decision on a few faces)."""
    import ctypes
    import os

    import htx
    from conftest import ROOT
    from htx.ic import quadrant_workload, shock_blob_workload
    from htx.learn import default_classifier

    w, nx, ny = (shock_blob_workload(4096), 512, 128) if case == "shock_blobs" else (quadrant_workload(256), 256, 256)
    prm, thr = default_classifier()
    print(f"[case] face flags vs standalone inference and CPU tiling, {case} {nx}x{ny}", flush=True)
    s = htx.System2D(w["xl"], w["xr"], nx, w["yb"], w["yt"], ny, w["cfl"], w["tmax"], 0.0, w["gamma"],
                     flux="hllc", bc=w["bc"], precision="learned", classifier=(prm, thr))
    s.init_workload(w)
    r_gpu = s.rhs()
    fx, fy = s.classifier_flags()
    for d, f in ((0, fx[:, 1:nx + 1]), (1, fy[1:ny + 1, :])):
        lg = s.classify(d)
        ref = (lg > thr).astype(np.uint8)
        bad = (f != ref) & (np.abs(lg - thr) > 1e-3)
        print(f"dir {d}: flagged {f.mean():.3f} (standalone {ref.mean():.3f}), mismatches {int(bad.sum())}",
              flush=True)
        assert bad.sum() == 0
    lib_path = os.path.join(ROOT, "tests", "host", "libfv2d_host_fast.so")
    if not os.path.exists(lib_path):
        pytest.skip("tests/host/libfv2d_host_fast.so not built")
    lib = ctypes.CDLL(lib_path)
    D, DP, UP = ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_ubyte)
    lib.fv2dh_rhs_flags.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, DP, DP, UP, UP]
    lib.fv2dh_rhs_prec.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, DP, DP]
    q = np.ascontiguousarray(s.get_state())
    r_cpu = np.empty_like(q)
    lib.fv2dh_rhs_flags(nx, ny, w["xl"], w["xr"], w["yb"], w["yt"], w["gamma"], 0, 1, 1, q.ctypes.data_as(DP),
                        r_cpu.ctypes.data_as(DP), fx.ctypes.data_as(UP), fy.ctypes.data_as(UP))
    s.close()
    # the same comparison for plain FP32-mixed measures how many cells the
    # rounding of FMA contraction alone moves (FP32 cut-off decisions on
    # nearly flat stencils)
    m = htx.System2D(w["xl"], w["xr"], nx, w["yb"], w["yt"], ny, w["cfl"], w["tmax"], 0.0, w["gamma"],
                     flux="hllc", bc=w["bc"], precision="fp32_mixed")
    m.set_state(q)
    rm_gpu = m.rhs()
    m.close()
    rm_cpu = np.empty_like(q)
    lib.fv2dh_rhs_prec(nx, ny, w["xl"], w["xr"], w["yb"], w["yt"], w["gamma"], 0, 1, 1, 1, q.ctypes.data_as(DP),
                       rm_cpu.ctypes.data_as(DP))
    frac = lambda a, b: float((np.abs(a - b) > 1e-9 * (np.abs(b) + 1.0)).mean())  # noqa: E731
    dl, dm = frac(r_gpu, r_cpu), frac(rm_gpu, rm_cpu)
    print(f"residual GPU vs CPU tiling, cells differing: learned {dl:.5f}, fp32_mixed {dm:.5f}", flush=True)
    assert dl <= dm + 0.005


# the features' noise floor on states advanced in FP32 bulk arithmetic
# (learned_fp32), relative to the face's density, pressure and sound speed
FP32_FLOOR = 1e-5


@pytest.mark.parametrize("precision", ["learned", "learned_fp32"])
@pytest.mark.parametrize("bc", ["zero_grad", "periodic"])
def test_face_flags_every_face_vs_numpy(bc, precision):
    """Every flag of the per-step flag kernel, the faces one cell past each
    boundary and the partial work blocks of a 100 x 72 grid included, equals
    the numpy features run through the bf16-rounded forward pass against the
    threshold (1 on a stencil of six equal states), except for logits within
    the inference's rounding band (twice the largest difference between the
    standalone matrix-core logits and that forward pass on the interior
    faces); with periodic boundaries each face past the boundary carries its
    wrapped image's flag.  In learned_fp32 the state carries relative noise of
    3e-6 (the size of FP32-bulk rounding noise) and the numpy features zero
    every difference within the noise floor, which decides some faces
    differently from the features without it."""
    import htx
    from htx.ic import SplitMix64, piecewise_state, quadrant_params
    from htx.learn import default_classifier

    nx, ny = 100, 72
    xs, ys, st = quadrant_params(SplitMix64(31))
    w = dict(kind="quadrant", xl=0.0, xr=1.0, yb=0.0, yt=1.0, gamma=GAMMA, xs=xs, ys=ys, states=st)
    q = np.ascontiguousarray(piecewise_state(w, nx, ny))
    floor = FP32_FLOOR if precision == "learned_fp32" else 0.0
    if floor:
        q = np.ascontiguousarray(q * (1.0 + 3e-6 * np.random.default_rng(5).standard_normal(q.shape)))
    prm, thr = default_classifier()
    print(f"[case] face flags of every face vs numpy, quadrant {nx}x{ny}, {bc}, {precision}", flush=True)
    s = htx.System2D(0.0, 1.0, nx, 0.0, 1.0, ny, 0.5, 10.0, 0.0, GAMMA, bc=bc, precision=precision,
                     classifier=(prm, thr))
    s.set_state(q)
    s.rhs()
    fx, fy = s.classifier_flags()
    for d, f in ((0, fx), (1, fy)):
        stn = L.stencils(q, bc, d, flag_faces=True)
        raw = L.forward_bf16(prm, L.stencil_features(stn, GAMMA, d).reshape(-1, 12)).reshape(f.shape)
        ref = raw
        if floor:  # the floor matters on this state
            ref = L.forward_bf16(prm, L.stencil_features(stn, GAMMA, d, floor).reshape(-1, 12)).reshape(f.shape)
            assert ((raw > thr) != (ref > thr)).sum() > 0
        flat = np.all(stn == stn[0:1], axis=(0, 1))
        # the inference's rounding band from the standalone logits, whose
        # features have no floor
        lg = s.classify(d)
        inner = raw[:, 1:nx + 1] if d == 0 else raw[1:ny + 1, :]
        band = 2.0 * float(np.abs(lg - inner).max()) + 1e-6
        expect = (flat | (ref > thr)).astype(np.uint8)
        decided = flat | (np.abs(ref - thr) > band)
        bad = (f != expect) & decided
        print(f"dir {d}: band {band:.2e}, decided {decided.mean():.4f}, flagged {f.mean():.3f}, "
              f"mismatches {int(bad.sum())}", flush=True)
        assert bad.sum() == 0
        assert decided.mean() > 0.9
    s.close()
    if bc == "periodic":
        assert np.array_equal(fx[:, 0], fx[:, nx]) and np.array_equal(fx[:, nx + 1], fx[:, 1])
        assert np.array_equal(fy[0, :], fy[ny, :]) and np.array_equal(fy[ny + 1, :], fy[1, :])
