"""Batches of independent problems: the persistent TENO5 launch, the direct
per-instance launches (strict FP64, FP32 bulk, learned) and the batched
Hermite launches reproduce single-instance solvers, the strict FP64 batch
reproduces the oracle, every built tile shape gives the same result, graph
replays equal eager launches, the autotuner restores the state it times
from."""
import numpy as np
import pytest

from conftest import rel_l1

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def instances():
    from htx.ic import SplitMix64, quadrant_params

    rng = SplitMix64(77)
    out = []
    for nx, ny, scheme in [(40, 24, "teno5"), (64, 64, "teno5"), (17, 33, "teno5"), (32, 32, "hteno"),
                           (24, 40, "hteno"), (100, 20, "teno5")]:
        xs, ys, st = quadrant_params(rng)
        m = max(nx, ny)
        out.append(dict(nx=nx, ny=ny, scheme=scheme, xr=nx / m, yt=ny / m, xs=xs * nx / m, ys=ys * ny / m,
                        states=st, flux="hllc" if scheme == "teno5" else "rusanov"))
    return out


def make_batch(precision="fp64_fast", tmax=0.05, herm_precision="fp64_fast"):
    from htx.batch import Batch, params_for

    inst = instances()
    ps = [params_for(0.0, d["xr"], d["nx"], 0.0, d["yt"], d["ny"], 0.5, tmax, GAMMA, scheme=d["scheme"],
                     flux=d["flux"], precision=precision if d["scheme"] == "teno5" else herm_precision)
          for d in inst]
    b = Batch(ps)
    for i, d in enumerate(inst):
        b.quadrant(i, d["xs"], d["ys"], d["states"])
    return b, inst


def single(d, precision, nsteps, herm_precision="fp64_fast"):
    import htx

    s = htx.System2D(0.0, d["xr"], d["nx"], 0.0, d["yt"], d["ny"], 0.5, 10.0, 0.0, GAMMA, scheme=d["scheme"],
                     flux=d["flux"], bc="zero_grad", precision=precision if d["scheme"] == "teno5" else herm_precision)
    s.quadrant(d["xs"], d["ys"], d["states"])
    s.steps_async(nsteps)
    return s.get_state(), s.time


def test_batch_matches_single_instances():
    print("[case] batch of 6 (4 TENO5, 2 Hermite) vs single solvers, 5 steps", flush=True)
    b, inst = make_batch(tmax=10.0)
    b.steps(5)
    for i, d in enumerate(inst):
        q_ref, t_ref = single(d, "fp64_fast", 5)
        q = b.get_state(i)
        t, n = b.time(i)
        err = rel_l1(q, q_ref)
        print(f"instance {i} {d['scheme']} {d['nx']}x{d['ny']}: rel L1 {err:.3e}, t {t:.6g} (single {t_ref:.6g})",
              flush=True)
        assert n == 5
        assert t == pytest.approx(t_ref, rel=1e-12)
        assert err < 1e-12


def test_batch_fp32_mixed_matches_single_instances():
    """Both schemes in the FP32-mixed precision: the batched launches run the
    single-instance kernels' bodies, so each instance matches its solver run
    alone (5 steps)."""
    print("[case] batch of 6 in fp32_mixed vs single solvers, 5 steps", flush=True)
    b, inst = make_batch("fp32_mixed", tmax=10.0, herm_precision="fp32_mixed")
    b.steps(5)
    for i, d in enumerate(inst):
        q_ref, t_ref = single(d, "fp32_mixed", 5, herm_precision="fp32_mixed")
        q = b.get_state(i)
        t, n = b.time(i)
        err = rel_l1(q, q_ref)
        print(f"instance {i} {d['scheme']} {d['nx']}x{d['ny']}: rel L1 {err:.3e}", flush=True)
        assert n == 5
        assert t == pytest.approx(t_ref, rel=1e-12)
        assert err < 1e-12
    b.close()


def test_batch_fp32_bulk_matches_single_instances():
    """Both schemes in the FP32 bulk precision (TENO5 one launch per
    instance, Hermite-TENO in the multi-instance kernels with FP32 face values
    and fluxes): each instance matches its solver run alone (5 steps), the
    TENO5 ones bit for bit (the same kernel), the Hermite ones to FP32
    rounding (the same bodies compiled into other kernels, where the FP32
    contractions may fall differently)."""
    print("[case] batch of 6 in fp32 vs single solvers, 5 steps", flush=True)
    b, inst = make_batch("fp32", tmax=10.0, herm_precision="fp32")
    b.steps(5)
    for i, d in enumerate(inst):
        q_ref, t_ref = single(d, "fp32", 5, herm_precision="fp32")
        q = b.get_state(i)
        t, n = b.time(i)
        err = rel_l1(q, q_ref)
        print(f"instance {i} {d['scheme']} {d['nx']}x{d['ny']}: rel L1 {err:.3e}", flush=True)
        assert np.isfinite(q).all()
        assert n == 5
        assert t == pytest.approx(t_ref, rel=1e-9)
        assert err < (1e-6 if d["scheme"] == "hteno" else 1e-12)
    b.close()


def test_batch_tile_shapes_agree():
    """The tile shape changes where faces are computed, not what: every shape
    gives the same state bit for bit."""
    print("[case] batch tile shapes 16x16, 32x8, 8x32", flush=True)
    ref = None
    for tx, ty in [(16, 16), (32, 8), (8, 32)]:
        b, inst = make_batch(tmax=10.0)
        b.set_tile(tx, ty)
        b.steps(4)
        qs = [b.get_state(i) for i in range(len(inst))]
        if ref is None:
            ref = qs
        else:
            for q, r in zip(qs, ref):
                np.testing.assert_array_equal(q, r)
        b.close()


@pytest.mark.parametrize("precision", ["fp64_fast", "learned"])
def test_batch_graph_replays_match_eager(precision):
    """A captured step replayed several times, with eager work and
    synchronisation between the replays, equals eager launches after every
    replay."""
    print(f"[case] batch graph replay vs eager ({precision})", flush=True)
    g, inst = make_batch(precision, tmax=10.0)
    e, _ = make_batch(precision, tmax=10.0)
    if precision.startswith("learned"):
        g.load_classifier()
        e.load_classifier()
    g.use_graph(True)
    for k in range(4):
        g.steps(2)
        e.steps(2)
        g.sync()
        e.sync()
        for i in range(len(inst)):
            np.testing.assert_array_equal(g.get_state(i), e.get_state(i))
            assert g.time(i) == e.time(i)


def test_batch_advance_and_tune():
    print("[case] batch advance to per-instance end times, autotuner", flush=True)
    b, inst = make_batch(tmax=0.02)
    before = [b.get_state(i) for i in range(len(inst))]
    shape, ms = b.tune(steps=2)
    print(f"autotuner: {shape}, ms per step {ms}", flush=True)
    assert shape in [(16, 16), (32, 8), (8, 32)]
    assert all(m > 0 for m in ms)
    for i in range(len(inst)):
        np.testing.assert_array_equal(b.get_state(i), before[i])
        assert b.time(i) == (0.0, 0)
    n = b.advance()
    for i in range(len(inst)):
        t, _ = b.time(i)
        assert t == pytest.approx(0.02, rel=1e-14)
        assert np.isfinite(b.get_state(i)).all()
    assert n >= 1


@pytest.mark.parametrize("precision", ["fp64_fast", "fp32_mixed", "learned", "learned_fp32", "fp32"])
def test_batch_many_tiles_per_workgroup(precision):
    """More TENO5 tiles than the persistent grid holds (3 x 19 x 17 tiles of
    16 x 16 against 2 workgroups per CU), so each workgroup walks several
    tiles; the tables are rebuilt by the autotuner in between.  Every
    instance equals its single-instance solver."""
    import htx
    from htx.batch import Batch, params_for
    from htx.ic import SplitMix64, quadrant_params

    print(f"[case] batch with several tiles per workgroup ({precision}), tune, 3 steps", flush=True)
    rng = SplitMix64(5)
    inst = []
    for nx, ny, scheme in [(300, 260, "teno5"), (128, 96, "hteno"), (290, 270, "teno5"), (310, 250, "teno5")]:
        xs, ys, st = quadrant_params(rng)
        m = max(nx, ny)
        inst.append(dict(nx=nx, ny=ny, scheme=scheme, xr=nx / m, yt=ny / m, xs=xs * nx / m, ys=ys * ny / m,
                         states=st, flux="hllc" if scheme == "teno5" else "rusanov"))
    ps = [params_for(0.0, d["xr"], d["nx"], 0.0, d["yt"], d["ny"], 0.5, 10.0, GAMMA, scheme=d["scheme"],
                     flux=d["flux"], precision=precision if d["scheme"] == "teno5" else "fp64_fast") for d in inst]
    b = Batch(ps)
    for i, d in enumerate(inst):
        b.quadrant(i, d["xs"], d["ys"], d["states"])
    if precision.startswith("learned"):
        b.load_classifier()
    shape, _ = b.tune(steps=1)
    print(f"autotuner: {shape}", flush=True)
    b.set_tile(16, 16)
    b.steps(3)
    for i, d in enumerate(inst):
        s = htx.System2D(0.0, d["xr"], d["nx"], 0.0, d["yt"], d["ny"], 0.5, 10.0, 0.0, GAMMA, scheme=d["scheme"],
                         flux=d["flux"], bc="zero_grad",
                         precision=precision if d["scheme"] == "teno5" else "fp64_fast")
        s.quadrant(d["xs"], d["ys"], d["states"])
        s.steps_async(3)
        q_ref, t_ref = s.get_state(), s.time
        s.close()
        q = b.get_state(i)
        t, n = b.time(i)
        err = rel_l1(q, q_ref)
        print(f"instance {i} {d['scheme']} {d['nx']}x{d['ny']}: rel L1 {err:.3e}", flush=True)
        assert n == 3
        assert t == pytest.approx(t_ref, rel=1e-12)
        # the batched and the single-instance kernels are separate builds of
        # the same tile code; FP64 rounding-level differences between them
        # can flip an FP32 cut-off decision of the mixed precision
        assert err < (1e-10 if precision == "fp32_mixed" and d["scheme"] == "teno5" else 1e-12)
    b.close()


def test_batch_strict_fp64_reproduces_the_oracle(orc):
    """The drop-in precision in a batch: TENO5 instances of uneven size run
    the strict stage kernel one launch per instance and equal the oracle's
    SSP-RK3 steps (own dt per instance) bit for bit; strict Hermite
    instances equal their single-instance strict solver bit for bit."""
    import htx
    from htx.batch import Batch, params_for
    from htx.ic import SplitMix64, quadrant_params
    from oracle import params

    print("[case] strict FP64 batch (3 TENO5, 2 Hermite) vs oracle / single strict solvers, 3 steps", flush=True)
    rng = SplitMix64(91)
    inst = []
    for nx, ny, scheme in [(40, 24, "teno5"), (17, 33, "teno5"), (36, 36, "hteno"), (64, 20, "teno5"),
                           (20, 28, "hteno")]:
        xs, ys, st = quadrant_params(rng)
        m = max(nx, ny)
        inst.append(dict(nx=nx, ny=ny, scheme=scheme, xr=nx / m, yt=ny / m, xs=xs * nx / m, ys=ys * ny / m,
                         states=st, flux="hllc" if scheme == "teno5" else "rusanov"))
    ps = [params_for(0.0, d["xr"], d["nx"], 0.0, d["yt"], d["ny"], 0.5, 10.0, GAMMA, scheme=d["scheme"],
                     flux=d["flux"], precision="fp64") for d in inst]
    b = Batch(ps)
    q0 = []
    for i, d in enumerate(inst):
        b.quadrant(i, d["xs"], d["ys"], d["states"])
        q0.append(b.get_state(i))
    b.use_graph(True)
    b.steps(3)
    for i, d in enumerate(inst):
        q = b.get_state(i)
        t, n = b.time(i)
        assert n == 3
        if d["scheme"] == "teno5":
            prm = params(d["nx"], d["ny"], 0.0, d["xr"], 0.0, d["yt"], GAMMA, "zero_grad", "hllc", True)
            qc, tc, k = orc.fv2d_advance(prm, q0[i], 0.5, 0.0, 10.0, 3)
            print(f"instance {i} teno5 {d['nx']}x{d['ny']}: max |q - oracle| {np.abs(q - qc).max():.3e}", flush=True)
            assert t == tc and k == 3
            np.testing.assert_array_equal(q, qc)
        else:
            s = htx.System2D(0.0, d["xr"], d["nx"], 0.0, d["yt"], d["ny"], 0.5, 10.0, 0.0, GAMMA, scheme="hteno",
                             flux="rusanov", bc="zero_grad", precision="fp64")
            s.set_state(q0[i])
            s.steps_async(3)
            q_ref, t_ref = s.get_state(), s.time
            s.close()
            print(f"instance {i} hteno {d['nx']}x{d['ny']}: max |q - single| {np.abs(q - q_ref).max():.3e}",
                  flush=True)
            assert t == t_ref
            np.testing.assert_array_equal(q, q_ref)
    b.close()


def test_batch_serving_sweep_against_fp64():
    """The serving-style sweep (seeded four-quadrant problems, uneven
    log-uniform sizes, TENO5 + HLLC or Hermite-TENO + Rusanov with
    probability 1/2 each) at reduced count, sizes and end time, run in the
    sweep's precision: TENO5 learned_fp32 (FP32 bulk arithmetic plus the
    learned indicator), Hermite fp32 (FP32 bulk).  Every instance reaches its end
    time and stays within 1e-3 relative L1 of the oracle-pinned
    single-instance FP64 solver."""
    import htx
    from htx.batch import Batch, params_for
    from htx.ic import batch_workload

    ws = batch_workload(count=12, lo=32, hi=160, tend=0.03)
    print(f"[case] serving sweep: {len(ws)} instances, learned_fp32 / fp32 vs strict FP64", flush=True)
    ps = [params_for(w["xl"], w["xr"], w["nx"], w["yb"], w["yt"], w["ny"], w["cfl"], w["tmax"], w["gamma"],
                     scheme=w["scheme"], flux=w["flux"], bc=w["bc"],
                     precision="learned_fp32" if w["scheme"] == "teno5" else "fp32") for w in ws]
    b = Batch(ps)
    b.load_classifier()
    for i, w in enumerate(ws):
        b.quadrant(i, w["xs"], w["ys"], w["states"])
    b.use_graph(True)
    b.advance()
    worst = 0.0
    for i, w in enumerate(ws):
        t, n = b.time(i)
        assert t == pytest.approx(w["tmax"], rel=1e-14)
        s = htx.System2D(w["xl"], w["xr"], w["nx"], w["yb"], w["yt"], w["ny"], w["cfl"], w["tmax"], 0.0, w["gamma"],
                         scheme=w["scheme"], flux=w["flux"], bc=w["bc"], precision="fp64")
        s.quadrant(w["xs"], w["ys"], w["states"])
        s.timeStepping(w["tmax"])
        q_ref = s.get_state()
        s.close()
        err = rel_l1(b.get_state(i), q_ref)
        worst = max(worst, err)
        print(f"instance {i} {w['scheme']} {w['nx']}x{w['ny']}: {n} steps, rel L1 vs FP64 {err:.3e}", flush=True)
        assert err < 1e-3
    print(f"worst rel L1 {worst:.3e}", flush=True)
    b.close()
