"""Batch bookkeeping on the CPU: the library's batch layout and launch tables
(htx_batch_tables.h) walked the way the batched launches walk them
(tests/host/batch_host.hip) -- a persistent TENO5 grid much smaller than the
tile count, so every workgroup runs several tiles, across instance
boundaries; Hermite workgroups found through the block -> instance maps;
one control block per instance.  Each instance's result must equal the same
instance run alone (bit for bit: the tile bodies are the same code) and the
FP64 oracle's steps of that instance (within the fast build's division
rounding)."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import params

LIB = os.path.join(ROOT, "tests", "host", "libbatch_host.so")
DP = ctypes.POINTER(ctypes.c_double)
GAMMA = 1.4


@pytest.fixture(scope="module")
def harness():
    if not os.path.exists(LIB):
        pytest.skip("tests/host/libbatch_host.so not built (run __graft_entry__.build())")
    from htx import _lib

    lib = ctypes.CDLL(LIB)
    lib.batchh_steps.argtypes = [ctypes.c_int, ctypes.POINTER(_lib.HtxParams), ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, DP, DP]
    return lib


def instances():
    from htx.ic import SplitMix64, quadrant_params

    rng = SplitMix64(91)
    out = []
    # ragged sides (not multiples of any tile side), both schemes interleaved
    for nx, ny, scheme in [(37, 21, "teno5"), (24, 40, "hteno"), (9, 50, "teno5"), (19, 13, "hteno"),
                           (50, 9, "teno5"), (16, 16, "teno5")]:
        xs, ys, st = quadrant_params(rng)
        m = max(nx, ny)
        out.append(dict(kind="quadrant", nx=nx, ny=ny, scheme=scheme, xl=0.0, xr=nx / m, yb=0.0, yt=ny / m,
                        xs=xs * nx / m, ys=ys * ny / m, states=st, gamma=GAMMA,
                        flux="hllc" if scheme == "teno5" else "rusanov"))
    return out


def run(lib, ws, precision, tx, ty, nblk, nsteps, tmax=10.0, herm_precision="fp64_fast"):
    from htx import _lib
    from htx.batch import params_for
    from htx.ic import piecewise_state

    ps = [params_for(w["xl"], w["xr"], w["nx"], w["yb"], w["yt"], w["ny"], 0.5, tmax, GAMMA, scheme=w["scheme"],
                     flux=w["flux"], precision=precision if w["scheme"] == "teno5" else herm_precision) for w in ws]
    arr = (_lib.HtxParams * len(ps))(*ps)
    qs = [piecewise_state(w, w["nx"], w["ny"], hermite=w["scheme"] == "hteno") for w in ws]
    q = np.ascontiguousarray(np.concatenate([a.reshape(-1) for a in qs]))
    t = np.zeros(len(ws))
    assert lib.batchh_steps(len(ws), arr, tx, ty, nblk, nsteps, q.ctypes.data_as(DP), t.ctypes.data_as(DP)) == 0
    out, pos = [], 0
    for a in qs:
        out.append(q[pos:pos + a.size].reshape(a.shape))
        pos += a.size
    return out, t, qs


@pytest.mark.parametrize("tile", [(16, 16), (32, 8), (8, 32)])
def test_batch_walk_matches_instances_alone(harness, tile):
    ws = instances()
    qb, tb, _ = run(harness, ws, "fp64_fast", *tile, nblk=3, nsteps=2)
    for i, w in enumerate(ws):
        qa, ta, _ = run(harness, [w], "fp64_fast", *tile, nblk=1, nsteps=2)
        np.testing.assert_array_equal(qb[i], qa[0], err_msg=f"instance {i} {w['scheme']} {w['nx']}x{w['ny']}")
        assert tb[i] == ta[0]


def test_batch_walk_fp32_mixed_matches_instances_alone(harness):
    """Both schemes in the FP32-mixed precision."""
    ws = instances()
    qb, tb, _ = run(harness, ws, "fp32_mixed", 16, 16, nblk=4, nsteps=2, herm_precision="fp32_mixed")
    for i, w in enumerate(ws):
        qa, ta, _ = run(harness, [w], "fp32_mixed", 16, 16, nblk=2, nsteps=2, herm_precision="fp32_mixed")
        np.testing.assert_array_equal(qb[i], qa[0])
        assert tb[i] == ta[0]


def test_batch_walk_matches_oracle(orc, harness):
    """Every instance against two oracle SSP-RK3 steps with its own dt; the
    fast build strength-reduces divisions, so within 1e-12 relative L1."""
    ws = instances()
    qb, tb, q0 = run(harness, ws, "fp64_fast", 16, 16, nblk=5, nsteps=2)
    for i, w in enumerate(ws):
        prm = params(w["nx"], w["ny"], w["xl"], w["xr"], w["yb"], w["yt"], GAMMA, "zero_grad", w["flux"], True)
        if w["scheme"] == "teno5":
            qr, tr, n = orc.fv2d_advance(prm, q0[i], 0.5, 0.0, 10.0, 2)
        else:
            qr, tr, n = orc.fvh2d_advance(prm, q0[i], 0.5, 0.0, 10.0, 2)
        assert n == 2
        err = np.abs(qb[i][:4] - qr[:4]).sum() / np.abs(qr[:4]).sum()
        assert err < 1e-12, (i, w["scheme"], err)
        assert abs(tb[i] - tr) <= 1e-14 * tr
