"""The learned precision with the shipped classifier: it stays finite and
positive on shock data, and its solution error against a 4x finer FP64
solution stays within 10 % of plain FP64 TENO5's, for the density and for
all four conserved fields together (the matched-accuracy target; bench.py
reports the same figures at larger sizes)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["shock_blobs", "quadrant", "vortex"])
def test_learned_matched_accuracy(case):
    from htx.ic import quadrant_workload, shock_blob_workload, vortex_workload
    from htx.study import matched_accuracy

    if case == "shock_blobs":
        w, nx, ny = shock_blob_workload(4096), 256, 64
    elif case == "quadrant":
        w, nx, ny = quadrant_workload(128), 128, 128
    else:
        w, nx, ny = vortex_workload(128), 128, 128
    print(f"[case] learned matched accuracy {case} {nx}x{ny}", flush=True)
    r = matched_accuracy(w, nx, ny, precisions=("fp64", "learned"))
    for p in ("fp64", "learned"):
        print(f"{p}: {r[p]}", flush=True)
    assert r["learned"]["finite"]
    assert r["learned"]["l1_rho_vs_fp64"] < 1.1
    assert r["learned"]["l1_all"] < 1.1 * r["fp64"]["l1_all"]
