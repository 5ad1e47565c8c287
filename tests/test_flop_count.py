"""The flop count F per cell-update that bench.py prices the stage kernel
against is the op count of the FP64 oracle (oracle/flop_count.py): every
add, sub, mul, div, sqrt and pow of one SSP-RK3 step, extrapolated to an
unbounded grid so that each face counts once."""
import importlib.util
import os

import numpy as np
import pytest

from conftest import ROOT

flop_count = pytest.importorskip("flop_count")


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(flop_count.LIB_COUNT):
        pytest.skip("oracle/liboracle_count.so not built (run __graft_entry__.build())")


def test_counting_build_computes_what_the_oracle_computes(orc, lib):
    """The counting build's arithmetic is the oracle's: one step bit for bit."""
    from oracle import params

    _, cnt = flop_count._oracle()
    n = 24
    prm = params(n, n, 0.0, 10.0, 0.0, 10.0, 1.4, "periodic", "hllc", True)
    q = orc.ic_vortex(prm, (5.0, 5.0), None)
    np.testing.assert_array_equal(cnt.fv2d_rhs(prm, q), orc.fv2d_rhs(prm, q))
    np.testing.assert_array_equal(cnt.fv2d_rk3(prm, q, 0.01), orc.fv2d_rk3(prm, q, 0.01))


def test_ghost_work_is_quadratic_in_one_over_n(lib):
    """The fit a + b/n + c/n^2 through 64, 128 and 256 predicts the count at
    32 to rounding: the oracle's redundant ghost-layer work is polynomial in
    n, so the fitted a is the per-cell count of an unbounded grid."""
    ns = np.array([64.0, 128.0, 256.0])
    C = np.array([flop_count.counts_per_cell_update(int(n))[:5].sum() for n in ns])
    A = np.stack([np.ones(3), 1 / ns, 1 / ns ** 2], axis=1)
    a, b, c = np.linalg.solve(A, C)
    direct = flop_count.counts_per_cell_update(32)[:5].sum()
    assert abs(a + b / 32 + c / 32 ** 2 - direct) < 1.0


@pytest.mark.parametrize("flux", ["hllc", "rusanov"])
def test_bench_prices_the_counted_flops(lib, flux):
    F, cats = flop_count.flops_inf("teno5", flux)
    assert _bench().FLOPS_PER_CELL_UPDATE[flux] == pytest.approx(F, abs=0.1)
    # a characteristic TENO5 step divides a lot: the division count is a
    # fixed share of the work
    assert 0.1 < cats["div"] / F < 0.2
