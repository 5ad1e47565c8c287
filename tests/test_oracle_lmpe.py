"""The LMPE troubled-cell detector (FVHermite_System1D.cpp:239-257): the
oracle's literal sequential sweep pinned by hand-computed cases, and the
kernel's run-parity form of the same sweep (htx_fv1d.h fv1d_lmpe, built for
the CPU in tests/host/fv1d_host.hip) equal to the oracle on many states."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "tests", "host", "libfv1d_host.so")


def _flat(nx, u0):
    q = np.empty((3, nx))
    q[0], q[1], q[2] = u0, 0.0, 2.5
    return q


def test_jump_flags_the_three_cells_around_it(orc):
    """u = 1 left of cell m, 0.2 from m on (dx = 1/40): the test at m-1 gives
    A = |-1 + 2 - 0.2| / 2 = 0.4 > dx/2 and xc = (-2 + 3 - 0.2) / 0.8 dx = dx,
    so m-2, m-1, m are flagged; m is then skipped and m+1 sees a flat
    stencil.  With zero-gradient ghosts nothing else is flagged."""
    nx, m = 40, 17
    u = np.where(np.arange(nx) < m, 1.0, 0.2)
    f = orc.fv1d_lmpe(nx, 0.0, 1.0, _flat(nx, u), bc="zero_grad")
    assert set(np.flatnonzero(f[0])) == {m - 2, m - 1, m}
    assert not f[1:].any()


def test_skip_rule_changes_the_flags(orc):
    """u = 1 on cells 5 and 6, 0 elsewhere (dx = 1/20, periodic): the tests
    at 4, 5, 6 and 7 all succeed (A = 0.5; xc = dx, 2 dx, dx, 2 dx).  The
    sweep acts at 4 (flagging 3, 4, 5), skips 5, acts at 6 (flagging 5, 6,
    7) and skips 7, so cell 8 stays clear; without the skip rule 7 would
    flag it."""
    nx = 20
    u = np.zeros(nx)
    u[5] = u[6] = 1.0
    f = orc.fv1d_lmpe(nx, 0.0, 1.0, _flat(nx, u), bc="periodic")
    assert list(np.flatnonzero(f[0])) == [3, 4, 5, 6, 7]


def test_periodic_ghosts_are_swept(orc):
    """The sweep starts two ghost cells left of cell 0: a jump across the
    periodic seam (u = 0.2 on the last cell only) is tested at the ghost
    copies too and flags cells 0 and 1 from the left ghost's test."""
    nx = 30
    u = np.ones(nx)
    u[-1] = 0.2
    fp = orc.fv1d_lmpe(nx, 0.0, 1.0, _flat(nx, u), bc="periodic")
    fz = orc.fv1d_lmpe(nx, 0.0, 1.0, _flat(nx, u), bc="zero_grad")
    assert fp[0, 0] == 1 and fz[0, 0] == 0
    assert fp[0, nx - 2] == 1 and fz[0, nx - 2] == 1


def test_smooth_data_is_not_flagged(orc):
    """A smooth wave: A ~ dx^2 |u''| / 2 stays below dx / 2."""
    nx = 200
    x = (np.arange(nx) + 0.5) / nx
    q = np.empty((3, nx))
    q[0] = 1.0 + 0.5 * np.sin(2 * np.pi * x)
    q[1] = q[0] * 0.3
    q[2] = 2.5 + 0.1 * np.cos(2 * np.pi * x)
    assert not orc.fv1d_lmpe(nx, 0.0, 1.0, q, bc="periodic").any()


@pytest.fixture(scope="module")
def harness():
    if not os.path.exists(LIB):
        pytest.skip("tests/host/libfv1d_host.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    lib.fv1dh_lmpe.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_ubyte)]
    return lib


@pytest.mark.parametrize("bc", ["periodic", "zero_grad"])
def test_run_parity_form_equals_the_sweep(orc, harness, bc):
    """The kernel's form (a cell's own test acts when it succeeds at an even
    offset from the start of its run of successful tests) against the
    oracle's literal sweep, on staircases, noise and long runs of successful
    tests (alternating data, where every test succeeds)."""
    rng = np.random.default_rng(11)
    nx = 160
    states = []
    for _ in range(20):
        steps = np.cumsum(rng.integers(0, 2, nx) * rng.uniform(-0.6, 0.6, nx))
        states.append(np.stack([1.5 + steps, rng.uniform(-0.5, 0.5) + 0.2 * steps, 2.0 + np.abs(steps)]))
    alt = np.where(np.arange(nx) % 2 == 0, 0.3, 1.0)
    states.append(np.stack([alt, alt, alt]))
    mixed = np.where((np.arange(nx) // 7) % 2 == 0, alt, 1.0)
    states.append(np.stack([mixed, 1.0 + 0.0 * mixed, mixed[::-1].copy()]))
    for q in states:
        q = np.ascontiguousarray(q)
        want = orc.fv1d_lmpe(nx, 0.0, 1.0, q, bc=bc)
        got = np.zeros((3, nx), dtype=np.uint8)
        harness.fv1dh_lmpe(nx, 0.0, 1.0, {"zero_grad": 0, "periodic": 1}[bc],
                           q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           got.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
        np.testing.assert_array_equal(got, want)
