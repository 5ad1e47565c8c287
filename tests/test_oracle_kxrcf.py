"""The oracle's KXRCF troubled-cell detector (FV_System1D.cpp:166-192)
pinned by what its reconstructions fix: the two 3-cell reconstructions of a
face agree on cell averages of a quadratic (no flags away from the
zero-gradient ghosts), and a single density
jump flags exactly the cells whose 4-cell window straddles it (the window
leans upwind of the momentum's sign); the reference's loop (i = s+1 ..
nxx-s-2 with s = 2, FV_System1D.cpp:15,166) visits every cell, the first
and last ones against the zero-gradient ghosts."""
import numpy as np
import pytest

N = 400  # the 1D configuration's grid on [0, 1]


def _state(rho, mom):
    q = np.zeros((3, rho.size))
    q[0] = rho
    q[1] = mom
    q[2] = 2.5 + 0.5 * mom * mom / rho
    return q


def test_quadratic_profile_is_never_flagged(orc):
    x0 = np.arange(N) / N
    x1 = x0 + 1.0 / N
    # exact cell averages of rho(x) = 1 + 0.3 x - 0.2 x^2
    prim = lambda x: x + 0.15 * x * x - 0.2 / 3.0 * x ** 3  # noqa: E731
    rho = (prim(x1) - prim(x0)) * N
    # (cells 1 and N-2 reach a zero-gradient ghost, which is no quadratic)
    for mom in (0.7, -0.7):
        assert not orc.fv1d_kxrcf(N, 0.0, 1.0, _state(rho, np.full(N, mom)))[2:N - 2].any()


@pytest.mark.parametrize("k", [1, 2, 57, 200, 398, 399])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_single_jump_flags_its_window(orc, k, sign):
    rho = np.where(np.arange(N) < k, 1.0, 0.125)
    f = orc.fv1d_kxrcf(N, 0.0, 1.0, _state(rho, np.full(N, 0.3 * sign)))
    # momentum >= 0: window i-1 .. i+2; < 0: window i-2 .. i+1
    lo, hi = (k - 2, k) if sign > 0 else (k - 1, k + 1)
    expect = np.zeros(N, dtype=np.uint8)
    expect[max(lo, 0):min(hi, N - 1) + 1] = 1
    assert np.array_equal(f, expect), np.nonzero(f)[0]


def test_sod_initial_state(orc):
    q = orc.ic_sod(N)
    f = orc.fv1d_kxrcf(N, 0.0, 1.0, q)
    assert list(np.nonzero(f)[0]) == [198, 199, 200]


def test_first_and_last_cells_see_the_ghosts(orc):
    """Cells 0 and nx-1 are visited and read the zero-gradient ghosts
    (rho_-1 = rho_0, rho_nx = rho_nx-1): a jump between cells 0 and 1 flags
    cells 0 and 1 (momentum >= 0: window i-1 .. i+2), one between nx-2 and
    nx-1 flags nx-3 .. nx-1."""
    rho = np.ones(N)
    rho[0], rho[-1] = 3.0, 0.2
    f = orc.fv1d_kxrcf(N, 0.0, 1.0, _state(rho, np.zeros(N)))
    assert list(np.nonzero(f)[0]) == [0, 1, N - 3, N - 2, N - 1]
    # momentum < 0: window i-2 .. i+1, so the jump at 0|1 flags cells 0 .. 2
    # and the jump at nx-2|nx-1 flags nx-2 and nx-1
    f = orc.fv1d_kxrcf(N, 0.0, 1.0, _state(rho, np.full(N, -0.5)))
    assert list(np.nonzero(f)[0]) == [0, 1, 2, N - 2, N - 1]


@pytest.mark.parametrize("jump, cells", [(5e-4, [198, 199, 200]), (2e-4, [199])])
def test_threshold_radius(orc, jump, cells):
    """A small jump A -> B at cell 200 (momentum >= 0): cells 198 and 200 see
    |h1 - h2| = |A - B| / 6, cell 199 sees |A - B| / 3, against radius |rho|
    with radius = (dx/2)^1.5 = 4.42e-5 at N = 400: a jump of 5e-4 flags all
    three, 2e-4 only cell 199 (radius dx^1.5 would flag only 199 at 5e-4)."""
    rho = np.where(np.arange(N) < 200, 1.0, 1.0 - jump)
    f = orc.fv1d_kxrcf(N, 0.0, 1.0, _state(rho, np.full(N, 0.1)))
    assert list(np.nonzero(f)[0]) == cells
