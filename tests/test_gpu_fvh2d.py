"""Parity of the HIP FP64 Hermite-TENO path (scheme "hteno") against the FP64
CPU oracle: residual, limited state, SSP-RK3 steps and the device-side
vortex generator with derivative moments."""
import numpy as np
import pytest

from conftest import rel_l1
from oracle import params

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def make(n, bc="periodic", flux="rusanov", quirks=True, tmax=10.0, precision="fp64"):
    import htx

    print(f"[case] hteno n={n} bc={bc} flux={flux} quirks={quirks} precision={precision}", flush=True)
    s = htx.System2D(0.0, 10.0, n, 0.0, 10.0, n, 0.4, tmax, 0.0, GAMMA, scheme="hteno", flux=flux, bc=bc,
                     ref_quirks=quirks, precision=precision)
    return s, params(n, n, 0.0, 10.0, 0.0, 10.0, GAMMA, bc, flux, quirks)


@pytest.mark.parametrize("precision", ["fp64", "fp64_fast"])
@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
@pytest.mark.parametrize("bc", ["periodic", "zero_grad"])
def test_hermite_rhs_parity(orc, flux, bc, precision):
    from htx.ic import vortex_modes

    s, prm = make(32, bc, flux, precision=precision)
    q = orc.ic_vortex_hermite(prm, (5.0, 5.0), vortex_modes())
    s.set_state(q)
    r = s.rhs()
    rc = orc.fvh2d_rhs(prm, q)
    for g in range(0, 12, 4):
        err = rel_l1(r[g:g + 4], rc[g:g + 4])
        print(f"fields {g}..{g + 3}: rel L1 {err:.3e}", flush=True)
        assert err < 1e-12


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp64_fast", 1e-9)])
def test_hermite_step_parity(orc, precision, tol):
    """Five steps from the same start; the RK base of the derivative moments
    is the limited state.  The FP64 drop-in is held to 1e-12; fp64_fast to
    1e-9 (its rounding can move HTENO cut-off decisions at their threshold)."""
    from htx.ic import vortex_modes

    s, prm = make(32, "periodic", "hllc", precision=precision)
    q = orc.ic_vortex_hermite(prm, (5.0, 5.0), vortex_modes())
    s.set_state(q)
    qc = q.copy()
    t = 0.0
    for k in range(5):
        dt = orc.fv2d_dt(prm, qc, 0.4, t, 10.0)
        assert s.step() == pytest.approx(dt, rel=1e-13)
        qc = orc.fvh2d_rk3(prm, qc, dt)
        t += dt
        err = rel_l1(s.get_state(), qc)
        print(f"step {k}: rel L1 {err:.3e}", flush=True)
        assert err < tol


def test_hermite_ic_generator(orc):
    from htx.ic import vortex_modes

    s, prm = make(24)
    s.isentropicVortex((5.0, 5.0), vortex_modes())
    np.testing.assert_allclose(s.get_state(), orc.ic_vortex_hermite(prm, (5.0, 5.0), vortex_modes()),
                               rtol=1e-12, atol=1e-14)


def test_hermite_conservation_512():
    """Periodic vortex at 512^2: all 12 fields conserved over 3 steps."""
    import htx
    from htx.ic import vortex_modes

    n = 512
    s = htx.System2D(0.0, 10.0, n, 0.0, 10.0, n, 0.5, 2.0, 0.0, GAMMA, scheme="hteno", flux="hllc",
                     bc="periodic")
    s.isentropicVortex((5.0, 5.0), vortex_modes())
    q0 = s.get_state()
    s.steps_async(3)
    q1 = s.get_state()
    assert np.isfinite(q1).all()
    for g in range(4):
        assert abs(q1[g].sum() - q0[g].sum()) <= 1e-12 * np.abs(q0[g]).sum()
