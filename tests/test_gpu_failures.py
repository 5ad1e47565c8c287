"""Failure detection on the GPU (the reference has none: its CFL loop takes
fmax over the cells, which skips NaN, and it steps on).  A NaN or infinite
conserved value, met by the CFL kernel when a state is set or by a stage-3
epilogue, and a non-finite or non-positive dt, become HTX_ENONFINITE in
the blocking calls and stop the stepping; states the reference handles
(negative density or pressure, a zero maxSpeed) keep the reference's
semantics bit for bit."""
import numpy as np
import pytest

from oracle import params

pytestmark = pytest.mark.gpu

GAMMA = 1.4
ENONFINITE = -7


def _vortex(orc, n):
    prm = params(n, n, 0.0, 10.0, 0.0, 10.0, GAMMA, "periodic", "hllc", True)
    return prm, orc.ic_vortex(prm, (5.0, 5.0), None)


def _solver(n, tmax=2.0, bc="periodic", precision="fp64", scheme="teno5"):
    import htx

    return htx.System2D(0.0, 10.0, n, 0.0, 10.0, n, 0.5, tmax, 0.0, GAMMA, flux="hllc", bc=bc,
                        precision=precision, scheme=scheme)


@pytest.mark.parametrize("precision,scheme", [("fp64", "teno5"), ("fp32_mixed", "teno5"), ("fp64", "hteno")])
def test_nan_state_is_refused(orc, precision, scheme):
    import htx

    n = 24
    _, q = _vortex(orc, n)
    if scheme == "hteno":
        q = np.concatenate([q, np.zeros((8, n, n))])
    q[2, 7, 11] = np.nan
    print(f"[case] NaN cell, {precision} {scheme}", flush=True)
    s = _solver(n, precision=precision, scheme=scheme)
    with pytest.raises(htx.HtxError) as ei:
        s.set_state(q)
    assert ei.value.code == ENONFINITE and "NaN or infinite" in str(ei.value)
    # stepping stays stopped: every blocking call reports it, t does not move
    for call in (s.step, s.compute_dt, lambda: s.timeStepping(1.0), s.sync):
        with pytest.raises(htx.HtxError) as ei:
            call()
        assert ei.value.code == ENONFINITE
    assert s.time == 0.0 and s.nsteps == 0
    # a good state clears the record
    _, q = _vortex(orc, n)
    if scheme == "hteno":
        q = np.concatenate([q, np.zeros((8, n, n))])
    s.set_state(q)
    assert s.step() > 0.0
    s.close()


def test_overflow_during_a_step_is_reported(orc):
    """A finite state whose fluxes overflow (an energy of 1e308 in one cell):
    setting it succeeds, the first step's stage-3 epilogue meets the
    infinities, and the step reports HTX_ENONFINITE."""
    import htx

    n = 24
    _, q = _vortex(orc, n)
    q[3, 12, 12] = 1e308
    print("[case] overflow in a step", flush=True)
    s = _solver(n)
    s.set_state(q)
    with pytest.raises(htx.HtxError) as ei:
        s.step()
    assert ei.value.code == ENONFINITE
    with pytest.raises(htx.HtxError):
        s.timeStepping(1.0)
    s.close()


def test_negative_density_and_pressure_follow_the_reference(orc):
    """The reference takes sqrt(|.|) in its sound speeds and steps on through
    negative density or pressure: the drop-in does the same, bit for bit with
    the oracle, and reports nothing."""
    n = 20
    prm, q = _vortex(orc, n)
    q[0, 5, 5] = -0.3  # negative density
    q[3, 9, 14] = 0.5 * (q[1, 9, 14] ** 2 + q[2, 9, 14] ** 2) / q[0, 9, 14] - 0.2  # negative pressure
    print("[case] negative rho and p", flush=True)
    s = _solver(n)
    s.set_state(q)
    r = s.rhs()
    np.testing.assert_array_equal(r, orc.fv2d_rhs(prm, q))
    dt = s.step()
    dt_c = orc.fv2d_dt(prm, q, 0.5, 0.0, 2.0)
    qc = orc.fv2d_rk3(prm, q.copy(), dt_c)
    assert dt == dt_c
    assert np.isfinite(qc).all()
    np.testing.assert_array_equal(s.get_state(), qc)
    s.close()


def test_zero_wave_speed_steps_to_the_end_time(orc):
    """Uniform flow with u = v = -c: the reference's maxSpeed
    max(u - c, |u + c|) (Equations_Euler2D.cpp:218, no fabs on the first
    term) is zero to rounding in both directions, so dt = cfl h / s exceeds
    the time left and is clamped to the end time (FV_System2D.cpp:633-645);
    the uniform state does not change.  Without the quirk the speed is 2c
    and the step is an ordinary one."""
    n = 16
    c = np.sqrt(GAMMA)
    q = np.empty((4, n, n))
    q[0], q[1], q[2], q[3] = 1.0, -c, -c, 1.0 / (GAMMA - 1) + c * c
    prm = params(n, n, 0.0, 10.0, 0.0, 10.0, GAMMA, "periodic", "hllc", True)
    assert orc.fv2d_dt(prm, q, 0.5, 0.25, 2.0) == 1.75
    print("[case] zero wave speed (maxSpeed quirk)", flush=True)
    s = _solver(n)
    s.set_state(q)
    assert s.step() == 2.0
    assert s.time == 2.0 and s.nsteps == 1
    np.testing.assert_array_equal(s.get_state(), q)
    s.close()
    import htx

    t = htx.System2D(0.0, 10.0, n, 0.0, 10.0, n, 0.5, 2.0, 0.0, GAMMA, flux="hllc", bc="periodic", ref_quirks=False)
    t.set_state(q)
    assert t.step() == pytest.approx(0.5 * (10.0 / n) / (2 * c), rel=1e-12)
    t.close()


def test_1d_nan_is_refused():
    import htx

    nx = 64
    q = np.ones((3, nx))
    q[2] = 2.5
    q[0, 10] = np.inf
    s = htx.System1D(0.0, 1.0, nx, 0.5, 0.2, 0.0, GAMMA, scheme="teno5", flux="hllc")
    with pytest.raises(htx.HtxError) as ei:
        s.set_state(q)
    assert ei.value.code == ENONFINITE
    s.close()


def test_batch_names_the_failing_instance():
    import htx
    from htx.batch import Batch, params_for

    ps = [params_for(0.0, 1.0, n, 0.0, 1.0, n, 0.5, 0.1, GAMMA, precision="fp32_mixed") for n in (32, 40, 24)]
    b = Batch(ps)
    st = [[1.0, 0.0, 0.0, 1.0], [0.5, 0.1, 0.0, 0.4], [0.8, 0.0, 0.2, 0.6], [1.2, -0.1, 0.0, 1.1]]
    for i in range(3):
        b.quadrant(i, 0.5, 0.5, st)
    bad = np.array(st)
    bad[1, 3] = np.nan
    with pytest.raises(htx.HtxError) as ei:
        b.quadrant(1, 0.5, 0.5, bad.tolist())
    assert ei.value.code == ENONFINITE and "instance 1" in str(ei.value)
    with pytest.raises(htx.HtxError) as ei:
        b.advance()
    assert ei.value.code == ENONFINITE and "instance 1" in str(ei.value)
    b.close()
