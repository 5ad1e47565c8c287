"""The 1D path (FV::System1D, dim = 1) on the GPU against the FP64 CPU
oracle: residual bit for bit for both reconstructions (the shipped TENO5i
and TENO5) and both fluxes, the first steps' dt and states, and the 1D
configuration (Sod, N = 400, TENO5 + HLLC, CFL 0.5 to t = 0.2) end to end."""
import numpy as np
import pytest

from conftest import rel_l1

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def _sod_state(nx):
    x = (np.arange(nx) + 0.5) / nx
    rho = np.where(x < 0.5, 1.0, 0.125)
    p = np.where(x < 0.5, 1.0, 0.1)
    return np.stack([rho, np.zeros(nx), p / (GAMMA - 1.0)])


def _supersonic_state(nx):
    """Sod's densities and pressures moving right at u = 2.5, faster than
    the sound speed on both sides: every wave speed is positive and no cell
    is at rest (the reference's CUFlux divides 0 by 0 in the middle field
    where u = 0, Equations_Euler1D.cpp:144-150)."""
    x = (np.arange(nx) + 0.5) / nx
    rho = np.where(x < 0.5, 1.0, 0.125)
    p = np.where(x < 0.5, 1.0, 0.1)
    u = 2.5
    return np.stack([rho, rho * u, p / (GAMMA - 1.0) + 0.5 * rho * u * u])


def _smooth_state(nx, seed=5):
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x + rng.uniform(0, 6))
    u = 0.4 * np.cos(4 * np.pi * x)
    p = 1.0 + 0.2 * np.sin(6 * np.pi * x)
    return np.stack([rho, rho * u, p / (GAMMA - 1.0) + 0.5 * (rho * u) * u])


@pytest.mark.parametrize("recon", ["teno5i", "teno5", "weno5js", "weno5z"])
@pytest.mark.parametrize("flux", ["rusanov", "hllc", "cu"])
@pytest.mark.parametrize("case", ["sod", "smooth", "ragged"])
def test_rhs_parity(orc, recon, flux, case):
    import htx

    nx = {"sod": 400, "smooth": 256, "ragged": 77}[case]
    if case == "sod":
        q = _supersonic_state(nx) if flux == "cu" else _sod_state(nx)
    else:
        q = _smooth_state(nx)
    print(f"[case] 1d rhs {case} nx={nx} {recon} {flux}", flush=True)
    s = htx.System1D(0.0, 1.0, nx, 0.5, 0.2, 0.0, GAMMA, scheme=recon, flux=flux)
    s.set_state(q)
    r = s.rhs()
    rc = orc.fv1d_rhs(nx, 0.0, 1.0, GAMMA, q, recon=recon, flux=flux)
    assert np.isfinite(r).all()
    np.testing.assert_array_equal(r, rc)
    s.close()


def test_cu_flux_at_rest_is_nonfinite_like_the_reference(orc):
    """CUFlux where the gas is at rest (Sod's initial state): the middle
    field's speeds |u| are 0 on both sides, a+ = a- = 0, and the reference's
    (a+ f_L - a- f_R) / (a+ - a-) is 0 / 0 (Equations_Euler1D.cpp:144-150).
    The GPU residual is NaN in exactly the oracle's cells, and a step
    reports the non-finite state as HTX_ENONFINITE instead of running on."""
    import htx
    from htx._lib import HtxError

    nx = 200
    q = _sod_state(nx)
    print("[case] 1d cu flux on a state at rest", flush=True)
    s = htx.System1D(0.0, 1.0, nx, 0.5, 0.2, 0.0, GAMMA, scheme="teno5", flux="cu")
    s.set_state(q)
    r = s.rhs()
    rc = orc.fv1d_rhs(nx, 0.0, 1.0, GAMMA, q, recon="teno5", flux="cu")
    assert np.isnan(rc).any()
    np.testing.assert_array_equal(np.isnan(r), np.isnan(rc))
    with pytest.raises(HtxError, match="HTX_ENONFINITE"):
        s.step()
    s.close()


VARIANTS = ["weno5n", "teno5n", "teno5a", "weno5is", "weno5is_ext", "weno5ismod", "weno5z_diff", "teno5_diff",
            "linear"]


@pytest.mark.parametrize("recon", VARIANTS)
@pytest.mark.parametrize("case", ["sod", "smooth"])
def test_rhs_parity_remaining_reconstructions(orc, recon, case):
    """Every other 5-point routine of Schemes::Lagrange as the 1D
    reconstruction (diff = dx): residual equal to the oracle's, bit for bit
    (1e-13 relative where the routine takes a fractional power, which is
    the device's pow)."""
    import htx

    nx = {"sod": 400, "smooth": 256}[case]
    q = _sod_state(nx) if case == "sod" else _smooth_state(nx)
    print(f"[case] 1d rhs {case} nx={nx} {recon}", flush=True)
    s = htx.System1D(0.0, 1.0, nx, 0.5, 0.2, 0.0, GAMMA, scheme=recon, flux="hllc")
    s.set_state(q)
    r = s.rhs()
    rc = orc.fv1d_rhs(nx, 0.0, 1.0, GAMMA, q, recon=recon, flux="hllc")
    if recon in ("weno5n", "weno5z_diff"):
        assert rel_l1(r, rc) < 1e-13
    else:
        np.testing.assert_array_equal(r, rc)
    s.close()


@pytest.mark.parametrize("recon,flux", [("teno5i", "rusanov"), ("teno5", "hllc"), ("weno5js", "hllc"),
                                        ("weno5z", "rusanov"), ("teno5n", "hllc"), ("weno5is_ext", "rusanov"),
                                        ("teno5a", "rusanov"), ("teno5", "cu"), ("teno5i", "cu")])
def test_steps_parity(orc, recon, flux):
    """Ten SSP-RK3 steps of the Sod problem (moving right supersonically for
    the central-upwind flux, which fails where the gas is at rest): dt and
    state of every step equal the oracle's."""
    import htx

    nx = 400
    q = _supersonic_state(nx) if flux == "cu" else _sod_state(nx)
    print(f"[case] 1d steps nx={nx} {recon} {flux}", flush=True)
    s = htx.System1D(0.0, 1.0, nx, 0.5, 0.2, 0.0, GAMMA, scheme=recon, flux=flux)
    s.set_state(q)
    qc, t = q.copy(), 0.0
    for k in range(10):
        s.step()
        qn, tn, n = orc.fv1d_advance(nx, 0.0, 1.0, GAMMA, qc, 0.5, t, 0.2, recon=recon, flux=flux, max_steps=1)
        assert n == 1
        assert s.time == tn
        assert np.isfinite(qn).all()
        np.testing.assert_array_equal(s.get_state(), qn)
        qc, t = qn, tn
    s.close()


def test_sod_configuration_end_to_end(orc):
    """The 1D configuration to t = 0.2: the GPU trajectory equals the
    oracle's (relative L1 below 1e-12; bit for bit unless rounding moves a
    TENO cut-off decision at its threshold)."""
    import htx

    nx = 400
    s = htx.System1D(0.0, 1.0, nx, 0.5, 0.2, 0.0, GAMMA, scheme="teno5", flux="hllc")
    s.sodShock()
    q0 = s.get_state()
    np.testing.assert_array_equal(q0, orc.ic_sod(nx, 0.0, 1.0, GAMMA))
    n = s.timeStepping()
    qr, tr, nr = orc.fv1d_advance(nx, 0.0, 1.0, GAMMA, q0, 0.5, 0.0, 0.2, recon="teno5", flux="hllc")
    err = rel_l1(s.get_state(), qr)
    print(f"sod: {n} steps (oracle {nr}), t {s.time} ({tr}), rel L1 {err:.3e}, "
          f"identical {np.array_equal(s.get_state(), qr)}", flush=True)
    assert n == nr and s.time == tr
    assert err < 1e-12
    s.close()


def test_kxrcf_flags_match_oracle(orc):
    """The GPU KXRCF detector equals the oracle's flags bit for bit on the
    Sod initial state (cells 198-200), on the states of a Sod run (each of
    40 steps) and on a seeded rough state with momenta of both signs; the
    2D solver refuses the call."""
    import htx
    from htx._lib import HtxError

    nx = 400
    print("[case] KXRCF flags, Sod N=400 and a seeded state", flush=True)
    s = htx.System1D(0.0, 1.0, nx, 0.5, 0.2, 0.0, GAMMA, scheme="teno5", flux="hllc")
    s.sodShock()
    f = s.troubledCells()
    assert list(np.nonzero(f)[0]) == [198, 199, 200]
    rates = []
    for _ in range(40):
        s.step()
        q = s.get_state()
        f = s.troubledCells()
        np.testing.assert_array_equal(f, orc.fv1d_kxrcf(nx, 0.0, 1.0, q))
        rates.append(f.mean())
    print(f"flag rate over the run: {min(rates):.4f} .. {max(rates):.4f}", flush=True)
    rng = np.random.default_rng(5)
    q = np.empty((3, nx))
    q[0] = 1.0 + 0.5 * rng.random(nx) * (rng.random(nx) < 0.1)
    q[1] = rng.normal(0.0, 0.5, nx)
    q[2] = 2.5 + 0.5 * q[1] ** 2 / q[0]
    s.set_state(q)
    f = s.troubledCells()
    np.testing.assert_array_equal(f, orc.fv1d_kxrcf(nx, 0.0, 1.0, q))
    assert 0 < f.sum() < nx
    # jumps next to both boundaries: the first and last cells see the ghosts
    for mom in (0.3, -0.3):
        q = np.empty((3, nx))
        q[0] = 1.0
        q[0, 0], q[0, -1] = 3.0, 0.2
        q[1] = mom
        q[2] = 2.5 + 0.5 * mom * mom / q[0]
        s.set_state(q)
        f = s.troubledCells()
        np.testing.assert_array_equal(f, orc.fv1d_kxrcf(nx, 0.0, 1.0, q))
        assert f[0] == 1 and f[-1] == 1
    s.close()
    t = htx.System2D(0.0, 1.0, 16, 0.0, 1.0, 16, 0.5, 1.0, 0.0, GAMMA)
    with pytest.raises(HtxError):
        htx._lib.check(htx._lib.load().htx_kxrcf_flags(t._h, np.zeros(16, np.uint8).ctypes.data))
    t.close()


@pytest.mark.parametrize("bc", ["zero_grad", "periodic"])
def test_lmpe_flags_match_oracle(orc, bc):
    """The GPU LMPE detector (run-parity form of the reference's sequential
    sweep) equals the oracle's literal sweep bit for bit, per field, on the
    Sod states of a 40-step run and on seeded staircases with long runs of
    successful tests; the 2D solver refuses the call.  The flag rates of LMPE
    and KXRCF over the Sod run are printed side by side."""
    import htx
    from htx._lib import HtxError

    nx = 400
    print(f"[case] LMPE flags, Sod N=400 and seeded staircases ({bc})", flush=True)
    s = htx.System1D(0.0, 1.0, nx, 0.5, 0.2, 0.0, GAMMA, scheme="teno5", flux="hllc", bc=bc)
    s.sodShock()
    lm, kx = [], []
    for k in range(41):
        q = s.get_state()
        f = s.lmpeCells()
        np.testing.assert_array_equal(f, orc.fv1d_lmpe(nx, 0.0, 1.0, q, bc=bc))
        lm.append(f.any(axis=0).mean())
        kx.append(s.troubledCells().mean())
        if k < 40:
            s.step()
    print(f"flag rate over the Sod run: LMPE {min(lm):.4f} .. {max(lm):.4f}, KXRCF {min(kx):.4f} .. {max(kx):.4f}",
          flush=True)
    rng = np.random.default_rng(9)
    for _ in range(5):
        steps = np.cumsum(rng.integers(0, 2, nx) * rng.uniform(-0.6, 0.6, nx))
        q = np.stack([2.0 + steps, 0.1 * steps, 3.0 + np.abs(steps)])
        alt = np.where(np.arange(nx) % 2 == 0, 0.5, 1.0)
        q[1, 100:180] = alt[100:180]
        s.set_state(q)
        np.testing.assert_array_equal(s.lmpeCells(), orc.fv1d_lmpe(nx, 0.0, 1.0, q, bc=bc))
    s.close()
    t = htx.System2D(0.0, 1.0, 16, 0.0, 1.0, 16, 0.5, 0.1, 0.0, GAMMA)
    with pytest.raises(HtxError) as ei:
        _lib_lmpe(t)
    assert ei.value.code == -5
    t.close()


def _lib_lmpe(solver2d):
    from htx import _lib
    from htx._lib import check

    f = np.zeros(3 * 16, dtype=np.uint8)
    check(_lib.load().htx_lmpe_flags(solver2d._h, f.ctypes.data))
