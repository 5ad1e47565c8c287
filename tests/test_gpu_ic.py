"""Device generators of the piecewise initial states against their host
restatement (htx.ic.piecewise_state)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme", ["teno5", "hteno"])
def test_quadrant_generator(scheme):
    import htx
    from htx.ic import piecewise_state, quadrant_workload

    n = 96
    w = quadrant_workload(n)
    print(f"[case] quadrant generator {scheme} n={n}", flush=True)
    s = htx.System2D(w["xl"], w["xr"], n, w["yb"], w["yt"], n, w["cfl"], w["tmax"], 0.0, w["gamma"], scheme=scheme,
                     bc="zero_grad")
    s.init_workload(w)
    np.testing.assert_array_equal(s.get_state(), piecewise_state(w, n, n, hermite=scheme == "hteno"))


def test_shock_blob_generator():
    import htx
    from htx.ic import piecewise_state, shock_blob_workload

    nx, ny = 256, 64
    w = shock_blob_workload(nx)
    print(f"[case] shock-blob generator {nx}x{ny}", flush=True)
    s = htx.System2D(w["xl"], w["xr"], nx, w["yb"], w["yt"], ny, w["cfl"], w["tmax"], 0.0, w["gamma"],
                     bc="zero_grad")
    s.init_workload(w)
    np.testing.assert_allclose(s.get_state(), piecewise_state(w, nx, ny), rtol=2e-15, atol=0)


@pytest.mark.parametrize("scheme", ["teno5", "hteno"])
def test_entropy_wave_generator_matches_oracle(orc, scheme):
    """The device entropy wave (cell averages and, for Hermite, the x and y
    derivative moments) against the oracle's 9x9 Gauss averages; the two use
    different sin/cos implementations, so agreement is to rounding."""
    import htx
    from oracle import params

    nx, ny = 40, 24
    print(f"[case] entropy-wave generator {scheme} {nx}x{ny}", flush=True)
    s = htx.System2D(0.0, 2.0, nx, 0.0, 1.5, ny, 0.5, 1.0, 0.0, 1.4, scheme=scheme, bc="periodic")
    s.entropyWave(1.0, 0.7, -0.4, 0.9, 0.3)
    prm = params(nx, ny, 0.0, 2.0, 0.0, 1.5, 1.4, "periodic", "hllc", True)
    ref = orc.ic_entropy(prm, 1.0, 0.7, -0.4, 0.9, 0.3, hermite=scheme == "hteno")
    np.testing.assert_allclose(s.get_state(), ref, rtol=1e-14, atol=1e-14)
    s.close()


def test_l1_density_error_on_device(orc):
    """printResult's numerical error reduced on the device: zero for the
    state just generated against its own exact field (same generator, same
    operations), the oracle's mean |rho - rho_exact| for a shifted exact
    solution, and HTX_ESTATE once a non-analytic state is set."""
    import htx
    from oracle import params

    n = 48
    print(f"[case] device L1 density error, entropy wave {n}^2", flush=True)
    s = htx.System2D(0.0, 2.0, n, 0.0, 2.0, n, 0.5, 1.0, 0.0, 1.4, bc="periodic")
    s.entropyWave()
    assert s.l1_density_error(0.0) == 0.0
    prm = params(n, n, 0.0, 2.0, 0.0, 2.0, 1.4, "periodic", "hllc", True)
    q = s.get_state()
    for t in (0.013, 0.25):
        ex = orc.ic_entropy(prm, shift=(t, t))
        ref = np.abs(q[0] - ex[0]).mean()
        got = s.l1_density_error(t)
        print(f"shift {t}: device {got:.16e}, oracle {ref:.16e}", flush=True)
        assert got == pytest.approx(ref, rel=1e-12)
    s.quadrant(1.0, 1.0, [[1.0, 0.0, 0.0, 1.0]] * 4)
    with pytest.raises(htx.HtxError) as ei:
        s.l1_density_error(0.0)
    assert ei.value.code == -4
    s.close()


def test_entropy_wave_error_converges():
    """The reference's entropy-wave case (rho = 1 + 0.5 sin(pi (x + y)),
    u = v = p = 1) on the periodic [0, 2]^2 advected to t = 0.2 with the
    drop-in TENO5 + HLLC: the device L1 density error against the advected
    exact solution falls with the grid (at least second order between 32^2,
    64^2 and 128^2) and is small."""
    import htx

    errs = []
    for n in (32, 64, 128):
        s = htx.System2D(0.0, 2.0, n, 0.0, 2.0, n, 0.5, 0.2, 0.0, 1.4, bc="periodic")
        s.entropyWave()
        s.timeStepping(0.2)
        assert s.time == 0.2
        errs.append(s.l1_density_error())
        s.close()
    print(f"[case] entropy wave L1 at t = 0.2: {errs}, ratios {errs[0] / errs[1]:.2f} {errs[1] / errs[2]:.2f}",
          flush=True)
    assert errs[0] / errs[1] > 4.0 and errs[1] / errs[2] > 4.0
    assert errs[2] < 1e-5
