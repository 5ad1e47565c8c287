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


@pytest.mark.parametrize("precision,tol_step,tol_free", [("fp64", 1e-12, 1e-9), ("fp64_fast", 1e-9, 1e-8)])
def test_hermite_step_parity(orc, precision, tol_step, tol_free):
    """Five SSP-RK3 steps; the RK base of the derivative moments is the
    limited state.  Per step from identical inputs (the GPU state is reset to
    the oracle's before each step) and free-running from the same start.  The
    FP64 drop-in is held to the parity tolerances (1e-12 per step, 1e-9
    free-running); fp64_fast to 1e-9 / 1e-8, because its rounding moves
    HTENO cut-off decisions that sit at their threshold."""
    from htx.ic import vortex_modes

    s, prm = make(32, "periodic", "hllc", precision=precision)
    free, _ = make(32, "periodic", "hllc", precision=precision)
    q = orc.ic_vortex_hermite(prm, (5.0, 5.0), vortex_modes())
    free.set_state(q)
    qc = q.copy()
    t = 0.0
    for k in range(5):
        s.set_state(qc)
        dt = orc.fv2d_dt(prm, qc, 0.4, t, 10.0)
        assert s.step() == pytest.approx(dt, rel=1e-13)
        qn = orc.fvh2d_rk3(prm, qc, dt)
        g = s.get_state()
        err = rel_l1(g, qn)
        free.step()
        err_free = rel_l1(free.get_state(), qn)
        print(f"step {k}: per-step rel L1 {err:.3e} (max |diff| {np.abs(g - qn).max():.3e}), "
              f"free-running {err_free:.3e}", flush=True)
        assert err < tol_step
        assert err_free < tol_free
        qc = qn
        t += dt


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


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_hermite_fp32_mixed_residual(orc, flux):
    """FP32-mixed Hermite residual (HTENO / TENO5 cut-off rules and transverse
    WENO weights in FP32, values, projections and fluxes in FP64): within
    2e-8 relative L1 of the FP64 oracle on the vortex (the CPU build of the
    same bodies measures 5.2e-9 at 32 x 28)."""
    from htx.ic import vortex_modes

    s, prm = make(48, "periodic", flux, precision="fp32_mixed")
    q = orc.ic_vortex_hermite(prm, (5.0, 5.0), vortex_modes())
    s.set_state(q)
    err = rel_l1(s.rhs(), orc.fvh2d_rhs(prm, q))
    print(f"hermite fp32-mixed residual rel L1 {err:.3e}", flush=True)
    assert err < 2e-8


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
@pytest.mark.parametrize("bc", ["periodic", "zero_grad"])
def test_hermite_fp32_bulk_residual(orc, flux, bc):
    """FP32 bulk Hermite residual (FP32 face values, transverse quadrature,
    back-projection, u and moment fluxes; FP64 Roe eigensystem, Gauss sums,
    divergence): each of the u, v, w field groups within 1e-4 relative L1 of
    the FP64 oracle on the vortex, the u fields conserved on the periodic
    grid, and a uniform state's residual exactly zero."""
    from htx.ic import vortex_modes

    s, prm = make(48, bc, flux, precision="fp32")
    q = orc.ic_vortex_hermite(prm, (5.0, 5.0), vortex_modes())
    s.set_state(q)
    r = s.rhs()
    rc = orc.fvh2d_rhs(prm, q)
    assert np.isfinite(r).all()
    for g in range(0, 12, 4):
        err = rel_l1(r[g:g + 4], rc[g:g + 4])
        print(f"hermite fp32 bulk fields {g}..{g + 3}: rel L1 {err:.3e}", flush=True)
        assert 0.0 < err < 1e-4
    if bc == "periodic":
        for e in range(4):
            assert abs(r[e].sum()) <= 1e-12 * np.abs(r[e]).sum()
    u = np.zeros((12, 48, 48))
    u[:4] = np.array([1.2, 0.3, -0.4, 2.5])[:, None, None]
    s.set_state(u)
    assert np.abs(s.rhs()).max() == 0.0
    s.close()


def _hermite_run(n, q0, dom, bc, flux, tend, precision):
    import htx

    s = htx.System2D(dom[0], dom[1], n, dom[2], dom[3], n, 0.4, tend, 0.0, GAMMA, scheme="hteno", flux=flux, bc=bc,
                     precision=precision)
    if q0 is None:
        from htx.ic import vortex_modes

        s.isentropicVortex((5.0, 5.0), vortex_modes())
    else:
        s.set_state(q0)
    s.timeStepping(tend)
    out = s.get_state()
    s.close()
    return out


@pytest.mark.parametrize("precision", ["fp32_mixed", "fp32"])
@pytest.mark.parametrize("case", ["quadrant", "vortex"])
def test_hermite_fp32_mixed_accuracy(case, precision):
    """Hermite FP32-mixed and FP32 bulk at matched accuracy: against a 4x finer FP64
    solution averaged back onto the grid, the L1 error of the cell averages
    is within 10 % of the plain FP64 scheme's error at the same resolution
    (density, and all four conserved fields).  Quadrant: seeded Riemann data
    with zero derivative moments, split 4 x 4 for the finer run; vortex: the
    generator at both resolutions."""
    from test_gpu_fv2d import _block_average, quadrant_state

    n = 48 if case == "quadrant" else 64
    flux = "rusanov"
    if case == "quadrant":
        dom, bc, tend = (0.0, 1.0, 0.0, 1.0), "zero_grad", 0.02
        u = quadrant_state(n, n, seed=4)
        q0 = np.concatenate([u, np.zeros((8, n, n))])
        uf = np.repeat(np.repeat(u, 4, axis=1), 4, axis=2)
        q0f = np.ascontiguousarray(np.concatenate([uf, np.zeros((8, 4 * n, 4 * n))]))
    else:
        dom, bc, tend = (0.0, 10.0, 0.0, 10.0), "periodic", 0.5
        q0 = q0f = None
    print(f"[case] hermite {precision} accuracy {case} n={n} t={tend}", flush=True)
    q64 = _hermite_run(n, q0, dom, bc, flux, tend, "fp64")[:4]
    q32 = _hermite_run(n, q0, dom, bc, flux, tend, precision)[:4]
    qf = _block_average(_hermite_run(4 * n, q0f, dom, bc, flux, tend, "fp64")[:4], 4)
    assert np.isfinite(q32).all()
    for name, sl in [("rho", slice(0, 1)), ("all", slice(0, 4))]:
        e64 = rel_l1(q64[sl], qf[sl])
        e32 = rel_l1(q32[sl], qf[sl])
        print(f"{name}: L1 error vs 4x finer FP64: fp64 {e64:.4e}, {precision} {e32:.4e}, ratio {e32 / e64:.4f}; "
              f"{precision} - fp64 {rel_l1(q32[sl], q64[sl]):.3e}", flush=True)
        assert e32 <= 1.10 * e64


@pytest.mark.parametrize("recon", ["hweno", "htenoi", "hlinear"])
@pytest.mark.parametrize("precision", ["fp64", "fp64_fast"])
def test_hermite_variants_parity(orc, recon, precision):
    """hweno, htenoi or the linear Hermite scheme in hteno's place
    (FVHermite_System2D.cpp:459, :486): residual within 1e-12 and 3 SSP-RK3
    steps within 1e-9 of the oracle's (the parity tolerances of the Hermite
    path), both builds."""
    import htx
    from htx.ic import vortex_modes

    n = 32
    print(f"[case] {recon} n={n} {precision}", flush=True)
    s = htx.System2D(0.0, 10.0, n, 0.0, 10.0, n, 0.4, 10.0, 0.0, GAMMA, scheme=recon, flux="rusanov", bc="periodic",
                     precision=precision)
    prm = params(n, n, 0.0, 10.0, 0.0, 10.0, GAMMA, "periodic", "rusanov", True, recon=recon)
    q = orc.ic_vortex_hermite(prm, (5.0, 5.0), vortex_modes())
    s.set_state(q)
    r, rc = s.rhs(), orc.fvh2d_rhs(prm, q)
    print(f"residual rel L1 {rel_l1(r, rc):.3e}, bit-identical {np.array_equal(r, rc)}", flush=True)
    assert rel_l1(r, rc) < 1e-12
    for _ in range(3):
        s.step()
    qc, tc, k = orc.fvh2d_advance(prm, q, 0.4, 0.0, 10.0, 3)
    g = s.get_state()
    print(f"3 steps rel L1 {rel_l1(g, qc):.3e}, bit-identical {np.array_equal(g, qc)}", flush=True)
    assert k == 3 and s.time == pytest.approx(tc, rel=1e-13)
    assert rel_l1(g, qc) < 1e-9
    s.close()
