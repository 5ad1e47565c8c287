"""The 1D path's bodies (htx_fv1d.h) built for the CPU with the strict
evaluation (tests/host/fv1d_host.hip): residual and SSP-RK3 stages bit for
bit against the FP64 oracle's FV::System1D restatement, for the shipped
TENO5i + characteristic Rusanov, the 1D configuration's TENO5 + HLLC and
every other 5-point reconstruction of Schemes::Lagrange (diff = dx)."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import RECON

LIB = os.path.join(ROOT, "tests", "host", "libfv1d_host.so")
DP = ctypes.POINTER(ctypes.c_double)
D = ctypes.c_double
GAMMA = 1.4


@pytest.fixture(scope="module")
def harness():
    if not os.path.exists(LIB):
        pytest.skip("tests/host/libfv1d_host.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    lib.fv1dh_stage.argtypes = [ctypes.c_int, D, D, D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, D, DP,
                                DP, DP]
    return lib


def _stage(lib, q, recon, flux, mode=0, dt=0.0, q0=None, quirks=1):
    nx = q.shape[1]
    q = np.ascontiguousarray(q)
    q0 = np.ascontiguousarray(q if q0 is None else q0)
    out = np.empty_like(q)
    lib.fv1dh_stage(nx, 0.0, 1.0, GAMMA, RECON[recon], {"rusanov": 0, "hllc": 1, "cu": 2}[flux], quirks,
                    mode, dt, q.ctypes.data_as(DP), q0.ctypes.data_as(DP), out.ctypes.data_as(DP))
    return out


def _states():
    nx = 120
    x = (np.arange(nx) + 0.5) / nx
    rho = np.where(x < 0.5, 1.0, 0.125)
    p = np.where(x < 0.5, 1.0, 0.1)
    sod = np.stack([rho, np.zeros(nx), p / (GAMMA - 1.0)])
    rng = np.random.default_rng(8)
    r2 = 1.0 + 0.3 * np.sin(2 * np.pi * x + 1.0) + 0.05 * rng.random(nx)
    u2 = 0.5 * np.cos(4 * np.pi * x)
    p2 = 1.0 + 0.2 * np.sin(6 * np.pi * x)
    smooth = np.stack([r2, r2 * u2, p2 / (GAMMA - 1.0) + 0.5 * (r2 * u2) * u2])
    # Sod moving right at u = 2.5 (supersonic on both sides, no cell at rest)
    sup = np.stack([rho, 2.5 * rho, p / (GAMMA - 1.0) + 0.5 * rho * 2.5 * 2.5])
    return {"sod": sod, "smooth": smooth, "supersonic": sup}


@pytest.mark.parametrize("recon,flux", [("teno5i", "rusanov"), ("teno5", "hllc"), ("teno5", "rusanov"),
                                        ("teno5i", "hllc"), ("weno5js", "hllc"), ("weno5z", "rusanov"),
                                        ("weno5n", "hllc"), ("teno5n", "rusanov"), ("teno5a", "hllc"),
                                        ("weno5is", "rusanov"), ("weno5is_ext", "hllc"), ("weno5ismod", "hllc"),
                                        ("weno5z_diff", "rusanov"), ("teno5_diff", "hllc"), ("linear", "rusanov"),
                                        ("teno5", "cu"), ("teno5i", "cu")])
def test_residual_bitwise(orc, harness, recon, flux):
    """Residuals equal to the oracle's bit for bit, and finite (the central-
    upwind flux is NaN where the gas is at rest, as the reference's
    CUFlux is, Equations_Euler1D.cpp:144-150: Sod's initial state only
    checks that both give NaN in the same cells)."""
    for name, q in _states().items():
        r = _stage(harness, q, recon, flux)
        rc = orc.fv1d_rhs(q.shape[1], 0.0, 1.0, GAMMA, q, recon=recon, flux=flux)
        np.testing.assert_array_equal(r, rc, err_msg=name)
        if flux == "cu" and name == "sod":
            assert np.isnan(rc).any()
        else:
            assert np.isfinite(r).all(), name


def test_steps_bitwise(orc, harness):
    """Five full SSP-RK3 steps of Sod (dt from the oracle's CFL rule) with
    the harness's stage bodies equal the oracle's steps."""
    q = _states()["sod"]
    nx = q.shape[1]
    t = 0.0
    for k in range(5):
        qn, tn, n = orc.fv1d_advance(nx, 0.0, 1.0, GAMMA, q, 0.5, t, 0.2, recon="teno5", flux="hllc", max_steps=1)
        # dt = cfl dx / max |u -+ c| (FV_System1D.cpp:270-278)
        u = q[1] / q[0]
        c = np.sqrt(np.abs(GAMMA * ((q[2] - 0.5 * q[1] * u) * (GAMMA - 1.0)) / q[0]))
        dt = 0.5 * (1.0 / nx) / np.max(np.maximum(np.abs(u - c), np.abs(u + c)))
        assert t + dt == tn
        u1 = _stage(harness, q, "teno5", "hllc", 1, dt, q)
        u2 = _stage(harness, u1, "teno5", "hllc", 2, dt, q)
        u3 = _stage(harness, u2, "teno5", "hllc", 3, dt, q)
        np.testing.assert_array_equal(u3, qn)
        q, t = qn, tn
