"""htx_step_noalloc (one SSP-RK3 step on caller-owned device memory, with its
intermediate states in the library's fixed device array) against htx_step on
a solver handle: the same state after the step, bit for bit, and the same
dt, on two calls in a row, the second on a smaller grid (and close to its
end time, so that dt is clamped); a grid above HTX_NOALLOC_MAX_CELLS is
refused."""
import ctypes

import numpy as np
import pytest

from oracle import params

pytestmark = pytest.mark.gpu

GAMMA = 1.4


class _Hip:
    """hipMalloc / hipMemcpy / hipFree from the HIP runtime the library uses
    (the test's own device buffers; the entry point allocates none)."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, a):
        assert self.lib.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.lib.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0

    def free(self, p):
        assert self.lib.hipFree(p) == 0


def _case(orc, hip, nx, ny, precision, flux, bc, t0, tmax, kind):
    import htx
    from htx import _lib
    from htx.batch import params_for
    from htx.ic import quadrant_workload, vortex_modes

    if kind == "vortex":
        dom = (0.0, 10.0, 0.0, 10.0)
        q = orc.ic_vortex(params(nx, ny, *dom, GAMMA, bc, flux, True), (5.0, 5.0), vortex_modes())
    else:
        from htx.ic import piecewise_state

        w = quadrant_workload(64)
        dom = (w["xl"], w["xr"], w["yb"], w["yt"])
        q = piecewise_state(dict(w, nx=nx, ny=ny), nx, ny)
    q = np.ascontiguousarray(q, dtype=np.float64)
    print(f"[case] noalloc {kind} {nx}x{ny} {precision} {flux} {bc} t0={t0} tmax={tmax}", flush=True)
    s = htx.System2D(dom[0], dom[1], nx, dom[2], dom[3], ny, 0.5, tmax, t0, GAMMA, flux=flux, bc=bc,
                     precision=precision)
    s.set_state(q)
    dt_ref = s.step()
    q_ref = s.get_state()
    s.close()

    p = params_for(dom[0], dom[1], nx, dom[2], dom[3], ny, 0.5, tmax, GAMMA, flux=flux, bc=bc, precision=precision)
    dq, ddt = hip.alloc(q.nbytes), hip.alloc(8)
    hip.h2d(dq, q)
    _lib.check(_lib.load().htx_step_noalloc(ctypes.byref(p), dq, t0, ddt, None))
    hip.sync()
    got, dt = np.empty_like(q), np.zeros(1)
    hip.d2h(got, dq)
    hip.d2h(dt, ddt)
    hip.free(dq)
    hip.free(ddt)
    print(f"dt {dt[0]!r} vs {dt_ref!r}; max |dq| {np.abs(got - q_ref).max():.3e}", flush=True)
    assert dt[0] == dt_ref and dt_ref > 0.0
    assert got.tobytes() == q_ref.tobytes()
    return dt_ref


def test_step_noalloc_matches_htx_step(orc):
    hip = _Hip()
    # a 96 x 80 periodic vortex, then a smaller zero-gradient quadrant problem
    # 0.001 before its end time (the second call reuses the array the first
    # left behind)
    _case(orc, hip, 96, 80, "fp64", "hllc", "periodic", 0.0, 2.0, "vortex")
    dt = _case(orc, hip, 40, 36, "fp64", "hllc", "zero_grad", 0.249, 0.25, "quadrant")
    assert dt == pytest.approx(0.001, rel=1e-9)
    _case(orc, hip, 64, 64, "fp32_mixed", "rusanov", "periodic", 0.0, 2.0, "vortex")
    _case(orc, hip, 24, 31, "fp64_fast", "hllc", "zero_grad", 0.0, 0.25, "quadrant")


def test_step_noalloc_refuses_oversize_and_bad_params():
    from htx import _lib
    from htx.batch import params_for

    lib = _lib.load()
    p = params_for(0.0, 1.0, 1025, 0.0, 1.0, 1024, 0.5, 1.0, GAMMA, precision="fp64")
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    assert lib.htx_step_noalloc(ctypes.byref(p), dummy, 0.0, None, None) == -1
    assert b"HTX_NOALLOC_MAX_CELLS" in lib.htx_last_error()
    p = params_for(0.0, 1.0, 32, 0.0, 1.0, 32, 0.5, 1.0, GAMMA, scheme="hteno", precision="fp64")
    assert lib.htx_step_noalloc(ctypes.byref(p), dummy, 0.0, None, None) == -5
    p = params_for(0.0, 1.0, 32, 0.0, 1.0, 32, 0.5, 1.0, GAMMA, precision="learned")
    assert lib.htx_step_noalloc(ctypes.byref(p), dummy, 0.0, None, None) == -5
