"""Batches of independent 2D Euler problems (include/htx.h, "Batches"): one
persistent TENO5 launch per RK stage for every TENO5 instance, batched
Hermite launches for the Hermite instances, one step capturable as a
hipGraph, and a tile-shape autotuner."""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr, fptr


def params_for(xl, xr, nx, yb, yt, ny, cfl, tmax, gamma, *, scheme="teno5", flux="hllc", bc="zero_grad",
               precision="fp32_mixed", ref_quirks=True, tinit=0.0):
    lib = _lib.load()
    p = _lib.HtxParams()
    check(lib.htx_params_default(ctypes.byref(p)))
    p.xl, p.xr, p.nx = float(xl), float(xr), int(nx)
    p.yb, p.yt, p.ny = float(yb), float(yt), int(ny)
    p.cfl, p.tmax, p.tinit, p.gamma = float(cfl), float(tmax), float(tinit), float(gamma)
    p.scheme = _lib.SCHEMES[scheme]
    p.flux = _lib.FLUXES[flux]
    p.bc = _lib.BCS[bc]
    p.precision = _lib.PRECISIONS[precision]
    p.ref_quirks = 1 if ref_quirks else 0
    return p


class Batch:
    def __init__(self, params):
        lib = _lib.load()
        arr = (_lib.HtxParams * len(params))(*params)
        h = ctypes.c_void_p()
        check(lib.htx_batch_create(arr, len(params), ctypes.byref(h)))
        self._h = h
        self.params = list(params)
        self.n = len(params)

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().htx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nfield(self, i):
        return 12 if self.params[i].scheme in [_lib.SCHEMES[k] for k in _lib.HERMITE_SCHEMES] else 4

    def set_state(self, i, q):
        p = self.params[i]
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(self.nfield(i), p.ny, p.nx)
        check(_lib.load().htx_batch_set_state(self._h, i, dptr(q), p.nx))

    def get_state(self, i):
        p = self.params[i]
        q = np.empty((self.nfield(i), p.ny, p.nx))
        check(_lib.load().htx_batch_get_state(self._h, i, dptr(q), p.nx))
        return q

    def quadrant(self, i, xs, ys, states):
        st = np.ascontiguousarray(np.asarray(states, dtype=np.float64).reshape(16))
        check(_lib.load().htx_batch_init_quadrant(self._h, i, float(xs), float(ys), dptr(st)))

    def load_classifier(self, params=None, threshold=None):
        from .learn import default_classifier

        if params is None:
            params, threshold = default_classifier()
        p = np.ascontiguousarray(params, dtype=np.float32)
        check(_lib.load().htx_batch_load_classifier(self._h, fptr(p), float(threshold)))

    def steps(self, n):
        check(_lib.load().htx_batch_steps(self._h, int(n)))

    def advance(self, max_steps=1 << 30):
        n = ctypes.c_int()
        check(_lib.load().htx_batch_advance(self._h, int(max_steps), ctypes.byref(n)))
        return n.value

    def time(self, i):
        t, n = ctypes.c_double(), ctypes.c_int()
        check(_lib.load().htx_batch_get_time(self._h, i, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def sync(self):
        check(_lib.load().htx_batch_sync(self._h))

    def use_graph(self, enable=True):
        check(_lib.load().htx_batch_use_graph(self._h, 1 if enable else 0))

    def set_tile(self, tx, ty):
        check(_lib.load().htx_batch_set_tile(self._h, int(tx), int(ty)))

    def tune(self, steps=3):
        tx, ty = ctypes.c_int(), ctypes.c_int()
        ms = (ctypes.c_double * 3)()
        check(_lib.load().htx_batch_tune(self._h, int(steps), ctypes.byref(tx), ctypes.byref(ty), ms))
        return (tx.value, ty.value), list(ms)
