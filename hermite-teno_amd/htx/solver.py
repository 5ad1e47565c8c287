"""Host-side mirror of the reference's 2D solver interface.

``System2D`` keeps the constructor and method names of the reference's
``FV::System2D`` (FV_System2D.h:12-20) and ``FVHermite::System2D``
(FVHermite_System2D.h:16-24); the numerics run in libhtx.so on the GPU.
Method choices that the reference makes by commenting code in or out
(flux, boundary condition, scheme) are constructor keywords here.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import check, dptr


class System2D:
    """FV::System2D(xl, xr, nx, yb, yt, ny, cflnumber, tmax, tInit, specificHeatRatio)."""

    def __init__(self, xl, xr, nx, yb, yt, ny, cflnumber, tmax, tInit, specificHeatRatio, *,
                 scheme="teno5", flux="hllc", bc="zero_grad", precision="fp64", ref_quirks=True,
                 tile=(0, 0), classifier=None):
        lib = _lib.load()
        p = _lib.HtxParams()
        check(lib.htx_params_default(ctypes.byref(p)))
        p.xl, p.xr, p.nx = float(xl), float(xr), int(nx)
        p.yb, p.yt, p.ny = float(yb), float(yt), int(ny)
        p.cfl, p.tmax, p.tinit, p.gamma = float(cflnumber), float(tmax), float(tInit), float(specificHeatRatio)
        p.scheme = _lib.SCHEMES[scheme]
        p.flux = _lib.FLUXES[flux]
        p.bc = _lib.BCS[bc]
        p.precision = _lib.PRECISIONS[precision]
        p.ref_quirks = 1 if ref_quirks else 0
        p.tile_x, p.tile_y = tile
        self.params = p
        self._h = None
        h = ctypes.c_void_p()
        check(lib.htx_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self.nx, self.ny = int(nx), int(ny)
        self.nfield = 12 if scheme in _lib.HERMITE_SCHEMES else 4
        self.dx = (p.xr - p.xl) / p.nx
        self.dy = (p.yt - p.yb) / p.ny
        self.gamma = p.gamma
        self.tFinal = p.tmax
        self.exact = None
        if precision in ("learned", "learned_fp32"):
            from .learn import default_classifier

            prm, thr = classifier if classifier is not None else default_classifier()
            self.load_classifier(prm, thr)

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            _lib.load().htx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- geometry ---------------------------------------------------------
    def cell_centres(self):
        """Cell centres as FV_System2D.cpp:37-38 (interior cells)."""
        p = self.params
        x = (np.arange(self.nx, dtype=np.float64) + 0.5) * self.dx + p.xl
        y = (np.arange(self.ny, dtype=np.float64) + 0.5) * self.dy + p.yb
        return x, y

    # -- state ------------------------------------------------------------
    def set_state(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(self.nfield, self.ny, self.nx)
        check(_lib.load().htx_set_state(self._h, dptr(q), self.nx))

    def get_state(self):
        q = np.empty((self.nfield, self.ny, self.nx), dtype=np.float64)
        check(_lib.load().htx_get_state(self._h, dptr(q), self.nx))
        return q

    # -- initial conditions (reference method names) ----------------------
    def isentropicVortex(self, centre=(0.0, 0.0), modes=None):
        """FV::System2D::isentropicVortex (FV_System2D.cpp:76-99).  ``centre``
        moves the vortex off the origin; ``modes`` is an (m, 6) array of
        density modes (k, l, amplitude, phase, Lx, Ly)."""
        m = np.zeros((0, 6)) if modes is None else np.ascontiguousarray(modes, dtype=np.float64)
        mm = np.ascontiguousarray(m.reshape(-1), dtype=np.float64)
        ptr = dptr(mm) if mm.size else None
        check(_lib.load().htx_init_vortex(self._h, float(centre[0]), float(centre[1]), ptr, int(m.shape[0])))

    def entropyWave(self, rho=1.0, u=1.0, v=1.0, p=1.0, amplitude=0.5):
        """FV::System2D::entropyWave / FVHermite::System2D::entropyWave
        (FV_System2D.cpp:49-74, FVHermite_System2D.cpp:52-86), generated on
        the device: rho + amplitude sin(pi (x + y)) at constant (u, v, p);
        the reference's case is the default."""
        check(_lib.load().htx_init_entropy_wave(self._h, float(rho), float(u), float(v), float(p), float(amplitude)))

    def l1_density_error(self, t_shift=None):
        """printResult's numerical error on the device (FV_System2D.cpp:
        653-666): mean |rho - rho_exact| against the last analytic initial
        condition advected by t_shift (default: the current time)."""
        err = ctypes.c_double()
        t = self.time if t_shift is None else float(t_shift)
        check(_lib.load().htx_l1_density_error(self._h, t, ctypes.byref(err)))
        return err.value

    def quadrant(self, xs, ys, states):
        """Four-quadrant Riemann data (primitive rows (rho, u, v, p) for
        k = (x >= xs) + 2 (y >= ys)), point-sampled on the device as
        FV::System2D::Riemann2D samples its quadrant template."""
        st = np.ascontiguousarray(np.asarray(states, dtype=np.float64).reshape(16))
        check(_lib.load().htx_init_quadrant(self._h, float(xs), float(ys), dptr(st)))

    def shock_blobs(self, xs, ys, nrm, post, ambient, blobs):
        """Planar shock (post-shock primitive state where (x - xs, y - ys) . nrm < 0)
        into ``ambient`` carrying Gaussian density blobs (cx, cy, A, r)."""
        b = np.ascontiguousarray(np.asarray(blobs, dtype=np.float64).reshape(-1))
        po = np.ascontiguousarray(post, dtype=np.float64)
        am = np.ascontiguousarray(ambient, dtype=np.float64)
        check(_lib.load().htx_init_shock_blobs(self._h, float(xs), float(ys), float(nrm[0]), float(nrm[1]), dptr(po),
                                               dptr(am), dptr(b) if b.size else None, int(b.size // 4)))

    def init_workload(self, w):
        """Initial state of a workload dict from htx.ic (vortex, quadrant or
        shock_blobs)."""
        kind = w.get("kind", "vortex")
        if kind == "vortex":
            self.isentropicVortex(w["centre"], w["modes"])
        elif kind == "quadrant":
            self.quadrant(w["xs"], w["ys"], w["states"])
        else:
            self.shock_blobs(w["xs"], w["ys"], w["nrm"], w["post"], w["ambient"], w["blobs"])

    # -- learned classifier ----------------------------------------------
    def load_classifier(self, params, threshold=0.0):
        """Classifier parameters (htx.learn layout) and logit threshold for
        precision "learned"."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        check(_lib.load().htx_load_classifier(self._h, _lib.fptr(p), float(threshold)))

    def classifier_flags(self):
        """(x-face flags [ny, nx + 2], y-face flags [ny + 2, nx]), uint8, 1 =
        smooth, of the last step's state: x face i | i+1 of row j at
        [j, i + 1] for i = -1 .. nx, y face j | j+1 of column i at [j + 1, i]."""
        fx = np.zeros((self.ny, self.nx + 2), dtype=np.uint8)
        fy = np.zeros((self.ny + 2, self.nx), dtype=np.uint8)
        check(_lib.load().htx_classifier_flags(self._h, fx.ctypes.data, fy.ctypes.data))
        return fx, fy

    def classifier_stats(self, reset=False):
        """(faces reconstructed, flagged smooth, took the linear scheme) since
        the last reset, each face once per stage."""
        f, g, l = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        check(_lib.load().htx_classifier_stats(self._h, ctypes.byref(f), ctypes.byref(g), ctypes.byref(l),
                                               1 if reset else 0))
        return f.value, g.value, l.value

    def classify(self, direction, features=False):
        """Logits of every x (direction 0) or y (1) face of the current state,
        indexed [j, i] (face of cell (i, j) with its +1 neighbour), from the
        bf16 matrix-core inference; optionally the 12 features too."""
        lg = np.empty((self.ny, self.nx), dtype=np.float32)
        ft = np.empty((self.ny, self.nx, 12), dtype=np.float32) if features else None
        check(_lib.load().htx_classify(self._h, int(direction), _lib.fptr(lg),
                                       _lib.fptr(ft) if ft is not None else None))
        return (lg, ft) if features else lg

    # -- the hot path -----------------------------------------------------
    def rhs(self):
        """One residual evaluation L(q) (FV::System2D::rhs)."""
        r = np.empty((self.nfield, self.ny, self.nx), dtype=np.float64)
        check(_lib.load().htx_rhs(self._h, dptr(r), self.nx))
        return r

    def stage(self, k, dt):
        """One SSP-RK3 stage (1, 2 or 3) with the given dt (htx_stage)."""
        check(_lib.load().htx_stage(self._h, int(k), float(dt)))

    def stage_state(self, k):
        """u1 (k = 1) or u2 (k = 2) of the last stages run."""
        q = np.empty((self.nfield, self.ny, self.nx))
        check(_lib.load().htx_get_stage_state(self._h, int(k), dptr(q), self.nx))
        return q

    def compute_dt(self):
        d = ctypes.c_double()
        check(_lib.load().htx_compute_dt(self._h, ctypes.byref(d)))
        return d.value

    def step(self):
        """One SSP-RK3 step (FV::System2D::RungeKutta3); returns dt taken."""
        d = ctypes.c_double()
        check(_lib.load().htx_step(self._h, ctypes.byref(d)))
        return d.value

    def steps_async(self, n):
        check(_lib.load().htx_steps_async(self._h, int(n)))

    def sync(self):
        check(_lib.load().htx_sync(self._h))

    def timeStepping(self, tfinal=None):
        """FV::System2D::timeStepping: advance to tFinal; returns steps taken."""
        n = ctypes.c_int()
        t = self.tFinal if tfinal is None else float(tfinal)
        check(_lib.load().htx_advance(self._h, t, ctypes.byref(n)))
        self.tFinal = t
        return n.value

    @property
    def time(self):
        t = ctypes.c_double()
        n = ctypes.c_int()
        check(_lib.load().htx_get_time(self._h, ctypes.byref(t), ctypes.byref(n)))
        return t.value

    @property
    def nsteps(self):
        t = ctypes.c_double()
        n = ctypes.c_int()
        check(_lib.load().htx_get_time(self._h, ctypes.byref(t), ctypes.byref(n)))
        return n.value

    def save(self, path):
        """Snapshot of the state, time, step count and configuration
        (htx_save; htx.snapshot.read reads it without a GPU)."""
        check(_lib.load().htx_save(self._h, os.fsencode(path)))

    def load(self, path):
        """Restore state, time and step count from a snapshot of a solver with
        the same grid and scheme (htx_load)."""
        check(_lib.load().htx_load(self._h, os.fsencode(path)))

    def timing_enable(self, on=True):
        check(_lib.load().htx_timing_enable(self._h, 1 if on else 0))

    def timing_read(self):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        check(_lib.load().htx_timing_read(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def printResult(self):
        """FV::System2D::printResult (FV_System2D.cpp:653-666): mean |rho -
        rho_exact| over interior cells, when an exact field has been set."""
        if self.exact is None:
            return None
        q = self.get_state()
        return float(np.abs(q[0] - self.exact[0]).sum() / (self.nx * self.ny))
