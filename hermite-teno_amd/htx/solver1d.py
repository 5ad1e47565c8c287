"""Host-side mirror of the reference's 1D solver, FV::System1D
(FV_System1D.h:12-26): constructor (xl, xr, nx, cflnumber, tmax, tInit,
specificHeatRatio) and entropyWave / blastWave / timeStepping /
printResult; the residual, SSP-RK3 stages and CFL reduction run in
libhtx.so on the GPU (dim = 1).  The reference ships TENO5i + characteristic
Rusanov with zero-gradient boundaries (FV_System1D.cpp:161, 199-200, 220);
those are the defaults.  sodShock() is the Sod problem of
Equations_Euler1D.cpp:317-332, point-sampled at the cell centres.
"""
import ctypes
import math
import os

import numpy as np

from . import _lib
from ._lib import check, dptr

# Gauss-Legendre 9-point nodes and weights on [-1, 1] (Schemes_Lagrange.cpp:22-42)
_GX = np.array([0.0, -0.8360311073266358, 0.8360311073266358, -0.9681602395076261, 0.9681602395076261,
                -0.3242534234038089, 0.3242534234038089, -0.6133714327005904, 0.6133714327005904])
_GW = np.array([0.3302393550012598, 0.1806481606948574, 0.1806481606948574, 0.08127438836157441,
                0.08127438836157441, 0.31234707704000286, 0.31234707704000286, 0.26061069640293544,
                0.26061069640293544])


class System1D:
    """FV::System1D(xl, xr, nx, cflnumber, tmax, tInit, specificHeatRatio)."""

    def __init__(self, xl, xr, nx, cflnumber, tmax, tInit, specificHeatRatio, *, scheme="teno5i", flux="rusanov",
                 bc="zero_grad", ref_quirks=True):
        lib = _lib.load()
        p = _lib.HtxParams()
        check(lib.htx_params_default(ctypes.byref(p)))
        p.dim = 1
        p.xl, p.xr, p.nx = float(xl), float(xr), int(nx)
        p.cfl, p.tmax, p.tinit, p.gamma = float(cflnumber), float(tmax), float(tInit), float(specificHeatRatio)
        p.scheme = _lib.SCHEMES[scheme]
        p.flux = _lib.FLUXES[flux]
        p.bc = _lib.BCS[bc]
        p.precision = _lib.PRECISIONS["fp64"]
        p.ref_quirks = 1 if ref_quirks else 0
        self.params = p
        self._h = None
        h = ctypes.c_void_p()
        check(lib.htx_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self.nx = int(nx)
        self.dx = (p.xr - p.xl) / p.nx
        self.gamma = p.gamma
        self.tFinal = p.tmax
        self.exact = None

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().htx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cell_centres(self):
        """FV_System1D.cpp:25 (interior cells)."""
        return (np.arange(self.nx, dtype=np.float64) + 0.5) * self.dx + self.params.xl

    def set_state(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(3, self.nx)
        check(_lib.load().htx_set_state(self._h, dptr(q), self.nx))

    def get_state(self):
        q = np.empty((3, self.nx), dtype=np.float64)
        check(_lib.load().htx_get_state(self._h, dptr(q), self.nx))
        return q

    def _conservative(self, rho, u, p):
        return np.stack([rho, rho * u, p / (self.gamma - 1.0) + 0.5 * (rho * u) * u])

    # -- initial conditions (reference method names) ----------------------
    def entropyWave(self):
        """FV::System1D::entropyWave (FV_System1D.cpp:35-55): rho = 1 +
        0.2 sin(pi x), u = p = 1, 9-point Gauss cell averages; exact = the
        wave advected by u t."""
        x = self.cell_centres()

        def avg(xc):
            acc = np.zeros((3, xc.size))
            for g, w in zip(_GX, _GW):
                xs = xc + self.dx * g / 2.0
                acc += w * self._conservative(1.0 + 0.2 * np.sin(math.pi * xs), np.ones_like(xs),
                                              np.ones_like(xs)) / 2.0
            return acc

        self.set_state(avg(x))
        self.exact = avg(x - self.tFinal)[0]

    def blastWave(self):
        """FV::System1D::blastWave (FV_System1D.cpp:57-65): the Shu-Osher
        state (Equations_Euler1D.cpp:334-345) point-sampled."""
        x = self.cell_centres()
        left = x < -4.0
        rho = np.where(left, 3.857143, 1.0 + 0.2 * np.sin(5.0 * x))
        u = np.where(left, 2.629369, 0.0)
        p = np.where(left, 10.333333, 1.0)
        self.set_state(self._conservative(rho, u, p))
        self.exact = None

    def sodShock(self):
        """Euler1D::sodShock (Equations_Euler1D.cpp:317-332) point-sampled."""
        x = self.cell_centres()
        left = x < 0.5
        self.set_state(self._conservative(np.where(left, 1.0, 0.125), np.zeros_like(x), np.where(left, 1.0, 0.1)))
        self.exact = None

    def troubledCells(self):
        """KXRCF troubled-cell flags (uint8, nx; 1 = discontinuous) of the
        current state, as FV::System1D::rhs computes them
        (FV_System1D.cpp:166-192); the first and last cells read the ghost cells."""
        f = np.zeros(self.nx, dtype=np.uint8)
        check(_lib.load().htx_kxrcf_flags(self._h, f.ctypes.data))
        return f

    def lmpeCells(self):
        """LMPE troubled-cell flags (uint8, (3, nx); 1 = discontinuous) of the
        current state per field, the detector of FVHermite::System1D::rhs
        (FVHermite_System1D.cpp:239-257) swept from clear flags."""
        f = np.zeros((3, self.nx), dtype=np.uint8)
        check(_lib.load().htx_lmpe_flags(self._h, f.ctypes.data))
        return f

    # -- time stepping ----------------------------------------------------
    def rhs(self):
        r = np.empty((3, self.nx), dtype=np.float64)
        check(_lib.load().htx_rhs(self._h, dptr(r), self.nx))
        return r

    def step(self):
        dt = ctypes.c_double()
        check(_lib.load().htx_step(self._h, ctypes.byref(dt)))
        return dt.value

    def timeStepping(self, tfinal=None):
        """FV::System1D::timeStepping (FV_System1D.cpp:264-283)."""
        n = ctypes.c_int()
        t = self.tFinal if tfinal is None else float(tfinal)
        check(_lib.load().htx_advance(self._h, t, ctypes.byref(n)))
        self.tFinal = t
        return n.value

    @property
    def time(self):
        t, n = ctypes.c_double(), ctypes.c_int()
        check(_lib.load().htx_get_time(self._h, ctypes.byref(t), ctypes.byref(n)))
        return t.value

    @property
    def nsteps(self):
        t, n = ctypes.c_double(), ctypes.c_int()
        check(_lib.load().htx_get_time(self._h, ctypes.byref(t), ctypes.byref(n)))
        return n.value

    def printResult(self):
        """FV::System1D::printResult (FV_System1D.cpp:285-300): L1 and Linf
        of the density against the exact solution."""
        if self.exact is None:
            return None
        d = np.abs(self.get_state()[0] - self.exact)
        print(f"Numerical error:{d.sum() / self.nx}")
        print(f"L inf: {d.max()}")
        return d.sum() / self.nx, d.max()

    def save(self, path):
        check(_lib.load().htx_save(self._h, os.fsencode(path)))

    def load(self, path):
        check(_lib.load().htx_load(self._h, os.fsencode(path)))
