"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Arrays are numpy float64,
field-major, interior cells only: q[f, j, i].
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
LIB_OMP = os.path.join(_HERE, "liboracle_omp.so")

_D = ctypes.c_double
_I = ctypes.c_int
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", _I), ("ny", _I),
        ("xl", _D), ("xr", _D), ("yb", _D), ("yt", _D),
        ("gamma", _D),
        ("bc", _I), ("flux", _I), ("quirks", _I), ("recon", _I),
    ]


_SIG = {
    "orc_teno5": (None, [_DP, _DP]),
    "orc_linear5": (None, [_DP, _DP]),
    "orc_teno5_flags": (None, [_DP, _IP]),
    "orc_weno5js3q": (None, [_DP, _DP]),
    "orc_primitive": (None, [_D, _DP, _DP]),
    "orc_conservative": (None, [_D, _DP, _DP]),
    "orc_eigen": (None, [_D, _I, _DP, _DP, _DP, _DP]),
    "orc_rusanov_char": (None, [_D, _I, _DP, _DP, _DP, _DP, _DP]),
    "orc_hllc": (None, [_D, _I, _I, _DP, _DP, _DP]),
    "orc_max_speed": (_D, [_D, _I, _I, _DP]),
    "orc_rk3_combine": (None, [_I, _DP, _DP, _DP, _D, _DP, ctypes.c_long]),
    "orc_gauss9": (None, [_DP, _DP]),
    "orc_fv2d_rhs": (None, [ctypes.POINTER(OrcParams), _DP, _DP]),
    "orc_fv2d_rhs_linear": (None, [ctypes.POINTER(OrcParams), _DP, _DP]),
    "orc_weno5js3q_linear": (None, [_DP, _DP]),
    "orc_hteno_flags": (None, [_D, _DP, _DP, _IP]),
    "orc_fv2d_max_speed": (None, [ctypes.POINTER(OrcParams), _DP, _DP, _DP]),
    "orc_face_labels": (None, [ctypes.POINTER(OrcParams), _DP, _I, ctypes.POINTER(ctypes.c_ubyte)]),
    "orc_fv2d_dt": (_D, [ctypes.POINTER(OrcParams), _DP, _D, _D, _D]),
    "orc_fv2d_rk3": (None, [ctypes.POINTER(OrcParams), _DP, _D]),
    "orc_fv2d_advance": (_I, [ctypes.POINTER(OrcParams), _DP, _D, _D, _D, _I, _DP]),
    "orc_ic_vortex": (None, [ctypes.POINTER(OrcParams), _D, _D, _DP, _I, _DP]),
    "orc_ic_entropy": (None, [ctypes.POINTER(OrcParams), _D, _D, _D, _D, _D, _D, _D, _I, _DP]),
    "orc_hteno": (None, [_D, _DP, _DP, _DP]),
    "orc_hermite_limiter": (_D, [_D, _DP, _DP]),
    "orc_dflux": (None, [_D, _I, _DP, _DP, _DP]),
    "orc_fvh2d_rhs": (None, [ctypes.POINTER(OrcParams), _DP, _DP]),
    "orc_fvh2d_rk3": (None, [ctypes.POINTER(OrcParams), _DP, _D]),
    "orc_fvh2d_rhs_lim": (None, [ctypes.POINTER(OrcParams), _DP, _DP, _DP]),
    "orc_fvh2d_advance": (_I, [ctypes.POINTER(OrcParams), _DP, _D, _D, _D, _I, _DP]),
    "orc_ic_vortex_hermite": (None, [ctypes.POINTER(OrcParams), _D, _D, _DP, _I, _DP]),
    "orc_teno5i": (None, [_DP, _DP]),
    "orc_e1_hllc": (None, [_D, _I, _DP, _DP, _DP]),
    "orc_fv1d_rhs": (None, [_I, _D, _D, _D, _I, _I, _I, _DP, _DP]),
    "orc_fv1d_advance": (_I, [_I, _D, _D, _D, _I, _I, _I, _DP, _D, _D, _D, _I, _DP]),
    "orc_ic_sod": (None, [_I, _D, _D, _D, _DP]),
    "orc_fv1d_kxrcf": (None, [_I, _D, _D, _DP, ctypes.POINTER(ctypes.c_ubyte)]),
    "orc_fv1d_lmpe": (None, [_I, _D, _D, _I, _DP, ctypes.POINTER(ctypes.c_ubyte)]),
    "orc_weno5": (None, [_DP, _I, _DP]),
    "orc_teno5_indicators": (None, [_DP, _DP, _DP]),
    "orc_hteno_indicators": (None, [_DP, _DP, _DP, _DP, _DP]),
    "orc_recon": (None, [_I, _DP, _D, _DP]),
    "orc_e1_cuflux": (None, [_D, _DP, _DP, _DP, _DP, _DP]),
    "orc_weno3js": (None, [_DP, _DP]),
    "orc_weno3js_2q": (None, [_DP, _DP]),
    "orc_quad2": (None, [_I, _DP, _DP]),
    "orc_quad4_lobatto": (None, [_I, _DP, _DP]),
    "orc_quad4_legendre": (None, [_I, _DP, _DP]),
    "orc_herm_recon": (None, [_I, _D, _DP, _DP, _DP]),
    "orc_hweno_glq": (None, [_I, _D, _DP, _DP, _DP]),
}

_libs = {}

# reconstruction names: the 5-point routines a 1D or 2D solver can take
# (ORC_RECON_*), and the Hermite face reconstructions (ORC_HRECON_*)
RECON = {"teno5": 0, "teno5i": 1, "weno5js": 2, "weno5z": 3, "weno5n": 4, "teno5n": 5, "teno5a": 6,
         "weno5is": 7, "weno5is_ext": 8, "weno5ismod": 9, "weno5z_diff": 10, "teno5_diff": 11, "linear": 12}
HRECON = {"hteno": 0, "hweno": 1, "htenoi": 2, "hlinear": 3}


def load(omp=False):
    path = LIB_OMP if omp else LIB
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIG.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _libs[path] = lib
    return lib


def _p(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_DP)


def params(nx, ny, xl=0.0, xr=1.0, yb=0.0, yt=1.0, gamma=1.4, bc="zero_grad", flux="hllc", quirks=True,
           recon="teno5"):
    """recon: a RECON name for the TENO5 solver, an HRECON name for the
    Hermite solver (orc_fvh2d_*)."""
    return OrcParams(nx, ny, xl, xr, yb, yt, gamma, {"zero_grad": 0, "periodic": 1}[bc],
                     {"rusanov": 0, "hllc": 1}[flux], 1 if quirks else 0,
                     HRECON[recon] if recon in HRECON else RECON[recon])


class Oracle:
    def __init__(self, omp=False):
        self.lib = load(omp)

    # scalar kernels
    def teno5(self, u):
        u, pu = _p(u)
        out = np.empty(2)
        self.lib.orc_teno5(pu, out.ctypes.data_as(_DP))
        return out

    def linear5(self, u):
        u, pu = _p(u)
        out = np.empty(2)
        self.lib.orc_linear5(pu, out.ctypes.data_as(_DP))
        return out

    def teno5_flags(self, u):
        u, pu = _p(u)
        f = np.zeros(6, dtype=np.int32)
        self.lib.orc_teno5_flags(pu, f.ctypes.data_as(_IP))
        return f

    def weno3q(self, u):
        u, pu = _p(u)
        out = np.empty(3)
        self.lib.orc_weno5js3q(pu, out.ctypes.data_as(_DP))
        return out

    def weno3q_linear(self, u):
        u, pu = _p(u)
        out = np.empty(3)
        self.lib.orc_weno5js3q_linear(pu, out.ctypes.data_as(_DP))
        return out

    def primitive(self, gamma, q):
        q, pq = _p(q)
        w = np.empty(4)
        self.lib.orc_primitive(gamma, pq, w.ctypes.data_as(_DP))
        return w

    def conservative(self, gamma, w):
        w, pw = _p(w)
        q = np.empty(4)
        self.lib.orc_conservative(gamma, pw, q.ctypes.data_as(_DP))
        return q

    def eigen(self, gamma, direction, qa, qb):
        qa, pa = _p(qa)
        qb, pb = _p(qb)
        L = np.empty(16)
        R = np.empty(16)
        self.lib.orc_eigen(gamma, direction, pa, pb, L.ctypes.data_as(_DP), R.ctypes.data_as(_DP))
        return L.reshape(4, 4), R.reshape(4, 4)

    def rusanov_char(self, gamma, direction, ql, qr, L, R):
        ql, pl = _p(ql)
        qr, pr = _p(qr)
        L, pL = _p(np.asarray(L).reshape(16))
        R, pR = _p(np.asarray(R).reshape(16))
        f = np.empty(4)
        self.lib.orc_rusanov_char(gamma, direction, pl, pr, pL, pR, f.ctypes.data_as(_DP))
        return f

    def hllc(self, gamma, direction, ql, qr, quirks=True):
        ql, pl = _p(ql)
        qr, pr = _p(qr)
        f = np.empty(4)
        self.lib.orc_hllc(gamma, direction, 1 if quirks else 0, pl, pr, f.ctypes.data_as(_DP))
        return f

    def max_speed(self, gamma, direction, w, quirks=True):
        w, pw = _p(w)
        return self.lib.orc_max_speed(gamma, direction, 1 if quirks else 0, pw)

    def rk3_combine(self, stage, q0, qs, rhs, dt):
        q0, p0 = _p(q0)
        qs, ps = _p(qs)
        rhs, pr = _p(rhs)
        out = np.empty_like(q0)
        self.lib.orc_rk3_combine(stage, p0, ps, pr, dt, out.ctypes.data_as(_DP), q0.size)
        return out

    def gauss9(self):
        x = np.empty(9)
        w = np.empty(9)
        self.lib.orc_gauss9(x.ctypes.data_as(_DP), w.ctypes.data_as(_DP))
        return x, w

    # 2D FV
    def fv2d_rhs(self, prm, q):
        q, pq = _p(q)
        r = np.empty_like(q)
        self.lib.orc_fv2d_rhs(ctypes.byref(prm), pq, r.ctypes.data_as(_DP))
        return r

    def fv2d_rhs_linear(self, prm, q):
        """The residual with linear reconstruction and linear transverse
        quadrature on every face."""
        q, pq = _p(q)
        r = np.empty_like(q)
        self.lib.orc_fv2d_rhs_linear(ctypes.byref(prm), pq, r.ctypes.data_as(_DP))
        return r

    def face_labels(self, prm, q, direction):
        """TENO5 cut-off label (1 = nothing cut) of every x (direction 1) or
        y (direction 2) face, indexed [j, i]: the face of cell (i, j) with
        its +1 neighbour."""
        q, pq = _p(q)
        out = np.empty((prm.ny, prm.nx), dtype=np.uint8)
        self.lib.orc_face_labels(ctypes.byref(prm), pq, int(direction),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
        return out

    def fv2d_max_speed(self, prm, q):
        q, pq = _p(q)
        sx, sy = _D(), _D()
        self.lib.orc_fv2d_max_speed(ctypes.byref(prm), pq, ctypes.byref(sx), ctypes.byref(sy))
        return sx.value, sy.value

    def fv2d_dt(self, prm, q, cfl, t, tfinal):
        q, pq = _p(q)
        return self.lib.orc_fv2d_dt(ctypes.byref(prm), pq, cfl, t, tfinal)

    def fv2d_rk3(self, prm, q, dt):
        q = np.array(q, dtype=np.float64, copy=True, order="C")
        self.lib.orc_fv2d_rk3(ctypes.byref(prm), q.ctypes.data_as(_DP), dt)
        return q

    def fv2d_advance(self, prm, q, cfl, t0, tfinal, max_steps=0):
        q = np.array(q, dtype=np.float64, copy=True, order="C")
        t = _D()
        n = self.lib.orc_fv2d_advance(ctypes.byref(prm), q.ctypes.data_as(_DP), cfl, t0, tfinal, max_steps,
                                      ctypes.byref(t))
        return q, t.value, n

    def ic_entropy(self, prm, rho=1.0, u=1.0, v=1.0, p=1.0, amp=0.5, shift=(0.0, 0.0), hermite=False):
        """Entropy-wave cell averages (4 fields, or 12 with the derivative
        moments), centres shifted by -shift (the exact solution at t is
        shift = (u t, v t))."""
        q = np.empty((12 if hermite else 4, prm.ny, prm.nx))
        self.lib.orc_ic_entropy(ctypes.byref(prm), rho, u, v, p, amp, shift[0], shift[1], 1 if hermite else 0,
                                q.ctypes.data_as(_DP))
        return q

    def ic_vortex(self, prm, centre=(0.0, 0.0), modes=None):
        m = np.zeros((0, 6)) if modes is None else np.asarray(modes, dtype=np.float64).reshape(-1, 6)
        mm = np.ascontiguousarray(m.reshape(-1))
        q = np.empty((4, prm.ny, prm.nx))
        self.lib.orc_ic_vortex(ctypes.byref(prm), centre[0], centre[1],
                               mm.ctypes.data_as(_DP) if mm.size else None, m.shape[0], q.ctypes.data_as(_DP))
        return q

    # Hermite-TENO
    def hteno(self, dif, u, d):
        u, pu = _p(u)
        d, pd = _p(d)
        out = np.empty(4)
        self.lib.orc_hteno(dif, pu, pd, out.ctypes.data_as(_DP))
        return out

    def hteno_flags(self, dif, u, d):
        """Cut-off flags (1 = kept): u^- [0:3], u^+ [3:6], d^- [6:9], d^+ [9:12]."""
        u, pu = _p(u)
        d, pd = _p(d)
        f = np.zeros(12, dtype=np.int32)
        self.lib.orc_hteno_flags(dif, pu, pd, f.ctypes.data_as(_IP))
        return f

    def hermite_limiter(self, dif, u, d):
        u, pu = _p(u)
        d, pd = _p(d)
        return self.lib.orc_hermite_limiter(dif, pu, pd)

    def dflux(self, gamma, direction, q, w):
        q, pq = _p(q)
        w, pw = _p(w)
        M = np.empty(16)
        self.lib.orc_dflux(gamma, direction, pq, pw, M.ctypes.data_as(_DP))
        return M.reshape(4, 4)

    def fvh2d_rhs(self, prm, q):
        q, pq = _p(q)
        r = np.empty_like(q)
        self.lib.orc_fvh2d_rhs(ctypes.byref(prm), pq, r.ctypes.data_as(_DP))
        return r

    def fvh2d_rhs_lim(self, prm, q):
        q, pq = _p(q)
        r = np.empty_like(q)
        ql = np.empty_like(q)
        self.lib.orc_fvh2d_rhs_lim(ctypes.byref(prm), pq, r.ctypes.data_as(_DP), ql.ctypes.data_as(_DP))
        return r, ql

    def fvh2d_rk3(self, prm, q, dt):
        q = np.array(q, dtype=np.float64, copy=True, order="C")
        self.lib.orc_fvh2d_rk3(ctypes.byref(prm), q.ctypes.data_as(_DP), dt)
        return q

    def fvh2d_advance(self, prm, q, cfl, t0, tfinal, max_steps=0):
        q = np.array(q, dtype=np.float64, copy=True, order="C")
        t = _D()
        n = self.lib.orc_fvh2d_advance(ctypes.byref(prm), q.ctypes.data_as(_DP), cfl, t0, tfinal, max_steps,
                                       ctypes.byref(t))
        return q, t.value, n

    def ic_vortex_hermite(self, prm, centre=(0.0, 0.0), modes=None):
        m = np.zeros((0, 6)) if modes is None else np.asarray(modes, dtype=np.float64).reshape(-1, 6)
        mm = np.ascontiguousarray(m.reshape(-1))
        q = np.empty((12, prm.ny, prm.nx))
        self.lib.orc_ic_vortex_hermite(ctypes.byref(prm), centre[0], centre[1],
                                       mm.ctypes.data_as(_DP) if mm.size else None, m.shape[0],
                                       q.ctypes.data_as(_DP))
        return q

    # 1D
    def teno5i(self, u):
        u, pu = _p(u)
        out = np.empty(2)
        self.lib.orc_teno5i(pu, out.ctypes.data_as(_DP))
        return out

    def e1_hllc(self, gamma, ql, qr, quirks=True):
        ql, pl = _p(ql)
        qr, pr = _p(qr)
        f = np.empty(3)
        self.lib.orc_e1_hllc(gamma, 1 if quirks else 0, pl, pr, f.ctypes.data_as(_DP))
        return f

    def fv1d_rhs(self, nx, xl, xr, gamma, q, recon="teno5", flux="hllc", quirks=True):
        q, pq = _p(q)
        r = np.empty_like(q)
        self.lib.orc_fv1d_rhs(nx, xl, xr, gamma, RECON[recon], {"rusanov": 0, "hllc": 1, "cu": 2}[flux],
                              1 if quirks else 0, pq, r.ctypes.data_as(_DP))
        return r

    def fv1d_advance(self, nx, xl, xr, gamma, q, cfl, t0, tfinal, recon="teno5", flux="hllc", quirks=True,
                     max_steps=0):
        q = np.array(q, dtype=np.float64, copy=True, order="C")
        t = _D()
        n = self.lib.orc_fv1d_advance(nx, xl, xr, gamma, RECON[recon],
                                      {"rusanov": 0, "hllc": 1, "cu": 2}[flux], 1 if quirks else 0, q.ctypes.data_as(_DP),
                                      cfl, t0, tfinal, max_steps, ctypes.byref(t))
        return q, t.value, n

    def weno5(self, u, z=False):
        """weno5JS (z False) or weno5Z of a 6-value window: (u-, u+) at the
        window's centre face."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.empty(2)
        self.lib.orc_weno5(u.ctypes.data_as(_DP), 1 if z else 0, out.ctypes.data_as(_DP))
        return out

    def fv1d_kxrcf(self, nx, xl, xr, q):
        """KXRCF troubled-cell flags (uint8, nx) of a 1D state (3, nx)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        f = np.zeros(nx, dtype=np.uint8)
        self.lib.orc_fv1d_kxrcf(nx, xl, xr, q.ctypes.data_as(_DP), f.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
        return f

    def fv1d_lmpe(self, nx, xl, xr, q, bc="periodic"):
        """LMPE troubled-cell flags (uint8, (3, nx)) of a 1D state (3, nx),
        from a reset state."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        f = np.zeros((3, nx), dtype=np.uint8)
        self.lib.orc_fv1d_lmpe(nx, xl, xr, {"zero_grad": 0, "periodic": 1}[bc], q.ctypes.data_as(_DP),
                               f.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
        return f

    def ic_sod(self, nx, xl=0.0, xr=1.0, gamma=1.4):
        q = np.empty((3, nx))
        self.lib.orc_ic_sod(nx, xl, xr, gamma, q.ctypes.data_as(_DP))
        return q

    # indicators (test hooks)
    def teno5_indicators(self, u):
        u, pu = _p(u)
        SI = np.empty(6)
        b = np.empty(2)
        self.lib.orc_teno5_indicators(pu, SI.ctypes.data_as(_DP), b.ctypes.data_as(_DP))
        return SI, b

    def hteno_indicators(self, u, ud):
        u, pu = _p(u)
        ud, pd = _p(ud)
        SI, SId, b = np.empty(3), np.empty(3), np.empty(2)
        self.lib.orc_hteno_indicators(pu, pd, SI.ctypes.data_as(_DP), SId.ctypes.data_as(_DP), b.ctypes.data_as(_DP))
        return SI, SId, b

    # the remaining Lagrange / Hermite routines
    def recon(self, kind, u, diff=0.0):
        """A 5-point reconstruction (RECON name) of a 6-value window:
        (u-, u+) at its centre face; diff for the routines that take one."""
        u, pu = _p(u)
        out = np.empty(2)
        self.lib.orc_recon(RECON[kind], pu, diff, out.ctypes.data_as(_DP))
        return out

    def weno3js(self, u):
        u, pu = _p(u)
        out = np.empty(2)
        self.lib.orc_weno3js(pu, out.ctypes.data_as(_DP))
        return out

    def weno3js_2q(self, u):
        u, pu = _p(u)
        out = np.empty(2)
        self.lib.orc_weno3js_2q(pu, out.ctypes.data_as(_DP))
        return out

    def quad2(self, kind, u):
        """kind 0 weno5JS2PointsQuadrature, 1 weno5Z2..., 2 teno5th2..."""
        u, pu = _p(u)
        out = np.empty(2)
        self.lib.orc_quad2(kind, pu, out.ctypes.data_as(_DP))
        return out

    def quad4(self, kind, u, lobatto=True):
        u, pu = _p(u)
        out = np.empty(4)
        (self.lib.orc_quad4_lobatto if lobatto else self.lib.orc_quad4_legendre)(kind, pu, out.ctypes.data_as(_DP))
        return out

    def herm_recon(self, kind, dif, u, d):
        """A Hermite face reconstruction (HRECON name): (u-, u+, d-, d+)."""
        u, pu = _p(u)
        d, pd = _p(d)
        out = np.empty(4)
        self.lib.orc_herm_recon(HRECON[kind], dif, pu, pd, out.ctypes.data_as(_DP))
        return out

    def hweno_glq(self, npts, dif, u, d):
        u, pu = _p(u)
        d, pd = _p(d)
        out = np.empty(4 if npts == 2 else 8)
        self.lib.orc_hweno_glq(npts, dif, pu, pd, out.ctypes.data_as(_DP))
        return out

    def e1_cuflux(self, gamma, L, R, ql, qr):
        """Euler1D::CUFlux with row-major 3x3 L, R."""
        L, pL = _p(L)
        R, pR = _p(R)
        ql, pl = _p(ql)
        qr, pr = _p(qr)
        f = np.empty(3)
        self.lib.orc_e1_cuflux(gamma, pL, pR, pl, pr, f.ctypes.data_as(_DP))
        return f
