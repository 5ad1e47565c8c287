"""Snapshot files of htx_save / htx_load (include/htx.h) on the host, no GPU:
read a snapshot into numpy, or write one (for instance to start a solver
from a state computed elsewhere).

Layout, little-endian: char magic[8] "HTXSNAP1"; int32 version (1); int32
nfield, nx, ny, scheme, flux, bc, precision, ref_quirks; double xl, xr, yb,
yt, cfl, tmax, tinit, gamma, t; int64 nsteps; double q[nfield][ny][nx];
uint64 checksum sum_i w_i (2 i + 1) mod 2^64 over the words w_i of q.
"""
import struct

import numpy as np

MAGIC = b"HTXSNAP1"
VERSION = 1
_INTS = ("nfield", "nx", "ny", "scheme", "flux", "bc", "precision", "ref_quirks")
_DOUBLES = ("xl", "xr", "yb", "yt", "cfl", "tmax", "tinit", "gamma", "t")
_HEAD = struct.Struct("<8si8i9dq")


def checksum(q):
    """sum_i w_i (2 i + 1) mod 2^64 over the 64-bit words w_i of the state."""
    w = np.frombuffer(np.ascontiguousarray(q, dtype="<f8").tobytes(), dtype="<u8")
    k = np.arange(w.size, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    with np.errstate(over="ignore"):
        return int(np.sum(w * k, dtype=np.uint64))


def read(path):
    """dict with the configuration fields, t, nsteps and q (nfield, ny, nx)."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < _HEAD.size or raw[:8] != MAGIC:
        raise ValueError(f"{path} is not a snapshot")
    vals = _HEAD.unpack_from(raw, 0)
    if vals[1] != VERSION:
        raise ValueError(f"unsupported snapshot version {vals[1]}")
    out = dict(zip(_INTS, vals[2:10]))
    out.update(zip(_DOUBLES, vals[10:19]))
    out["nsteps"] = vals[19]
    nq = out["nfield"] * out["ny"] * out["nx"]
    end = _HEAD.size + 8 * nq
    if len(raw) != end + 8:
        raise ValueError(f"{path}: truncated snapshot")
    q = np.frombuffer(raw[_HEAD.size:end], dtype="<f8").reshape(out["nfield"], out["ny"], out["nx"]).copy()
    (h,) = struct.unpack_from("<Q", raw, end)
    if h != checksum(q):
        raise ValueError(f"{path}: snapshot checksum mismatch")
    out["q"] = q
    return out


def write(path, q, t=0.0, nsteps=0, **cfg):
    """Write a snapshot of the field-major state q (nfield, ny, nx); cfg holds
    the configuration fields (scheme, flux, bc, precision, ref_quirks, xl, xr,
    yb, yt, cfl, tmax, tinit, gamma; zero when absent)."""
    q = np.ascontiguousarray(q, dtype="<f8")
    nfield, ny, nx = q.shape
    ints = [nfield, nx, ny] + [int(cfg.get(k, 0)) for k in _INTS[3:]]
    dbl = [float(cfg.get(k, 0.0)) for k in _DOUBLES[:-1]] + [float(t)]
    qb = q.tobytes()
    with open(path, "wb") as f:
        f.write(_HEAD.pack(MAGIC, VERSION, *ints, *dbl, int(nsteps)))
        f.write(qb)
        f.write(struct.pack("<Q", checksum(q)))
