"""Snapshot files (include/htx.h htx_save / htx_load) on the host: the
numpy reader and writer of hermite-teno_amd/htx/snapshot.py, the checksum
against a word-by-word loop, and the rejection of corrupt, truncated and
foreign files."""
import os
import struct

import numpy as np
import pytest


def _q(shape, seed=3):
    return np.random.default_rng(seed).normal(size=shape)


def test_round_trip(tmp_path):
    from htx import snapshot

    q = _q((12, 7, 9))
    p = tmp_path / "a.snap"
    snapshot.write(p, q, t=0.125, nsteps=42, scheme=1, flux=1, bc=1, precision=3, ref_quirks=1, xl=-1.0, xr=2.0,
                   yb=0.5, yt=1.5, cfl=0.4, tmax=3.0, tinit=0.0, gamma=1.4)
    r = snapshot.read(p)
    assert (r["nfield"], r["ny"], r["nx"]) == (12, 7, 9)
    assert (r["scheme"], r["flux"], r["bc"], r["precision"], r["ref_quirks"]) == (1, 1, 1, 3, 1)
    assert (r["xl"], r["xr"], r["yb"], r["yt"], r["cfl"], r["tmax"], r["gamma"]) == (-1.0, 2.0, 0.5, 1.5, 0.4, 3.0,
                                                                                      1.4)
    assert r["t"] == 0.125 and r["nsteps"] == 42
    np.testing.assert_array_equal(r["q"], q)
    # header 124 bytes, state, 8-byte checksum
    assert os.path.getsize(p) == 124 + 8 * q.size + 8


def test_checksum_matches_word_loop():
    """The vectorised checksum equals sum_i w_i (2 i + 1) mod 2^64 summed
    word by word with Python integers (special values included)."""
    from htx import snapshot

    q = np.array([[0.0, -0.0, 1.0, np.inf], [np.nan, -2.5, 1e-310, 1e308]])
    words = struct.unpack("<8Q", q.astype("<f8").tobytes())
    ref = sum(w * (2 * i + 1) for i, w in enumerate(words)) % (1 << 64)
    assert snapshot.checksum(q) == ref
    big = _q((3, 50, 40))
    words = np.frombuffer(big.tobytes(), dtype="<u8").tolist()
    assert snapshot.checksum(big) == sum(w * (2 * i + 1) for i, w in enumerate(words)) % (1 << 64)


def test_rejects_corrupt_truncated_and_foreign(tmp_path):
    from htx import snapshot

    q = _q((4, 5, 6))
    p = tmp_path / "b.snap"
    snapshot.write(p, q)
    raw = bytearray(p.read_bytes())
    bad = bytearray(raw)
    bad[124 + 17] ^= 0x10  # one bit of the state
    (tmp_path / "flip.snap").write_bytes(bytes(bad))
    with pytest.raises(ValueError, match="checksum"):
        snapshot.read(tmp_path / "flip.snap")
    (tmp_path / "short.snap").write_bytes(bytes(raw[:-9]))
    with pytest.raises(ValueError, match="truncated"):
        snapshot.read(tmp_path / "short.snap")
    (tmp_path / "other.snap").write_bytes(b"NOTASNAP" + bytes(raw[8:]))
    with pytest.raises(ValueError, match="not a snapshot"):
        snapshot.read(tmp_path / "other.snap")
