"""Snapshots through the C ABI on the GPU: htx_save writes what the numpy
reader finds (state bit for bit, time, step count, configuration);
continuing from htx_load reproduces the uninterrupted run bit for bit;
files written on the host load; mismatched grids and damaged files are
refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(scheme, n=40, precision="fp64"):
    import htx

    s = htx.System2D(0.0, 10.0, n, 0.0, 10.0, n, 0.5, 10.0, 0.0, 1.4, scheme=scheme, flux="hllc", bc="periodic",
                     precision=precision)
    s.isentropicVortex((5.0, 5.0), None)
    return s


@pytest.mark.parametrize("scheme,precision", [("teno5", "fp64"), ("teno5", "fp32_mixed"), ("hteno", "fp64")])
def test_save_read_load_continue(tmp_path, scheme, precision):
    import htx
    from htx import snapshot

    print(f"[case] snapshot {scheme} {precision}", flush=True)
    a = _solver(scheme, precision=precision)
    a.steps_async(3)
    p = tmp_path / "s.snap"
    a.save(p)
    r = snapshot.read(p)
    np.testing.assert_array_equal(r["q"], a.get_state())
    assert r["t"] == a.time and r["nsteps"] == 3
    assert r["scheme"] == htx._lib.SCHEMES[scheme] and r["nx"] == 40 and r["gamma"] == 1.4
    a.steps_async(4)
    b = _solver(scheme, precision=precision)
    b.load(p)
    assert b.time == r["t"] and b.nsteps == 3
    b.steps_async(4)
    np.testing.assert_array_equal(b.get_state(), a.get_state())
    assert b.time == a.time and b.nsteps == 7


def test_load_host_written_and_refusals(tmp_path):
    import htx
    from htx import snapshot

    s = _solver("teno5", n=32)
    q = s.get_state() * 1.01
    p = tmp_path / "h.snap"
    cfg = dict(scheme=0, flux=1, bc=1, precision=0, ref_quirks=1, xl=0.0, xr=10.0, yb=0.0, yt=10.0, cfl=0.5,
               tmax=10.0, tinit=0.0, gamma=1.4)
    snapshot.write(p, q, t=0.25, nsteps=5, **cfg)
    s.load(p)
    np.testing.assert_array_equal(s.get_state(), q)
    assert s.time == 0.25 and s.nsteps == 5
    # another precision, CFL number and end time may continue the run; other
    # physics may not (the message names the field)
    snapshot.write(tmp_path / "p.snap", q, t=0.25, nsteps=5, **dict(cfg, precision=1, cfl=0.3, tmax=1.0))
    s.load(tmp_path / "p.snap")
    for field, value in (("bc", 0), ("gamma", 5.0 / 3.0), ("flux", 0), ("xr", 12.0), ("ref_quirks", 0)):
        snapshot.write(tmp_path / "x.snap", q, t=0.25, nsteps=5, **dict(cfg, **{field: value}))
        with pytest.raises(htx.HtxError) as ei:
            s.load(tmp_path / "x.snap")
        assert ei.value.code == -1 and field.rstrip("r") in str(ei.value)
    other = _solver("teno5", n=24)
    with pytest.raises(htx.HtxError) as ei:
        other.load(p)
    assert ei.value.code == -1
    raw = bytearray(p.read_bytes())
    raw[200] ^= 1
    (tmp_path / "bad.snap").write_bytes(bytes(raw))
    with pytest.raises(htx.HtxError) as ei:
        s.load(tmp_path / "bad.snap")
    assert ei.value.code == -6
    with pytest.raises(htx.HtxError) as ei:
        s.load(tmp_path / "missing.snap")
    assert ei.value.code == -6
