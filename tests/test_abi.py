"""The C ABI library loads and exports every symbol include/htx.h declares;
parameter validation runs on the host and needs no GPU."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "htx.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(htx_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from htx import _lib

    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes table covers the same set
    assert sorted(_lib.SIGNATURES) == syms


def test_abi_version_and_params_struct():
    from htx import _lib

    lib = _lib.load()
    assert lib.htx_abi_version() == 1
    p = _lib.HtxParams()
    assert lib.htx_params_default(ctypes.byref(p)) == 0
    assert p.struct_size == ctypes.sizeof(_lib.HtxParams)
    assert (p.dim, p.nx, p.ny, p.gamma, p.cfl, p.ref_quirks) == (2, 64, 64, 1.4, 0.5, 1)


@pytest.mark.parametrize("field,value", [("nx", 0), ("ny", -3), ("gamma", 1.0), ("cfl", 0.0),
                                         ("flux", 7), ("bc", 9), ("scheme", 17), ("scheme", -1), ("struct_size", 4)])
def test_invalid_params_rejected_without_exit(field, value):
    """Where the reference calls exit(1) (Schemes_Hermite.cpp:16-25 on a zero
    grid size) or prints and continues (bad dir codes), the ABI returns
    HTX_EINVAL with a message and the process lives on."""
    from htx import _lib

    lib = _lib.load()
    p = _lib.HtxParams()
    lib.htx_params_default(ctypes.byref(p))
    setattr(p, field, value)
    h = ctypes.c_void_p()
    rc = lib.htx_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1
    assert h.value is None
    assert lib.htx_last_error()


def test_null_arguments():
    from htx import _lib

    lib = _lib.load()
    assert lib.htx_create(None, None) == -1
    assert lib.htx_step(None, None) == -1
    assert lib.htx_sync(None) == -1
    lib.htx_destroy(None)  # ignored


def test_python_mirror_raises_typed_error():
    import htx

    with pytest.raises(htx.HtxError) as ei:
        htx.System2D(0.0, 1.0, 0, 0.0, 1.0, 8, 0.5, 1.0, 0.0, 1.4)
    assert ei.value.code == -1


def test_batch_rejects_unbuilt_combinations():
    """The TENO5 instances of a batch share flux and precision, and the
    persistent launch (fp64_fast, fp32_mixed) is built for the reference's
    HLLC and the characteristic Rusanov flux only: both are refused
    (HTX_ENOTSUP) before any device work."""
    import htx
    from htx.batch import Batch, params_for

    ps = [params_for(0.0, 1.0, 32, 0.0, 1.0, 32, 0.5, 0.1, 1.4, precision="learned"),
          params_for(0.0, 1.0, 40, 0.0, 1.0, 24, 0.5, 0.1, 1.4, precision="fp32_mixed")]
    with pytest.raises(htx.HtxError) as ei:
        Batch(ps)
    assert ei.value.code == -5
    assert "share flux and precision" in str(ei.value)
    ps = [params_for(0.0, 1.0, 32, 0.0, 1.0, 32, 0.5, 0.1, 1.4, precision="fp32_mixed", ref_quirks=False)]
    with pytest.raises(htx.HtxError) as ei:
        Batch(ps)
    assert ei.value.code == -5
    assert "reference quirks" in str(ei.value)


def test_hermite_rejects_learned_precision():
    """Hermite-TENO is built in fp64, fp64_fast, fp32_mixed and fp32; the learned
    precision (a TENO5 classifier) is refused before any device work, for a
    solver and for batched Hermite instances, as are Hermite instances of
    one batch that mix precisions."""
    import htx
    from htx.batch import Batch, params_for

    with pytest.raises(htx.HtxError) as ei:
        htx.System2D(0.0, 1.0, 16, 0.0, 1.0, 16, 0.5, 1.0, 0.0, 1.4, scheme="hteno", precision="learned")
    assert ei.value.code == -5
    assert "TENO5 scheme only" in str(ei.value)
    ps = [params_for(0.0, 1.0, 32, 0.0, 1.0, 32, 0.5, 0.1, 1.4, scheme="hteno", flux="rusanov",
                     precision="learned")]
    with pytest.raises(htx.HtxError) as ei:
        Batch(ps)
    assert ei.value.code == -5
    ps = [params_for(0.0, 1.0, 32, 0.0, 1.0, 32, 0.5, 0.1, 1.4, scheme="hteno", flux="rusanov", precision=p)
          for p in ("fp64_fast", "fp32_mixed")]
    with pytest.raises(htx.HtxError) as ei:
        Batch(ps)
    assert ei.value.code == -5
    assert "share flux and precision" in str(ei.value)


def test_one_dimensional_parameters():
    """dim = 1 (FV::System1D): FP64 only, TENO5 or TENO5i; TENO5i is refused
    in 2D; dim other than 1 and 2 is a bad argument.  All checked before
    any device work."""
    import htx
    from htx import _lib

    lib = _lib.load()
    import ctypes

    def create(**kw):
        p = _lib.HtxParams()
        assert lib.htx_params_default(ctypes.byref(p)) == 0
        for k, v in kw.items():
            setattr(p, k, v)
        h = ctypes.c_void_p()
        return lib.htx_create(ctypes.byref(p), ctypes.byref(h))

    assert create(dim=1, precision=_lib.PRECISIONS["fp32_mixed"]) == -5
    assert create(dim=1, scheme=_lib.SCHEMES["hteno"]) == -1
    # the 2D reconstruction variants are built in the FP64 precisions only
    for sch in ("teno5i", "weno5js", "weno5z"):
        assert create(dim=2, scheme=_lib.SCHEMES[sch], precision=_lib.PRECISIONS["fp32_mixed"]) == -5
        assert create(dim=2, scheme=_lib.SCHEMES[sch], precision=_lib.PRECISIONS["learned"]) == -5
    assert create(dim=3) == -1
    with pytest.raises(htx.HtxError):
        htx.System1D(0.0, 1.0, 0, 0.5, 0.2, 0.0, 1.4)
