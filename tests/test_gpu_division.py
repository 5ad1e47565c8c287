"""The strict build's shared-reciprocal quotients (htx_numerics.h: div_by /
quot for one divisor shared by several numerators, quot_const for constant
divisors) equal the hardware's correctly rounded FP64 division bit for bit
over the operand ranges the kernels use them on: 2^28 random divisors in
[2^-300, 2^300] with four numerators each (|x| in [2^-400, 2^400] and
signed zeros; mantissas random, all ones or all zeros), and 2^28 numerators
over the whole finite range for the constant divisors 5040, 6, 24, 6720 and
1680 of the TENO, quadrature and Hermite indicators."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "tests", "gpu_harness", "libdivcheck.so")


def test_shared_reciprocal_quotients_are_correctly_rounded():
    if not os.path.exists(LIB):
        pytest.skip("tests/gpu_harness/libdivcheck.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    lib.divcheck_run.argtypes = [ctypes.c_ulonglong, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p]
    bad = np.zeros(2, dtype=np.uint64)
    ex = np.zeros(4)
    print("[case] 2^28 random division cases on the GPU", flush=True)
    assert lib.divcheck_run(0x48544E4F00 + 77, 1 << 28, bad.ctypes.data, ex.ctypes.data) == 0
    print(f"mismatches: shared {bad[0]}, constant {bad[1]}; examples {ex}", flush=True)
    assert bad[0] == 0 and bad[1] == 0
