"""The short forms of the correctly rounded sqrt and 1/x on the sweeps' hit paths
(sqrt_rn_short on [2^-96, +inf], rcp_rn_short on +-[2^-64, 2^64], morphogenesis_amd/csrc/
mg_device.h) give the bits of the compiler's own sqrtf and 1.0f / x (IEEE division and square
root) for every float of those ranges. The sweeps take them only when every active lane of the
wavefront is in range, so this check together with the bit-exact sweep parity tests pins them.
About 4e9 inputs, a few seconds on one MI355X."""
import ctypes
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hip", "lib", "librounding_check.so")


def f32_bits(x):
    import numpy as np

    return int(np.array([x], np.float32).view(np.uint32)[0])


@pytest.mark.parametrize("which,lo,hi,name", [
    (0, 0x0F800000, 0x7F800000, "sqrt on [2^-96, +inf]"),
    (1, 0x1F800000, 0x5F800000, "1/x on +-[2^-64, 2^64]"),
])
def test_short_forms_match_compiler_expansion(built, which, lo, hi, name):
    assert os.path.exists(LIB), "tests/hip/lib/librounding_check.so is not built (build())"
    assert lo == f32_bits(2.0 ** (-96 if which == 0 else -64))
    lib = ctypes.CDLL(LIB)
    lib.mg_rounding_check.restype = ctypes.c_int
    lib.mg_rounding_check.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.POINTER(ctypes.c_ulonglong)]
    out = (ctypes.c_ulonglong * 3)()
    print(f"case: {name}", flush=True)
    sys.stdout.flush()
    rc = lib.mg_rounding_check(which, lo, hi, out)
    assert rc == 0, f"HIP error {rc}"
    print(f"{name}: {out[0]} inputs, {out[1]} differ", flush=True)
    assert out[0] == (hi - lo + 1) * (1 if which == 0 else 2)
    assert out[1] == 0, f"{name}: {out[1]} inputs differ, first pattern 0x{out[2]:08x}"
