"""GPU self-test of the device arithmetic the kernels rely on for bit-exact fp32 results.

div_rn_fast (the kernels' VCC-free division, eden_amd/csrc/device_model.hpp) must give the
same bits as the compiler's correctly rounded division whenever it reports its operands in
range, and the kernels' div_rn falls back to the compiler's division otherwise: checked on
2^30 hashed operand pairs (half raw 32-bit patterns, half with exponents around the fast
range) and on every pair of a list of special values (zeros, range boundaries, denormals,
infinities, NaN).
"""
import ctypes as C

import numpy as np
import pytest

from eden_amd import _native as N

pytestmark = pytest.mark.gpu


def _special_values():
    vals = [0.0, 1.0, 3.0, 1.0 / 3.0, np.inf, np.nan, 1e-45, 1e-40, 1.17549435e-38, 3.4028235e38]
    for e in [1, 2, 3, 22, 23, 24, 25, 26, 27, 28, 85, 86, 87, 88, 126, 127, 128, 165, 166, 167, 168, 200,
              248, 249, 250, 251, 252, 253, 254]:
        for m in (1.0, 1.5, 1.9999999):
            vals.append(m * 2.0 ** (e - 127))
    vals = np.array(vals, np.float32)
    return np.concatenate([vals, -vals])


def test_fast_division_matches_compiler_division():
    lib = N.load_library()
    sp = _special_values()
    bad = C.c_int64(-1)
    print(f"[arith] divide check: 2^30 hashed pairs + {len(sp)}^2 special pairs", flush=True)
    rc = lib.edenhip_diag_divide(20261202, 1 << 30, N.ptr(sp), len(sp), C.byref(bad))
    assert rc == 0
    assert bad.value == 0
