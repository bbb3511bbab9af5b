"""tmx_norm_noalloc against tmx_norm through the C ABI: the entry point that keeps its partial
sums in the library's fixed device array instead of pool memory.  Two calls in a row, the
second on a smaller input (the array keeps the first call's partials), each bit-identical to
tmx_norm on the same data; the pool's usage does not change across the calls."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _norm(eng, fn, ptr, n):
    r = C.c_double()
    eng.check(getattr(eng.lib(), fn)(ptr, n, C.byref(r), None))
    return r.value


def test_norm_noalloc_matches_norm(eng):
    rng = np.random.default_rng(7)
    big = rng.standard_normal(3_000_001) * 1e3
    buf = eng.DeviceBuffer.from_array(big)
    for n in (big.size, 1000, 1):  # the second and third calls are smaller
        used = eng.pool_info()[1]
        got = _norm(eng, "tmx_norm_noalloc", buf.ptr, n)
        assert eng.pool_info()[1] == used
        ref = _norm(eng, "tmx_norm", buf.ptr, n)
        assert got == ref, (n, got, ref)
        assert abs(got - np.linalg.norm(big[:n])) <= 1e-12 * got
    assert _norm(eng, "tmx_norm_noalloc", buf.ptr, 0) == 0.0
