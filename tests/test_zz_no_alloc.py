"""GPU: tbs_norm2_noalloc, whose partials live in the library's static device
scratch array instead of an allocation, against tbs_norm2 through the C ABI:
bitwise the same norm on two calls in a row (the second on a smaller vector,
so stale partials of the first stay in the scratch), and a vector beyond its
limit (2^29 values) refused before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_norm2_noalloc_matches_norm2(tbs):
    lib = tbs.load()
    ctx = tbs.Context(0, 9)
    for n in (13_824_000, 1_000, 8_193):        # 1,536,000 x 9 values, then smaller vectors
        print(f"case norm2_noalloc n={n}", flush=True)
        x = np.random.default_rng(n).uniform(-1.0, 1.0, n)
        d = ctx.array(x)
        a, b = C.c_double(), C.c_double()
        tbs._capi.check(lib.tbs_norm2(ctx.handle, n, C.c_void_p(d.ptr), C.byref(a)))
        tbs._capi.check(lib.tbs_norm2_noalloc(ctx.handle, n, C.c_void_p(d.ptr), C.byref(b)))
        assert a.value == b.value, (n, a.value, b.value)
        assert abs(b.value - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)
        d.close()
    small = ctx.array(np.ones(16))
    out = C.c_double(-1.0)
    rc = lib.tbs_norm2_noalloc(ctx.handle, (1 << 29) + 1, C.c_void_p(small.ptr), C.byref(out))
    assert rc == -1                              # TBS_ERR_ARG
    assert b"maximum" in lib.tbs_last_error()
    assert out.value == -1.0
    rc = lib.tbs_norm2_noalloc(ctx.handle, 16, C.c_void_p(small.ptr), C.byref(out))
    assert rc == 0 and out.value == 4.0
    small.close()
    ctx.close()
