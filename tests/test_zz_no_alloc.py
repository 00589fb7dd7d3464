"""The allocation-free updateRR entry point against the allocating one (GPU).

lpmx_update_rr_noalloc runs lpmx_update_rr's kernels with every intermediate in the
library's static device array, so its norms, reaction count and residual must be
the allocating entry point's bit for bit.  Two calls in a row, the second on a
smaller model, check that what the array keeps from the last call does not leak
into the next.
"""
import ctypes as C

import numpy as np
import pytest

from lpm_harness import P, get, push

pytestmark = pytest.mark.gpu


def deformed_model(k, seed):
    L = P.lib()
    n, x, t, c = P.build_sc_model(P.sc_box_for(k))
    rng = np.random.default_rng(seed)
    push("XYZ", x + 1e-3 * rng.standard_normal(x.shape))
    push("PEX", 10.0 * rng.standard_normal(3 * n))
    dbc = np.ones(3 * n, np.int32)
    dbc[rng.uniform(size=3 * n) < 0.2] = 0
    push("DISPBC_INDEX", dbc)
    P.check(L.lpmx_bond_force(6), "elastic force")
    return n, int((dbc == 0).sum())


def both(n):
    L = P.lib()
    fn = L.lpmx_update_rr_noalloc
    fn.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)]
    fn.restype = C.c_int
    a, b, c = C.c_double(), C.c_double(), C.c_int()
    P.check(L.lpmx_update_rr(C.byref(a), C.byref(b), C.byref(c)), "update_rr")
    r1 = get("RESIDUAL", (3 * n,)).copy()
    push("RESIDUAL", np.zeros(3 * n))
    a2, b2, c2 = C.c_double(), C.c_double(), C.c_int()
    P.check(fn(C.byref(a2), C.byref(b2), C.byref(c2)), "update_rr_noalloc")
    r2 = get("RESIDUAL", (3 * n,))
    return (a.value, b.value, c.value, r1), (a2.value, b2.value, c2.value, r2)


def test_update_rr_noalloc_matches_update_rr():
    for k, seed in ((12, 1), (7, 2)):  # 1,728 particles, then 343
        n, nfix = deformed_model(k, seed)
        (a, b, c, r1), (a2, b2, c2, r2) = both(n)
        print(f"{n} particles: |R| {a:.6e} / {a2:.6e}, |reaction| {b:.6e} / {b2:.6e}, count {c} / {c2}", flush=True)
        assert c == nfix and c2 == c
        assert a > 0 and b > 0
        assert a2 == a and b2 == b
        np.testing.assert_array_equal(r2, r1)
