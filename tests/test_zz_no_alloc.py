"""nttm_mul_noalloc, the entry point that allocates nothing: through the C ABI its products equal nttm_mul's on two
calls in a row (the second smaller, so stale contents of the shared workspace from the first call must not leak), at
the largest size it takes, for an empty operand, and it refuses a product above its workspace with NTTM_EINVAL before
any launch.  Runs alone in a fresh process and after the other tests (the workspace is per GPU, held to the end)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu


def noalloc(L, c, a, b):
    return L.nttm_mul_noalloc(C.c_void_p(c.data_ptr()), C.c_void_p(a.data_ptr()) if a.numel() else None, a.numel(),
                              C.c_void_p(b.data_ptr()) if b.numel() else None, b.numel(), None)


def test_noalloc_equals_mul_twice_then_smaller(nttm):
    import torch
    L = nttm.lib()
    for (na, nb), seed in (((1 << 19, (1 << 19) - 7), 0x5EED0301), ((3001, 1999), 0x5EED0302), ((1, 1), 0x5EED0303)):
        a, s = O.gen_limbs(na, seed); b, _ = O.gen_limbs(nb, s)
        da, db = dev(a), dev(b)
        ref = host(nttm.mul(da, db))
        c = torch.full((na + nb,), -1, dtype=torch.int64, device="cuda")
        say("noalloc", na, nb)
        assert noalloc(L, c, da, db) == 0, L.nttm_last_error()
        got = host(c)
        assert np.array_equal(got, ref)
        if na * nb <= 6_000_000:
            assert np.array_equal(got, O.mul_school(a, b))
    # an empty operand: c = 0
    c = torch.full((5,), -1, dtype=torch.int64, device="cuda")
    e = torch.empty(0, dtype=torch.int64, device="cuda")
    assert noalloc(L, c, dev(O.gen_limbs(5, 1)[0]), e) == 0
    assert not host(c).any()


def test_noalloc_refuses_above_its_workspace(nttm):
    import torch
    L = nttm.lib()
    a = torch.zeros((1 << 20) + 1, dtype=torch.int64, device="cuda")
    c = torch.zeros((2 << 20) + 2, dtype=torch.int64, device="cuda")
    assert noalloc(L, c, a, a) == -1
    assert b"maximum" in L.nttm_last_error()
    h = np.zeros(16, dtype=np.uint64)
    assert L.nttm_mul_noalloc(C.c_void_p(c.data_ptr()), C.c_void_p(h.ctypes.data), 8, C.c_void_p(h.ctypes.data), 8,
                              None) == -1
