"""The drop-in boundary on the GPU: products whose operands and results live in host memory (pageable numpy arrays,
pinned torch tensors, any mix with device tensors) are staged through HBM and come out bit-exact against GMP, as the
reference's rns_mul_batch takes registered host memory and stages the rest (E/rns_mul.h:42-47, E/rns_mul.c:807-817);
a batch larger than the staging chunk runs in sub-batches; entry points that take device memory only refuse host
memory, another allocation's tail, and stale plans with an error code before any launch; a plan keeps its operands
alive after the caller drops them."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu
B = O.B


def host_batch(seed, nprod, max_len, with_x=True, shared_b=True):
    """(a, b, x) numpy operands: log-uniform lengths, every third with x, every second sharing the previous B array"""
    rng = np.random.default_rng(seed)
    ops, s = [], seed
    for i in range(nprod):
        na = int(np.exp(rng.uniform(0, np.log(max_len))))
        nb = int(np.exp(rng.uniform(0, np.log(max_len))))
        a, s = O.gen_limbs(na, s)
        b, s = O.gen_limbs(nb, s)
        if shared_b and i % 2 == 1:
            b = ops[-1][1]
        x = None
        if with_x and i % 3 == 0:
            x, s = O.gen_limbs(int(rng.integers(1, len(a) + len(b) + 1)), s)
        ops.append((a, b, x))
    return ops


def check(a, b, x, c, ncn):
    nc = len(a) + len(b) + (1 if x is not None else 0)
    c = np.asarray(c, dtype=np.uint64)[:nc]
    if O.have_gmp():
        assert O.gmp_check(a, b, c, x)
    else:
        assert np.array_equal(c, O.mul_school(a, b, x))
    nz = np.nonzero(c)[0]
    assert ncn == (int(nz[-1]) + 1 if len(nz) else 0)


def test_batch_host_pageable_vs_gmp(nttm):
    """a, b, x and c all pageable numpy arrays: 64 products up to 30000 limbs, with x and shared B"""
    ops = host_batch(0x5EED0101, 64, 30000)
    outs = [np.full(len(a) + len(b) + (1 if x is not None else 0), 7, dtype=np.uint64) for a, b, x in ops]
    prods = nttm.make_prods([(a, b, x, c) for (a, b, x), c in zip(ops, outs)])
    say("batch of 64, host pageable operands")
    nttm.mul_batch(prods)
    for (a, b, x), c, pr in zip(ops, outs, prods):
        check(a, b, x, c, pr.ncn)


def test_batch_host_mixed_and_pinned(nttm):
    """every combination of residences across the batch: device tensors, pinned CPU tensors and pageable arrays for a,
    b and c, including pinned operands of 2^20 limbs (8 MiB, above the bounce half: copied by DMA directly)"""
    import torch
    ops = host_batch(0x5EED0102, 24, 9000, with_x=True, shared_b=False)
    a_big, s = O.gen_limbs(1 << 20, 0x5EED0103)
    b_big, _ = O.gen_limbs((1 << 20) - 5, s)
    ops.append((a_big, b_big, None))
    kinds = ["dev", "pin", "page"]

    def place(arr, k):
        if k == "dev":
            return dev(arr)
        if k == "pin":
            t = torch.empty(len(arr), dtype=torch.int64, pin_memory=True)
            t.copy_(torch.from_numpy(np.ascontiguousarray(arr, dtype=np.uint64).view(np.int64)))
            return t
        return np.ascontiguousarray(arr, dtype=np.uint64).copy()

    items, placed = [], []
    for i, (a, b, x) in enumerate(ops):
        ka, kb, kc = kinds[i % 3], kinds[(i // 3) % 3], kinds[(i // 9) % 3]
        if i == len(ops) - 1:
            ka, kb, kc = "pin", "pin", "pin"
        nc = len(a) + len(b) + (1 if x is not None else 0)
        c = place(np.zeros(nc, dtype=np.uint64), kc)
        items.append((place(a, ka), place(b, kb), None if x is None else place(x, kinds[i % 2 + 1]), c))
        placed.append((ka, kb, kc))
    prods = nttm.make_prods(items)
    say("batch of 25, mixed residences")
    nttm.mul_batch(prods)
    for (a, b, x), it, pr in zip(ops, items, prods):
        c = it[3]
        c = host(c) if hasattr(c, "is_cuda") and c.is_cuda else (c.numpy().view(np.uint64) if hasattr(c, "numpy") and not isinstance(c, np.ndarray) else c)
        check(a, b, x, c, pr.ncn)


def test_batch_host_split_into_sub_batches(nttm):
    """a 256 KiB staging chunk: the 40-product host batch runs as many sub-batches (shared B copies per sub-batch);
    results equal the same batch on device memory"""
    import torch
    ops = host_batch(0x5EED0104, 40, 6000)
    outs = [np.zeros(len(a) + len(b) + (1 if x is not None else 0), dtype=np.uint64) for a, b, x in ops]
    nttm.shutdown()
    os.environ["NTTM_STAGE_CHUNK"] = str(256 << 10)
    try:
        nttm.init(0)
        prods = nttm.make_prods([(a, b, x, c) for (a, b, x), c in zip(ops, outs)])
        say("host batch in 256 KiB sub-batches")
        nttm.mul_batch(prods)
        dev_items = [(dev(a), dev(b), None if x is None else dev(x),
                      torch.zeros(len(c), dtype=torch.int64, device="cuda")) for (a, b, x), c in zip(ops, outs)]
        dprods = nttm.make_prods(dev_items)
        nttm.mul_batch(dprods)
        for c, it, pr, dp in zip(outs, dev_items, prods, dprods):
            assert np.array_equal(c, host(it[3])) and pr.ncn == dp.ncn
        for (a, b, x), c, pr in zip(ops[:8], outs, prods):
            check(a, b, x, c, pr.ncn)
    finally:
        torch.cuda.synchronize()
        del os.environ["NTTM_STAGE_CHUNK"]
        nttm.shutdown(); nttm.init(0)


def c_mul(L, c, a, b, x=None):
    """nttm_mul / nttm_mul_add through the C ABI on numpy arrays (host memory): the staged single product"""
    p = lambda v: None if v is None else C.c_void_p(v.ctypes.data)
    if x is None:
        return L.nttm_mul(p(c), p(a), len(a), p(b), len(b), None)
    return L.nttm_mul_add(p(c), p(a), len(a), p(b), len(b), p(x), len(x), None)


@pytest.mark.skipif(not O.have_gmp(), reason="GMP absent")
def test_single_products_host_vs_gmp(nttm):
    """nttm_mul and nttm_mul_add on pageable host arrays: 2^20 x 2^20 limbs (16 MiB of operands, many bounce halves)
    and a small one with x; nttm_mul_cut's low limbs"""
    L = nttm.lib()
    a, s = O.gen_limbs(1 << 20, 0x5EED0001)
    b, _ = O.gen_limbs(1 << 20, 0x5EED0002)
    c = np.zeros(2 << 20, dtype=np.uint64)
    say("nttm_mul 2^20 x 2^20 host")
    assert c_mul(L, c, a, b) == 0, L.nttm_last_error()
    assert O.gmp_check(a, b, c)
    a2, s = O.gen_limbs(3001, 5); b2, s = O.gen_limbs(1999, s); x2, _ = O.gen_limbs(4000, s)
    c2 = np.zeros(5001, dtype=np.uint64)
    say("nttm_mul_add host")
    assert c_mul(L, c2, a2, b2, x2) == 0, L.nttm_last_error()
    assert O.gmp_check(a2, b2, c2, x2)
    c3 = np.zeros(3000, dtype=np.uint64)
    say("nttm_mul_cut host")
    rc = L.nttm_mul_cut(C.c_void_p(c3.ctypes.data), C.c_void_p(a2.ctypes.data), 3001, C.c_void_p(b2.ctypes.data), 1999,
                        0, 3000, None)
    assert rc == 0, L.nttm_last_error()
    assert np.array_equal(c3, O.gmp_mul(a2, b2)[:3000])


def test_each_host_and_async(nttm):
    """nttm_mul_each on host arrays; nttm_mul_each_async (launch all, synchronize once) on device arrays; both equal
    the batch"""
    import torch
    ops = host_batch(0x5EED0105, 30, 5000)
    outs = [np.zeros(len(a) + len(b) + (1 if x is not None else 0), dtype=np.uint64) for a, b, x in ops]
    prods = nttm.make_prods([(a, b, x, c) for (a, b, x), c in zip(ops, outs)])
    say("each, host")
    nttm.mul_each(prods)
    items = [(dev(a), dev(b), None if x is None else dev(x), torch.full((len(c),), -1, dtype=torch.int64, device="cuda"))
             for (a, b, x), c in zip(ops, outs)]
    aprods = nttm.make_prods(items)
    say("each async, device")
    nttm.mul_each_async(aprods)
    for (a, b, x), c, it, pr, ap in zip(ops, outs, items, prods, aprods):
        check(a, b, x, c, pr.ncn)
        assert np.array_equal(host(it[3]), c) and ap.ncn == pr.ncn


def test_device_only_entries_refuse_host_memory(nttm):
    """transform primitives, the tree and plans take device memory: host arrays give NTTM_EINVAL before any launch,
    and a device range past the end of its hipMalloc allocation too"""
    import torch
    L = nttm.lib()
    h = np.zeros(1 << 10, dtype=np.uint64)
    hp = C.c_void_p(h.ctypes.data)
    assert L.nttm_ntt_fwd(0, hp, 10, 1, None) == -1
    assert b"host memory" in L.nttm_last_error()
    assert L.nttm_ntt_inv(0, hp, 10, 1, None) == -1
    assert L.nttm_pw(0, hp, hp, 1024, None) == -1
    d = torch.zeros(1 << 10, dtype=torch.int64, device="cuda")
    assert L.nttm_ntt_inv_pw(0, C.c_void_p(d.data_ptr()), hp, 0, 10, 1, None) == -1
    assert L.nttm_load(0, C.c_void_p(d.data_ptr()), hp, 100, 1024, None) == -1
    res = (C.c_uint64 * 8)()
    q = (C.c_uint64 * 1)(4611686018427388039)
    assert L.nttm_limbs_mod(res, hp, 1024, q, 1, None) == -1
    leaf = (nttm.BsNode * 1)(nttm.BsNode(h.ctypes.data, 4, h.ctypes.data, 4))
    pn, qn = C.c_size_t(), C.c_size_t()
    assert L.nttm_bs_tree(leaf, 1, C.c_void_p(d.data_ptr()), 1024, C.byref(pn), C.c_void_p(d.data_ptr()), 1024,
                          C.byref(qn), None) == -1
    with pytest.raises(nttm.NttmError) as e:
        nttm.BatchPlan(nttm.make_prods([(h[:4], h[4:8], None, h[8:16])]), autotune=False)
    assert e.value.code == -1
    # a range past the end of its allocation (hipMalloc of 1 MiB, a transform of 8 MiB), through the HIP runtime this
    # process has loaded (the one the library is bound to)
    path = next(l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64.so" in l)
    hip = C.CDLL(path)
    ptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(ptr), C.c_size_t(1 << 20)) == 0
    try:
        assert L.nttm_ntt_fwd(0, ptr, 20, 1, None) == -1
        assert b"past the end" in L.nttm_last_error()
    finally:
        hip.hipFree(ptr)
    torch.cuda.synchronize()


def test_e_fraction_refuses_device_output(nttm):
    """nttm_e_fraction writes its output from the host (host threads first-touch the pages, the fraction arrives by
    DMA): a device array is refused with NTTM_EINVAL before any work"""
    import torch
    L = nttm.lib()
    d = torch.zeros(64, dtype=torch.int64, device="cuda")
    assert L.nttm_e_fraction(C.cast(d.data_ptr(), C.POINTER(C.c_uint64)), 64, None) == -1
    assert b"device memory" in L.nttm_last_error()
    assert int(nttm.e_fraction(64)[63]) == 718281828459045235   # host memory: e = 2.718281828459045235...


def test_stale_plan_refused_after_reinit(nttm):
    """a plan made before nttm_shutdown + nttm_init returns NTTM_ESTATE from run and ncn (its captured tables are
    gone); destroying it is still safe"""
    import torch
    a, s = O.gen_limbs(900, 11); b, _ = O.gen_limbs(700, s)
    c = torch.zeros(1600, dtype=torch.int64, device="cuda")
    plan = nttm.BatchPlan(nttm.make_prods([(dev(a), dev(b), None, c)]), autotune=False)
    say("plan before re-init")
    plan.run(); plan.ncn()
    assert np.array_equal(host(c), O.mul_school(a, b))
    torch.cuda.synchronize()
    nttm.shutdown(); nttm.init(0)
    with pytest.raises(nttm.NttmError) as e:
        plan.run()
    assert e.value.code == -5
    with pytest.raises(nttm.NttmError) as e:
        plan.ncn()
    assert e.value.code == -5
    plan.close()


def test_plan_keeps_operands_alive(nttm):
    """the caller's tensor list dropped after planning: the plan still holds the arrays, a run equals the eager batch"""
    import gc
    import torch
    ops = host_batch(0x5EED0106, 12, 4000, with_x=True, shared_b=False)
    items = [(dev(a), dev(b), None if x is None else dev(x),
              torch.zeros(len(a) + len(b) + (1 if x is not None else 0), dtype=torch.int64, device="cuda"))
             for a, b, x in ops]
    outs = [it[3] for it in items]
    plan = nttm.BatchPlan(nttm.make_prods(items), autotune=False)
    del items
    gc.collect()
    junk = [torch.full((5000,), -3, dtype=torch.int64, device="cuda") for _ in range(40)]   # reuse freed blocks
    say("plan after the caller dropped its operands")
    plan.run()
    ncn = plan.ncn()
    for (a, b, x), c, n in zip(ops, outs, ncn):
        check(a, b, x, host(c), n)
    del junk
    plan.close()
