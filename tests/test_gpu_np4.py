"""Products on four primes end to end (the path above the three-prime bound, E/crt.c:91-99: four primes when the CRT
bound exceeds the three primes' product): the library re-initialised with nprimes = 4 forms a 2^20 x 2^20 product, a fused
+ x, a mixed batch, a graph plan and a grid product, each exact against GMP or the oracle; the k_crt launches carry
four planes (the live profile's byte count); and a product above the three-prime bound gives NTTM_EBOUND under a
fixed three-prime set, before any launch."""
import numpy as np
import pytest

import oracle as O
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu


@pytest.fixture
def np4(nttm):
    import torch
    torch.cuda.synchronize()
    nttm.shutdown()
    nttm.init(0, nprimes=4)
    yield nttm
    torch.cuda.synchronize()
    nttm.shutdown()
    nttm.init(0)


@pytest.mark.skipif(not O.have_gmp(), reason="GMP absent")
def test_np4_plain_2e20_vs_gmp(np4):
    a, _ = O.gen_limbs(1 << 20, 0x5EED0001)
    b, _ = O.gen_limbs(1 << 20, 0x5EED0002)
    say("np4 2^20 x 2^20")
    np4.profile(True)
    c = host(np4.mul(dev(a), dev(b)))
    stats = np4.profile_read()
    np4.profile(False)
    assert O.gmp_check(a, b, c)
    ncoef = (2 << 20) - 1
    assert stats["k_crt"][2] == 8.0 * 4 * ncoef + 8.0 * (2 << 20), stats["k_crt"]   # four residue planes read


@pytest.mark.skipif(not O.have_gmp(), reason="GMP absent")
def test_np4_mul_add_and_carries(np4):
    a, s = O.gen_limbs(70001, 3); b, s = O.gen_limbs(50003, s); x, _ = O.gen_limbs(120004, s)
    say("np4 mul_add")
    c = host(np4.mul_add(dev(a), dev(b), dev(x)))
    assert O.gmp_check(a, b, c, x)
    ones = np.full(40000, O.B - 1, dtype=np.uint64)
    say("np4 (B^n - 1)^2")
    c = host(np4.mul(dev(ones), dev(ones)))
    assert O.gmp_check(ones, ones, c)


def test_np4_batch_equals_each_and_oracle(np4):
    import torch
    rng = np.random.default_rng(0x5EED0201)
    ops, s = [], 0x5EED0201
    for i in range(40):
        na = int(np.exp(rng.uniform(0, np.log(20000)))); nb = int(np.exp(rng.uniform(0, np.log(20000))))
        a, s = O.gen_limbs(na, s); b, s = O.gen_limbs(nb, s)
        if i % 2:
            b = ops[-1][1]
        x = None
        if i % 3 == 0:
            x, s = O.gen_limbs(int(rng.integers(1, len(a) + len(b) + 1)), s)
        ops.append((a, b, x))
    def run(each):
        items = [(dev(a), dev(b), None if x is None else dev(x),
                  torch.full((len(a) + len(b) + (1 if x is not None else 0),), -1, dtype=torch.int64, device="cuda"))
                 for a, b, x in ops]
        pr = np4.make_prods(items)
        (np4.mul_each if each else np4.mul_batch)(pr)
        return [(host(it[3]), p.ncn) for it, p in zip(items, pr)]
    say("np4 batch 40")
    gb = run(False)
    say("np4 each 40")
    ge = run(True)
    for (a, b, x), (cb, nb_), (ce, ne_) in zip(ops, gb, ge):
        w = O.gmp_mul(a, b) if (x is None and O.have_gmp()) else O.mul_school(a, b, x) if len(a) * len(b) <= 4_000_000 \
            else O.mul_ntt(a, b, [np4.prime(i) for i in range(4)], x)
        if x is not None and O.have_gmp():
            assert O.gmp_check(a, b, cb, x)
        else:
            assert np.array_equal(cb, w)
        assert np.array_equal(cb, ce) and nb_ == ne_


def test_np4_plan_equals_eager(np4):
    import torch
    ops, s = [], 0x5EED0202
    for n in (100, 3000, 7000, 12000, 33000, 900):
        a, s = O.gen_limbs(n, s); b, s = O.gen_limbs(n + 17, s)
        ops.append((a, b))
    items = [(dev(a), dev(b), None, torch.zeros(len(a) + len(b), dtype=torch.int64, device="cuda")) for a, b in ops]
    say("np4 plan")
    plan = np4.BatchPlan(np4.make_prods(items), autotune=False)
    for rep in range(3):
        for it in items:
            it[3].fill_(-1)
        plan.run()
        ncn = plan.ncn()
        for (a, b), it, n in zip(ops, items, ncn):
            c = host(it[3])
            assert np.array_equal(c, O.mul_ntt(a, b, [np4.prime(i) for i in range(4)])) and n == len(a) + len(b) - (c[-1] == 0)
    plan.close()


@pytest.mark.skipif(not O.have_gmp(), reason="GMP absent")
def test_np4_grid_product_vs_gmp(nttm):
    """four primes under a 6 MiB workspace cap: a 2^18 x 2^18-limb product formed as a grid of pieces"""
    import torch
    torch.cuda.synchronize()
    nttm.shutdown()
    nttm.init(0, nprimes=4, workspace_bytes=6 << 20)
    try:
        a, s = O.gen_limbs(1 << 18, 0x5EED0203); b, _ = O.gen_limbs((1 << 18) + 333, s)
        say("np4 grid 2^18")
        c = host(nttm.mul(dev(a), dev(b)))
        assert O.gmp_check(a, b, c)
    finally:
        torch.cuda.synchronize()
        nttm.shutdown(); nttm.init(0)


def test_np3_bound_gives_ebound(nttm):
    """a fixed three-prime set refuses a product of max_terms(3) + 1 terms with NTTM_EBOUND (E/crt.c:91-99) before any
    launch: one device allocation of 2 (max_terms + 1) limbs holds a, b and c (never read)"""
    import torch
    torch.cuda.synchronize()
    nttm.shutdown()
    nttm.init(0, nprimes=3)
    try:
        m = nttm.max_terms(3) + 1
        buf = torch.empty(2 * m, dtype=torch.int64, device="cuda")
        pr = (nttm.Prod * 1)(nttm.Prod(buf.data_ptr(), m, buf.data_ptr(), m, buf.data_ptr(), None, 0, 0))
        say("np3 bound")
        with pytest.raises(nttm.NttmError) as e:
            nttm.mul_batch(pr)
        assert e.value.code == -4
        del buf
    finally:
        torch.cuda.synchronize()
        nttm.shutdown(); nttm.init(0)
