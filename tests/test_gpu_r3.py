"""GPU parity of the transforms of 3 2^k points and of products on those lengths, through the C ABI (bit-exact).
The reference's counterparts are ntt_fwd3 / ntt_inv3 / ntt_inv3_pw_y (E/ntt.c:671-779) and its length rule
(E/rns_mul.c:270-277); the oracle restates them (orc_dft3, orc_ntt3_fwd / _inv), pinned in test_oracle.py."""
import contextlib
import os

import numpy as np
import pytest

import oracle as O
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu
B = O.B


def rand_res(n, p, seed):
    a, _ = O.gen_limbs(n, seed)
    return a % np.uint64(p)


def to_int(limbs):
    v = 0
    for x in reversed([int(t) for t in limbs]):
        v = v * B + x
    return v


@contextlib.contextmanager
def config(nttm, **cfg):
    """the library re-initialised with cfg for the block, then restored to the session's configuration"""
    import torch
    session = dict(tile_log=int(os.environ.get("NTTM_TEST_TILE_LOG", "0")), pts=int(os.environ.get("NTTM_TEST_PTS", "0")))
    torch.cuda.synchronize()
    nttm.shutdown()
    nttm.init(0, **dict(session, **cfg))
    try:
        yield nttm
    finally:
        torch.cuda.synchronize()
        nttm.shutdown()
        nttm.init(0, **session)


@pytest.mark.parametrize("prime", range(4))
@pytest.mark.parametrize("logk", [1, 2, 3, 5, 8])
def test_fwd3_equals_dft3(nttm, prime, logk):
    p, w = nttm.prime(prime), nttm.root3(prime, logk)
    n = 3 << logk
    for batch in (1, 2):
        xs = [rand_res(n, p, 10 * logk + prime + b) for b in range(batch)]
        say("fwd3-dft", prime, logk, batch)
        t = dev(np.concatenate(xs))
        nttm.ntt_fwd3(prime, t, logk, batch)
        got = host(t).reshape(batch, -1)
        for b in range(batch):
            assert np.array_equal(got[b], O.dft3(xs[b], p, w)), (prime, logk, b)


@pytest.mark.parametrize("logk", [10, 12, 13, 17, 21])
def test_fwd3_inv3_equal_host(nttm, logk):
    prime = logk % 4
    p, w = nttm.prime(prime), nttm.root3(prime, logk)
    x = rand_res(3 << logk, p, 1000 + logk)
    say("fwd3-host", prime, logk)
    t = dev(x)
    nttm.ntt_fwd3(prime, t, logk)
    X = host(t)
    assert np.array_equal(X, O.ntt3_fwd(x, p, w))
    say("inv3-host", prime, logk)
    nttm.ntt_inv3(prime, t, logk)
    assert np.array_equal(host(t), x)


@pytest.mark.parametrize("ymode", [0, 1, 2])
@pytest.mark.parametrize("logk", [4, 12, 14])
def test_inv3_pw_modes(nttm, ymode, logk):
    prime = 2
    p, w = nttm.prime(prime), nttm.root3(prime, logk)
    n, batch = 3 << logk, 3
    xs = [rand_res(n, p, 70 + b) for b in range(batch)]
    ny = batch if ymode == nttm.Y_FULL else 1 if ymode == nttm.Y_BCAST else (batch + 1) // 2
    ys = [rand_res(n, p, 90 + b) for b in range(ny)]
    say("inv3-pw", ymode, logk)
    t, y = dev(np.concatenate(xs)), dev(np.concatenate(ys))
    nttm.ntt_inv3_pw(prime, t, y, ymode, logk, batch)
    got = host(t).reshape(batch, -1)
    for b in range(batch):
        yb = ys[b] if ymode == nttm.Y_FULL else ys[0] if ymode == nttm.Y_BCAST else ys[b // 2]
        prod = np.array([(int(u) * int(v)) % p for u, v in zip(xs[b], yb)], dtype=np.uint64)
        assert np.array_equal(got[b], O.ntt3_inv(prod, p, w)), (ymode, logk, b)


# sizes around the 3 2^k boundaries: single-pass and two-pass engines, odd lengths, one limb against many
R3_SIZES = [(5, 7), (13, 12), (700, 68), (1024, 512), (1100, 1000), (2047, 1025), (3000, 3000), (6000, 143), (9000, 3287),
            (20000, 4577), (40000, 8000)]


def test_products_on_radix3_lengths(nttm):
    with config(nttm, radix3=nttm.R3_FIT):
        for na, nb in R3_SIZES:
            a, s = O.gen_limbs(na, na * 7 + nb)
            b, _ = O.gen_limbs(nb, s)
            logn, r3 = nttm.debug_product_len(na + nb - 1, nttm.R3_FIT)
            say("r3 product", na, nb, logn, r3)
            c = host(nttm.mul(dev(a), dev(b)))
            want = O.mul_school(a, b) if na * nb <= 4_000_000 else O.mul_ntt(a, b, [nttm.prime(i) for i in range(3)])
            assert np.array_equal(c, want), (na, nb)
            x, _ = O.gen_limbs(na + nb - 3, 5)
            say("r3 product + x", na, nb)
            c = host(nttm.mul_add(dev(a), dev(b), dev(x)))
            assert to_int(c) == to_int(a) * to_int(b) + to_int(x)


def test_radix3_closed_forms(nttm):
    with config(nttm, radix3=nttm.R3_FIT):
        n = 3000                                                 # 5999 coefficients -> 3 2^11
        ones = np.full(n, B - 1, dtype=np.uint64)
        say("r3 all ones", n)
        c = host(nttm.mul(dev(ones), dev(ones)))                 # (B^n - 1)^2 = B^2n - 2 B^n + 1
        assert int(c[0]) == 1 and not c[1:n].any() and int(c[n]) == B - 2 and (c[n + 1:] == B - 1).all()


def test_batch_mixed_lengths_radix3(nttm):
    """a batch whose products fall in both 2^k and 3 2^k groups, every product exact"""
    rng = np.random.default_rng(33)
    import torch
    ops, s = [], 4
    for i in range(40):
        na = int(rng.integers(1, 3000)); nb = int(rng.integers(1, 3000))
        a, s = O.gen_limbs(na, s); b, s = O.gen_limbs(nb, s)
        x = None
        if i % 4 == 0:
            x, s = O.gen_limbs(int(rng.integers(1, na + nb + 1)), s)
        ops.append((a, b, x))
    with config(nttm, radix3=nttm.R3_FIT):
        items = []
        for a, b, x in ops:
            nc = len(a) + len(b) + (1 if x is not None else 0)
            items.append((dev(a), dev(b), None if x is None else dev(x),
                          torch.full((nc,), -1, dtype=torch.int64, device="cuda")))
        say("r3 batch 40")
        prods = nttm.make_prods(items)
        nttm.mul_batch(prods)
        kinds = set()
        for (a, b, x), it, pr in zip(ops, items, prods):
            kinds.add(nttm.debug_product_len(len(a) + len(b) - 1, nttm.R3_FIT)[1])
            assert np.array_equal(host(it[3]), O.mul_school(a, b, x))
        assert kinds == {0, 1}


@pytest.mark.skipif(not O.have_gmp(), reason="GMP absent")
def test_auto_radix3_products_vs_gmp(nttm):
    """the default rule takes 3 2^k at 2^13, 2^14 and 2^22 products: exact against GMP"""
    for na, nb in ((3000, 3000), (6100, 6000), (1_500_000, 1_500_000)):
        logn, r3 = nttm.debug_product_len(na + nb - 1)
        assert r3 == 1
        a, s = O.gen_limbs(na, na + 1)
        b, _ = O.gen_limbs(nb, s)
        say("auto r3 vs gmp", na, nb, logn)
        c = host(nttm.mul(dev(a), dev(b)))
        assert O.gmp_check(a, b, c)
