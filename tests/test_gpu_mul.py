"""GPU parity of the product (nttm_mul, nttm_mul_add, nttm_mul_batch, nttm_mul_each) against the oracle: the schoolbook
and GMP exactly, closed forms, edge sizes and operand kinds, batch == one-by-one, and the two headline shapes (2^20 and
2^27 limbs per operand).  Mirrors the reference's t_mul sections 1, 2, 4 and 4b (E/tests/t_mul.c:116-328)."""
import numpy as np
import pytest

import oracle as O
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu
B = O.B


def to_int(limbs):
    v = 0
    for x in reversed([int(t) for t in limbs]):
        v = v * B + x
    return v


def operands(na, nb, seed, kind=O.GEN_UNIFORM):
    a, s = O.gen_limbs(na, seed, kind)
    b, _ = O.gen_limbs(nb, s, kind)
    return a, b


EDGE = [(1, 1), (1, 2), (2, 1), (3, 5), (17, 1), (1, 17), (64, 64), (511, 512), (512, 511), (384 * 5, 5), (1000, 999),
        (2048, 2048), (4097, 4095), (8192, 1)]


@pytest.mark.parametrize("na,nb", EDGE)
def test_mul_edges_vs_school(nttm, na, nb):
    a, b = operands(na, nb, 1000 * na + nb)
    say("mul", na, nb)
    c = host(nttm.mul(dev(a), dev(b)))
    assert np.array_equal(c, O.mul_school(a, b))


@pytest.mark.parametrize("kind", range(5))
def test_mul_kinds(nttm, kind):
    a, b = operands(3000, 2999, 77 + kind, kind)
    say("mul kind", kind)
    c = host(nttm.mul(dev(a), dev(b)))
    assert np.array_equal(c, O.mul_school(a, b))


@pytest.mark.skipif(not O.have_gmp(), reason="GMP absent")
@pytest.mark.parametrize("na,nb", [(1 << 14, 1 << 14), ((1 << 16) + 1, (1 << 16) - 1), (1 << 18, 3), (123457, 65537),
                                   ((1 << 19) - 7, 1 << 19)])
def test_mul_vs_gmp(nttm, na, nb):
    a, b = operands(na, nb, na ^ nb)
    say("mul gmp", na, nb)
    c = host(nttm.mul(dev(a), dev(b)))
    assert len(c) == na + nb and (c < B).all()
    assert O.gmp_check(a, b, c)


def test_mul_closed_forms(nttm):
    n = 1 << 17
    ones = np.full(n, B - 1, dtype=np.uint64)
    say("all ones", n)
    c = host(nttm.mul(dev(ones), dev(ones)))                     # (B^n - 1)^2 = B^2n - 2 B^n + 1
    assert int(c[0]) == 1 and not c[1:n].any() and int(c[n]) == B - 2 and (c[n + 1:] == B - 1).all()
    a, _ = O.gen_limbs(5000, 3)
    say("times one / zero-padded powers")
    c = host(nttm.mul(dev(a), dev(np.array([1], dtype=np.uint64))))
    assert np.array_equal(c[:5000], a) and int(c[5000]) == 0
    pa = np.zeros(7, dtype=np.uint64); pa[6] = 1
    pb = np.zeros(4, dtype=np.uint64); pb[3] = 1
    c = host(nttm.mul(dev(pa), dev(pb)))                         # B^6 B^3 = B^9
    assert int(c[9]) == 1 and c.sum() == 1


def test_mul_add(nttm):
    for na, nb, nx in ((1, 1, 1), (100, 37, 137), (2049, 2048, 10), (5000, 4000, 9000)):
        a, b = operands(na, nb, na + 3 * nb)
        x, _ = O.gen_limbs(nx, nx)
        say("mul_add", na, nb, nx)
        c = host(nttm.mul_add(dev(a), dev(b), dev(x)))
        assert len(c) == na + nb + 1
        assert np.array_equal(c, O.mul_school(a, b, x))
    ones = np.full(3000, B - 1, dtype=np.uint64)                 # the final carry limb: (B^n - 1)^2 + (B^2n - 1)
    x = np.full(6000, B - 1, dtype=np.uint64)
    c = host(nttm.mul_add(dev(ones), dev(ones), dev(x)))
    assert to_int(c) == (B ** 3000 - 1) ** 2 + B ** 6000 - 1 and int(c[6000]) == 1


def test_mul_cut(nttm):
    """the cut product (rns_mul_low_db / high_db / band_db): the low min(na+nb, w) limbs of the FULL product"""
    for na, nb in ((1, 1), (37, 100), (3000, 2500), (20000, 7)):
        a, s = O.gen_limbs(na, na + 5 * nb)
        b, _ = O.gen_limbs(nb, s)
        full = O.mul_school(a, b)
        da, db = dev(a), dev(b)
        for w in sorted({1, 2, 7, nb, na, na + nb - 1, na + nb, na + nb + 9, nttm.WHOLE}):
            for lowcut in (0, min(w, na + nb) // 2):
                say("mul_cut", na, nb, lowcut, w)
                c = host(nttm.mul_cut(da, db, lowcut, w))
                assert np.array_equal(c, full[:min(w, na + nb)]), (na, nb, lowcut, w)
    # huge operands, small window: only the low w limbs of each operand enter the transform
    n = 1 << 22
    a, s = O.gen_limbs(n, 91)
    b, _ = O.gen_limbs(n, s)
    say("mul_cut low 1000 of 2^22 x 2^22")
    c = host(nttm.mul_cut(dev(a), dev(b), 0, 1000))
    assert np.array_equal(c, O.mul_school(a[:1000], b[:1000])[:1000])


def test_grid_above_one_transform(nttm):
    """products that no transform within the workspace cap holds are formed as a grid of piece products
    (E/rns_dist.c:610-700): exact against the oracle, with x, with the closed form's carry ripples, and cut"""
    import torch
    cap = 4 << 20                                                # 4 MiB: 2^17-point products need a grid
    nttm.shutdown(); nttm.init(0, workspace_bytes=cap)
    try:
        P3 = [nttm.prime(i) for i in range(3)]
        a, s = O.gen_limbs(60000, 61)
        b, _ = O.gen_limbs(50000, s)
        full = O.mul_ntt(a, b, P3)
        say("grid product")
        c = host(nttm.mul(dev(a), dev(b)))
        assert np.array_equal(c, full)
        x, _ = O.gen_limbs(110000, 62)
        say("grid product + x")
        c = host(nttm.mul_add(dev(a), dev(b), dev(x)))
        assert np.array_equal(c, O.mul_ntt(a, b, P3, x))
        ones = np.full(65000, B - 1, dtype=np.uint64)                # (B^n - 1)^2: carries through every piece
        say("grid all ones")
        c = host(nttm.mul(dev(ones), dev(ones)))
        n = 65000
        assert int(c[0]) == 1 and not c[1:n].any() and int(c[n]) == B - 2 and (c[n + 1:] == B - 1).all()
        say("grid low cut")
        c = host(nttm.mul_cut(dev(a), dev(b), 0, 70000))
        assert np.array_equal(c, full[:70000])
        say("grid high cut")
        lowcut = 60000
        c = host(nttm.mul_cut(dev(a), dev(b), lowcut))
        hi_got, hi_want = to_int(c[lowcut:]), to_int(full[lowcut:])
        assert 0 <= hi_want - hi_got <= 64 * 64                  # short by the skipped pieces' carries at most
    finally:
        torch.cuda.synchronize(); nttm.shutdown(); nttm.init(0)


def test_zero_operands(nttm):
    import torch
    a, _ = O.gen_limbs(10, 1)
    empty = torch.empty(0, dtype=torch.int64, device="cuda")
    say("zero operands")
    c = host(nttm.mul(dev(a), empty))
    assert len(c) == 10 and not c.any()
    x, _ = O.gen_limbs(5, 2)
    c = host(nttm.mul_add(empty, dev(a), dev(x)))
    assert np.array_equal(c[:5], x) and not c[5:].any()


def make_batch(nprod, seed, max_len=5000, with_x=True, shared_b=True):
    rng = np.random.default_rng(seed)
    items, host_ops = [], []
    s = seed
    for i in range(nprod):
        na = int(np.exp(rng.uniform(0, np.log(max_len))))
        nb = int(np.exp(rng.uniform(0, np.log(max_len))))
        a, s = O.gen_limbs(na, s)
        b, s = O.gen_limbs(nb, s)
        if shared_b and i % 2 == 1:                           # the tree's P1 Q2 + P2 and Q1 Q2 share Q2
            b = host_ops[-1][1]
        x = None
        if with_x and i % 3 == 0:
            x, s = O.gen_limbs(int(rng.integers(1, len(a) + len(b) + 1)), s)
        host_ops.append((a, b, x))
    return host_ops


def run_batch(nttm, host_ops, each=False):
    import torch
    items = []
    for a, b, x in host_ops:
        nc = len(a) + len(b) + (1 if x is not None else 0)
        c = torch.full((max(nc, 1),), -1, dtype=torch.int64, device="cuda")
        items.append((dev(a), dev(b), None if x is None else dev(x), c))
    prods = nttm.make_prods(items)
    (nttm.mul_each if each else nttm.mul_batch)(prods)
    out = []
    for (a, b, x), it, pr in zip(host_ops, items, prods):
        nc = len(a) + len(b) + (1 if x is not None else 0)
        out.append((host(it[3])[:nc], pr.ncn))
    return out


def test_batch_equals_each_equals_oracle(nttm):
    ops = make_batch(96, 11)
    say("batch 96")
    got_b = run_batch(nttm, ops)
    say("each 96")
    got_e = run_batch(nttm, ops, each=True)
    for (a, b, x), (cb, nb_), (ce, ne_) in zip(ops, got_b, got_e):
        want = O.mul_school(a, b, x) if len(a) * len(b) <= 4_000_000 else O.mul_ntt(a, b, [nttm.prime(i) for i in range(3)], x)
        assert np.array_equal(cb, want) and np.array_equal(ce, want)
        nz = np.nonzero(want)[0]
        assert nb_ == ne_ == (int(nz[-1]) + 1 if len(nz) else 0)


def test_batch_carry_chains(nttm):
    """Products whose carries run through whole blocks of limbs, batched with random ones: (B^n - 1)^2 + x with x of
    all limbs B - 1, so the carry chains of neighbouring products in one launch must not leak into each other."""
    ops = []
    s = 77
    for i, n in enumerate((3000, 1, 2048, 4097, 1500, 2049)):
        ones = np.full(n, B - 1, dtype=np.uint64)
        if i % 2:
            a, s = O.gen_limbs(n, s)
            ops.append((a, ones, None))
        else:
            ops.append((ones, ones, np.full(2 * n, B - 1, dtype=np.uint64)))
    say("batch carry chains")
    got = run_batch(nttm, ops)
    for (a, b, x), (c, ncn) in zip(ops, got):
        want = to_int(a) * to_int(b) + (to_int(x) if x is not None else 0)
        assert to_int(c) == want and (c < B).all()


def binsplit_level(n_pairs, seed, lo, hi):
    """one level of a binary-splitting tree: P = P1 Q2 + P2 and Q = Q1 Q2 for n_pairs nodes, the two products of a
    node sharing Q2 (E/binsplit.c:1583-1700); P1 and Q1 of a node have one length, as in the tree"""
    rng = np.random.default_rng(seed)
    ops, s = [], seed
    for _ in range(n_pairs):
        l1, lq2, lp2 = (int(rng.integers(lo, hi)) for _ in range(3))
        p1, s = O.gen_limbs(l1, s); q1, s = O.gen_limbs(l1, s)
        q2, s = O.gen_limbs(lq2, s); p2, s = O.gen_limbs(lp2, s)
        ops.append((p1, q2, p2))
        ops.append((q1, q2, None))
    return ops


@pytest.mark.parametrize("radix3", [1, 2])
def test_batch_shared_b(nttm, radix3):
    """products sharing their B operand transform it once (half the B loads: the profile's bytes of the limb-reading
    kernels count (products + distinct B) transforms); results equal the oracle.  2 x 6 products of 18000-20000
    coefficients: 2^15 points (radix3 NEVER) or 3 2^13 (FIT), one group each."""
    import torch
    ops = binsplit_level(6, 17 + radix3, 9000, 10000)
    items = []
    for a, b, x in ops:
        nc = len(a) + len(b) + (1 if x is not None else 0)
        items.append((dev(a), None, None if x is None else dev(x), torch.full((nc,), -1, dtype=torch.int64, device="cuda")))
    for k in range(0, len(items), 2):                             # the pair's products point at one device copy of Q2
        qb = dev(ops[k][1])
        items[k] = (items[k][0], qb, items[k][2], items[k][3])
        items[k + 1] = (items[k + 1][0], qb, items[k + 1][2], items[k + 1][3])
    nttm.shutdown(); nttm.init(0, radix3=radix3)
    try:
        say("shared-B batch", radix3)
        prods = nttm.make_prods(items)
        nttm.profile(True)
        nttm.mul_batch(prods)
        stats = nttm.profile_read()
        nttm.profile(False)
        for (a, b, x), it in zip(ops, items):
            want = O.mul_ntt(a, b, [nttm.prime(i) for i in range(3)], x)
            assert np.array_equal(host(it[3]), want)
        cnt, cnt_b = len(ops), len(ops) // 2
        name, pts = ("k_pass<fwd,limb>", 1 << 15) if radix3 == 1 else ("k_r3_fwd<limb>", 3 << 13)
        assert stats[name][2] == 3 * (cnt + cnt_b) * pts * 8.0, stats
    finally:
        torch.cuda.synchronize(); nttm.shutdown(); nttm.init(0)


def test_batch_leaf_shapes(nttm):
    """the benchmarked leaf workload itself (bench.leaf_workload: 4096 products, log-uniform lengths in [64, 65536]
    limbs, unequal pairs, splitmix streams 0x5EED0005 / 0x5EED0006), every product checked exactly against GMP (the
    residue form where GMP is absent) and its normalized length against the result"""
    import bench
    ops = [(a, b, None) for a, b in bench.leaf_workload()]
    say("batch 4096 leaves (the bench's workload)")
    got = run_batch(nttm, ops)
    if O.have_gmp():
        ok = O.gmp_check_batch([(a, b, c) for (a, b, _), (c, _n) in zip(ops, got)])
    else:
        ok = [O.tref_res(a, b, c) for (a, b, _), (c, _n) in zip(ops, got)]
    assert all(ok), [i for i, v in enumerate(ok) if not v][:10]
    for c, ncn in got:
        nz = np.nonzero(c)[0]
        assert ncn == int(nz[-1]) + 1


@pytest.mark.skipif(not O.have_gmp(), reason="GMP absent")
def test_plain_product_2e20_vs_gmp(nttm):
    """2^20 x 2^20 limbs, seeds 0x5EED0001 / 0x5EED0002, transform length 2^21, exact against GMP."""
    a, _ = O.gen_limbs(1 << 20, 0x5EED0001)
    b, _ = O.gen_limbs(1 << 20, 0x5EED0002)
    say("2^20 x 2^20")
    c = host(nttm.mul(dev(a), dev(b)))
    assert O.gmp_check(a, b, c)


def test_large_product_2e27_residue_form(nttm):
    """2^27 x 2^27 limbs (transform 2^28), seeds 0x5EED0003 / 0x5EED0004: limbs < B, length, residues mod three
    62-bit primes and both 64-limb edges exactly (the reference checker's form above 2^21 limbs)."""
    import torch
    n = 1 << 27
    a, _ = O.gen_limbs(n, 0x5EED0003)
    b, _ = O.gen_limbs(n, 0x5EED0004)
    say("2^27 x 2^27")
    da, db = dev(a), dev(b)
    c = nttm.mul(da, db)
    torch.cuda.synchronize()
    del da, db
    ch = host(c)
    del c
    assert O.tref_res(a, b, ch)
    if O.have_gmp():
        assert O.tref_edges(a, b, ch)
