"""GPU parity of the transform primitives, the load, the pointwise product and the CRT against the oracle, through the
C ABI (bit-exact: integer residues).  Mirrors the reference's t_ntt sections 1, 2, 4, 4b and t_crt / t_crtk
(E/tests/t_ntt.c:182-249,302-366; E/tests/t_crt.c; E/tests/t_crtk.c)."""
import numpy as np
import pytest

import oracle as O
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu
B = O.B


def rand_res(n, p, seed):
    a, _ = O.gen_limbs(n, seed)
    return a % np.uint64(p)


@pytest.mark.parametrize("prime", range(4))
@pytest.mark.parametrize("logn", [1, 2, 3, 4, 7, 9])
def test_fwd_equals_dft(nttm, prime, logn):
    p, w = nttm.prime(prime), nttm.root(prime, logn)
    for batch in (1, 3):
        xs = [rand_res(1 << logn, p, 100 * logn + 7 * prime + b) for b in range(batch)]
        say("fwd-dft", prime, logn, batch)
        t = dev(np.concatenate(xs))
        nttm.ntt_fwd(prime, t, logn, batch)
        got = host(t).reshape(batch, -1)
        for b in range(batch):
            assert np.array_equal(got[b], O.dft_bitrev(xs[b], p, w)), (prime, logn, b)


@pytest.mark.parametrize("logn", [10, 13, 14, 16, 17, 20, 21, 22, 23])
def test_fwd_equals_host_ntt(nttm, logn):
    for prime in (0, 3):
        p, w = nttm.prime(prime), nttm.root(prime, logn)
        batch = 2 if logn <= 21 else 1
        xs = [rand_res(1 << logn, p, 31 * logn + b) for b in range(batch)]
        say("fwd-host", prime, logn, batch)
        t = dev(np.concatenate(xs))
        nttm.ntt_fwd(prime, t, logn, batch)
        got = host(t).reshape(batch, -1)
        for b in range(batch):
            assert np.array_equal(got[b], O.ntt_fwd(xs[b], p, w)), (prime, logn, b)


@pytest.mark.parametrize("logn", [1, 5, 12, 13, 14, 18, 21, 24])
def test_inverse_roundtrip_and_host(nttm, logn):
    prime = logn % 4
    p, w = nttm.prime(prime), nttm.root(prime, logn)
    x = rand_res(1 << logn, p, 7 + logn)
    say("roundtrip", prime, logn)
    t = dev(x)
    nttm.ntt_fwd(prime, t, logn)
    nttm.ntt_inv(prime, t, logn)
    assert np.array_equal(host(t), x)
    y = rand_res(1 << logn, p, 8 + logn)                        # inverse alone vs the host DIT
    t = dev(y)
    say("inverse", prime, logn)
    nttm.ntt_inv(prime, t, logn)
    assert np.array_equal(host(t), O.ntt_inv(y, p, w))


@pytest.mark.parametrize("logn", [6, 15])
@pytest.mark.parametrize("ymode", [0, 1, 2])
def test_inv_pw_layouts(nttm, logn, ymode):
    prime = 1
    p, w = nttm.prime(prime), nttm.root(prime, logn)
    batch, n = 4, 1 << logn
    xs = [rand_res(n, p, 50 + b) for b in range(batch)]
    ny = {0: batch, 1: 1, 2: batch // 2}[ymode]
    ys = [rand_res(n, p, 90 + b) for b in range(ny)]
    say("inv_pw", logn, ymode)
    t = dev(np.concatenate(xs))
    nttm.ntt_inv_pw(prime, t, dev(np.concatenate(ys)), ymode, logn, batch)
    got = host(t).reshape(batch, -1)
    for b in range(batch):
        yb = ys[{0: b, 1: 0, 2: b // 2}[ymode]]
        prod = np.array([(int(u) * int(v)) % p for u, v in zip(xs[b], yb)], dtype=np.uint64)
        assert np.array_equal(got[b], O.ntt_inv(prod, p, w)), (ymode, b)


def test_pointwise(nttm):
    for prime in range(4):
        p = nttm.prime(prime)
        x, y = rand_res(10000, p, 1), rand_res(10000, p, 2)
        x[:3] = [0, p - 1, p - 1]; y[:3] = [p - 1, p - 1, 1]
        t = dev(x)
        say("pw", prime)
        nttm.pw(prime, t, dev(y))
        assert [int(v) for v in host(t)] == [(int(a) * int(b)) % p for a, b in zip(x, y)]


def test_load(nttm):
    a, _ = O.gen_limbs(3000, 77)
    a[:4] = [0, B - 1, 1, 2 ** 59]
    for prime in range(4):
        p = nttm.prime(prime)
        t = dev(np.zeros(4096, dtype=np.uint64) + np.uint64(5))
        say("load", prime)
        nttm.load(prime, t, dev(a))
        got = host(t)
        assert np.array_equal(got[:3000], a % np.uint64(p)) and not got[3000:].any()


def crt_case(nttm, np_, n, nc, planes, x=None):
    say("crt", np_, n, nc, x is not None)
    c = nttm.crt([dev(q) for q in planes], n, nc, None if x is None else dev(x))
    want, ov = O.crt_carry(planes, [nttm.prime(i) for i in range(np_)], nc, x)
    assert ov == 0
    assert np.array_equal(host(c), want)


@pytest.mark.parametrize("np_", [3, 4])
def test_crt_random_planes(nttm, np_):
    for n in (1, 5, 2047, 2048, 2049, 30000):
        planes = [rand_res(n, nttm.prime(i), 1000 * i + n) for i in range(np_)]
        crt_case(nttm, np_, n, n + np_, planes)
        x, _ = O.gen_limbs(n + np_ - 1, n)
        crt_case(nttm, np_, n, n + np_ + 1, planes, x)


@pytest.mark.parametrize("np_", [3, 4])
def test_crt_adversarial(nttm, np_):
    # every coefficient at the largest CRT value (residues p - 1): maximal digits and carries everywhere
    n = 9000
    planes = [np.full(n, nttm.prime(i) - 1, dtype=np.uint64) for i in range(np_)]
    crt_case(nttm, np_, n, n + np_, planes)
    x = np.full(n + np_ - 1, B - 1, dtype=np.uint64)
    crt_case(nttm, np_, n, n + np_ + 1, planes, x)


def test_crt_long_carry_ripple(nttm):
    # coefficient values B - 1 everywhere and x = 1: one carry runs across every limb and every block
    n, np_ = 100000, 3
    planes = [np.full(n, (B - 1) % nttm.prime(i), dtype=np.uint64) for i in range(np_)]
    x = np.zeros(n, dtype=np.uint64); x[0] = 1
    crt_case(nttm, np_, n, n + 1, planes, x)
    say("crt ripple result")
    c = nttm.crt([dev(q) for q in planes], n, n + 1, dev(x))
    got = host(c)
    assert not got[:n].any() and int(got[n]) == 1


@pytest.mark.parametrize("hole", [False, True])
def test_crt_carry_across_many_blocks(nttm, hole):
    # 2^25 + 12345 coefficients of value B - 1 (more blocks than one pass of the block-carry scan covers) and x = 1;
    # with a hole (value B - 2 at h) the value is B^n - B^h: limbs below h are 0, limbs h .. n-1 are B - 1
    import torch
    n, np_ = (1 << 25) + 12345, 3
    h = n // 2 + 777 if hole else n
    planes = []
    for i in range(np_):
        p = nttm.prime(i)
        t = torch.full((n,), (B - 1) % p, dtype=torch.int64, device="cuda")
        if hole:
            t[h] = (B - 2) % p
        planes.append(t)
    x = torch.zeros(n, dtype=torch.int64, device="cuda"); x[0] = 1
    say("crt many blocks", n, hole)
    got = host(nttm.crt(planes, n, n + 1, x))
    assert not got[:h].any()
    assert (got[h:n] == B - 1).all()
    assert int(got[n]) == (0 if hole else 1)


def test_crt_overflow_is_reported(nttm):
    import ntt_amd
    planes = [np.full(10, nttm.prime(i) - 1, dtype=np.uint64) for i in range(3)]
    say("crt overflow")
    with pytest.raises(ntt_amd.NttmError) as e:
        nttm.crt([dev(q) for q in planes], 10, 10)             # the value needs 12 limbs
    assert e.value.code == -6


def test_pass_split_independence(nttm):
    """the forward transform and the product do not depend on the pass split, tile, points per thread, the fixed-shape
    or streaming kernels, or the length rule (E/tests/t_ntt.c:302-327, section 4)"""
    import os
    import torch
    logn, prime = 24, 1
    p, w = nttm.prime(prime), nttm.root(prime, logn)
    x = rand_res(1 << logn, p, 2024)
    a, s = O.gen_limbs(3 << 19, 5)
    b, _ = O.gen_limbs(5 << 18, s)
    configs = [dict(), dict(tile_log=12), dict(pts=8), dict(tile_log=12, pts=8), dict(radix3=1), dict(radix3=2),
               dict(env="NTTM_GENERIC_PASSES"), dict(env="NTTM_STREAM")]
    ref_X = ref_c = None
    try:
        for cfg in configs:
            cfg = dict(cfg)
            env = cfg.pop("env", None)
            torch.cuda.synchronize(); nttm.shutdown()
            if env:
                os.environ[env] = "1" if env == "NTTM_GENERIC_PASSES" else "0"
            try:
                nttm.init(0, **cfg)
                say("split independence", cfg, env)
                t = dev(x)
                nttm.ntt_fwd(prime, t, logn)
                X = host(t)
                nttm.ntt_inv(prime, t, logn)
                assert np.array_equal(host(t), x), (cfg, env)
                c = host(nttm.mul(dev(a), dev(b)))
            finally:
                if env:
                    del os.environ[env]
                torch.cuda.synchronize(); nttm.shutdown()
            if ref_X is None:
                ref_X, ref_c = X, c
                assert O.tref_res(a, b, c)
            assert np.array_equal(X, ref_X) and np.array_equal(c, ref_c), (cfg, env)
    finally:
        nttm.init(0)
