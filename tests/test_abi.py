"""The C ABI library on the CPU (no GPU calls): it loads, exports every entry point include/nttm.h declares, its host
queries agree with the oracle and with the mathematics, and the pass planner's tiles cover every element of every
batch exactly once inside the buffer (the kernels' own index map, evaluated on the host before any GPU run)."""
import os
import re

import numpy as np
import pytest

import ntt_amd
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "nttm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nttm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ntt_amd.lib()
    syms = declared_symbols()
    assert len(syms) >= 18
    for s in syms:
        assert hasattr(L, s), s


def test_primes():
    L = O.lib()
    ps = [ntt_amd.prime(i) for i in range(4)]
    assert len(set(ps)) == 4
    for p in ps:
        assert L.orc_is_prime(p) and p < 2 ** 50 and (p - 1) % (1 << 32) == 0 and ((p - 1) >> 32) % 3 == 0
    assert ntt_amd.prime(4) == 0 and ntt_amd.prime(-1) == 0


@pytest.mark.parametrize("i", range(4))
def test_roots_exact_order_and_chain(i):
    L = O.lib()
    p = ntt_amd.prime(i)
    prev = None
    for k in range(0, 33):
        w = ntt_amd.root(i, k)
        assert w == L.orc_root2k(p, k)
        assert L.orc_powmod(w, 1 << k, p) == 1
        if k:
            assert L.orc_powmod(w, 1 << (k - 1), p) == p - 1
        if prev is not None:
            assert L.orc_mulmod(w, w, p) == prev
        prev = w
    assert ntt_amd.root(i, 33) == 0


@pytest.mark.parametrize("i", range(4))
def test_roots3_exact_order_and_chain(i):
    p = ntt_amd.prime(i)
    prev = None
    for k in range(0, 33):
        w = ntt_amd.root3(i, k)
        n = 3 << k
        assert w == O.root3k(p, k)
        assert pow(w, n, p) == 1 and pow(w, n // 3, p) != 1 and (k == 0 or pow(w, n // 2, p) == p - 1)
        assert pow(w, 3, p) == ntt_amd.root(i, k)                 # the thirds' transforms use w^3
        if prev is not None:
            assert w * w % p == prev
        prev = w
    assert ntt_amd.root3(i, 33) == 0


def test_product_length_rules():
    # the reference's rule (FIT): 3 2^(L-2) whenever ncoef fits it; AUTO takes it only where it moves fewer bytes
    # (the 3 2^k transform's engine needs one pass fewer than the 2^L transform: L = 13, 14, 22, 23)
    for L in range(6, 31):
        n = 1 << L
        for nc in (n // 2 + 1, 3 * n // 4, 3 * n // 4 + 1, n):
            fits = nc <= 3 * n // 4 and L - 2 >= 4
            assert ntt_amd.debug_product_len(nc, ntt_amd.R3_NEVER) == (L, 0)
            assert ntt_amd.debug_product_len(nc, ntt_amd.R3_FIT) == ((L - 2, 1) if fits else (L, 0))
            auto = ntt_amd.debug_product_len(nc, ntt_amd.R3_AUTO)
            assert auto == ((L - 2, 1) if fits and L in (13, 14, 22, 23) else (L, 0)), (L, nc, auto)
    assert ntt_amd.debug_product_len(1, ntt_amd.R3_FIT) == (4, 0)


def test_max_terms():
    ps = [ntt_amd.prime(i) for i in range(4)]
    for np_ in (3, 4):
        t = ntt_amd.max_terms(np_)
        assert t == O.max_terms(ps[:np_])
        M = 1
        for p in ps[:np_]:
            M *= p
        D = (10 ** 18 - 1) ** 2
        if (M - 1) // D >= 2 ** 64:
            assert t == 2 ** 64 - 1                               # saturated: the bound exceeds a word
        else:
            assert t * D < M <= (t + 1) * D
    assert ntt_amd.max_terms(3) >= 1 << 30                      # a 2^30-term product fits three primes


def check_plan(logn, batch, full, last_log=0, tile_log=13):
    passes = ntt_amd.debug_plan(logn, batch, last_log, tile_log)
    assert sum(m for _, m, _, _ in passes) == logn
    top = logn
    total = batch << logn
    for lo, m, logw, logh in passes:                              # contiguous bit ranges, high to low
        assert lo + m == top and m >= 1 and logw <= lo and m + logw + logh <= (tile_log if lo else (last_log or tile_log))
        if lo == 0 and last_log:
            assert m + logh <= last_log
        if lo > 0 and logn >= 17:
            assert logw >= 4                                      # strided passes read >= 128-byte runs
        top = lo
        T = 1 << (m + logw + logh)
        assert total % T == 0
        tiles = total // T
        if full:
            seen = np.zeros(total, dtype=np.int8)
            for t in range(tiles):
                idx = np.array(ntt_amd.debug_tile_indices(lo, m, logw, logh, t), dtype=np.uint64)
                assert idx.max() < total
                seen[idx] += 1
            assert (seen == 1).all()
        else:
            for t in (0, 1, tiles // 2, tiles - 1):
                idx = ntt_amd.debug_tile_indices(lo, m, logw, logh, t)
                assert max(idx) < total and len(set(idx)) == T
    assert top == 0


@pytest.mark.parametrize("logn", list(range(1, 19)))
@pytest.mark.parametrize("batch", [1, 2, 3, 8])
@pytest.mark.parametrize("tiles", [(13, 0), (13, 12), (12, 12)])
def test_plan_covers_every_element_once(logn, batch, tiles):
    check_plan(logn, batch, full=(batch << logn) <= (1 << 16), tile_log=tiles[0], last_log=tiles[1])


@pytest.mark.parametrize("logn", [20, 21, 24, 27, 28, 30, 32])
@pytest.mark.parametrize("tiles", [(13, 0), (13, 12), (12, 12)])
def test_plan_large_bounds(logn, tiles):
    check_plan(logn, 1, full=False, tile_log=tiles[0], last_log=tiles[1])


def test_product_plan_shape():
    # the 2^28 transform of the 2^27-limb product: two strided 8-bit passes (32-column tiles) and a 12-bit last pass
    assert ntt_amd.debug_plan(28, 1, 12, 13) == [(20, 8, 5, 0), (12, 8, 5, 0), (0, 12, 0, 0)]
    assert ntt_amd.debug_plan(28, 1, 12, 12) == [(20, 8, 4, 0), (12, 8, 4, 0), (0, 12, 0, 0)]


def test_product_plans():
    """the product transforms: 2^26 .. 2^28 points put 8 bits into each strided pass (the 8 x 32 streamed shape) and
    10 to 12 into the last, in 4096-point tiles; 2^29 and 2^30 take 9-bit strided passes (the 9 x 16 streamed shape);
    other lengths keep a 12-bit last pass; every plan covers each element of the batch once (the kernels' own index
    map, sampled tiles at the large sizes)"""
    P = ntt_amd.debug_product_plan
    assert P(26) == [(18, 8, 5, 0), (10, 8, 5, 0), (0, 10, 0, 2)]
    assert P(27) == [(19, 8, 5, 0), (11, 8, 5, 0), (0, 11, 0, 1)]
    assert P(28) == [(20, 8, 5, 0), (12, 8, 5, 0), (0, 12, 0, 0)]
    assert P(29) == [(20, 9, 4, 0), (12, 8, 5, 0), (0, 12, 0, 0)]
    assert P(30) == [(21, 9, 4, 0), (12, 9, 4, 0), (0, 12, 0, 0)]
    assert P(21)[-1] == (0, 12, 0, 0) and P(32)[-1] == (0, 12, 0, 0) and P(17, 64)[-1][:2] == (0, 12)
    for logn in range(20, 33):
        for batch in (1, 3):
            passes = P(logn, batch)
            assert sum(m for _, m, _, _ in passes) == logn
            top, total = logn, batch << logn
            for lo, m, logw, logh in passes:
                assert lo + m == top and m + logw + logh <= 13 and (lo == 0 or logw >= 4)
                if lo == 0:
                    assert m + logh <= 12 and m >= 10 or logn <= 12
                top = lo
                T = 1 << (m + logw + logh)
                assert total % T == 0
                tiles = total // T
                for t in (0, 1, tiles // 3, tiles - 1):
                    idx = ntt_amd.debug_tile_indices(lo, m, logw, logh, t)
                    assert max(idx) < total and len(set(idx)) == T
    full = P(14, 8)                                               # small enough to check every element
    seen = np.zeros(8 << 14, dtype=np.int8)
    for lo, m, logw, logh in full:
        seen[:] = 0
        T = 1 << (m + logw + logh)
        for t in range((8 << 14) // T):
            seen[np.array(ntt_amd.debug_tile_indices(lo, m, logw, logh, t), dtype=np.int64)] += 1
        assert (seen == 1).all()


def test_limbs_to_ascii_formats_18_digits_per_limb():
    """the digit formatter of the wrapper (the text form of e's digits in the tests and the bench): top limb first, 18
    zero-padded digits per limb, against Python's own formatting"""
    rng = np.random.default_rng(4)
    limbs = np.concatenate([np.array([0, 1, 10 ** 17, 10 ** 18 - 1, 999999999, 10 ** 9], dtype=np.uint64),
                            rng.integers(0, 10 ** 18, 1000, dtype=np.uint64)])
    want = "".join("%018d" % int(v) for v in limbs[::-1]).encode()
    assert ntt_amd.limbs_to_ascii(limbs) == want
    assert ntt_amd.limbs_to_ascii(np.zeros(0, dtype=np.uint64)) == b""


REF_ECALC = "/root/reference/ecalc"
SHIM = os.path.join(ROOT, "ntt_amd", "ecalc_shim", "rns_nttm.c")


def _compile_shim(src, out):
    import subprocess
    cmd = ["gcc", "-std=c11", "-Wall", "-Werror", "-c", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + REF_ECALC, "-I" + os.path.join(ROOT, "include"), src, "-o", out]
    return subprocess.run(cmd, capture_output=True, text=True)


@pytest.mark.skipif(not os.path.isdir(REF_ECALC), reason="the reference tree is absent")
def test_ecalc_shim_compiles_against_the_reference_headers(tmp_path):
    """the drop-in shim (rns_init / rns_mul_batch / rns_local_core / the cut products over libnttm) compiles with the
    reference's own rns_mul.h, dbig.h and fatal.h; its static assertions prove rns_prod and nttm_prod have the same
    size and field offsets, so the batch passes the caller's records through unchanged"""
    r = _compile_shim(SHIM, str(tmp_path / "shim.o"))
    assert r.returncode == 0, r.stderr
    # the layout check has teeth: a shim that compares a field against the wrong one does not compile
    bad = tmp_path / "bad.c"
    bad.write_text(open(SHIM).read().replace("offsetof(rns_prod, ncn) == offsetof(nttm_prod, ncn)",
                                             "offsetof(rns_prod, ncn) == offsetof(nttm_prod, nx)"))
    r = _compile_shim(str(bad), str(tmp_path / "bad.o"))
    assert r.returncode != 0 and "field ncn" in r.stderr


def test_wrappers_refuse_bad_tensors():
    """the Python wrappers check residence, dtype, contiguity and output size before any library call"""
    import torch
    a = torch.zeros(8, dtype=torch.int64)
    with pytest.raises(ValueError, match="CUDA"):
        ntt_amd.mul(a, a)
    with pytest.raises(ValueError, match="CUDA"):
        ntt_amd.mul_add(a, a, a)
    with pytest.raises(ValueError, match="CUDA"):
        ntt_amd.mul_cut(a, a, 0, 4)
    with pytest.raises(ValueError, match="torch tensor"):
        ntt_amd.mul(np.zeros(4, dtype=np.uint64), a)
    h = np.zeros(8, dtype=np.uint64)
    with pytest.raises(ValueError, match="at least 17"):
        ntt_amd.make_prods([(h, h, h[:3], np.zeros(16, dtype=np.uint64))])
    with pytest.raises(ValueError, match="int64"):
        ntt_amd.make_prods([(a.to(torch.int32), a, None, np.zeros(16, dtype=np.uint64))])
    with pytest.raises(ValueError, match="contiguous"):
        ntt_amd.make_prods([(torch.zeros(16, dtype=torch.int64)[::2], a, None, np.zeros(16, dtype=np.uint64))])
    with pytest.raises(ValueError, match="at most na \\+ nb"):
        ntt_amd.make_prods([(h[:2], h[:2], h, np.zeros(16, dtype=np.uint64))])
    arr = ntt_amd.make_prods([(h, h, None, np.zeros(16, dtype=np.uint64))])
    assert arr._keep is not None and arr[0].na == 8 and arr[0].nb == 8
