"""e's digits through the product library (nttm_e_fraction: leaves folded on the device, binary splitting, Newton
reciprocal, corrected quotient, every product on the device), pinned by exact Python integers at small sizes, by the
reference's own 50-digit windows of e (E/verify.c:197-201) at 10^6 and 10^8 digits, and at 10^9 digits by the
reference's sha1 of the whole digit output (TESTING.md:31) and its window at 10^9 (tests/golden/e_windows.json)."""
import hashlib
import json
import math
import os

import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu
B = 10 ** 18
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "e_windows.json")))
WIN = GOLD["windows"]


def e_fraction_python(m):
    """floor((e - 2) B^m) from sum_{k<=N} 1/k! with N! > B^(m+2), exact integers"""
    need = 18 * (m + 2) * math.log(10)
    n = 2
    while math.lgamma(n + 1) < need:
        n += 1
    q = math.factorial(n)
    p = sum(q // math.factorial(k) for k in range(1, n + 1))
    return ((p - q) * B ** m) // q


@pytest.mark.parametrize("m", [1, 7, 40, 300])
def test_e_fraction_exact(nttm, m):
    say("e fraction", m)
    got = nttm.e_fraction(m)
    want = e_fraction_python(m)
    assert sum(int(v) * B ** i for i, v in enumerate(got)) == want


@pytest.mark.parametrize("m", [7, 300, 55600])
def test_e_fraction_full_check(nttm, m):
    """the quotient through the remainder A B^m - Y Q (NTTM_E_FULL_CHECK=1, the path of the rare case the guard limb
    leaves undecided) gives the same limbs as the one-product quotient, and both the exact value at small sizes"""
    import numpy as np
    say("e fraction full check", m)
    fast = nttm.e_fraction(m)
    os.environ["NTTM_E_FULL_CHECK"] = "1"
    try:
        full = nttm.e_fraction(m)
    finally:
        os.environ.pop("NTTM_E_FULL_CHECK", None)
    assert np.array_equal(fast, full)
    if m <= 300:
        assert sum(int(v) * B ** i for i, v in enumerate(full)) == e_fraction_python(m)


def limbs_int(v):
    """the integer of base-10^18 limbs (least significant first), by halves"""
    v = [int(t) for t in v]

    def rec(lo, hi):
        if hi - lo <= 64:
            r = 0
            for t in reversed(v[lo:hi]):
                r = r * B + t
            return r
        mid = (lo + hi) // 2
        return rec(lo, mid) + rec(mid, hi) * B ** (mid - lo)
    return rec(0, len(v)) if v else 0


@pytest.mark.parametrize("L,na,nb", [(16, 16, 16), (48, 48, 17), (1024, 1024, 1024), (1024, 700, 1000),
                                      (3 * 4096, 3 * 4096, 9000), (1 << 16, 1 << 16, 40000), (3 << 14, 20000, 3 << 14)])
def test_mul_cyclic(nttm, orc, L, na, nb):
    """the cyclic product of length L (2^k and 3 2^k, operands up to L limbs, so the convolution wraps): congruent to
    a b modulo B^L - 1, lowest limb c_0 = a_0 b_0 + sum a_i b_(L-i) mod B, at most L + 3 limbs; every coefficient
    against the cyclic convolution in exact integers up to L = 1024"""
    import numpy as np
    from gpu_util import dev, host
    a, s = orc.gen_limbs(na, 5 * L + na)
    b, s = orc.gen_limbs(nb, s)
    say("cyclic product", L, na, nb)
    v = host(nttm.mul_cyclic(dev(a), dev(b), L))
    ai = [int(t) for t in a] + [0] * (L - na)
    bi = [int(t) for t in b] + [0] * (L - nb)
    V, A, Bv = limbs_int(v), limbs_int(ai), limbs_int(bi)
    M = B ** L - 1
    assert V % M == (A * Bv) % M
    c0 = ai[0] * bi[0] + sum(ai[i] * bi[L - i] for i in range(1, L))
    assert int(v[0]) == c0 % B
    assert all(int(t) < B for t in v)
    if L <= 1024:
        c = [sum(ai[i] * bi[(k - i) % L] for i in range(L)) for k in range(L)]
        assert V == sum(ck * B ** k for k, ck in enumerate(c))


def test_e_fraction_wrapped_newton(nttm):
    """the Newton reciprocal with its product formed modulo B^L - 1 from 2^15 limbs up (the default) gives the same
    limbs as with whole products (NTTM_E_NO_WRAP=1); 5.4 million digits, so three levels wrap"""
    import numpy as np
    m = 300000
    say("e wrapped newton", m)
    wrapped = nttm.e_fraction(m)
    os.environ["NTTM_E_NO_WRAP"] = "1"
    try:
        whole = nttm.e_fraction(m)
    finally:
        os.environ.pop("NTTM_E_NO_WRAP", None)
    assert np.array_equal(wrapped, whole)
    d = nttm.limbs_to_ascii(wrapped)
    assert d[49:99].decode() == WIN["50"] and d[999999:1000049].decode() == WIN["1000000"]


@pytest.mark.parametrize("m", [1, 7, 300, 55600, 5555560])
def test_e_leaf_length_bounds(nttm, m):
    """the tree is planned from upper bounds of the leaves' lengths computed from the series (lgamma), not read back:
    NTTM_E_LEAF_CHECK=1 reads the schoolbook levels' true lengths and fails unless every bound lies within one limb
    above them; the digits stay exact"""
    say("e leaf bounds", m)
    os.environ["NTTM_E_LEAF_CHECK"] = "1"
    try:
        got = nttm.e_fraction(m)
    finally:
        os.environ.pop("NTTM_E_LEAF_CHECK", None)
    if m <= 300:
        assert sum(int(v) * B ** i for i, v in enumerate(got)) == e_fraction_python(m)
    else:
        d = nttm.limbs_to_ascii(got)
        assert d[49:99].decode() == WIN["50"] and d[999999:1000049].decode() == WIN["1000000"]


def test_e_million_digits_windows(nttm):
    m = 55600                                                    # 1,000,800 digits
    say("e 1e6 digits")
    d = nttm.e_digits(m)
    assert d.startswith("718281828459045235360287471352662497757")
    for off in ("50", "1000000"):
        o = int(off)
        assert d[o - 1:o + 49] == WIN[off], off


def test_e_1e8_digits_windows(nttm):
    m = 5555560                                                  # 100,000,080 digits: the 10^8 window fits
    say("e 1e8 digits")
    d = nttm.limbs_to_ascii(nttm.e_fraction(m))
    for off in ("50", "1000000", "100000000"):
        o = int(off)
        assert d[o - 1:o + 49].decode() == WIN[off], off


def test_e_1e9_digits_reference_hash(nttm, orc, tmp_path):
    """the reference's production check at 10^9 digits: sha1 of "2." + digits + newline, and the last window; then
    the same digits through the packed file (nttm_e_write_packed) read back by the oracle's reader"""
    m = 55555556                                                 # 1,000,000,008 digits
    say("e 1e9 digits")
    frac = nttm.e_fraction(m)
    d = nttm.limbs_to_ascii(frac)
    say("e 1e9 phases", nttm.e_last_phase_ms())
    assert d[999999950:10 ** 9].decode() == WIN["999999951"]
    h = hashlib.sha1()
    h.update(b"2.")
    h.update(memoryview(d)[:10 ** 9])
    h.update(b"\n")
    assert h.hexdigest()[:16] == GOLD["sha1_16"]["1000000000"]
    del d
    say("e 1e9 packed file")
    path = tmp_path / "e1e9.packed"
    nttm.e_write_packed(str(path), frac, 10 ** 9)
    del frac
    hdr, limbs = orc.read_packed(str(path))
    os.remove(path)
    txt = orc.packed_ascii(hdr, limbs)
    assert len(txt) == 10 ** 9 + 3
    assert hashlib.sha1(txt).hexdigest()[:16] == GOLD["sha1_16"]["1000000000"]
