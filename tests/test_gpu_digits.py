"""The packed digit output on the device (nttm_pack60, nttm_limbs_mod, nttm_e_write_packed) against the oracle's
restatement of the reference's 60-bit stream (E/packed_fmt.h:50-88, pinned by the reference's own bytes in
tests/golden) and against exact integers; the e file read back by the oracle's reader gives the reference's digits."""
import json
import os

import numpy as np
import pytest

from gpu_util import dev, say

pytestmark = pytest.mark.gpu
B = 10 ** 18
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "e_windows.json")))


def to_int(limbs):
    v = 0
    for x in reversed([int(t) for t in limbs]):
        v = v * B + x
    return v


@pytest.mark.parametrize("n", [1, 2, 3, 7, 8, 9, 15, 16, 17, 64, 1001, 100003])
def test_pack60_matches_oracle(nttm, orc, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, B, n, dtype=np.uint64)
    x[0] = B - 1
    x[-1] = 0 if n > 2 else x[-1]
    say("pack60", n)
    got = nttm.pack60(dev(x)).cpu().numpy().tobytes()
    assert got == orc.pack60(x[::-1])


@pytest.mark.parametrize("lo,hi", [(0, 1), (3, 4), (5, 13), (1, 1000), (999, 1000), (17, 90017)])
def test_pack60_ranges(nttm, orc, lo, hi):
    rng = np.random.default_rng(lo * 7 + hi)
    x = rng.integers(0, B, 90017 if hi > 1000 else 1000, dtype=np.uint64)
    say("pack60 range", lo, hi)
    got = nttm.pack60(dev(x), lo, hi).cpu().numpy().tobytes()
    assert got == orc.pack60(x[lo:hi][::-1])


@pytest.mark.parametrize("n", [1, 5, 300, 70001, 2000003])
def test_limbs_mod_t1(nttm, orc, n):
    rng = np.random.default_rng(n + 1)
    x = rng.integers(0, B, n, dtype=np.uint64)
    x[-1] = B - 1
    say("limbs_mod", n)
    got = nttm.limbs_mod(dev(x))
    if n <= 70001:
        v = to_int(x)
        assert got == [v % q for q in orc.T1_PRIMES]
    else:
        assert got == [orc.limbs_mod(x, q) for q in orc.T1_PRIMES]


def test_limbs_mod_other_moduli(nttm):
    x = np.array([B - 1] * 40 + [12345], dtype=np.uint64)
    qs = [10 ** 18 + 9, 2 ** 63 - 25, 4611686018427388039]        # odd moduli in (10^18, 2^63)
    say("limbs_mod other moduli")
    assert nttm.limbs_mod(dev(x), qs) == [to_int(x) % q for q in qs]


def test_e_packed_file_million_digits(nttm, orc, tmp_path):
    m, d_out = 55600, 1000790                                     # d = 1,000,800 computed; the text stops 10 short
    say("e packed 1e6")
    frac = nttm.e_fraction(m)
    path = tmp_path / "e.packed"
    nttm.e_write_packed(str(path), frac, d_out)
    h, limbs = orc.read_packed(str(path))
    assert h["magic"] == b"ECPACK18" and h["version"] == 2 and h["hdr_bytes"] == 4096 and h["reserved0"] == 60
    assert (h["limb_bytes"], h["limb_digits"], h["order"], h["part"], h["nparts"]) == (8, 18, 1, 0, 1)
    assert (h["d"], h["d_out"], h["nl"], h["pad"], h["lo"], h["hi"]) == (18 * m, d_out, m + 1, 17, 0, m + 1)
    assert (h["k0"], h["k1"], h["complete"], h["nq"]) == (0, 18 * m + 1, 1, 8)
    assert h["q"] == list(orc.T1_PRIMES)
    assert os.path.getsize(path) == 4096 + 15 * ((m + 1) // 2) + 8 * ((m + 1) & 1)
    x = np.append(frac, np.uint64(2))                             # X = 2 B^m + fraction, least significant first
    assert np.array_equal(limbs, x[::-1])
    assert h["dres"] == [orc.limbs_mod(x, q) for q in orc.T1_PRIMES]
    txt = orc.packed_ascii(h, limbs)
    assert len(txt) == d_out + 3 and txt.startswith(b"2.71828182845904523536") and txt.endswith(b"\n")
    w = GOLD["windows"]
    for off in ("50", "1000000"):
        o = int(off)
        assert txt[o + 1:o + 51].decode() == w[off], off
