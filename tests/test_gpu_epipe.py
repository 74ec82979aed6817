"""e's digits through the product library (nttm_e_fraction: binary splitting, Newton reciprocal, corrected quotient,
every product on the device), pinned by exact Python integers at small sizes and by the reference's own 50-digit
windows of e (E/verify.c:197-201, tests/golden/e_windows.json) at a million digits."""
import json
import math
import os

import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu
B = 10 ** 18
HERE = os.path.dirname(os.path.abspath(__file__))
WIN = json.load(open(os.path.join(HERE, "golden", "e_windows.json")))["windows"]


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


def test_e_million_digits_windows(nttm):
    m = 55600                                                    # 1,000,800 digits
    say("e 1e6 digits")
    d = nttm.e_digits(m)
    assert d.startswith("718281828459045235360287471352662497757")
    for off in ("50", "1000000"):
        o = int(off)
        assert d[o - 1:o + 49] == WIN[off], off
