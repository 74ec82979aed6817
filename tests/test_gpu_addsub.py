"""The e pipeline's device addition and subtraction (k_addsub with the CRT's block carry maps, scan and fix) against
exact integers: carries and borrows rippling through whole 2048-limb blocks, shifted operands, the sign flag of a
subtraction, and sizes from one limb to about 150 blocks."""
import numpy as np
import pytest

from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu
B = 10 ** 18


def to_int(limbs):
    v = 0
    for x in reversed([int(t) for t in limbs]):
        v = v * B + x
    return v


def ref(a, b, nout, sa, sb, sub):
    """limb by limb with a sequential carry (linear time, unlike converting 10^5-limb numbers to integers): the limbs
    of (a B^sa +- b B^sb) mod B^nout and the carry out of the top limb"""
    A = np.zeros(nout, dtype=object)
    Bv = np.zeros(nout, dtype=object)
    A[sa:sa + len(a)] = [int(v) for v in a[:max(0, nout - sa)]]
    Bv[sb:sb + len(b)] = [int(v) for v in b[:max(0, nout - sb)]]
    out, c = [], 0
    for i in range(nout):
        v = int(A[i]) + (-int(Bv[i]) if sub else int(Bv[i])) + c
        c = v // B                                                     # floor: -1 borrows, +1 carries
        out.append(v - c * B)
    return out, (c >= 0) if sub else (c > 0)


CASES = [
    # (na, nb, sa, sb, kind)
    (1, 1, 0, 0, "rand"), (5, 3, 0, 2, "rand"), (2048, 2048, 0, 0, "ones"), (4096, 1, 0, 0, "ones"),
    (6000, 6000, 0, 0, "rand"), (5000, 3, 0, 4999, "rand"), (20000, 1, 0, 0, "zeros_top"), (9000, 8999, 3, 4, "rand"),
    (300000, 300000, 0, 0, "rand"), (300000, 1, 0, 0, "ones"),
]


def operands(na, nb, kind, rng):
    if kind == "ones":                           # all B - 1: a +1 carries through every limb
        a = np.full(na, B - 1, dtype=np.uint64)
        b = np.ones(nb, dtype=np.uint64)
    elif kind == "zeros_top":                    # 1 followed by zeros: -1 borrows through every limb
        a = np.zeros(na, dtype=np.uint64); a[-1] = 1
        b = np.ones(nb, dtype=np.uint64)
    else:
        a = rng.integers(0, B, na, dtype=np.uint64)
        b = rng.integers(0, B, nb, dtype=np.uint64)
    return a, b


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("sub", [False, True])
def test_addsub_vs_integers(nttm, case, sub):
    na, nb, sa, sb, kind = CASES[case]
    rng = np.random.default_rng(case * 2 + sub)
    a, b = operands(na, nb, kind, rng)
    nout = max(na + sa, nb + sb) + (0 if sub else 1)
    say("addsub", CASES[case], sub)
    r, flag = nttm.debug_addsub(dev(a), dev(b), nout, sa, sb, sub)
    want, wflag = ref(a, b, nout, sa, sb, sub)
    got = host(r)
    assert (got < B).all()
    assert [int(v) for v in got] == want and flag == wflag
    if nout <= 8000:                                                   # small enough for exact integers
        va, vb = to_int(a) * B ** sa, to_int(b) * B ** sb
        v = va - vb if sub else va + vb
        assert to_int(got) == v % B ** nout


def test_subtract_both_signs(nttm):
    """a - b and b - a for close numbers (equal top limbs): one flag set, the other clear, the true difference kept"""
    rng = np.random.default_rng(9)
    a = rng.integers(0, B, 7000, dtype=np.uint64)
    b = a.copy()
    b[17] = (int(b[17]) + 1) % B if b[17] != B - 1 else 0
    for x, y in ((a, b), (b, a)):
        say("addsub signs")
        r, ge = nttm.debug_addsub(dev(x), dev(y), 7000, sub=True)
        assert ge == (to_int(x) >= to_int(y))
        if ge:
            assert to_int(host(r)) == to_int(x) - to_int(y)
