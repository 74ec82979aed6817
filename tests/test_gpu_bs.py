"""GPU parity of the binary-splitting product tree (nttm_bs_tree, the caller of the batched product in the reference,
E/binsplit.c:1583-1700) against exact Python integers.  The node combination (P_L Q_R + P_R, Q_L Q_R) is associative,
so any tree shape gives the same root; the check folds the leaves in a balanced recursion.  For e's series the root
has closed forms: Q = N!, P = sum_{k=1..N} N!/k!."""
import math

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


def from_int(v):
    out = []
    while v:
        v, r = divmod(v, B)
        out.append(r)
    return np.array(out, dtype=np.uint64)


def combine(nodes):
    if len(nodes) == 1:
        return nodes[0]
    m = len(nodes) // 2
    (pl, ql), (pr, qr) = combine(nodes[:m]), combine(nodes[m:])
    return pl * qr + pr, ql * qr


@pytest.mark.parametrize("n", [1, 2, 5000, (1 << 15) + 7])
def test_bs_tree_e_series(nttm, n):
    """leaves P_k = 1, Q_k = k + 1: the root is (sum_{k=1..n} n!/k!, n!)"""
    one = dev(np.array([1], dtype=np.uint64))
    leaves = [(one, dev(np.array([k + 1], dtype=np.uint64))) for k in range(n)]
    say("bs tree e", n)
    P, Q = nttm.bs_tree(leaves)
    q = to_int(host(Q))
    assert q == math.factorial(n)
    p_want, q_want = combine([(1, k + 1) for k in range(n)])
    assert q_want == q and to_int(host(P)) == p_want


def test_bs_tree_random_nodes(nttm):
    """multi-limb leaves of random lengths with P < Q (as in a convergent series; some P = 0), an odd count at several
    levels"""
    rng = np.random.default_rng(7)
    leaves, ints, s = [], [], 11
    for i in range(777):
        lq = int(rng.integers(1, 40))
        lp = int(rng.integers(0, lq))                            # fewer limbs than Q: P < Q
        p, s = O.gen_limbs(lp, s) if lp else (np.zeros(0, dtype=np.uint64), s)
        q, s = O.gen_limbs(lq, s)
        leaves.append((dev(p), dev(q)))
        ints.append((to_int(p), to_int(q)))
    say("bs tree random 777")
    P, Q = nttm.bs_tree(leaves)
    p_want, q_want = combine(ints)
    assert np.array_equal(host(Q), from_int(q_want)) and np.array_equal(host(P), from_int(p_want))


def test_bs_tree_planned_equals_levelwise(nttm):
    """the tree planned on the host from unnormalized lengths (no synchronization between levels) gives the same root,
    limb for limb, as the level-by-level tree that normalizes every level's lengths (NTTM_TREE_LEVELWISE=1); leaves of
    up to 600 limbs so that the upper levels run two-pass transforms and shared-B groups"""
    import os
    import torch
    rng = np.random.default_rng(23)
    leaves, s = [], 5
    for i in range(301):
        lq = int(rng.integers(1, 600))
        lp = int(rng.integers(0, lq))
        p, s = O.gen_limbs(lp, s) if lp else (np.zeros(0, dtype=np.uint64), s)
        q, s = O.gen_limbs(lq, s)
        leaves.append((p, q))
    say("bs tree planned")
    P1, Q1 = nttm.bs_tree([(dev(p), dev(q)) for p, q in leaves])
    p1, q1 = host(P1), host(Q1)
    try:
        torch.cuda.synchronize(); nttm.shutdown()
        os.environ["NTTM_TREE_LEVELWISE"] = "1"
        nttm.init(0)
        say("bs tree levelwise")
        P2, Q2 = nttm.bs_tree([(dev(p), dev(q)) for p, q in leaves])
        p2, q2 = host(P2), host(Q2)
    finally:
        os.environ.pop("NTTM_TREE_LEVELWISE", None)
        torch.cuda.synchronize(); nttm.shutdown(); nttm.init(0)
    assert np.array_equal(p1, p2) and np.array_equal(q1, q2)
    assert p1[-1] != 0 and q1[-1] != 0
    if O.have_gmp() and len(q1) < 200000:
        pw, qw = combine([(to_int(p), to_int(q)) for p, q in leaves])
        assert np.array_equal(q1, from_int(qw)) and np.array_equal(p1, from_int(pw))


def test_bs_tree_root_caps(nttm):
    """the root written straight into the caller's buffers (room for its unnormalized lengths) and through the pool
    (room for the normalized lengths only): the same limbs; two leaves (one level, operands in the caller's leaves)
    and 77 leaves"""
    rng = np.random.default_rng(31)
    for nleaf in (2, 77):
        leaves, s = [], 3
        for i in range(nleaf):
            lq = int(rng.integers(1, 300))
            lp = int(rng.integers(0, lq))
            p, s = O.gen_limbs(lp, s) if lp else (np.zeros(0, dtype=np.uint64), s)
            q, s = O.gen_limbs(lq, s)
            leaves.append((dev(p), dev(q)))
        say("bs tree caps", nleaf)
        P1, Q1 = nttm.bs_tree(leaves)
        p1, q1 = host(P1), host(Q1)
        P2, Q2 = nttm.bs_tree(leaves, caps=(len(p1), len(q1)))
        assert np.array_equal(p1, host(P2)) and np.array_equal(q1, host(Q2))
        pw, qw = combine([(to_int(host(p)), to_int(host(q))) for p, q in leaves])
        assert np.array_equal(q1, from_int(qw)) and np.array_equal(p1, from_int(pw))


@pytest.mark.parametrize("n", [1, 5, 4095, 4096, 4097, 3 * 4096 + 17, 100003])
def test_top_nonzero(nttm, n):
    """the top-down scan with early exit: all zero, the top limb set, the top non-zero limb several chunks (4096 limbs)
    below the top, only limb 0 set, non-zero limbs scattered below the top one"""
    import torch
    rng = np.random.default_rng(n)
    cases = [np.zeros(n, dtype=np.uint64)]
    x = np.zeros(n, dtype=np.uint64); x[n - 1] = 5; cases.append(x)
    x = np.zeros(n, dtype=np.uint64); x[0] = 1; cases.append(x)
    for top in {n // 2, max(0, n - 3 * 4096 - 5), max(0, n - 4097), max(0, n - 4096)}:
        x = np.zeros(n, dtype=np.uint64)
        x[: top + 1] = rng.integers(0, 3, size=top + 1).astype(np.uint64)
        x[top] = 7
        cases.append(x)
    for x in cases:
        want = int(np.flatnonzero(x)[-1]) + 1 if x.any() else 0
        say("top nonzero", n, want)
        assert nttm.top_nonzero(dev(x)) == want
