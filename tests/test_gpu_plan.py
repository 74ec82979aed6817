"""Batch plans (nttm_batch_plan_*): one hipGraph per batch with an autotuned schedule.  Every replay is compared with
the eager batch (nttm_mul_batch, itself checked against the oracle and GMP in test_gpu_mul.py) on the same operands,
with eager work and synchronizations between the replays; operands rewritten between runs are read anew; products
with an empty operand, a fused + x and a shared B operand are included; the schedule is reported per group."""
import os
import re

import numpy as np
import pytest

import oracle as O
from gpu_util import dev, host, say

pytestmark = pytest.mark.gpu
B = O.B


def mixed_ops(seed, nprod=48, max_len=20000):
    """(a, b, x, share_prev_b): log-uniform lengths (several length classes, some fitting 3 2^k points), every third
    with x, every second sharing the previous product's B device buffer, two with an empty operand"""
    rng = np.random.default_rng(seed)
    ops, s = [], seed
    for i in range(nprod):
        na = int(np.exp(rng.uniform(0, np.log(max_len))))
        nb = int(np.exp(rng.uniform(0, np.log(max_len))))
        a, s = O.gen_limbs(na, s)
        b, s = O.gen_limbs(nb, s)
        share = i % 2 == 1
        if share:
            b = ops[-1][1]
        x = None
        if i % 3 == 0:
            x, s = O.gen_limbs(int(rng.integers(1, len(a) + len(b) + 1)), s)
        ops.append((a, b, x, share))
    x0, s = O.gen_limbs(min(7, len(ops[0][1])), s)                    # nx <= na + nb
    ops.append((np.zeros(0, dtype=np.uint64), ops[0][1], x0, False))   # a empty, c = x
    ops.append((ops[1][0], np.zeros(0, dtype=np.uint64), None, False))  # b empty, c = 0
    return ops


def device_items(ops):
    import torch
    items = []
    for a, b, x, share in ops:
        nc = len(a) + len(b) + (1 if x is not None else 0)
        db = items[-1][1] if share else dev(b)
        items.append((dev(a), db, None if x is None else dev(x), torch.full((max(nc, 1),), -1, dtype=torch.int64, device="cuda")))
    return items


def outputs(ops, items):
    return [host(it[3])[:len(a) + len(b) + (1 if x is not None else 0)] for (a, b, x, _), it in zip(ops, items)]


def eager(nttm, ops, items):
    """the same products through nttm_mul_batch into fresh outputs"""
    import torch
    twin = [(a, b, x, torch.full_like(c, -1)) for a, b, x, c in items]
    prods = nttm.make_prods(twin)
    nttm.mul_batch(prods)
    return outputs(ops, twin), [p.ncn for p in prods]


@pytest.mark.parametrize("autotune", [True, False])
def test_plan_replays_equal_eager(nttm, autotune):
    import torch
    ops = mixed_ops(31 + autotune)
    items = device_items(ops)
    say("plan create", autotune)
    plan = nttm.BatchPlan(nttm.make_prods(items), autotune=autotune)
    try:
        want, want_ncn = eager(nttm, ops, items)
        for rep in range(3):
            say("plan replay", rep)
            for it in items:
                it[3].fill_(-1)
            torch.cuda.synchronize()
            plan.run()
            got_ncn = plan.ncn()
            got = outputs(ops, items)
            assert all(np.array_equal(g, w) for g, w in zip(got, want)), rep
            assert got_ncn == want_ncn, rep
            eager(nttm, ops, items)                               # eager work between the replays
            torch.cuda.synchronize()
        # operands rewritten in place: the next run reads them anew
        rng = np.random.default_rng(5)
        for (a, b, x, share), it in zip(ops, items):
            if len(a):
                it[0].copy_(dev(rng.integers(0, B, len(a), dtype=np.uint64)))
        want2, ncn2 = eager(nttm, ops, items)
        say("plan run after rewrite")
        plan.run()
        assert plan.ncn() == ncn2
        assert all(np.array_equal(g, w) for g, w in zip(outputs(ops, items), want2))
        info = plan.info()
        groups = [l for l in info.splitlines() if l.startswith("group")]
        assert groups and info.splitlines()[-1].startswith("plan products=%d empty=2" % len(ops))
        for l in groups:
            tile, pts = int(re.search(r"tile=(\d+)", l).group(1)), int(re.search(r"pts=(\d+)", l).group(1))
            assert tile in (12, 13) and pts in (8, 16)
        assert sum(int(re.search(r"products=(\d+)", l).group(1)) for l in groups) == len(ops) - 2
        if autotune:
            assert int(re.search(r"candidates=(\d+)", info).group(1)) >= 4 * len(groups)
    finally:
        plan.close()


def test_plan_reuses_remembered_choices(nttm):
    """autotune="reuse": a plan over products of the same classes as an autotuned one times nothing and takes the same
    schedule, and its replays equal the eager batch; after nttm_shutdown + nttm_init nothing is remembered"""
    import torch
    ops = mixed_ops(35)
    items = device_items(ops)
    sched = lambda info: sorted(re.sub(r" ms=\S+", "", l) for l in info.splitlines() if l.startswith("group"))
    count = lambda info, key: int(re.search(key + r"=(\d+)", info).group(1))
    say("plan tuned")
    p1 = nttm.BatchPlan(nttm.make_prods(items), autotune=True)
    items2 = device_items(ops)                                 # other buffers, the same shapes
    say("plan reuse")
    p2 = nttm.BatchPlan(nttm.make_prods(items2), autotune="reuse")
    try:
        i1, i2 = p1.info(), p2.info()
        assert count(i1, "candidates") > 0 and count(i1, "reused") == 0
        assert count(i2, "candidates") == 0 and count(i2, "reused") >= len(sched(i2)), i2
        assert sched(i2) == sched(i1)
        want, want_ncn = eager(nttm, ops, items2)
        for rep in range(2):
            for it in items2:
                it[3].fill_(-1)
            torch.cuda.synchronize()
            say("reuse plan replay", rep)
            p2.run()
            assert p2.ncn() == want_ncn
            assert all(np.array_equal(g, w) for g, w in zip(outputs(ops, items2), want)), rep
    finally:
        p1.close()
        p2.close()
    nttm.shutdown()
    nttm.init(0, tile_log=int(os.environ.get("NTTM_TEST_TILE_LOG", "0")), pts=int(os.environ.get("NTTM_TEST_PTS", "0")))
    say("plan reuse after re-init")
    p3 = nttm.BatchPlan(nttm.make_prods(items2), autotune="reuse")
    try:
        i3 = p3.info()
        assert count(i3, "reused") == 0 and count(i3, "candidates") > 0
    finally:
        p3.close()


def test_plan_leaf_workload_equals_eager(nttm):
    """the 4096-leaf workload (log-uniform lengths in [64, 65536], seeds 0x5EED0005/6) through an autotuned plan"""
    import torch
    rng = np.random.default_rng(0x5EED0005)
    ops, s = [], 0x5EED0006
    for _ in range(4096):
        na = int(np.exp(rng.uniform(np.log(64), np.log(65536))))
        nb = int(np.exp(rng.uniform(np.log(64), np.log(65536))))
        a, s = O.gen_limbs(na, s)
        b, s = O.gen_limbs(nb, s)
        ops.append((a, b, None, False))
    items = device_items(ops)
    say("leaf plan create")
    plan = nttm.BatchPlan(nttm.make_prods(items))
    try:
        want, want_ncn = eager(nttm, ops, items)
        for rep in range(2):
            for it in items:
                it[3].fill_(-1)
            torch.cuda.synchronize()
            say("leaf plan replay", rep)
            plan.run()
            assert plan.ncn() == want_ncn
            assert all(np.array_equal(g, w) for g, w in zip(outputs(ops, items), want))
    finally:
        plan.close()
