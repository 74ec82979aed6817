#!/usr/bin/env python3
"""bench.py - the product hot path on one MI355X.

Default workload: one 2^27 x 2^27-limb product (base-10^18 uint64 limbs drawn by splitmix64 from seeds 0x5EED0003 /
0x5EED0004, top limb nonzero), transform length 2^28, three primes, operands resident in HBM.  A step is one full
product c = a b through the C ABI (load, 9 transforms with the fused pointwise product, CRT with the device carry scan).
Also measured in the same run: the plain 2^20 x 2^20 product (median of 12 calls and of 12 replays of it as a
one-product plan, exact against GMP, GMP timed on the same operands); the batched binary-splitting leaves (4096 products, log-uniform lengths in [64, 65536] limbs) as one
graph replay, one batched call, one call per product (synchronizing, and launched back to back), and one batched call
on host-resident operands (PCIe included); one binary-splitting level (256 nodes, each P = P1 Q2 + P2 and
Q = Q1 Q2 sharing Q2) batched with Q2 shared and unshared; e to 10^9 digits through the whole pipeline (checked
against the reference's sha1 of its digit output); and, on the host, GMP mpz_mul on a bounded sample of the same
operands (2^22..2^24 limbs) as the CPU baseline, the full size modeled from it.  --cpu-full runs the one-off
full-size GMP measurement and exact check.

Prints one JSON line (rank 0).  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

B = 10 ** 18
METRIC = "2^27-limb product secs + % HBM roofline; batched vs per-launch leaf mults/s"
HBM_PEAK_GBS = 8000.0
MI300A_REF_S = 0.1263          # the reference's single-device 2^27 x 2^27 product (results/D13b_strategy_e0.txt:26)
G1, G2, G3 = np.uint64(0x9E3779B97F4A7C15), np.uint64(0xBF58476D1CE4E5B9), np.uint64(0x94D049BB133111EB)


def splitmix(seed, count, offset=0):
    """values offset+1 .. offset+count of the splitmix64 stream seeded with `seed` (vectorized)."""
    i = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * G1
        z = (z ^ (z >> np.uint64(30))) * G2
        z = (z ^ (z >> np.uint64(27))) * G3
    return z ^ (z >> np.uint64(31))


def gen_limbs(n, seed):
    """uniform base-10^18 limbs as the reference's operand generator makes them (E/tests/harness.h:55-81)."""
    out = np.empty(n, dtype=np.uint64)
    step = 1 << 24
    for o in range(0, n, step):
        k = min(step, n - o)
        out[o:o + k] = splitmix(seed, k, o) % np.uint64(B)
    if out[-1] == 0:
        out[-1] = 1
    return out


def leaf_workload(count=4096, lo=64, hi=65536):
    """lengths log-uniform in [lo, hi] from a stream seeded 0x5EED0005 (na, nb drawn independently), limbs from a stream
    seeded 0x5EED0006 (one stream cut in order: a_0, b_0, a_1, ...)"""
    u = (splitmix(0x5EED0005, 2 * count) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    lens = np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)
    lens = np.clip(lens, lo, hi)
    total = int(lens.sum())
    limbs = splitmix(0x5EED0006, total) % np.uint64(B)
    ops, off = [], 0
    for j in range(count):
        na, nb = int(lens[2 * j]), int(lens[2 * j + 1])
        a = limbs[off:off + na].copy(); off += na
        b = limbs[off:off + nb].copy(); off += nb
        if a[-1] == 0:
            a[-1] = 1
        if b[-1] == 0:
            b[-1] = 1
        ops.append((a, b))
    return ops


def to_dev(a):
    import torch
    return torch.from_numpy(a.view(np.int64)).cuda()


def host_info():
    model = "?"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(a, b, logs=(22, 23, 24), reps=(3, 3, 1)):
    """GMP mpz_mul (1 thread) on the first 2^k limbs of the same operands for each k of `logs` (median of reps[i]),
    conversions off the clock: about 15 s of CPU work on the GPU host.  The full operand length is modeled by a
    least-squares line through log2(seconds) against log2(limbs) over the measured points, and the one-off full-size
    measurement (bench.py --cpu-full, kept in profiles/cpu_gmp_2e27_full.json) is quoted beside it."""
    import oracle
    if not oracle.have_gmp():
        return None
    n_full = len(a)
    t = {}
    for lg, r in zip(logs, reps):
        k = 1 << lg
        if k > n_full:
            continue
        t[lg] = oracle.gmp_time_mul(a[:k], b[:k], r)
    xs = np.array(sorted(t), dtype=float)
    ys = np.log2([t[int(x)] for x in xs])
    slope, icpt = np.polyfit(xs, ys, 1) if len(xs) > 1 else (1.0, ys[0] - xs[0])
    lg_full = np.log2(n_full)
    est = float(2 ** (icpt + slope * lg_full))
    model, ncpu = host_info()
    out = {"value": round(est, 4),
           "unit": "s per 2^%d x 2^%d-limb product" % (int(lg_full), int(lg_full)),
           "cores": 1, "kind": "port",
           "sample": ("GMP mpz_mul (oracle GMP bridge), 1 thread, on the first 2^k limbs of the same seeded operands "
                      "for k in %s (medians of %s runs), conversions off the clock; value modeled at 2^%d limbs by a "
                      "least-squares fit of log2 seconds on log2 limbs (growth x%.3f per doubling); host %s, %s cpus"
                      % (list(map(int, xs)), list(reps), int(lg_full), 2 ** slope, model, ncpu)),
           "sample_s": {"2^%d" % k: round(v, 4) for k, v in sorted(t.items())},
           "value_is": "modeled"}
    try:
        full = json.load(open(os.path.join(ROOT, "profiles", "cpu_gmp_2e27_full.json")))
        if full.get("limbs") == n_full:
            out["full_size_measured_once"] = {k: full[k] for k in ("mpz_mul_s", "convert_s", "exact_equal", "host",
                                                                    "command") if k in full}
    except (OSError, ValueError):
        pass
    return out


def cpu_full(nttm, a, b):
    """one-off: the 2^27 x 2^27 product on the GPU, then on the host a and b converted to mpz, one mpz_mul timed on
    them (1 thread) and the GPU's product compared with GMP's exactly (minutes of CPU work)"""
    import oracle
    import torch
    da, db = to_dev(a), to_dev(b)
    c = nttm.mul(da, db)
    torch.cuda.synchronize()
    ch = c.cpu().numpy().view(np.uint64).copy()
    del c, da, db
    print("[bench] GPU product done; GMP on the host", file=sys.stderr, flush=True)
    eq, t_conv, t_mul, t_conv_c = oracle.gmp_full_check(a, b, ch)
    model, ncpu = host_info()
    return {"limbs": len(a), "mpz_mul_s": round(t_mul, 3), "convert_s": round(t_conv, 3),
            "convert_c_s": round(t_conv_c, 3), "exact_equal": eq, "threads_mul": 1,
            "host": "%s, %s cpus" % (model, ncpu),
            "command": "python bench.py --cpu-full (seeds 0x5EED0003 / 0x5EED0004)"}


def bench_config1(nttm, reps=12):
    """the plain product: two 2^20-limb operands (seeds 0x5EED0001 / 0x5EED0002), transform 2^21, three primes; the
    median of `reps` timed calls (each returns complete), its fraction of the HBM roofline on the algorithmic bytes
    352 2^21, the median of `reps` replays of the same product as a one-product batch plan, the product checked exactly
    against GMP, and GMP's mpz_mul on the same operands (1 thread)"""
    import oracle
    import torch
    n = 1 << 20
    a, b = gen_limbs(n, 0x5EED0001), gen_limbs(n, 0x5EED0002)
    da, db = to_dev(a), to_dev(b)
    dc = torch.empty(2 * n, dtype=torch.int64, device="cuda")
    for _ in range(3):
        nttm.mul(da, db, out=dc)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        nttm.mul(da, db, out=dc)
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    b_alg = 8 * 2 * n + 112 * 3 * (2 * n) + 8 * 2 * n
    out = {"workload": "2^20 x 2^20-limb product, transform 2^21, 3 primes", "runs": reps,
           "median_ms": round(1e3 * med, 4), "min_ms": round(1e3 * min(ts), 4), "max_ms": round(1e3 * max(ts), 4),
           "b_alg_bytes": b_alg, "hbm_GBs": round(b_alg / med / 1e9, 1),
           "hbm_frac": round(b_alg / med / 1e9 / HBM_PEAK_GBS, 4)}
    # the same product planned once as a one-product batch plan: each call is one graph launch and the read of its
    # length (it returns complete), so the host's per-kernel launch costs are paid at planning time
    dp = torch.empty(2 * n, dtype=torch.int64, device="cuda")
    plan = nttm.BatchPlan(nttm.make_prods([(da, db, None, dp)]), autotune=True)
    try:
        for _ in range(3):
            plan.run()
            plan.ncn()
        tp = []
        for _ in range(reps):
            t0 = time.perf_counter()
            plan.run()
            plan.ncn()
            tp.append(time.perf_counter() - t0)
        pmed = float(np.median(tp))
        out.update({"plan_median_ms": round(1e3 * pmed, 4), "plan_min_ms": round(1e3 * min(tp), 4),
                    "plan_hbm_frac": round(b_alg / pmed / 1e9 / HBM_PEAK_GBS, 4),
                    "plan_equals_call": bool(torch.equal(dp, dc))})
    finally:
        plan.close()
    if oracle.have_gmp():
        ch = dc.cpu().numpy().view(np.uint64)
        out["exact_vs_gmp"] = bool(oracle.gmp_check(a, b, ch))
        tg = oracle.gmp_time_mul(a, b, 5)
        out["cpu_gmp_1_thread_s"] = round(tg, 4)
        out["gpu_over_gmp"] = round(tg / med, 1)
    return out


def cpu_leaves_baseline(ops, gpu_products_per_s):
    """GMP mpz_mul over the whole leaf workload on 1 thread and on every host thread (one product per thread at a
    time, dynamic schedule), conversions off the clock"""
    import oracle
    if not oracle.have_gmp():
        return None
    model, ncpu = host_info()
    threads = min(ncpu or 1, 16)
    t1 = oracle.gmp_time_batch(ops, 1, 1)
    tn = oracle.gmp_time_batch(ops, threads, 1)
    return {"kind": "port", "products": len(ops), "gmp_1_thread_s": round(t1, 4),
            "gmp_1_thread_products_per_s": round(len(ops) / t1, 1), "threads": threads,
            "gmp_all_threads_s": round(tn, 4), "gmp_all_threads_products_per_s": round(len(ops) / tn, 1),
            "gpu_graph_over_gmp_all_threads": round(gpu_products_per_s * tn / len(ops), 1),
            "sample": "the full 4096-product leaf workload, mpz_mul only, host %s" % model}


def dominant(ks):
    name = max(ks, key=lambda k: ks[k][1])
    launches, ms, byts = ks[name]
    return name, launches, ms, byts


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        rec = json.load(open(path)).get(name)
        return rec.get("bytes") if rec else None
    except (OSError, ValueError, AttributeError):
        return None


def bench_leaves(nttm, reps, each=True, cpu=False):
    import torch
    ops = leaf_workload()
    items = []
    for a, b in ops:
        c = torch.empty(len(a) + len(b), dtype=torch.int64, device="cuda")
        items.append((to_dev(a), to_dev(b), None, c))
    prods = nttm.make_prods(items)
    nttm.mul_batch(prods)                                    # warm (workspace, code)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        nttm.mul_batch(prods)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    ref = [it[3].clone() for it in items]
    # the same batch as an autotuned plan: one graph launch per batch
    t0 = time.perf_counter()
    plan = nttm.BatchPlan(prods, autotune=True)
    torch.cuda.synchronize()
    t_plan_make = time.perf_counter() - t0
    plan.run()                                               # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.run()
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / reps
    graph_same = all(torch.equal(r, it[3]) for r, it in zip(ref, items))
    sched = [l for l in plan.info().splitlines() if l.startswith("group")]
    plan.close()
    out = {"products": len(ops), "batched_products_per_s": round(len(ops) / t_batch, 1),
           "batched_ms": round(1e3 * t_batch, 3),
           "graph_products_per_s": round(len(ops) / t_graph, 1), "graph_ms": round(1e3 * t_graph, 3),
           "graph_plan_autotune_s": round(t_plan_make, 3), "graph_equals_batch": bool(graph_same),
           "graph_schedule": sched}
    if cpu:
        print("[bench] leaves cpu baseline (GMP)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_leaves_baseline(ops, len(ops) / t_graph)
    # per-kernel device time and algorithmic bytes of one eager batch (the graph replays the same launches)
    nttm.profile(True)
    nttm.mul_batch(prods)
    kst = nttm.profile_read()
    nttm.profile(False)
    out["batch_kernels"] = {k: {"launches": v[0], "ms": round(v[1], 4), "GB/s": round(v[2] / (v[1] * 1e-3) / 1e9, 1)}
                            for k, v in sorted(kst.items(), key=lambda kv: -kv[1][1])}
    # the same leaves with every operand and result in pageable host memory (staged through HBM: PCIe included)
    hops = [(np.ascontiguousarray(a), np.ascontiguousarray(b), None, np.zeros(len(a) + len(b), dtype=np.uint64))
            for a, b in ops]
    hprods = nttm.make_prods(hops)
    nttm.mul_batch(hprods)                                   # warm (staging buffers)
    t0 = time.perf_counter()
    for _ in range(reps):
        nttm.mul_batch(hprods)
    t_host = (time.perf_counter() - t0) / reps
    host_same = all(np.array_equal(h[3], r.cpu().numpy().view(np.uint64)) for h, r in zip(hops[:256], ref[:256]))
    in_bytes = 8 * sum(len(a) + len(b) for a, b in ops)
    out.update({"host_operands_products_per_s": round(len(ops) / t_host, 1), "host_operands_ms": round(1e3 * t_host, 3),
                "host_operands_bytes_each_way": [in_bytes, in_bytes],
                "host_operands_equal_device_first256": bool(host_same)})
    if not each:
        return out
    t0 = time.perf_counter()
    nttm.mul_each(prods)
    torch.cuda.synchronize()
    t_each = time.perf_counter() - t0
    same = all(torch.equal(r, it[3]) for r, it in zip(ref[:64], items[:64]))
    nttm.mul_each_async(prods)                               # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nttm.mul_each_async(prods)
    t_async = time.perf_counter() - t0
    same_async = all(torch.equal(r, it[3]) for r, it in zip(ref, items))
    out.update({"per_launch_products_per_s": round(len(ops) / t_each, 1), "speedup": round(t_each / t_batch, 2),
                "graph_speedup": round(t_each / t_graph, 2), "per_launch_ms": round(1e3 * t_each, 3),
                "batch_equals_each_first64": bool(same),
                "per_launch_async_products_per_s": round(len(ops) / t_async, 1),
                "per_launch_async_ms": round(1e3 * t_async, 3),
                "graph_speedup_vs_async": round(t_async / t_graph, 2),
                "batch_equals_each_async": bool(same_async),
                "per_launch_note": "per_launch: one nttm_mul call per product, each synchronizing and reading its "
                                   "length back; per_launch_async: the same calls launched back to back, one "
                                   "synchronization at the end"})
    return out


def bench_binsplit(nttm, reps, nodes=256, lo=12000, hi=16000, seed=0x5EED0007):
    """One level of a binary-splitting tree as one batch: per node P = P1 Q2 + P2 and Q = Q1 Q2, the two products
    sharing Q2 (E/binsplit.c:1583-1700).  Timed twice: Q2 shared (one device copy, transformed once per node) and
    unshared (a copy per product)."""
    import torch
    lens = lo + (splitmix(seed, 3 * nodes) % np.uint64(hi - lo)).astype(np.int64)
    shared, unshared = [], []
    for i in range(nodes):
        l1, lq2, lp2 = (int(v) for v in lens[3 * i:3 * i + 3])
        p1, q1 = gen_limbs(l1, seed + 4 * i + 1), gen_limbs(l1, seed + 4 * i + 2)
        q2, p2 = gen_limbs(lq2, seed + 4 * i + 3), gen_limbs(lp2, seed + 4 * i + 4)
        dp1, dq1, dq2, dp2 = to_dev(p1), to_dev(q1), to_dev(q2), to_dev(p2)
        cp = torch.empty(l1 + lq2 + 1, dtype=torch.int64, device="cuda")
        cq = torch.empty(l1 + lq2, dtype=torch.int64, device="cuda")
        shared += [(dp1, dq2, dp2, cp), (dq1, dq2, None, cq)]
        unshared += [(dp1, dq2.clone(), dp2, cp.clone()), (dq1, dq2.clone(), None, cq.clone())]
    out = {"nodes": nodes, "products": 2 * nodes, "limbs": [lo, hi]}
    for tag, items in (("shared_b", shared), ("unshared_b", unshared)):
        prods = nttm.make_prods(items)
        nttm.mul_batch(prods)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            nttm.mul_batch(prods)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        out[tag + "_ms"] = round(1e3 * t, 3)
        out[tag + "_products_per_s"] = round(2 * nodes / t, 1)
    out["same_result"] = all(torch.equal(a[3], b[3]) for a, b in zip(shared, unshared))
    out["speedup"] = round(out["unshared_b_ms"] / out["shared_b_ms"], 3)
    return out


PROF_NAMES = (("k_fused2", "k_fused2"), ("k_pass_stream<0", "k_pass<fwd>"), ("k_pass_stream<1", "k_pass<fwd,limb>"),
              ("k_pass_stream<2", "k_pass<inv>"), ("k_crt<", "k_crt"), ("k_crt_scan", "k_crt_scan"),
              ("k_crt_fix", "k_crt_fix"))


def bench_config4(nttm, da, db, reps=3):
    """the split-precision (INT8 matrix-core) radix-16 stage on the large product's shape: one prime's forward
    transform of 2^28 points with and without it, and the whole product with every forward transform through it
    (limbs -> nttm_ntt_fwd_mfma -> pointwise inverse -> CRT), checked bit-identical to nttm_mul's product"""
    import torch
    n = da.numel()
    logn = (2 * n - 1).bit_length()
    x = torch.empty(1 << logn, dtype=torch.int64, device="cuda")
    nttm.load(0, x, da)
    base = x.clone()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(f):
        best = 1e30
        for _ in range(reps):
            x.copy_(base)
            torch.cuda.synchronize()
            ev[0].record()
            f()
            ev[1].record()
            torch.cuda.synchronize()
            best = min(best, ev[0].elapsed_time(ev[1]))
        return best
    t_fp64 = timed(lambda: nttm.ntt_fwd(0, x, logn))
    y64 = x.clone()
    nttm.profile(True)
    t_mfma = timed(lambda: nttm.ntt_fwd_mfma(0, x, logn))
    ks = nttm.profile_read()
    nttm.profile(False)
    same_fwd = bool(torch.equal(x, y64))
    del y64
    stage = ks.get("k_dft16_mfma", (0, 0.0, 0.0))
    stage_ms = stage[1] / max(stage[0], 1)
    int8_ops = 30 * 2 * 16 * 16 * 64 * ((1 << logn) // 256)           # 30 MFMA per 256 points
    # the whole product with the stage in every forward transform
    planes = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(3):
        pa = torch.empty(1 << logn, dtype=torch.int64, device="cuda")
        pb = torch.empty_like(pa)
        nttm.load(i, pa, da)
        nttm.load(i, pb, db)
        nttm.ntt_fwd_mfma(i, pa, logn)
        nttm.ntt_fwd_mfma(i, pb, logn)
        nttm.ntt_inv_pw(i, pa, pb, 0, logn)
        del pb
        planes.append(pa)
    c = nttm.crt(planes, 2 * n - 1, 2 * n)
    torch.cuda.synchronize()
    t_prod = time.perf_counter() - t0
    del planes
    same_prod = bool(torch.equal(c, nttm.mul(da, db)))
    return {"transform_points": 1 << logn, "fwd_fp64_ms": round(t_fp64, 3), "fwd_mfma_ms": round(t_mfma, 3),
            "mfma_stage_kernel_ms": round(stage_ms, 3), "mfma_stage_int8_ops": int8_ops,
            "mfma_stage_int8_TOPs": round(int8_ops / (stage_ms * 1e-3) / 1e12, 1) if stage_ms else None,
            "fwd_bit_identical": same_fwd,
            "product_via_mfma_stage_s": round(t_prod, 4), "product_bit_identical_to_nttm_mul": same_prod,
            "note": "primitive path (canonical residue planes, separate load / forward / inverse / CRT calls): "
                    "compare fwd_mfma_ms with fwd_fp64_ms; the default product keeps the FP64 stage"}


REF_STRAT_A_S = {26: 0.0317, 27: 0.0631, 28: 0.1263, 29: 0.2578, 30: 0.5085, 31: 1.1028}   # results/D13b_strategy_e0.txt


def b_alg(na, nb, points, np_):
    """the algorithmic HBM bytes of one product: operand limbs read once, per prime three transforms of two passes
    (16 B per point each) plus the B plane read by the fused pointwise pass, the CRT reading every plane, result written"""
    return 8 * (na + nb) + 112 * np_ * points + 8 * (na + nb)


def bench_sweep(nttm, log_lo=24, log_hi=32, reps=3):
    """one product per transform length 2^log_lo .. 2^log_hi points (operands of 2^(log-1) limbs each, generated on the
    device from seeds 0x5EED0003 / 0x5EED0004), the largest the real-shape share of one GPU (2^31 x 2^31 limbs, four
    primes); each checked by the residues of a, b and c mod eight primes above 2^62 (c = a b exactly modulo each) and
    every limb below 10^18, all computed on the device; the reference's one-device strategy-A times beside them"""
    import torch
    out = []
    for lg in range(log_lo, log_hi + 1):
        n = 1 << (lg - 1)
        rec = {"points": 1 << lg, "limbs": n}
        try:
            print("[bench] sweep 2^%d points" % lg, file=sys.stderr, flush=True)
            da = nttm.gen_limbs_dev(n, 0x5EED0003)
            db = nttm.gen_limbs_dev(n, 0x5EED0004)
            dc = torch.empty(2 * n, dtype=torch.int64, device="cuda")
            nttm.mul(da, db, out=dc)                             # warm (workspace)
            torch.cuda.synchronize()
            ts = []
            for _ in range(reps if lg < 31 else 2):
                t0 = time.perf_counter()
                nttm.mul(da, db, out=dc)
                ts.append(time.perf_counter() - t0)
            sec = float(np.median(ts))
            np_ = 3 if n <= nttm.max_terms(3) else 4
            ra, rb, rc = nttm.limbs_mod(da), nttm.limbs_mod(db), nttm.limbs_mod(dc)
            res_ok = all((x * y) % q == z for x, y, z, q in zip(ra, rb, rc, nttm.T1_PRIMES))
            rng_ok = bool(((dc >= 0) & (dc < B)).all().item())
            top_ok = int(dc[-1].item()) != 0 or int(dc[-2].item()) != 0     # normalized length 2n or 2n - 1
            lim_ok = rng_ok and top_ok
            ba = b_alg(n, n, 1 << lg, np_)
            rec.update({"primes": np_, "ms": round(1e3 * sec, 3), "runs": len(ts), "b_alg_bytes": ba,
                        "hbm_frac": round(ba / sec / 1e9 / HBM_PEAK_GBS, 4), "residues_ok": bool(res_ok),
                        "limbs_ok": bool(lim_ok)})
            if lg in REF_STRAT_A_S:
                rec["reference_mi300a_s"] = REF_STRAT_A_S[lg]
                rec["speedup_vs_reference"] = round(REF_STRAT_A_S[lg] / sec, 2)
            del da, db, dc
        except Exception as e:                                  # noqa: BLE001 -- reported in the line
            rec["error"] = str(e)[:200]
        torch.cuda.empty_cache()
        out.append(rec)
    return out


def bench_e(nttm, digits):
    """e to `digits` fractional digits through nttm_e_fraction (one warm-up run, then one timed run, wall clock; the
    warm-up's time is reported as first_run_seconds), checked against the reference's
    sha1 of its ASCII output and its 50-digit windows (tests/golden/e_windows.json)"""
    import hashlib
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "e_windows.json")))
    m = (digits + 17) // 18
    t0 = time.perf_counter()
    f = nttm.e_fraction(m)                                        # warm-up: the process's first e (buffers grown)
    first = time.perf_counter() - t0
    del f
    t0 = time.perf_counter()
    f = nttm.e_fraction(m)
    sec = time.perf_counter() - t0
    phases = nttm.e_last_phase_ms()
    d = nttm.limbs_to_ascii(f)
    del f
    out = {"digits": digits, "limbs": m, "seconds": round(sec, 3), "first_run_seconds": round(first, 3),
           "digits_per_s": round(digits / sec, 1),
           "phases_ms": {k: round(v, 1) for k, v in phases.items()},
           "windows_ok": {o: d[int(o) - 1:int(o) + 49].decode() == w
                          for o, w in gold["windows"].items() if int(o) + 49 <= digits}}
    want = gold["sha1_16"].get(str(digits))
    if want:
        h = hashlib.sha1(b"2.")
        h.update(memoryview(d)[:digits])
        h.update(b"\n")
        out["sha1_16"] = h.hexdigest()[:16]
        out["sha1_matches_reference"] = out["sha1_16"] == want
    out["reference"] = "10^9 digits in 7.55-9.18 s on 4 MI300A APUs (results/MEMORY_IO.md:154,163,202)"
    return out


def trace_summary(db):
    """per-kernel calls / total / average / min / max duration (us) of a rocprofv3 --kernel-trace results database"""
    import sqlite3
    con = sqlite3.connect(db)
    rows = con.execute("""select name, count(*), sum(duration) / 1e3, avg(duration) / 1e3, min(duration) / 1e3,
                          max(duration) / 1e3, max(lds_size), max(grid_x), max(workgroup_x)
                          from kernels group by name order by sum(duration) desc""").fetchall()
    tot = sum(r[2] for r in rows)
    print("| kernel | calls | total us | avg us | min us | max us | % | lds B | grid | wg |")
    print("|---|---|---|---|---|---|---|---|---|---|")
    for r in rows:
        n = r[0] if len(r[0]) < 60 else r[0][:57] + "..."
        print("| %s | %d | %.1f | %.1f | %.1f | %.1f | %.1f | %s | %s | %s |"
              % (n, r[1], r[2], r[3], r[4], r[5], 100 * r[2] / tot, r[6], r[7], r[8]))


def replay_summary(db, replays=3):
    """the graph replays of a kernel trace -- the last `replays` runs of the period whose repetitions make the longest
    tail of the last burst of library kernels (k_*, no gap above 200 us) -- summarized per run: kernel time per family,
    and the span from the first start to the last end (the difference is the idle time between launches)"""
    import sqlite3
    con = sqlite3.connect(db)
    rows = con.execute("select name, start, end from kernels order by start").fetchall()
    ks = [r for r in rows if r[0].startswith("k_") or r[0].startswith("void k_")]
    bursts, cur = [], ks[:1]
    for a, b in zip(ks, ks[1:]):
        if b[1] - a[2] > 200000:
            bursts.append(cur)
            cur = [b]
        else:
            cur.append(b)
    bursts.append(cur)
    runs = None
    for bt in reversed(bursts):
        names, n = [r[0] for r in bt], len(bt)
        best, best_len = 0, 0
        for P in range(1, n // replays + 1):            # the period with the longest periodic tail (the shortest on a tie)
            L = P
            while L < n and names[n - 1 - L] == names[n - 1 - L + P]:
                L += 1
            if L >= replays * P and L > best_len:
                best, best_len = P, L
        if best:
            runs = [bt[n - (replays - i) * best:n - (replays - i - 1) * best] for i in range(replays)]
            break
    if not runs:
        print("no burst of %d repeated runs found" % replays)
        return
    period = len(runs[0])
    fam = {}
    for run in runs:
        for n, s0, e in run:
            f = fam.setdefault(n, [0, 0.0])
            f[0] += 1
            f[1] += (e - s0) / 1e3
    busy = sum(v[1] for v in fam.values()) / replays
    span = sum((run[-1][2] - run[0][1]) / 1e3 for run in runs) / replays
    print("%d kernels per run; per run: kernel time %.1f us, span %.1f us (%.1f%% idle between launches)"
          % (period, busy, span, 100 * (1 - busy / span)))
    print("| kernel | launches per run | us per run | % |")
    print("|---|---|---|---|")
    for n, (c, t) in sorted(fam.items(), key=lambda kv: -kv[1][1]):
        nm = n if len(n) < 60 else n[:57] + "..."
        print("| %s | %d | %.1f | %.1f |" % (nm, c // replays, t / replays, 100 * t / replays / busy))


def traffic_summary(fetch_db, write_db):
    """HBM bytes per dispatch of every kernel from a FETCH_SIZE and a WRITE_SIZE rocprofv3 --pmc run of
    `bench.py --copy-probe`, each scaled by what its counter reads for the probe's 1 GiB device copy (the first
    __amd_rocclr_copyBuffer dispatch of the run).  Prints a table and the JSON that bench.py reads back."""
    import collections
    import sqlite3

    def load(db):
        con = sqlite3.connect(db)
        per, meta = collections.defaultdict(float), {}
        for d, k, v in con.execute("select dispatch_id, kernel_name, value from counters_collection"):
            per[(d, k)] += v
            meta[d] = k
        return per, meta

    (fe, fm), (wr, wm) = load(fetch_db), load(write_db)
    probe = min(d for d, k in fm.items() if k == "__amd_rocclr_copyBuffer")
    cf, cw = fe[(probe, fm[probe])], wr[(probe, wm[probe])]
    ff, fw = float(1 << 20) / cf, float(1 << 20) / cw
    print("copy probe (1 GiB): FETCH_SIZE %.0f KiB, WRITE_SIZE %.0f KiB -> scale %.3f, %.3f" % (cf, cw, ff, fw))
    avg = collections.defaultdict(lambda: [0.0, 0.0, 0])
    for (d, k), v in fe.items():
        avg[k][0] += v; avg[k][2] += 1
    for (d, k), v in wr.items():
        avg[k][1] += v
    out = {}
    print("| kernel | dispatches | fetch KiB raw | write KiB raw | HBM read GB | HBM write GB |")
    print("|---|---|---|---|---|---|")
    for k, (f, w, n) in sorted(avg.items(), key=lambda kv: -kv[1][0]):
        f, w = f / n, w / n
        rd, wt = f * ff * 1024, w * fw * 1024
        print("| %s | %d | %.0f | %.0f | %.3f | %.3f |" % (k[:60], n, f, w, rd / 1e9, wt / 1e9))
        tag = next((t for pat, t in PROF_NAMES if pat in k), None)
        if tag:
            out[tag] = {"bytes": round(rd + wt), "read_bytes": round(rd), "write_bytes": round(wt)}
    out["_scale"] = {"fetch": ff, "write": fw, "probe_bytes": 1 << 30}
    print(json.dumps(out))


def counter_summary(db):
    """every counter of a rocprofv3 --pmc run averaged per dispatch of each kernel (summed over the counter's
    instances), one row per kernel"""
    import collections
    import sqlite3
    con = sqlite3.connect(db)
    per = collections.defaultdict(float)
    disp = collections.defaultdict(set)
    names = set()
    for d, k, c, v in con.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection"):
        per[(k, c)] += v
        disp[k].add(d)
        names.add(c)
    cols = sorted(names)
    print("| kernel | dispatches | " + " | ".join(cols) + " |")
    print("|---|---|" + "---|" * len(cols))
    for k in sorted(disp, key=lambda k: -per[(k, cols[0])] if cols else 0):
        n = len(disp[k])
        print("| %s | %d | " % (k[:60], n) + " | ".join("%.4g" % (per[(k, c)] / n) for c in cols) + " |")


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--counter-summary":
        counter_summary(sys.argv[2])
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--trace-summary":
        trace_summary(sys.argv[2])
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--replay-summary":
        replay_summary(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 3)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--traffic-summary":
        traffic_summary(sys.argv[2], sys.argv[3])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-limbs", type=int, default=27, help="operand length 2^k limbs (default 27)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="one-off: GMP on the full operands (conversions, one timed mpz_mul, exact compare); minutes")
    ap.add_argument("--no-config1", action="store_true", help="skip the 2^20 x 2^20 plain product")
    ap.add_argument("--config1-only", action="store_true", help="only the 2^20 x 2^20 plain product, one JSON object")
    ap.add_argument("--sweep-only", action="store_true",
                    help="only the size sweep: one product per transform length 2^24 .. 2^32 points")
    ap.add_argument("--sweep-max", type=int, default=32, help="largest transform of the sweep, log2 points")
    ap.add_argument("--no-sweep", action="store_true", help="skip the size sweep")
    ap.add_argument("--no-leaves", action="store_true")
    ap.add_argument("--leaf-reps", type=int, default=3)
    ap.add_argument("--e-digits", type=int, default=10 ** 9, help="digits of e for the pipeline line (0: skip)")
    ap.add_argument("--e-only", action="store_true", help="only the e pipeline, one JSON object")
    ap.add_argument("--no-config4", action="store_true", help="skip the INT8 matrix-core stage measurement")
    ap.add_argument("--config4-only", action="store_true", help="only the INT8 matrix-core stage measurement")
    ap.add_argument("--leaves-only", action="store_true", help="only the batched-leaf workloads, one JSON object")
    ap.add_argument("--batched-only", action="store_true", help="with --leaves-only: the 4096-leaf batch alone")
    ap.add_argument("--tile-log", type=int, default=0, help="strided-pass tile, log2 points (12 or 13; 0 = default)")
    ap.add_argument("--pts", type=int, default=0, help="transform points per thread (8 or 16; 0 = default)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first copy 1 GiB device to device once (the reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the product's multi-GPU forms are not built (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import ntt_amd as nttm

    if args.e_only:                                              # the e pipeline alone (for profiling)
        nttm.init(0, tile_log=args.tile_log, pts=args.pts)
        print(json.dumps({"e_pipeline": bench_e(nttm, args.e_digits)}))
        return
    if args.config1_only:                                        # the plain product alone (for profiling)
        nttm.init(0, tile_log=args.tile_log, pts=args.pts)
        print(json.dumps({"plain_product_2e20": bench_config1(nttm)}))
        return
    if args.sweep_only:
        nttm.init(0, tile_log=args.tile_log, pts=args.pts)
        print(json.dumps({"size_sweep": bench_sweep(nttm, log_hi=args.sweep_max)}))
        return
    if args.leaves_only:                                         # the batched-leaf workloads alone (for profiling)
        nttm.init(0, tile_log=args.tile_log, pts=args.pts)
        leaves = bench_leaves(nttm, args.leaf_reps, each=not args.batched_only)
        if not args.batched_only:
            leaves["binsplit_level"] = bench_binsplit(nttm, args.leaf_reps)
        print(json.dumps({"leaves": leaves}))
        return

    n = 1 << args.log_limbs
    print("[bench] generating 2 x 2^%d limbs" % args.log_limbs, file=sys.stderr, flush=True)
    a = gen_limbs(n, 0x5EED0003)
    b = gen_limbs(n, 0x5EED0004)
    nttm.init(0, tile_log=args.tile_log, pts=args.pts)
    if args.cpu_full:
        print(json.dumps(cpu_full(nttm, a, b)))
        return
    if args.config4_only:                                        # the matrix-core stage alone (for profiling)
        print(json.dumps({"split_precision_stage": bench_config4(nttm, to_dev(a), to_dev(b))}))
        return
    if args.copy_probe:
        src = torch.ones(1 << 27, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    da, db = to_dev(a), to_dev(b)
    dc = torch.empty(2 * n, dtype=torch.int64, device="cuda")
    print("[bench] warmup %d" % args.warmup, file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        nttm.mul(da, db, out=dc)
    torch.cuda.synchronize()
    first = dc.clone()

    nttm.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nttm.mul(da, db, out=dc)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ks = nttm.profile_read()
    nttm.profile(False)
    phase = nttm.last_phase_ms()
    deterministic = bool(torch.equal(first, dc))
    del first
    sec = (t1 - t0) / args.steps
    print("[bench] %.3f ms per product" % (1e3 * sec), file=sys.stderr, flush=True)

    ncoef = 2 * n - 1
    logt = int(np.ceil(np.log2(ncoef)))
    npts = 1 << logt
    np_ = 3
    b_alg = 8 * (2 * n) + 112 * np_ * npts + 8 * (2 * n)
    name, launches, kms, kbytes = dominant(ks)
    achieved = kbytes / (kms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(name), "kernel": name,
            "launches": launches, "avg_launch_ms": round(kms / launches, 4),
            "alg_bytes_per_launch": kbytes / launches}
    if name == "k_fused2":
        # what bounds it instead of HBM (rocprofv3 counters of the same product)
        roof["limit"] = ("vector issue: 178 VALU instructions per point (166.5 FP64) at 79% of the issue rate "
                         "(profiles/product_2e27_valu_issue.md)")

    cfg1 = None
    if not args.no_config1:
        print("[bench] plain product 2^20 x 2^20", file=sys.stderr, flush=True)
        cfg1 = bench_config1(nttm)

    cfg4 = None
    if not args.no_config4:
        print("[bench] split-precision stage (INT8 matrix cores)", file=sys.stderr, flush=True)
        cfg4 = bench_config4(nttm, da, db)

    leaves = None
    if not args.no_leaves:
        print("[bench] leaves", file=sys.stderr, flush=True)
        del dc
        torch.cuda.empty_cache()
        leaves = bench_leaves(nttm, args.leaf_reps, cpu=not args.no_cpu_baseline)
        print("[bench] binary-splitting level", file=sys.stderr, flush=True)
        leaves["binsplit_level"] = bench_binsplit(nttm, args.leaf_reps)

    epipe = None
    if args.e_digits:
        print("[bench] e to %d digits" % args.e_digits, file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
        epipe = bench_e(nttm, args.e_digits)

    sweep = None
    if not args.no_sweep:
        torch.cuda.empty_cache()
        sweep = bench_sweep(nttm, log_hi=args.sweep_max)

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (GMP)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(a, b)

    kernels = {k: {"launches": v[0], "ms": round(v[1] / args.steps, 4), "GB/s": round(v[2] / (v[1] * 1e-3) / 1e9, 1)}
               for k, v in ks.items()}
    line = {"metric": METRIC, "value": round(sec, 6), "unit": "s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(1e3 * sec, 4), "higher_is_better": False, "scaling": "weak",
            "vs_baseline": round(sec / MI300A_REF_S, 4) if args.log_limbs == 27 else None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "2^%d x 2^%d-limb product, transform 2^%d, %d primes"
                                   % (args.log_limbs, args.log_limbs, logt, np_),
                       "na": n, "nb": n, "transform_points": npts, "primes": np_,
                       "seeds": ["0x5EED0003", "0x5EED0004"], "limb_base": "10^18",
                       "residue_arithmetic": "exact integers below 2^50 carried in FP64 (Shoup/Barrett), CRT in u64",
                       "b_alg_bytes": b_alg, "product_hbm_GBs": round(b_alg / sec / 1e9, 1),
                       "product_hbm_frac": round(b_alg / sec / 1e9 / HBM_PEAK_GBS, 4),
                       "phase_ms": {"b_forward": round(phase[0], 3), "a_forward_pw_inverse": round(phase[1], 3),
                                    "crt": round(phase[2], 3)},
                       "kernels_per_step": kernels, "deterministic": deterministic,
                       "vs_baseline_ref": "time ratio to 0.1263 s, 1 MI300A (results/D13b_strategy_e0.txt:26)",
                       "plain_product_2e20": cfg1, "leaves": leaves, "e_pipeline": epipe,
                       "split_precision_stage": cfg4, "size_sweep": sweep},
            "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    nttm.shutdown()


if __name__ == "__main__":
    main()
