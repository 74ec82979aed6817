"""The product library's FP64 residue arithmetic (ntt_amd/csrc/arith.h), compiled for the host with the same
operation sequence, against exact integer arithmetic: congruence, the output bounds every kernel relies on
(|r| <= p for the stated input ranges), and the extreme inputs of those ranges."""
import ctypes as C
import os
import random
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "arith_host.c")
SAN = os.environ.get("NTTM_NATIVE_SANITIZE") == "1"        # the AddressSanitizer / UBSan build (tests/test_sanitize.py)
SO = os.path.join(HERE, "native", "arith_host_san.so" if SAN else "arith_host.so")
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]
B = 10 ** 18


@pytest.fixture(scope="module")
def A():
    if not os.path.exists(SO) or os.path.getmtime(SO) < max(os.path.getmtime(SRC), os.path.getmtime(
            os.path.join(HERE, "..", "ntt_amd", "csrc", "arith.h"))):
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared"] + (SAN_FLAGS if SAN else []) +
                       ["-o", SO, SRC, "-lm"], check=True)
    L = C.CDLL(SO)
    d = C.c_double
    L.t_mul_shoup.argtypes = [d, d, d, d]; L.t_mul_shoup.restype = d
    L.t_mul_barrett.argtypes = [d, d, d, d]; L.t_mul_barrett.restype = d
    L.t_reduce.argtypes = [d, d, d]; L.t_reduce.restype = d
    L.t_limb_to_res.argtypes = [C.c_uint64, d, d]; L.t_limb_to_res.restype = d
    L.t_res_canon.argtypes = [d, d, d]; L.t_res_canon.restype = C.c_uint64
    L.t_div_b.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]; L.t_div_b.restype = None
    L.t_mulmod_b.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.t_mulmod_b.restype = None
    L.t_f64_to_u52.argtypes = [d]; L.t_f64_to_u52.restype = C.c_uint64
    L.t_prime.argtypes = [C.c_int]; L.t_prime.restype = C.c_uint64
    L.t_divb_v.argtypes = []; L.t_divb_v.restype = C.c_uint64
    return L


def primes(A):
    return [A.t_prime(i) for i in range(4)]


def bal(v, p):
    v %= p
    return v - p if v > p // 2 else v


def test_primes_and_constant(A):
    for p in primes(A):
        assert p < 2 ** 50 and (p - 1) % 2 ** 32 == 0
    assert A.t_divb_v() == (2 ** 128 - 1) // (16 * B) - 2 ** 64


def edge_values(p, k):
    """k p, k p - 1, ..., symmetric, plus small values"""
    lim = k * p
    return [lim, lim - 1, -lim, -lim + 1, 0, 1, -1, p, -p, p - 1, -(p - 1), p // 2, -(p // 2), 2 * p, -2 * p]


def test_mul_shoup_bounds_and_congruence(A):
    rnd = random.Random(1)
    for p in primes(A):
        ws = [bal(w, p) for w in (1, p - 1, p // 2, p // 2 + 1, 2, 3)] + [bal(rnd.randrange(p), p) for _ in range(30)]
        for w in ws:
            wq = w / p
            for a in [x for x in edge_values(p, 4) if abs(x) <= 4 * p] + [rnd.randrange(-4 * p, 4 * p + 1) for _ in range(300)]:
                r = A.t_mul_shoup(float(a), float(w), wq, float(p))
                assert r == int(r) and abs(r) <= 3 * p // 4 + 1, (a, w, r)     # single-rounding quotient: 3p/4
                assert (int(r) - a * w) % p == 0
                if abs(a) <= 2 * p:
                    assert abs(r) <= 5 * p // 8 + 1


def test_mul_barrett_bounds_and_congruence(A):
    rnd = random.Random(2)
    for p in primes(A):
        vals = edge_values(p, 1) + [rnd.randrange(-p, p + 1) for _ in range(200)]
        vals = [v for v in vals if abs(v) <= p]
        for a in vals:
            for b in vals[:40] + [rnd.randrange(-p, p + 1) for _ in range(10)]:
                r = A.t_mul_barrett(float(a), float(b), float(p), 1.0 / p)
                assert r == int(r) and abs(r) <= 3 * p // 4 + 1, (a, b, r)
                assert (int(r) - a * b) % p == 0


def test_reduce_and_canon(A):
    rnd = random.Random(3)
    for p in primes(A):
        for s in [v for v in edge_values(p, 4) if abs(v) <= 4 * p] + [rnd.randrange(-4 * p, 4 * p + 1) for _ in range(2000)]:
            r = A.t_reduce(float(s), float(p), 1.0 / p)
            assert r == int(r) and abs(r) <= p // 2 + 1 and (int(r) - s) % p == 0
        for v in [v for v in edge_values(p, 1) if abs(v) <= p] + [rnd.randrange(-p, p + 1) for _ in range(2000)]:
            c = A.t_res_canon(float(v), float(p), 1.0 / p)
            assert 0 <= c < p and c == v % p


def test_limb_to_res(A):
    rnd = random.Random(4)
    for p in primes(A):
        for x in [0, 1, B - 1, B - 2, p - 1, p, p + 1, 2 ** 32 - 1, 2 ** 32, 2 ** 59, 2 ** 50] + [rnd.randrange(B) for _ in range(3000)]:
            r = A.t_limb_to_res(x, float(p), 1.0 / p)
            assert r == int(r) and abs(r) <= p // 2 + 1 and (int(r) - x) % p == 0


def test_div_b(A):
    rnd = random.Random(5)
    q, r = C.c_uint64(), C.c_uint64()
    cases = [(0, 0), (0, 1), (0, B - 1), (0, B), (B - 1, 2 ** 64 - 1), (B - 1, 0), (1, 0), (0, 2 ** 64 - 1)]
    cases += [(rnd.randrange(B), rnd.randrange(2 ** 64)) for _ in range(20000)]
    cases += [(rnd.randrange(2 ** 20), rnd.randrange(2 ** 64)) for _ in range(2000)]
    for hi, lo in cases:
        A.t_div_b(hi, lo, C.byref(q), C.byref(r))
        n = (hi << 64) | lo
        assert (q.value, r.value) == (n // B, n % B), (hi, lo)


def test_mulmod_b(A):
    """t M = a B + b exactly, for the digit split of the CRT: t below 2^64 (the CRT uses t < 2^50), M below B, with
    Mp = floor(M 2^64 / B); the extremes of both ranges and the products of the prime prefixes' base-B digits."""
    rnd = random.Random(11)
    a, b = C.c_uint64(), C.c_uint64()
    ps = primes(A)
    ms = [1, 2, B - 1, B - 2, 10 ** 17, 2 ** 59, ps[0], (ps[0] * ps[1]) % B, (ps[0] * ps[1]) // B,
          (ps[0] * ps[1] * ps[2]) % B, ((ps[0] * ps[1] * ps[2]) // B) % B, (ps[0] * ps[1] * ps[2]) // B ** 2]
    ms += [rnd.randrange(B) for _ in range(40)]
    ts = [0, 1, 2 ** 50 - 1, ps[0] - 1, ps[3] - 1, 2 ** 52, 2 ** 63, 2 ** 64 - 1] + [rnd.randrange(2 ** 50) for _ in range(300)]
    for m in ms:
        mp = (m << 64) // B
        for t in ts:
            A.t_mulmod_b(t, m, mp, C.byref(a), C.byref(b))
            assert (a.value, b.value) == divmod(t * m, B), (t, m)


def test_f64_to_u52(A):
    rnd = random.Random(12)
    for v in [0, 1, 2, 2 ** 32 - 1, 2 ** 32, 2 ** 50 - 1, 2 ** 52 - 1] + [rnd.randrange(2 ** 52) for _ in range(5000)]:
        assert A.t_f64_to_u52(float(v)) == v
