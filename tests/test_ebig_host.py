"""The e pipeline's host-side reciprocal and quotient (ntt_amd/csrc/ebig.h), built for the host with a schoolbook
product (tests/native/ebig_host.cpp), against exact Python integers: the long-division base case, the Newton reciprocal's
3-unit error bound at every precision, the exact corrected quotient, and the cost shape (products on truncated Q)."""
import ctypes as C
import json
import math
import os
import random
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "ebig_host.cpp")
HDR = os.path.join(HERE, "..", "ntt_amd", "csrc", "ebig.h")
SAN = os.environ.get("NTTM_NATIVE_SANITIZE") == "1"        # the AddressSanitizer / UBSan build (tests/test_sanitize.py)
SO = os.path.join(HERE, "native", "ebig_host_san.so" if SAN else "ebig_host.so")
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]
B = 10 ** 18
U64P = C.POINTER(C.c_uint64)


@pytest.fixture(scope="module")
def E():
    if not os.path.exists(SO) or os.path.getmtime(SO) < max(os.path.getmtime(SRC), os.path.getmtime(HDR)):
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared"] + (SAN_FLAGS if SAN else []) + ["-o", SO, SRC],
                       check=True)
    L = C.CDLL(SO)
    sz = C.c_size_t
    L.eb_t_recip.argtypes = [U64P, sz, sz, U64P, sz, C.c_int]; L.eb_t_recip.restype = C.c_long
    L.eb_t_wraps.restype = C.c_long
    L.eb_t_quotient.argtypes = [U64P, sz, U64P, sz, sz, U64P, sz, C.c_int]; L.eb_t_quotient.restype = C.c_long
    L.eb_t_div_pow.argtypes = [sz, U64P, sz, U64P, sz]; L.eb_t_div_pow.restype = C.c_long
    L.eb_t_products.restype = C.c_long
    L.eb_t_max_operand.restype = sz
    return L


def limbs(v):
    out = []
    while v:
        v, r = divmod(v, B)
        out.append(r)
    return out


def arr(v):
    l = limbs(v)
    return (C.c_uint64 * max(len(l), 1))(*l), len(l)


def value(buf, n):
    return sum(int(buf[i]) * B ** i for i in range(n))


def recip(E, q, m, wrap=1):
    qa, n = arr(q)
    cap = m + 4
    out = (C.c_uint64 * cap)()
    k = E.eb_t_recip(qa, n, m, out, cap, wrap)
    assert k >= 0
    return value(out, k)


def quotient(E, a, q, m, full=0):
    aa, na = arr(a)
    qa, n = arr(q)
    cap = m + 2
    out = (C.c_uint64 * cap)()
    k = E.eb_t_quotient(aa, na, qa, n, m, out, cap, full)
    assert k >= 0
    return value(out, k)


def q_kinds(rnd, n):
    """Q of n limbs: random, all B - 1, a power of B, top limb 1 or B - 1 over random low limbs"""
    r = rnd.randrange(B ** (n - 1), B ** n)
    return [r, B ** n - 1, B ** (n - 1), B ** (n - 1) + rnd.randrange(B ** (n - 1)),
            (B - 1) * B ** (n - 1) + rnd.randrange(B ** (n - 1))]


def test_div_pow_base_case(E):
    rnd = random.Random(3)
    for _ in range(60):
        nd = rnd.randint(1, 6)
        d = rnd.randrange(B ** (nd - 1), B ** nd)
        e = rnd.randint(nd - 1, nd + 9)
        da, n = arr(d)
        out = (C.c_uint64 * (e + 2))()
        k = E.eb_t_div_pow(e, da, n, out, e + 2)
        assert value(out, k) == B ** e // d


@pytest.mark.parametrize("n,m", [(1, 1), (1, 9), (2, 8), (3, 9), (5, 17), (12, 30), (40, 70), (90, 33), (7, 200)])
def test_recip_within_three_units(E, n, m):
    rnd = random.Random(n * 1000 + m)
    for q in q_kinds(rnd, n):
        x = recip(E, q, m)
        assert abs(x - B ** (n + m) // q) <= 3, (n, m)


@pytest.mark.parametrize("full", [0, 1])
@pytest.mark.parametrize("n,m", [(1, 1), (2, 3), (4, 9), (10, 10), (25, 60), (60, 25), (3, 120)])
def test_quotient_exact(E, n, m, full):
    """the one-product quotient (the guard limb decides) and the path through the remainder"""
    rnd = random.Random(7 * n + m)
    for q in q_kinds(rnd, n):
        for a in (0, 1, q - 1, rnd.randrange(q), q // 2):
            assert quotient(E, a, q, m, full) == a * B ** m // q, (n, m)


def quotient_products(E, a, q, m):
    """products that eb_quotient made for floor(a B^m / q) (1: the guard limb decided; 2: through the remainder)"""
    E.eb_t_reset()
    recip(E, q, m + 1)
    r = E.eb_t_products()
    E.eb_t_reset()
    assert quotient(E, a, q, m) == a * B ** m // q
    return E.eb_t_products() - r


@pytest.mark.parametrize("m", [4, 9, 30])
def test_quotient_exact_divisions(E, m):
    """Q a product of powers of 2 and 5 that divides B^m, so A B^m / Q is an integer for every A: the fraction sits
    on an integer, the guard limb cannot decide, and the quotient goes through the remainder (two products)"""
    for q in (2 ** 60, 5 ** 26, 2 ** 40 * 5 ** 30, 2 ** 100 * 5 ** 3, 1 << 59):
        if B ** m % q:
            continue
        for a in (1, 3, q - 1, q // 2, q // 3, (q * 2) // 3 + 1):
            assert quotient_products(E, a, q, m) == 2, (q, a, m)


def test_quotient_one_product(E):
    """random A < Q: the guard limb decides, and the quotient takes the one product A x"""
    rnd = random.Random(17)
    for n, m in ((3, 7), (10, 20), (40, 25)):
        q = rnd.randrange(B ** (n - 1), B ** n)
        for _ in range(8):
            assert quotient_products(E, rnd.randrange(q), q, m) == 1, (n, m)


@pytest.mark.parametrize("n,m", [(3, 40), (40, 70), (90, 33), (7, 200), (300, 250), (600, 1000)])
def test_recip_wrapped_product(E, n, m):
    """the Newton step's product formed modulo B^L - 1 (the wrap-around) gives the same reciprocal, limb for limb, as
    the whole product, and the wrapped path is taken"""
    rnd = random.Random(n * 7 + m)
    for q in q_kinds(rnd, n):
        x0 = recip(E, q, m, wrap=0)
        x1 = recip(E, q, m, wrap=1)
        assert E.eb_t_wraps() > 0
        assert x0 == x1, (n, m)
        assert abs(x1 - B ** (n + m) // q) <= 3


def test_reciprocal_products_on_truncated_q(E):
    """Q of 600 limbs, 40 limbs of reciprocal: every Newton product stays near 40 limbs, and their number grows with
    log m"""
    rnd = random.Random(11)
    q = rnd.randrange(B ** 599, B ** 600)
    E.eb_t_reset()
    x = recip(E, q, 40)
    assert abs(x - B ** 640 // q) <= 3
    assert E.eb_t_max_operand() <= 40 + 4
    assert E.eb_t_products() <= 2 * math.ceil(math.log2(40))


def test_e_fraction_from_series(E):
    """e's fraction through the host reciprocal and quotient from P, Q of the series; the reference's first window"""
    m = 8
    need = 18 * (m + 2) * math.log(10)
    n = 2
    while math.lgamma(n + 1) < need:
        n += 1
    q = math.factorial(n)
    p = sum(q // math.factorial(k) for k in range(1, n + 1))
    y = quotient(E, p - q, q, m)
    digits = "%0*d" % (18 * m, y)
    win = json.load(open(os.path.join(HERE, "golden", "e_windows.json")))["windows"]["50"]
    assert digits[49:99] == win
    assert digits.startswith("718281828459045235360287471352662497757")
