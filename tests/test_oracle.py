"""Pins of the oracle (CPU): every function against the reference's golden vectors (tests/golden, written by the
reference's own compiled code), against the reference build itself where it exists (oracle/_ref), against GMP, and
against closed forms and brute force chosen so that a dropped term, a wrong index or a transposed operand fails."""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))
B = O.B
REF_P = GOLD["primes"]


def u64(x):
    return np.array(x, dtype=np.uint64)


def to_int(limbs):
    v = 0
    for x in reversed([int(t) for t in limbs]):
        v = v * B + x
    return v


def from_int(v, n):
    out = []
    for _ in range(n):
        out.append(v % B); v //= B
    assert v == 0
    return u64(out)


# ---- generator ----------------------------------------------------------------------------------------------------
def test_generator_matches_reference_vectors():
    for g in GOLD["gen_limbs"]:
        a, s = O.gen_limbs(g["n"], g["seed"], g["kind"])
        assert [int(v) for v in a[:16]] == g["limbs"], g["kind"]
        assert [int(v) for v in a[-4:]] == g["tail"]
        assert int(sum(int(v) for v in a)) == g["sum"]
        assert s == g["seed_after"]


def test_generator_properties():
    a, _ = O.gen_limbs(5000, 0x5EED0003)
    assert (a < B).all() and a[-1] != 0
    assert len(set(int(v) for v in a)) == 5000            # no repeats at this size
    z, _ = O.gen_limbs(10, 1, O.GEN_ZEROS)
    assert list(z) == [0] * 9 + [1]
    o, _ = O.gen_limbs(10, 1, O.GEN_ONES)
    assert (o == B - 1).all()


# ---- schoolbook ------------------------------------------------------------------------------------------------------
def test_school_matches_reference_vectors():
    for case in GOLD["school"]:
        assert [int(v) for v in O.mul_school(u64(case["a"]), u64(case["b"]))] == case["c"]


@pytest.mark.parametrize("na,nb", [(1, 1), (1, 7), (13, 2), (64, 64), (100, 37)])
def test_school_vs_python_ints(na, nb):
    a, s = O.gen_limbs(na, na * 1000 + nb)
    b, _ = O.gen_limbs(nb, s)
    x, _ = O.gen_limbs(na + nb - 1, 99)
    assert to_int(O.mul_school(a, b)) == to_int(a) * to_int(b)
    c = O.mul_school(a, b, x)
    assert len(c) == na + nb + 1 and to_int(c) == to_int(a) * to_int(b) + to_int(x)
    assert (c < B).all()


def test_school_against_reference_build():
    R = O.ref()
    if R is None:
        pytest.skip("reference build absent (oracle/_ref)")
    import ctypes as C
    for na, nb in ((3, 5), (511, 512), (1000, 1)):
        a, _ = O.gen_limbs(na, na); b, _ = O.gen_limbs(nb, nb + 1, O.GEN_SPARSE)
        r = np.zeros(na + nb, dtype=np.uint64)
        R.limb_mul_school(r.ctypes.data_as(C.POINTER(C.c_uint64)), a.ctypes.data_as(C.POINTER(C.c_uint64)), na,
                          b.ctypes.data_as(C.POINTER(C.c_uint64)), nb)
        assert np.array_equal(r, O.mul_school(a, b))


# ---- modular helpers, roots, NTT ------------------------------------------------------------------------------------
def test_reference_roots_reproduced():
    L = O.lib()
    for i, p in enumerate(REF_P):
        assert L.orc_is_prime(p)
        assert L.orc_generator(p) == GOLD["generators"][i]
        assert L.orc_root2k(p, 33) == GOLD["w33"][i]
        for k, vals in GOLD["roots"].items():
            assert L.orc_root2k(p, int(k)) == vals[i]


def test_root_orders():
    L = O.lib()
    for p in REF_P:
        for k in (1, 5, 20, 33):
            w = L.orc_root2k(p, k)
            assert L.orc_powmod(w, 1 << k, p) == 1
            assert L.orc_powmod(w, 1 << (k - 1), p) == p - 1


def test_is_prime_brute_force():
    L = O.lib()
    sieve = [True] * 5000
    sieve[0] = sieve[1] = False
    for i in range(2, 5000):
        if sieve[i]:
            for j in range(i * i, 5000, i):
                sieve[j] = False
    assert all(bool(L.orc_is_prime(n)) == sieve[n] for n in range(5000))
    assert not L.orc_is_prime(4611686018427388039 * 3)
    assert L.orc_is_prime(4611686018427388039)               # t1_q[0] of the reference checker


@pytest.mark.parametrize("logn", [1, 2, 3, 5, 8])
def test_ntt_fwd_equals_dft(logn):
    p = REF_P[0]
    w = O.lib().orc_root2k(p, logn)
    x, _ = O.gen_limbs(1 << logn, logn)
    x = x % np.uint64(p)
    assert np.array_equal(O.ntt_fwd(x, p, w), O.dft_bitrev(x, p, w))


def test_dft_is_the_textbook_sum():
    # n = 4 by hand: X[k] = sum_i x[i] w^(ik), written at bitrev(k) = [0, 2, 1, 3]
    p = 17
    w = 4                                                    # order 4 mod 17
    x = [3, 1, 4, 1]
    X = [sum(x[i] * pow(w, i * k, p) for i in range(4)) % p for k in range(4)]
    assert [int(v) for v in O.dft_bitrev(u64(x), p, w)] == [X[0], X[2], X[1], X[3]]


@pytest.mark.parametrize("logn", [1, 4, 10, 14])
def test_ntt_roundtrip_and_convolution(logn):
    p = REF_P[1]
    w = O.lib().orc_root2k(p, logn)
    n = 1 << logn
    a, s = O.gen_limbs(n, 1000 + logn); b, _ = O.gen_limbs(n, s)
    a %= np.uint64(p); b %= np.uint64(p)
    assert np.array_equal(O.ntt_inv(O.ntt_fwd(a, p, w), p, w), a)
    if n <= 1024:                                             # cyclic convolution by brute force
        A, Bf = O.ntt_fwd(a, p, w), O.ntt_fwd(b, p, w)
        prod = u64([(int(u) * int(v)) % p for u, v in zip(A, Bf)])
        conv = O.ntt_inv(prod, p, w)
        ai, bi = [int(v) for v in a], [int(v) for v in b]
        want = [sum(ai[i] * bi[(k - i) % n] for i in range(n)) % p for k in range(n)]
        assert [int(v) for v in conv] == want


# ---- transforms of 3 2^k points -----------------------------------------------------------------------------------------
def test_reference_roots3_reproduced():
    # the reference's W3X33 table and ec_root3 (E/modarith.h:63,206-214) follow from the smallest generator
    for i, p in enumerate(REF_P):
        assert O.root3k(p, 33) == GOLD["w3x33"][i]
        for k, vals in GOLD["roots3"].items():
            assert O.root3k(p, int(k)) == vals[i]
        w = O.root3k(p, 20)
        n = 3 << 20
        assert pow(w, n, p) == 1 and pow(w, n // 2, p) != 1 and pow(w, n // 3, p) != 1


def test_dft3_hand_computed():
    # n = 3 mod 7 (w = 2, a cube root of unity): X = [6, 17, 57] mod 7 = [6, 3, 1], order X[3t + r] at r + 0
    assert [int(v) for v in O.dft3(u64([1, 2, 3]), 7, 2)] == [6, 3, 1]
    # n = 6 mod 7 (w = 3 of order 6), x = e_1: X[k] = 3^k = [1, 3, 2, 6, 4, 5]; position 2r + t holds X[3t + r]
    assert [int(v) for v in O.dft3(u64([0, 1, 0, 0, 0, 0]), 7, 3)] == [1, 6, 3, 4, 2, 5]
    # n = 12 mod 13 (w = 2 of order 12), x = e_1: X[k] = 2^k; position 4r + bitrev2(t) holds X[3t + r]
    assert [int(v) for v in O.dft3(u64([0, 1] + [0] * 10), 13, 2)] == [1, 12, 8, 5, 2, 11, 3, 10, 4, 9, 6, 7]


@pytest.mark.parametrize("logk", [0, 1, 2, 3, 6])
def test_ntt3_equals_dft3(logk):
    for p in (REF_P[2], 1125844072267777):
        w = O.root3k(p, logk)
        x, _ = O.gen_limbs(3 << logk, 77 + logk)
        x = x % np.uint64(p)
        assert np.array_equal(O.ntt3_fwd(x, p, w), O.dft3(x, p, w))


@pytest.mark.parametrize("logk", [1, 3, 7, 12])
def test_ntt3_roundtrip_and_convolution(logk):
    p = 1125818302464001
    w = O.root3k(p, logk)
    n = 3 << logk
    a, s = O.gen_limbs(n, 500 + logk); b, _ = O.gen_limbs(n, s)
    a %= np.uint64(p); b %= np.uint64(p)
    assert np.array_equal(O.ntt3_inv(O.ntt3_fwd(a, p, w), p, w), a)
    if n <= 400:                                              # cyclic convolution of length 3 2^k by brute force
        A, Bf = O.ntt3_fwd(a, p, w), O.ntt3_fwd(b, p, w)
        conv = O.ntt3_inv(u64([(int(u) * int(v)) % p for u, v in zip(A, Bf)]), p, w)
        ai, bi = [int(v) for v in a], [int(v) for v in b]
        assert [int(v) for v in conv] == [sum(ai[i] * bi[(k - i) % n] for i in range(n)) % p for k in range(n)]


# ---- Garner, digits, CRT ---------------------------------------------------------------------------------------------
def test_garner_matches_reference_vectors():
    L = O.lib()
    import ctypes as C
    Pp = u64(REF_P)
    for g in GOLD["garner"]:
        r = u64(g["r"])
        for np_, key in ((3, "g3"), (4, "g4")):
            out = np.zeros(np_, dtype=np.uint64)
            L.orc_garner(r.ctypes.data_as(C.POINTER(C.c_uint64)), Pp.ctypes.data_as(C.POINTER(C.c_uint64)), np_,
                         out.ctypes.data_as(C.POINTER(C.c_uint64)))
            assert [int(v) for v in out] == g[key]
            val = sum(int(v) << (64 * i) for i, v in enumerate(out))
            assert all(val % REF_P[i] == g["r"][i] for i in range(np_))     # the CRT property itself
            d = np.zeros(np_, dtype=np.uint64)
            L.orc_words_to_dec(out.ctypes.data_as(C.POINTER(C.c_uint64)), np_, d.ctypes.data_as(C.POINTER(C.c_uint64)), np_)
            assert [int(v) for v in d] == g["dec%d" % np_]
            assert sum(int(v) * B ** i for i, v in enumerate(d)) == val


def test_div1e18_matches_reference_vectors():
    L = O.lib()
    import ctypes as C
    q = C.c_uint64(); r = C.c_uint64()
    for hi, lo, wq, wr in GOLD["div1e18"]:
        L.orc_div1e18(hi, lo, C.byref(q), C.byref(r))
        assert (q.value, r.value) == (wq, wr)
        assert wq * B + wr == (hi << 64) + lo and wr < B


def test_crt_carry_matches_reference_vectors():
    for case in GOLD["crt_carry"]:
        np_ = case["np"]; n = len(case["planes"][0])
        limbs, ov = O.crt_carry([u64(q) for q in case["planes"]], REF_P[:np_], n + 4)
        assert [int(v) for v in limbs] == case["limbs"] and ov == 0


def test_crt_carry_with_x_by_python_ints():
    rng = np.random.default_rng(5)
    n = 50
    planes = [u64([int(rng.integers(0, p)) for _ in range(n)]) for p in REF_P[:3]]
    x = rng.integers(0, B, n + 2, dtype=np.uint64)
    L = O.lib()
    import ctypes as C
    Pp = u64(REF_P)
    total = to_int(x)
    for k in range(n):
        r = u64([q[k] for q in planes]); out = np.zeros(3, dtype=np.uint64)
        L.orc_garner(r.ctypes.data_as(C.POINTER(C.c_uint64)), Pp.ctypes.data_as(C.POINTER(C.c_uint64)), 3,
                     out.ctypes.data_as(C.POINTER(C.c_uint64)))
        total += sum(int(v) << (64 * i) for i, v in enumerate(out)) * B ** k
    nc = n + 4
    limbs, ov = O.crt_carry(planes, REF_P[:3], nc, x)
    assert ov == 0 and to_int(limbs) == total


def test_max_terms_reference_bound():
    assert O.max_terms(REF_P[:3]) == GOLD["np3_max_terms"] == 58424467928
    M = REF_P[0] * REF_P[1] * REF_P[2]
    t = GOLD["np3_max_terms"]
    assert t * (B - 1) ** 2 < M <= (t + 1) * (B - 1) ** 2


# ---- the whole product on the CPU -------------------------------------------------------------------------------------
OUR_P = [1125844072267777, 1125818302464001, 1125625028935681, 1125122517762049]


@pytest.mark.parametrize("na,nb,kind", [(1, 1, 0), (1, 9, 0), (511, 512, 0), (384, 1, 0), (300, 300, 1),
                                        (200, 77, 2), (128, 128, 3), (1000, 900, 4)])
def test_mul_ntt_vs_school(na, nb, kind):
    a, s = O.gen_limbs(na, 7 * na + nb, kind); b, _ = O.gen_limbs(nb, s, kind)
    for P in (OUR_P[:3], REF_P[:3], OUR_P):
        assert np.array_equal(O.mul_ntt(a, b, P), O.mul_school(a, b))


def test_mul_ntt_with_x_and_closed_forms():
    n = 257
    ones = np.full(n, B - 1, dtype=np.uint64)
    c = O.mul_ntt(ones, ones, OUR_P[:3])                    # (B^n - 1)^2 = B^2n - 2 B^n + 1
    assert to_int(c) == (B ** n - 1) ** 2
    assert int(c[0]) == 1 and int(c[n]) == B - 2 and all(int(v) == 0 for v in c[1:n]) and all(int(v) == B - 1 for v in c[n + 1:])
    a, _ = O.gen_limbs(100, 3)
    one = u64([1])
    assert np.array_equal(O.mul_ntt(a, one, OUR_P[:3])[:100], a)
    x, _ = O.gen_limbs(150, 4)
    b, _ = O.gen_limbs(60, 5)
    c = O.mul_ntt(a, b, OUR_P[:3], x)
    assert to_int(c) == to_int(a) * to_int(b) + to_int(x)
    pw = O.mul_ntt(np.array([0] * 5 + [1], dtype=np.uint64), np.array([0] * 3 + [1], dtype=np.uint64), OUR_P[:3])
    assert to_int(pw) == B ** 8                             # B^5 B^3


@pytest.mark.skipif(not O.have_gmp(), reason="GMP absent")
def test_gmp_bridge_and_checker():
    a, s = O.gen_limbs(3000, 11); b, _ = O.gen_limbs(2500, s)
    c = O.gmp_mul(a, b)
    assert np.array_equal(c, O.mul_school(a, b))
    assert O.gmp_check(a, b, c)
    x, _ = O.gen_limbs(10, 1)
    assert O.gmp_check(a, b, O.mul_school(a, b, x), x)
    assert O.tref_res(a, b, c) and O.tref_edges(a, b, c)
    assert O.gmp_time_mul(a, b, 3) > 0
    pairs = [O.gen_limbs(40 + 30 * i, 100 + i)[0:1] + O.gen_limbs(500 - 20 * i, 200 + i)[0:1] for i in range(20)]
    assert O.gmp_time_batch(pairs, 1, 3) > 0 and O.gmp_time_batch(pairs, 4, 1) > 0
    assert O.gmp_time_batch([], 2, 1) >= 0


@pytest.mark.skipif(not O.have_gmp(), reason="GMP absent")
def test_gmp_batch_and_full_checks():
    """the batch checker and the full-size checker agree with the schoolbook on exact products and reject a product
    with one limb changed by one, a limb >= B and a truncated result"""
    trip = []
    for i in range(12):
        a, s = O.gen_limbs(10 + 97 * i, 300 + i); b, _ = O.gen_limbs(700 - 31 * i, s)
        trip.append((a, b, O.mul_school(a, b)))
    assert all(O.gmp_check_batch(trip))
    bad = [(a, b, c.copy()) for a, b, c in trip[:3]]
    bad[0][2][5] += np.uint64(1)
    bad[1][2][0] = np.uint64(B)
    bad[2] = (bad[2][0], bad[2][1], bad[2][2][:-1])
    assert O.gmp_check_batch(bad + trip[3:4]) == [False, False, False, True]
    a, b, c = trip[7]
    ok, t_conv, t_mul, t_c = O.gmp_full_check(a, b, c)
    assert ok and min(t_conv, t_mul, t_c) >= 0
    d = c.copy(); d[len(d) // 2] ^= np.uint64(1)
    assert not O.gmp_full_check(a, b, d)[0]


def test_checker_catches_single_limb_faults():
    # the reference checker's self-test (E/tests/tref.h:105-126): low, middle and top limbs changed
    a, s = O.gen_limbs(700, 21); b, _ = O.gen_limbs(650, s)
    c = O.mul_school(a, b)
    assert O.tref_res(a, b, c)
    rng = np.random.default_rng(1)
    for pos in (1, len(c) // 2, len(c) - 1):
        d = c.copy()
        d[pos] = (int(d[pos]) + 1 + int(rng.integers(0, B - 1))) % B
        assert not O.tref_res(a, b, d) or (O.have_gmp() and not O.tref_edges(a, b, d))
        if O.have_gmp():
            assert not O.gmp_check(a, b, d)


def test_limbs_mod_matches_reference_vectors():
    lm = GOLD["limbs_mod"]
    a = u64(lm["a"])
    for q, r in zip(lm["q"], lm["r"]):
        assert O.limbs_mod(a, q) == r == to_int(a) % q


# ---- the packed digit stream and its header residues (E/packed_fmt.h, E/verify.c:18) ------------------------------
def test_pack60_matches_reference_vectors():
    for case in GOLD["pack60"]:
        l = u64(case["limbs"])
        got = O.pack60(l)
        assert len(got) == case["bytes"] == 15 * (len(l) // 2) + 8 * (len(l) & 1)
        assert got.hex() == case["hex"]
        assert np.array_equal(O.unpack60(got, len(l)), l)


def test_pack60_hand_example_and_bit_positions():
    # limbs 1 and 2 (file order): bit 0 and bit 61 of the stream
    assert O.pack60(u64([1, 2])) == bytes([1, 0, 0, 0, 0, 0, 0, 0x20]) + bytes(7)
    # a lone limb is 8 bytes, zero-extended; limb j starts at bit 60 j: 1 in the third limb is bit 120 = byte 15
    assert O.pack60(u64([B - 1])) == (B - 1).to_bytes(8, "little")
    assert O.pack60(u64([0, 0, 1])) == bytes(15) + bytes([1]) + bytes(7)
    # the stream as one integer: sum of limb_j 2^(60 j)
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 4, 7, 10, 31):
        l = rng.integers(0, B, n, dtype=np.uint64)
        s = O.pack60(l)
        assert int.from_bytes(s, "little") == sum(int(v) << (60 * j) for j, v in enumerate(l))
        assert np.array_equal(O.unpack60(s, n), l)


def test_t1_primes_and_residues_match_reference():
    t1 = GOLD["t1"]
    assert list(O.T1_PRIMES) == t1["q"]
    # the first eight primes above 2^62 (brute force over the gaps)
    found = [q for q in range(2 ** 62, 2 ** 62 + 400) if O.lib().orc_is_prime(q)][:8]
    assert found == t1["q"]
    a = u64(t1["a"])
    for q, r in zip(t1["q"], t1["r"]):
        assert O.limbs_mod(a, q) == r == to_int(a) % q


def test_packed_reader_round_trip(tmp_path):
    """a packed file assembled here by the header's documented layout reads back to the same fields and text form"""
    import struct
    x = u64([360287471352662497, 718281828459045235, 2])         # X = 2.718281828459045235360287471352662497
    nl, pad, d = 3, 17, 36
    hdr = bytearray(4096)
    hdr[:8] = b"ECPACK18"
    struct.pack_into("<9I", hdr, 8, 2, 4096, 0x01020304, 8, 18, 1, 0, 1, 60)
    struct.pack_into("<10Q", hdr, 48, d, 30, nl, pad, 0, nl, 0, d + 1, 1, 8)
    struct.pack_into("<8Q", hdr, 128, *O.T1_PRIMES)
    body = O.pack60(x[::-1])
    f = tmp_path / "e.packed"
    f.write_bytes(bytes(hdr) + body)
    h, limbs = O.read_packed(str(f))
    assert h["magic"] == b"ECPACK18" and h["version"] == 2 and h["reserved0"] == 60 and h["nl"] == 3
    assert np.array_equal(limbs, x[::-1])
    assert O.packed_ascii(h, limbs) == b"2.718281828459045235360287471352\n"         # d_out = 30 digits
