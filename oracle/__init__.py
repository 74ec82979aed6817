"""The oracle: a CPU restatement of the reference's product path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import this package, and only as
the checker (or as the timed CPU side of the baseline).  The product library never loads it.

``lib()`` loads ``oracle/liborc.so`` (built by ``oracle/Makefile``); ``ref()`` loads ``oracle/_ref/libecref.so``, the
reference's own host code compiled where it lies, when that build exists (only where the reference tree is present).
The numpy helpers below wrap the C entry points declared in ``oracle/oracle.h``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
B = 10 ** 18
GEN_UNIFORM, GEN_ONES, GEN_ZEROS, GEN_BIT, GEN_SPARSE = range(5)

_u64p = C.POINTER(C.c_uint64)
_lib = None
_ref = None


SANITIZE = os.environ.get("ORC_SANITIZE") == "1"       # load the AddressSanitizer / UBSan build of the oracle
LIBNAME = "liborc_san.so" if SANITIZE else "liborc.so"


def build(force=False):
    """Compile the oracle (and the reference build when the reference tree exists) with make."""
    args = ["make", "-s", "-C", HERE] + (["liborc_san.so"] if SANITIZE else [])
    if force:
        args.append("-B")
    subprocess.run(args, check=True)


def _p(a):
    return a.ctypes.data_as(_u64p) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, LIBNAME)
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_splitmix64.argtypes = [_u64p]; L.orc_splitmix64.restype = C.c_uint64
        L.orc_gen_limbs.argtypes = [_u64p, C.c_size_t, C.c_int, _u64p]; L.orc_gen_limbs.restype = None
        L.orc_mul_school.argtypes = [_u64p, _u64p, C.c_size_t, _u64p, C.c_size_t]; L.orc_mul_school.restype = None
        L.orc_mul_add_school.argtypes = [_u64p, _u64p, C.c_size_t, _u64p, C.c_size_t, _u64p, C.c_size_t]
        L.orc_mul_add_school.restype = None
        for f in ("orc_mulmod", "orc_powmod"):
            getattr(L, f).argtypes = [C.c_uint64] * 3; getattr(L, f).restype = C.c_uint64
        L.orc_invmod.argtypes = [C.c_uint64] * 2; L.orc_invmod.restype = C.c_uint64
        L.orc_is_prime.argtypes = [C.c_uint64]; L.orc_is_prime.restype = C.c_int
        L.orc_generator.argtypes = [C.c_uint64]; L.orc_generator.restype = C.c_uint64
        L.orc_root2k.argtypes = [C.c_uint64, C.c_int]; L.orc_root2k.restype = C.c_uint64
        for f in ("orc_ntt_fwd", "orc_ntt_inv"):
            getattr(L, f).argtypes = [_u64p, C.c_int, C.c_uint64, C.c_uint64]; getattr(L, f).restype = None
        L.orc_dft_bitrev.argtypes = [_u64p, _u64p, C.c_int, C.c_uint64, C.c_uint64]; L.orc_dft_bitrev.restype = None
        L.orc_root3k.argtypes = [C.c_uint64, C.c_int]; L.orc_root3k.restype = C.c_uint64
        L.orc_dft3.argtypes = [_u64p, _u64p, C.c_int, C.c_uint64, C.c_uint64]; L.orc_dft3.restype = None
        for f in ("orc_ntt3_fwd", "orc_ntt3_inv"):
            getattr(L, f).argtypes = [_u64p, C.c_int, C.c_uint64, C.c_uint64]; getattr(L, f).restype = None
        L.orc_garner.argtypes = [_u64p, _u64p, C.c_int, _u64p]; L.orc_garner.restype = None
        L.orc_div1e18.argtypes = [C.c_uint64, C.c_uint64, _u64p, _u64p]; L.orc_div1e18.restype = None
        L.orc_words_to_dec.argtypes = [_u64p, C.c_int, _u64p, C.c_int]; L.orc_words_to_dec.restype = None
        L.orc_crt_carry.argtypes = [C.POINTER(_u64p), _u64p, C.c_int, C.c_size_t, _u64p, C.c_size_t, _u64p, C.c_size_t]
        L.orc_crt_carry.restype = C.c_uint64
        L.orc_max_terms.argtypes = [_u64p, C.c_int]; L.orc_max_terms.restype = C.c_uint64
        L.orc_mul_ntt.argtypes = [_u64p, _u64p, C.c_size_t, _u64p, C.c_size_t, _u64p, C.c_size_t, _u64p, C.c_int]
        L.orc_mul_ntt.restype = C.c_int
        L.orc_limbs_mod.argtypes = [_u64p, C.c_size_t, C.c_uint64]; L.orc_limbs_mod.restype = C.c_uint64
        L.orc_tref_res.argtypes = [_u64p, C.c_size_t, _u64p, C.c_size_t, _u64p, C.c_size_t]; L.orc_tref_res.restype = C.c_int
        L.orc_have_gmp.argtypes = []; L.orc_have_gmp.restype = C.c_int
        L.orc_gmp_mul.argtypes = [_u64p, _u64p, C.c_size_t, _u64p, C.c_size_t]; L.orc_gmp_mul.restype = C.c_int
        L.orc_gmp_time_mul.argtypes = [_u64p, C.c_size_t, _u64p, C.c_size_t, C.c_int]; L.orc_gmp_time_mul.restype = C.c_double
        L.orc_gmp_tref_edges.argtypes = [_u64p, C.c_size_t, _u64p, C.c_size_t, _u64p, C.c_size_t]
        L.orc_gmp_tref_edges.restype = C.c_int
        L.orc_gmp_check.argtypes = [_u64p, C.c_size_t, _u64p, C.c_size_t, _u64p, C.c_size_t, _u64p, C.c_size_t]
        L.orc_gmp_check.restype = C.c_int
        L.orc_gmp_full_check.argtypes = [_u64p, C.c_size_t, _u64p, C.c_size_t, _u64p, C.c_size_t, C.POINTER(C.c_double)]
        L.orc_gmp_full_check.restype = C.c_int
        L.orc_gmp_check_batch.argtypes = [C.POINTER(_u64p), C.POINTER(C.c_size_t), C.POINTER(_u64p), C.POINTER(C.c_size_t),
                                          C.POINTER(_u64p), C.POINTER(C.c_size_t), C.c_size_t, C.POINTER(C.c_uint8)]
        L.orc_gmp_check_batch.restype = C.c_size_t
        L.orc_gmp_time_batch.argtypes = [C.POINTER(_u64p), C.POINTER(C.c_size_t), C.POINTER(_u64p), C.POINTER(C.c_size_t),
                                         C.c_size_t, C.c_int, C.c_int]
        L.orc_gmp_time_batch.restype = C.c_double
        L.orc_pack60_bytes.argtypes = [C.c_size_t]; L.orc_pack60_bytes.restype = C.c_uint64
        L.orc_pack60.argtypes = [_u64p, C.c_size_t, C.c_void_p]; L.orc_pack60.restype = None
        L.orc_unpack60.argtypes = [C.c_void_p, C.c_size_t, _u64p]; L.orc_unpack60.restype = None
        _lib = L
    return _lib


def ref():
    """The reference's own host code (oracle/_ref/libecref.so), or None where it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libecref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_gen_limbs.argtypes = [_u64p, C.c_size_t, C.c_int, _u64p]; R.ref_gen_limbs.restype = None
        R.ref_div1e18.argtypes = [C.c_uint64, C.c_uint64, _u64p, _u64p]; R.ref_div1e18.restype = None
        R.ref_words_to_dec3.argtypes = [_u64p, _u64p]; R.ref_words_to_dec3.restype = None
        R.ref_words_to_dec4.argtypes = [_u64p, _u64p]; R.ref_words_to_dec4.restype = None
        R.ref_primes.argtypes = [_u64p]; R.ref_primes.restype = None
        R.ref_np3_max_terms.argtypes = []; R.ref_np3_max_terms.restype = C.c_uint64
        R.ref_set_np.argtypes = [C.c_int]; R.ref_set_np.restype = None
        R.ref_tref_ok.argtypes = [_u64p, C.c_size_t, _u64p, C.c_size_t, _u64p, C.c_size_t, C.c_int]
        R.ref_tref_ok.restype = C.c_int
        R.ref_limbs_mod_u64.argtypes = [_u64p, C.c_size_t, C.c_uint64]; R.ref_limbs_mod_u64.restype = C.c_uint64
        R.ref_mpz_mul_limbs.argtypes = [_u64p, _u64p, C.c_size_t, _u64p, C.c_size_t]; R.ref_mpz_mul_limbs.restype = None
        R.limb_mul_school.argtypes = [_u64p, _u64p, C.c_size_t, _u64p, C.c_size_t]; R.limb_mul_school.restype = None
        R.crt_garner3.argtypes = [_u64p, _u64p]; R.crt_garner3.restype = None
        R.crt_garner4.argtypes = [_u64p, _u64p]; R.crt_garner4.restype = None
        R.crt_carry_par4.argtypes = [C.POINTER(_u64p), C.c_size_t, _u64p, C.c_int]; R.crt_carry_par4.restype = None
        _ref = R
    return _ref


# ---- numpy helpers -------------------------------------------------------------------------------------------------
def gen_limbs(n, seed, kind=GEN_UNIFORM):
    """n limbs of the given kind from a splitmix64 stream seeded with `seed`; returns (limbs, advanced seed)."""
    a = np.zeros(max(n, 1), dtype=np.uint64)
    s = np.array([seed], dtype=np.uint64)
    lib().orc_gen_limbs(_p(a), n, kind, _p(s))
    return a[:n], int(s[0])


def mul_school(a, b, x=None):
    a = np.ascontiguousarray(a, dtype=np.uint64); b = np.ascontiguousarray(b, dtype=np.uint64)
    n = len(a) + len(b) + (1 if x is not None else 0)
    c = np.zeros(max(n, 1), dtype=np.uint64)
    if x is None:
        lib().orc_mul_school(_p(c), _p(a), len(a), _p(b), len(b))
    else:
        x = np.ascontiguousarray(x, dtype=np.uint64)
        lib().orc_mul_add_school(_p(c), _p(a), len(a), _p(b), len(b), _p(x), len(x))
    return c[:n]


def mul_ntt(a, b, primes, x=None):
    a = np.ascontiguousarray(a, dtype=np.uint64); b = np.ascontiguousarray(b, dtype=np.uint64)
    P = np.ascontiguousarray(primes, dtype=np.uint64)
    n = len(a) + len(b) + (1 if x is not None else 0)
    c = np.zeros(max(n, 1), dtype=np.uint64)
    xa = None if x is None else np.ascontiguousarray(x, dtype=np.uint64)
    rc = lib().orc_mul_ntt(_p(c), _p(a), len(a), _p(b), len(b), _p(xa), 0 if xa is None else len(xa), _p(P), len(P))
    if rc:
        raise ValueError("orc_mul_ntt refused the product (prime bound or roots)")
    return c[:n]


def ntt_fwd(x, p, w):
    x = np.array(x, dtype=np.uint64); logn = int(len(x)).bit_length() - 1
    lib().orc_ntt_fwd(_p(x), logn, p, w)
    return x


def ntt_inv(x, p, w):
    x = np.array(x, dtype=np.uint64); logn = int(len(x)).bit_length() - 1
    lib().orc_ntt_inv(_p(x), logn, p, w)
    return x


def dft_bitrev(x, p, w):
    x = np.ascontiguousarray(x, dtype=np.uint64); X = np.zeros_like(x)
    lib().orc_dft_bitrev(_p(x), _p(X), int(len(x)).bit_length() - 1, p, w)
    return X


def _logk3(n):
    m = n // 3
    assert n == 3 * m and m & (m - 1) == 0, "length must be 3 2^k"
    return m.bit_length() - 1


def root3k(p, logk):
    return int(lib().orc_root3k(p, logk))


def dft3(x, p, w):
    """O(n^2) DFT of n = 3 2^k points; X[3t + r] at r 2^k + bitrev(t)."""
    x = np.ascontiguousarray(x, dtype=np.uint64); X = np.zeros_like(x)
    lib().orc_dft3(_p(x), _p(X), _logk3(len(x)), p, w)
    return X


def ntt3_fwd(x, p, w):
    x = np.array(x, dtype=np.uint64)
    lib().orc_ntt3_fwd(_p(x), _logk3(len(x)), p, w)
    return x


def ntt3_inv(x, p, w):
    x = np.array(x, dtype=np.uint64)
    lib().orc_ntt3_inv(_p(x), _logk3(len(x)), p, w)
    return x


def crt_carry(planes, primes, nc, x=None):
    """Limbs of sum_k CRT(planes[.][k]) B^k (+ x); returns (limbs, overflow word)."""
    P = np.ascontiguousarray(primes, dtype=np.uint64)
    pl = [np.ascontiguousarray(q, dtype=np.uint64) for q in planes]
    arr = (_u64p * len(pl))(*[_p(q) for q in pl])
    c = np.zeros(max(nc, 1), dtype=np.uint64)
    xa = None if x is None else np.ascontiguousarray(x, dtype=np.uint64)
    ov = lib().orc_crt_carry(arr, _p(P), len(P), len(pl[0]), _p(xa), 0 if xa is None else len(xa), _p(c), nc)
    return c[:nc], int(ov)


def limbs_mod(a, q):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return int(lib().orc_limbs_mod(_p(a), len(a), q))


def max_terms(primes):
    P = np.ascontiguousarray(primes, dtype=np.uint64)
    return int(lib().orc_max_terms(_p(P), len(P)))


# ---- the packed digit file (E/packed_fmt.h) --------------------------------------------------------------------------
# the T1 primes of the header residues: the first eight primes above 2^62 (E/verify.c:15-20)
T1_PRIMES = (4611686018427388039, 4611686018427388073, 4611686018427388081, 4611686018427388091,
             4611686018427388093, 4611686018427388097, 4611686018427388157, 4611686018427388181)
# the header (E/packed_fmt.h:32-48): magic, 9 x u32, 4 bytes of alignment, 10 x u64, q[8], dres[8]
_HDR_U32 = ("version", "hdr_bytes", "endian", "limb_bytes", "limb_digits", "order", "part", "nparts", "reserved0")
_HDR_U64 = ("d", "d_out", "nl", "pad", "lo", "hi", "k0", "k1", "complete", "nq")


def pack60(limbs_file_order):
    """the 60-bit stream of limbs given most significant first (E/packed_fmt.h:50-78), as bytes"""
    l = np.ascontiguousarray(limbs_file_order, dtype=np.uint64)
    out = np.zeros(max(int(lib().orc_pack60_bytes(len(l))), 1), dtype=np.uint8)
    lib().orc_pack60(_p(l), len(l), out.ctypes.data)
    return out[:int(lib().orc_pack60_bytes(len(l)))].tobytes()


def unpack60(buf, n):
    """n limbs (file order) from a 60-bit stream"""
    b = np.frombuffer(buf, dtype=np.uint8)
    if len(b) < int(lib().orc_pack60_bytes(n)):
        raise ValueError("stream shorter than %d limbs" % n)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    lib().orc_unpack60(b.ctypes.data, n, _p(out))
    return out[:n]


def read_packed(path):
    """(header dict, limbs in file order) of a single-part version-2 packed file"""
    import struct
    with open(path, "rb") as f:
        raw = f.read()
    h = {"magic": raw[:8]}
    h.update(zip(_HDR_U32, struct.unpack_from("<9I", raw, 8)))
    h.update(zip(_HDR_U64, struct.unpack_from("<10Q", raw, 48)))
    h["q"] = list(struct.unpack_from("<8Q", raw, 128))
    h["dres"] = list(struct.unpack_from("<8Q", raw, 192))
    h["text"] = raw[1024:h["hdr_bytes"]].split(b"\0", 1)[0].decode(errors="replace")
    n = h["hi"] - h["lo"]
    return h, unpack60(raw[h["hdr_bytes"]:], n)


def packed_ascii(h, limbs_file_order):
    """the text form of a whole packed number: "2." + digits 1..d_out + "\\n" (E/packed_fmt.h:6-14)"""
    d = np.asarray(limbs_file_order, dtype=np.uint64)
    out = np.empty((len(d), 18), dtype=np.uint8)
    for part, col in (((d // 10 ** 9).astype(np.uint32), 0), ((d % 10 ** 9).astype(np.uint32), 9)):
        for i in range(8, -1, -1):
            out[:, col + i] = (part % 10).astype(np.uint8) + 48
            part //= 10
    s = out.tobytes()[h["pad"]:]                                  # digit k at index k
    return s[:1] + b"." + s[1:h["d_out"] + 1] + b"\n"


def have_gmp():
    return bool(lib().orc_have_gmp())


def gmp_mul(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64); b = np.ascontiguousarray(b, dtype=np.uint64)
    c = np.zeros(len(a) + len(b), dtype=np.uint64)
    if lib().orc_gmp_mul(_p(c), _p(a), len(a), _p(b), len(b)):
        raise RuntimeError("oracle built without GMP")
    return c


def gmp_check(a, b, c, x=None):
    a = np.ascontiguousarray(a, dtype=np.uint64); b = np.ascontiguousarray(b, dtype=np.uint64)
    c = np.ascontiguousarray(c, dtype=np.uint64)
    xa = None if x is None else np.ascontiguousarray(x, dtype=np.uint64)
    r = lib().orc_gmp_check(_p(a), len(a), _p(b), len(b), _p(xa), 0 if xa is None else len(xa), _p(c), len(c))
    if r < 0:
        raise RuntimeError("oracle built without GMP")
    return bool(r)


def gmp_time_mul(a, b, reps=3):
    a = np.ascontiguousarray(a, dtype=np.uint64); b = np.ascontiguousarray(b, dtype=np.uint64)
    return float(lib().orc_gmp_time_mul(_p(a), len(a), _p(b), len(b), reps))


def gmp_time_batch(pairs, threads=1, reps=1):
    """seconds of mpz_mul over every (a, b) of pairs on `threads` OpenMP threads, conversions off the clock"""
    arrs = [(np.ascontiguousarray(a, dtype=np.uint64), np.ascontiguousarray(b, dtype=np.uint64)) for a, b in pairs]
    n = len(arrs)
    pa = (_u64p * n)(*[_p(a) for a, _ in arrs])
    pb = (_u64p * n)(*[_p(b) for _, b in arrs])
    na = (C.c_size_t * n)(*[len(a) for a, _ in arrs])
    nb = (C.c_size_t * n)(*[len(b) for _, b in arrs])
    return float(lib().orc_gmp_time_batch(pa, na, pb, nb, n, threads, reps))


def gmp_check_batch(triples):
    """[ok_i]: c_i == a_i b_i exactly for every (a, b, c) of triples, checked on every host thread"""
    arrs = [tuple(np.ascontiguousarray(v, dtype=np.uint64) for v in t) for t in triples]
    n = len(arrs)
    ptr = lambda k: (_u64p * n)(*[_p(t[k]) for t in arrs])
    ln = lambda k: (C.c_size_t * n)(*[len(t[k]) for t in arrs])
    ok = (C.c_uint8 * max(n, 1))()
    r = lib().orc_gmp_check_batch(ptr(0), ln(0), ptr(1), ln(1), ptr(2), ln(2), n, ok)
    if r == C.c_size_t(-1).value:
        raise RuntimeError("oracle built without GMP")
    return [bool(ok[i]) for i in range(n)]


def gmp_full_check(a, b, c):
    """(equal, seconds to convert a and b, seconds of one mpz_mul, seconds to convert c): the exact check of a large
    product against GMP with the mpz_mul timed on the same operands, progress on stderr"""
    a = np.ascontiguousarray(a, dtype=np.uint64); b = np.ascontiguousarray(b, dtype=np.uint64)
    c = np.ascontiguousarray(c, dtype=np.uint64)
    t = (C.c_double * 3)()
    r = lib().orc_gmp_full_check(_p(a), len(a), _p(b), len(b), _p(c), len(c), t)
    if r < 0:
        raise RuntimeError("oracle built without GMP")
    return bool(r), t[0], t[1], t[2]


def tref_res(a, b, c):
    a = np.ascontiguousarray(a, dtype=np.uint64); b = np.ascontiguousarray(b, dtype=np.uint64)
    c = np.ascontiguousarray(c, dtype=np.uint64)
    return bool(lib().orc_tref_res(_p(a), len(a), _p(b), len(b), _p(c), len(c)))


def tref_edges(a, b, c):
    a = np.ascontiguousarray(a, dtype=np.uint64); b = np.ascontiguousarray(b, dtype=np.uint64)
    c = np.ascontiguousarray(c, dtype=np.uint64)
    r = lib().orc_gmp_tref_edges(_p(a), len(a), _p(b), len(b), _p(c), len(c))
    if r < 0:
        raise RuntimeError("oracle built without GMP")
    return bool(r)


def check_product(a, b, c, x=None, exact_limit=1 << 21):
    """The reference checker's policy (E/tests/tref.h:1-21): exact by GMP up to exact_limit product limbs (or when
    GMP is absent, the schoolbook below 2^14 limbs), else the residue form (limbs, length, residues, both edges)."""
    na, nb = len(a), len(b)
    if have_gmp() and (na + nb <= exact_limit or x is not None):
        return gmp_check(a, b, c, x)
    if not have_gmp() and na + nb <= (1 << 14):
        return bool(np.array_equal(np.asarray(c, dtype=np.uint64), mul_school(a, b, x)))
    if x is not None:
        raise RuntimeError("residue form has no x term")
    ok = tref_res(a, b, c)
    if ok and have_gmp():
        ok = tref_edges(a, b, c)
    return ok
