"""Writes tests/golden/reference_vectors.json from the reference's OWN code (oracle/_ref/libecref.so: the reference's
bigint.c, crt.c, verify.c and header functions compiled where they lie) and from the reference's published constants.

Run where the reference tree and its build exist:  make -C oracle && python tests/golden/make_golden.py
The fixture is data only (inputs and the reference's outputs), so the tests can pin the oracle on machines where the
reference build is absent.
"""
import ctypes as C
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.json")
u64p = C.POINTER(C.c_uint64)


def p(a):
    return a.ctypes.data_as(u64p)


def main():
    R = O.ref()
    if R is None:
        sys.exit("oracle/_ref/libecref.so is missing: build it with make -C oracle where the reference tree exists")
    R.ref_tables.argtypes = [u64p] * 4
    R.ref_root.argtypes = [C.c_int, C.c_int]; R.ref_root.restype = C.c_uint64
    out = {"source": "cecoppola/ntt ecalc host code (bigint.c, crt.c, verify.c, modarith.h, tests/harness.h, tests/tref.h)"}

    # prime set and root tables (E/modarith.h:55-69), the three-prime bound (E/crt.c:28-44; ALGORITHM.md:113)
    P = np.zeros(4, dtype=np.uint64); R.ref_primes(p(P))
    tabs = [np.zeros(4, dtype=np.uint64) for _ in range(4)]
    R.ref_tables(*[p(t) for t in tabs])
    out["primes"] = [int(v) for v in P]
    out["generators"] = [int(v) for v in tabs[0]]
    out["w33"] = [int(v) for v in tabs[1]]
    out["roots"] = {str(k): [int(R.ref_root(i, k)) for i in range(4)] for k in (1, 2, 10, 20, 28, 33)}
    # roots of order 3 2^k (E/modarith.h:63,206-214)
    R.ref_root3.argtypes = [C.c_int, C.c_int]; R.ref_root3.restype = C.c_uint64
    out["w3x33"] = [int(v) for v in tabs[2]]
    out["roots3"] = {str(k): [int(R.ref_root3(i, k)) for i in range(4)] for k in (0, 1, 2, 10, 20, 33)}
    out["np3_max_terms"] = int(R.ref_np3_max_terms())

    # operand generator (E/tests/harness.h:55-81)
    gens = []
    for kind in range(5):
        for n, seed in ((1, 1), (5, 7), (64, 0x5EED0001), (1000, 0x5EED0002)):
            a = np.zeros(n, dtype=np.uint64); s = np.array([seed], dtype=np.uint64)
            R.ref_gen_limbs(p(a), n, kind, p(s))
            gens.append({"kind": kind, "n": n, "seed": seed, "limbs": [int(v) for v in a[:16]],
                         "tail": [int(v) for v in a[-4:]], "sum": int(sum(int(v) for v in a)), "seed_after": int(s[0])})
    out["gen_limbs"] = gens

    # schoolbook products (E/bigint.c:77)
    rng = np.random.default_rng(20261202)
    school = []
    for na, nb in ((1, 1), (2, 3), (7, 5), (33, 17)):
        a = rng.integers(0, O.B, na, dtype=np.uint64); b = rng.integers(0, O.B, nb, dtype=np.uint64)
        a[-1] = O.B - 1
        c = np.zeros(na + nb, dtype=np.uint64); R.limb_mul_school(p(c), p(a), na, p(b), nb)
        school.append({"a": [int(v) for v in a], "b": [int(v) for v in b], "c": [int(v) for v in c]})
    out["school"] = school

    # Garner (E/crt.c:114-144) and decimal digits (E/modarith.h:71-105)
    garner = []
    for t in range(8):
        r = np.array([int(rng.integers(0, int(q))) for q in P], dtype=np.uint64)
        if t == 0:
            r = P - np.uint64(1)
        o3 = np.zeros(3, dtype=np.uint64); R.crt_garner3(p(r), p(o3))
        o4 = np.zeros(4, dtype=np.uint64); R.crt_garner4(p(r), p(o4))
        d3 = np.zeros(3, dtype=np.uint64); R.ref_words_to_dec3(p(o3), p(d3))
        d4 = np.zeros(4, dtype=np.uint64); R.ref_words_to_dec4(p(o4), p(d4))
        garner.append({"r": [int(v) for v in r], "g3": [int(v) for v in o3], "g4": [int(v) for v in o4],
                       "dec3": [int(v) for v in d3], "dec4": [int(v) for v in d4]})
    out["garner"] = garner
    divs = []
    for t in range(16):
        hi = int(rng.integers(0, O.B)) if t else O.B - 1
        lo = int(rng.integers(0, 2 ** 63)) * 2 + (t & 1)
        q = np.zeros(1, dtype=np.uint64); rr = np.zeros(1, dtype=np.uint64)
        R.ref_div1e18(hi, lo, p(q), p(rr))
        divs.append([hi, lo, int(q[0]), int(rr[0])])
    out["div1e18"] = divs

    # the windowed CRT of residue planes (E/crt.c:146-199) with three and four primes
    crts = []
    for np_, n, adv in ((3, 40, False), (3, 40, True), (4, 33, False)):
        planes = []
        for i in range(4):
            if adv:
                planes.append(np.full(n, int(P[i]) - 1, dtype=np.uint64))
            else:
                planes.append(np.array([int(rng.integers(0, int(P[i]))) for _ in range(n)], dtype=np.uint64))
        R.ref_set_np(np_)
        arr = (u64p * 4)(*[p(q) for q in planes])
        res = np.zeros(n + 4, dtype=np.uint64)
        R.crt_carry_par4(arr, n, p(res), 3)
        crts.append({"np": np_, "planes": [[int(v) for v in q] for q in planes[:np_]], "limbs": [int(v) for v in res]})
    R.ref_set_np(3)
    out["crt_carry"] = crts

    # residue checks (E/tests/harness.h:259-264) with the T1 primes of the reference checker (E/verify.c:18)
    t1q = [4611686018427388039, 4611686018427388073, 4611686018427388081]
    a = rng.integers(0, O.B, 100, dtype=np.uint64)
    out["limbs_mod"] = {"a": [int(v) for v in a], "q": t1q, "r": [int(R.ref_limbs_mod_u64(p(a), 100, q)) for q in t1q]}

    # the packed digit stream (E/packed_fmt.h:56-78) and the header residues' primes and residues (E/verify.c:18,30)
    R.ref_ecp60_pack.argtypes = [u64p, C.c_size_t, C.c_void_p]; R.ref_ecp60_pack.restype = None
    R.ref_ecp60_bytes.argtypes = [C.c_uint64]; R.ref_ecp60_bytes.restype = C.c_uint64
    R.ref_t1_q.argtypes = [u64p]; R.ref_t1_q.restype = None
    R.vf_limbs_mod.argtypes = [u64p, C.c_size_t, C.c_uint64]; R.vf_limbs_mod.restype = C.c_uint64
    rng2 = np.random.default_rng(0xEC60)
    packs = []
    for n in (1, 2, 3, 5, 8, 33):
        l = rng2.integers(0, O.B, n, dtype=np.uint64)
        l[0] = O.B - 1
        if n > 2:
            l[1] = 0
        nb = int(R.ref_ecp60_bytes(n))
        buf = np.zeros(nb + 16, dtype=np.uint8)
        R.ref_ecp60_pack(p(l), n, buf.ctypes.data)
        packs.append({"limbs": [int(v) for v in l], "bytes": nb, "hex": buf[:nb].tobytes().hex()})
    out["pack60"] = packs
    T1 = np.zeros(8, dtype=np.uint64); R.ref_t1_q(p(T1))
    a = rng2.integers(0, O.B, 600, dtype=np.uint64)
    out["t1"] = {"q": [int(v) for v in T1], "a": [int(v) for v in a], "r": [int(R.vf_limbs_mod(p(a), 600, int(q))) for q in T1]}

    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
