/* oracle.h - CPU restatement of the reference's product path, for TESTS ONLY.
 *
 * This is test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker (or, for cpu_baseline, as the timed CPU side).  The product library never links or calls it.
 *
 * What it restates (reference paths relative to the reference repository root, E/ = ecalc/):
 *   - the operand generator: splitmix64, limb % 10^18, top limb forced nonzero     (E/tests/harness.h:55-81)
 *   - the base-10^18 schoolbook product with __int128 columns                        (E/bigint.c:77-110)
 *   - c = a b + x with the top carry limb                                            (E/rns_mul.h:48-51, E/rns_mul.c:403)
 *   - DIF forward (natural in, bit-reversed out) / DIT inverse (x n^-1) NTTs        (E/ntt.h:1-8, E/ntt.c:630-669)
 *   - transforms of 3 2^k points: radix-3 layer over the thirds + 2^k NTTs          (E/ntt.c:671-779)
 *   - Garner for 3 or 4 primes                                                       (E/crt.c:114-144)
 *   - base-10^18 digits of a CRT value and the windowed limb sum                     (E/modarith.h:71-105, E/crt.c:146-199,
 *                                                                                     E/rns_mul.c:332-405)
 *   - the largest term count np primes reconstruct                                   (E/crt.c:28-44)
 *   - residue checks of a limb array mod q (Horner over B)                           (E/tests/harness.h:259-264)
 *   - the reference product checker's residue form                                   (E/tests/tref.h:46-79)
 * GMP-backed pieces (exact product, timing of mpz_mul, the product-tree conversions of E/tests/harness.h:104-199) are in
 * orc_gmp.c and exist only when GMP was found at build time (orc_have_gmp()).
 *
 * The parity pins of every function (reference golden vectors, the reference's own host code built under oracle/_ref,
 * GMP, closed forms) are run by tests/test_oracle.py.
 */
#ifndef NTTM_ORACLE_H
#define NTTM_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define ORC_B 1000000000000000000ULL

enum { ORC_GEN_UNIFORM = 0, ORC_GEN_ONES = 1, ORC_GEN_ZEROS = 2, ORC_GEN_BIT = 3, ORC_GEN_SPARSE = 4 };

uint64_t orc_splitmix64(uint64_t *state);
/* a[0..n) of the given kind from a splitmix64 stream seeded with *state (advanced) */
void orc_gen_limbs(uint64_t *a, size_t n, int kind, uint64_t *state);

/* c[0..na+nb) = a b (schoolbook); c[0..na+nb+1) = a b + x when x is given (nx <= na + nb) */
void orc_mul_school(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb);
void orc_mul_add_school(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
                        const uint64_t *x, size_t nx);

/* exact modular helpers (unsigned __int128) */
uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t p);
uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t p);
uint64_t orc_invmod(uint64_t a, uint64_t p);
int orc_is_prime(uint64_t n);
/* the smallest generator of (Z/p)^* (p prime) */
uint64_t orc_generator(uint64_t p);
/* g^((p-1)/2^logn) for the smallest generator g: a root of exact order 2^logn (0 if 2^logn does not divide p-1) */
uint64_t orc_root2k(uint64_t p, int logn);

/* in place, residues canonical in [0,p); w is a primitive 2^logn-th root */
void orc_ntt_fwd(uint64_t *x, int logn, uint64_t p, uint64_t w);      /* natural -> bit-reversed */
void orc_ntt_inv(uint64_t *x, int logn, uint64_t p, uint64_t w);      /* bit-reversed -> natural, x n^-1 (uses w^-1) */
/* O(n^2) DFT X[k] = sum_i x[i] w^(ik), written at position bitrev(k) */
void orc_dft_bitrev(const uint64_t *x, uint64_t *X, int logn, uint64_t p, uint64_t w);

/* transforms of n = 3 2^logk points (E/ntt.c:671-779): w a primitive n-th root; output order: X[3t + r] at position
 * r 2^logk + bitrev(t) */
uint64_t orc_root3k(uint64_t p, int logk);           /* g^((p-1)/(3 2^logk)), 0 if 3 2^logk does not divide p-1 */
void orc_dft3(const uint64_t *x, uint64_t *X, int logk, uint64_t p, uint64_t w);    /* O(n^2) definition */
void orc_ntt3_fwd(uint64_t *x, int logk, uint64_t p, uint64_t w);                    /* radix-3 layer + 2^logk DIF */
void orc_ntt3_inv(uint64_t *x, int logk, uint64_t p, uint64_t w);                    /* the inverse, x n^-1 */

/* Garner: r[i] in [0,P[i]) -> the value < prod P as np little-endian words */
void orc_garner(const uint64_t *r, const uint64_t *P, int np, uint64_t *out);
/* q = floor((hi 2^64 + lo) / 10^18), r = remainder; requires hi < 10^18 */
void orc_div1e18(uint64_t hi, uint64_t lo, uint64_t *q, uint64_t *r);
/* the nw-word value w (little-endian, < 10^(18 nd)) into nd base-10^18 digits */
void orc_words_to_dec(const uint64_t *w, int nw, uint64_t *d, int nd);
/* planes[i][k] (k < ncoef) are the residues mod P[i] of convolution coefficient k; writes the nc limbs of
 * sum_k coef_k B^k + x (x optional).  Returns the final carry beyond c[nc-1] (0 for a valid product). */
uint64_t orc_crt_carry(const uint64_t *const *planes, const uint64_t *P, int np, size_t ncoef,
                       const uint64_t *x, size_t nx, uint64_t *c, size_t nc);
/* largest nterms with nterms (10^18-1)^2 < prod P (E/crt.c:28-44 for any np) */
uint64_t orc_max_terms(const uint64_t *P, int np);
/* the whole product path on the CPU with the given primes: load -> fwd(a), fwd(b) -> pointwise -> inv -> CRT.
 * c gets na+nb limbs (na+nb+1 with x).  logn chosen as ceil(log2(na+nb-1)), at least 1.  Returns 0, or -1 when the
 * primes cannot hold the coefficients or lack the 2^logn roots. */
int orc_mul_ntt(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *x, size_t nx,
                const uint64_t *P, int np);

/* value of a limb array mod q (any q < 2^63) */
uint64_t orc_limbs_mod(const uint64_t *a, size_t n, uint64_t q);
/* the residue form of the reference checker, parts (1) limbs < B and length, (2) residues mod three 62-bit primes;
 * returns 1 if c passes.  Parts (3)/(4) need GMP: orc_gmp_tref_edges. */
int orc_tref_res(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *c, size_t nc);

/* ---- packed digit output (E/packed_fmt.h:50-88): l in file order (most significant limb first) ---- */
uint64_t orc_pack60_bytes(size_t n);
/* the 60-bit little-endian bit stream of n limbs, written bit by bit */
void orc_pack60(const uint64_t *l, size_t n, uint8_t *out);
/* n limbs read back from the stream, two per 15 bytes */
void orc_unpack60(const uint8_t *in, size_t n, uint64_t *l);

/* ---- GMP bridge (orc_gmp.c) ---- */
int orc_have_gmp(void);
/* c = a b exactly via mpz_mul (conversions by product tree); c has room for na+nb limbs.  Returns 0 or -1 (no GMP) */
int orc_gmp_mul(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb);
/* seconds of mpz_mul alone on the converted operands, median of reps; -1 without GMP */
double orc_gmp_time_mul(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int reps);
/* seconds of n mpz_mul calls a[i] b[i] on `threads` OpenMP threads, conversions off the clock, median of reps; -1
 * without GMP */
double orc_gmp_time_batch(const uint64_t *const *a, const size_t *na, const uint64_t *const *b, const size_t *nb, size_t n,
                          int threads, int reps);
/* the residue form's parts (3) low 64 limbs exact and (4) top limbs within [AhBh, AhBh + Ah + Bh]; 1 if c passes */
int orc_gmp_tref_edges(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *c, size_t nc);
/* n products checked exactly on every host thread: ok[i] = 1 when c_i = a_i b_i; returns the failures ((size_t)-1 without
 * GMP) */
size_t orc_gmp_check_batch(const uint64_t *const *a, const size_t *na, const uint64_t *const *b, const size_t *nb,
                           const uint64_t *const *c, const size_t *nc, size_t n, uint8_t *ok);
/* a and b converted (t[0] s), one mpz_mul timed (t[1] s), c converted (t[2] s); 1 when c = a b, progress on stderr */
int orc_gmp_full_check(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *c, size_t nc, double *t);
/* 1 if c equals a b + x exactly (x optional), by GMP */
int orc_gmp_check(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *x, size_t nx,
                  const uint64_t *c, size_t nc);

#ifdef __cplusplus
}
#endif
#endif
