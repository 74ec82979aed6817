/* arith.h - exact modular arithmetic of the product library, host + device.
 *
 * Residues are FP64 values that hold integers exactly.  Primes are p = c 2^k + 1 < 2^50 (k >= 32), so every live value
 * lies in [-p, p] (|v| < 2^50) and sums of two stay below 2^52: FP64 add/sub/FMA on them are exact, and the full-rate
 * FP64 pipe of CDNA4 does the modular products.  Bounds below use u = 2^-53 and p < 2^50.
 *
 *  rint_prod(a, b)       a b rounded once to the nearest integer, for |a b| < 2^51: fma(a, b, C) - C with C = 1.5 2^52
 *                        (the FMA result lies in [2^52, 2^53), where the FP64 spacing is 1; the subtraction is exact).
 *                        One FMA and one add instead of a product and a rint, and a single rounding.
 *  mul_shoup(a, w, wq)   Shoup's product with a precomputed quotient factor wq = fl(w / p) (V. Shoup; D. Harvey,
 *                        "Faster arithmetic for number-theoretic transforms", J. Symb. Comput. 60, 2014), in FP64:
 *                        hi + lo = a w exactly (FMA), q = rint_prod(a, wq), r = (hi - q p) + lo.  For |a| <= 4p,
 *                        |w| <= p/2: |a w / p| <= 2p < 2^51 and |a wq - a w / p| <= 2^51 u = 1/4, so
 *                        |q - a w / p| <= 3/4 and the integer r has |r| <= 3p/4 (|r| <= 5p/8 for |a| <= 2p).
 *                        hi - q p = r - lo is an integer below 2^53, so the FMA and the final add are exact.
 *  mul_barrett(a, b)     both operands variable, |a|, |b| <= p: q = rint_prod(fl(a b), fl(1/p)); |fl(a b) fl(1/p) - a b / p|
 *                        < p 2u (1+u) < 1/4, so |r| <= 3p/4.
 *  reduce(s)             |s| <= 4p -> s - p rint_prod(s, 1/p), |result| <= p/2 + 1.
 *  The reference computes its modmul in FP64 as well, with 52-bit primes and a Barrett quotient of the rounded product
 *  (E/modarith.h:118-130); here twiddle products use Shoup's precomputed factor and every value stays balanced.
 */
#ifndef NTTM_ARITH_H
#define NTTM_ARITH_H
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define NTTM_HD __host__ __device__ __forceinline__
#else
#define NTTM_HD static inline
#endif

#ifndef NTTM_MAX_PRIMES
#define NTTM_MAX_PRIMES 4
#endif
#define NTTM_B 1000000000000000000ULL

/* the prime set: c 2^k + 1 < 2^50 with 2^32 | p - 1 and 3 | (p - 1) / 2^32 (3 2^k lengths possible); the first three
 * have the largest product below 2^150 among such primes with 3 | c */
static const uint64_t NTTM_PRIMES[NTTM_MAX_PRIMES] = {
    1125844072267777ULL,   /* 262131 * 2^32 + 1 */
    1125818302464001ULL,   /* 262125 * 2^32 + 1 */
    1125625028935681ULL,   /*   4095 * 2^38 + 1 */
    1125122517762049ULL,   /* 261963 * 2^32 + 1 */
};
#define NTTM_LOGN_MAX 32   /* the smallest 2-adic order in the set */

NTTM_HD double nttm_fma(double a, double b, double c) { return fma(a, b, c); }
#define NTTM_RND 6755399441055744.0          /* 1.5 2^52 */
NTTM_HD double rint_prod(double a, double b) { return nttm_fma(a, b, NTTM_RND) - NTTM_RND; }

NTTM_HD double mul_shoup(double a, double w, double wq, double p)
{
    double hi = a * w;
    double lo = nttm_fma(a, w, -hi);
    double q = rint_prod(a, wq);
    return nttm_fma(-q, p, hi) + lo;
}
NTTM_HD double mul_barrett(double a, double b, double p, double pinv)
{
    double hi = a * b;
    double lo = nttm_fma(a, b, -hi);
    double q = rint_prod(hi, pinv);
    return nttm_fma(-q, p, hi) + lo;
}
NTTM_HD double reduce(double s, double p, double pinv) { return nttm_fma(-rint_prod(s, pinv), p, s); }

/* a base-10^18 limb (< 2^60) to a residue in [-p/2 - 1, p/2 + 1] */
NTTM_HD double limb_to_res(uint64_t x, double p, double pinv)
{
    double h = (double)(uint32_t)(x >> 32) * 4294967296.0;     /* exact: < 2^60 with 28 significant bits */
    double l = (double)(uint32_t)x;
    double q = rint_prod(h + l, pinv);
    return nttm_fma(-q, p, h) + l;                               /* h - q p exact (|.| < p + 2^32), then + l exact */
}
/* [-p, p] -> canonical [0, p) */
NTTM_HD uint64_t res_canon(double v, double p, double pinv)
{
    double r = reduce(v, p, pinv);
    r = r < 0.0 ? r + p : r;
    return (uint64_t)r;
}
/* canonical (any value < 2^53) to FP64 */
NTTM_HD double u64_to_res(uint64_t x) { return (double)x; }

/* ---- 128-bit helpers for the CRT (integer) ---------------------------------------------------------------------- */
NTTM_HD uint64_t umulhi64(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

/* ---- the CRT's digit split: products modulo B = 10^18 -------------------------------------------------------------
 * mulmod_b(t, M, Mp): t M = a B + b with 0 <= b < B, for t < 2^64 and M < B, Mp = floor(M 2^64 / B) (Shoup's
 * precomputed quotient factor with B as the modulus): q = floor(t Mp / 2^64) satisfies q <= floor(t M / B) <= q + 1,
 * since t M / B - t Mp / 2^64 = t (M 2^64 - Mp B) / (B 2^64) < t / 2^64 <= 1; b = t M - q B is below 2B < 2^64, so
 * it is exact modulo 2^64, and one correction finishes. */
NTTM_HD void mulmod_b(uint64_t t, uint64_t M, uint64_t Mp, uint64_t *a, uint64_t *b)
{
    uint64_t q = umulhi64(t, Mp);
    uint64_t r = t * M - q * NTTM_B;
    const int c = r >= NTTM_B;
    *a = q + (uint64_t)c;
    *b = c ? r - NTTM_B : r;
}
/* an integer-valued double in [0, 2^52) to its integer: the low 52 bits of 2^52 + v */
NTTM_HD uint64_t f64_to_u52(double v)
{
    union { double d; uint64_t u; } x;
    x.d = v + 4503599627370496.0;
    return x.u & ((1ULL << 52) - 1);
}

/* division by B = 10^18 with a precomputed inverse of the normalized divisor D = 16 B (2^63 <= D < 2^64):
 * N. Moller and T. Granlund, "Improved division by invariant integers", IEEE Trans. Computers 60(2), 2011, Alg. 4.
 * v = floor((2^128 - 1) / D) - 2^64.  Requires hi < B; q = floor((hi 2^64 + lo) / B), r the remainder. */
#define NTTM_DIVB_D  (16ULL * NTTM_B)
#define NTTM_DIVB_V  0x2725dd1d243aba0eULL   /* floor((2^128 - 1) / (16 * 10^18)) - 2^64 (checked by the tests) */
NTTM_HD void div_b(uint64_t hi, uint64_t lo, uint64_t *q, uint64_t *r)
{
    const uint64_t d = NTTM_DIVB_D, v = NTTM_DIVB_V;
    uint64_t u1 = (hi << 4) | (lo >> 60), u0 = lo << 4;           /* N 16, u1 < d */
    uint64_t q0 = v * u1, q1 = umulhi64(v, u1);
    uint64_t s0 = q0 + u0;
    q1 = q1 + u1 + (s0 < q0);
    q0 = s0;
    q1 = q1 + 1;
    uint64_t rr = u0 - q1 * d;
    if (rr > q0) { q1 -= 1; rr += d; }
    if (rr >= d) { q1 += 1; rr -= d; }
    *q = q1; *r = rr >> 4;
}

#endif
