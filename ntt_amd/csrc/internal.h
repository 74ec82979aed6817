/* internal.h - shared declarations of the product library's translation units (not part of the C ABI). */
#ifndef NTTM_INTERNAL_H
#define NTTM_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include "arith.h"
#include "../../include/nttm.h"

#define NTTM_TILE_LOG 13            /* points per transform tile: 2^13 FP64 values = 64 KiB of LDS */
#define NTTM_STAGE_LOG 14           /* stage-major twiddle table: T[2^s + j] = w_{2^(s+1)}^j for s < 14 */
#define NTTM_BQ_LOG 8               /* boundary twiddles w_{2^32}^E from four tables of 2^8 entries, one per byte of E */
#define NTTM_THREADS 256

struct PrimeK {                     /* per-prime constants, passed by value */
    double p, pinv;
    uint64_t pu;
};

struct TwTables {                   /* device pointers, one set per prime and direction */
    const double2 *stage;           /* (w, wq) balanced, 2^NTTM_STAGE_LOG entries */
    const double2 *bq;              /* four-level boundary table: bq[k 2^8 + x] = w32^(x 2^(8k)), k < 4 (16 KiB) */
    const double2 *b3;              /* the same for the root W3 of order 3 2^32: b3[k 2^8 + x] = W3^(x 2^(8k)); entry
                                     * 4 2^8 is the FORWARD cube root of unity w3 = W3^(2^32) in both directions */
};

/* operand limbs of the products of a batch (device array, one per transform) */
struct LoadDesc {
    const uint64_t *a; uint64_t na;
    const uint64_t *b; uint64_t nb;
};

/* one pass over bits [lo, lo + m) of every transform of 2^logn points in a batch; a tile holds W low columns x 2^m
 * rows x H high blocks (log2 of each below) */
struct PassArgs {
    double *x;
    const double *y;                /* pointwise operand (inverse first pass), else null */
    int logn, lo, m, logw, logh;
    int ymode;
    int r3y;                        /* y of a radix-3 length: the transforms of the batch are thirds (3 per y item) */
    uint64_t batch;
    PrimeK pk;
    TwTables tw;                    /* forward or inverse tables of the pass's direction */
    TwTables twi;                   /* inverse tables (the fused forward + inverse pass) */
    double2 scale;                  /* (w, wq) of a final multiply (n^-1), used when do_scale */
    int do_scale;
    /* input from limbs instead of the plane (the first forward pass): transform b reads limbs[b][0 .. nl) */
    const uint64_t *limbs; uint64_t nlimbs;     /* batch of one */
    const LoadDesc *ldesc; int which;           /* batches: operand a (which = 0) or b of ldesc[b] */
    double2 in_scale; int do_in_scale;          /* multiply loaded limbs by in_scale (n^-1 folded into one operand) */
    /* B of the double-fused pass shared between products: transform t reads plane transform bmap[t] (null: t); only
     * for tiles inside one transform (m + logh <= logn) */
    const uint32_t *bmap;
    /* Deferred boundary twiddles between a transform's first two strided forward passes (both streamed):
     * dfr_out = 1 on the first pass stores its outputs without its boundary twiddles (and without in_scale);
     * dfr_in = 1 on the second multiplies its loads by them: element (h, c1) of the first pass (h its stored row,
     * c1 < 2^(dfr_lom - dfr_gm) its column) by w_(2^dfr_lom)^(brev_dfr_gm(h) c1), times dfr_scale when dfr_do_scale */
    int dfr_out, dfr_in, dfr_lom, dfr_gm, dfr_do_scale;
    double2 dfr_scale;
};

struct Ctx;
/* live kernel timing: when profiling is on, launches are bracketed by events (prof_begin before, prof_end after) */
void prof_begin(const char *name, double alg_bytes, hipStream_t s);
void prof_end(hipStream_t s);
/* host helpers shared between files */
int nttm_fail(int code, const char *fmt, ...);
int nttm_hip_check(hipError_t e, const char *what);
/* NTTM_EINVAL unless [p, p + bytes) is device memory of the library's device (null or empty ranges pass) */
int nttm_require_device(const void *p, size_t bytes, const char *what);
int nttm_require_host(const void *p, size_t bytes, const char *what);
#define HIPCK(call) do { int rc_ = nttm_hip_check((call), #call); if (rc_) return rc_; } while (0)
#define RCK(call) do { int rc_ = (call); if (rc_) return rc_; } while (0)

/* the CRT (crt.hip): planes hold np FP64 residue planes (values in [-p, p]); writes nc limbs of c */
struct CrtConst {
    PrimeK pk[NTTM_MAX_PRIMES];
    double cinv[NTTM_MAX_PRIMES];               /* (p_0 ... p_{i-1})^-1 mod p_i, balanced */
    double mmod[NTTM_MAX_PRIMES][NTTM_MAX_PRIMES];   /* (p_0 ... p_{j-1}) mod p_i, balanced (j < i) */
    uint64_t mprod[NTTM_MAX_PRIMES][NTTM_MAX_PRIMES];/* p_0 ... p_{i-1} as little-endian words */
    uint64_t mb[NTTM_MAX_PRIMES][NTTM_MAX_PRIMES];   /* p_0 ... p_{i-1} in base 10^18, digit j < i */
    uint64_t mbp[NTTM_MAX_PRIMES][NTTM_MAX_PRIMES];  /* floor(mb 2^64 / 10^18) (mulmod_b) */
};
#endif
