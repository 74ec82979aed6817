/* nttm.h - the C ABI of the MI355X multi-prime NTT product library (libnttm.so).
 *
 * Numbers are arrays of base-10^18 limbs (uint64_t, each < 10^18), least significant first, not necessarily
 * normalized -- the reference's convention (E/bigint.h:12, E/rns_mul.h:48-51).  The caller owns every array it passes.
 *
 * Memory: the products (nttm_mul, nttm_mul_add, nttm_mul_cut, nttm_mul_batch, nttm_mul_each) take a, b, x and c in
 * device memory of the library's device or in host memory -- pageable, pinned (hipHostMalloc) or managed -- as the
 * reference's rns_mul_batch does (E/rns_mul.h:42-47).  Host arrays are staged: copied into library-owned HBM through a
 * pinned bounce buffer before the product and copied back after it (E/rns_mul.c:807-817 stages the same way); device
 * arrays are used in place.  Every other entry point takes DEVICE memory (hipMalloc) for its limb and residue arrays
 * unless stated, and returns NTTM_EINVAL, before any launch, for host memory, memory of another device or a range
 * that runs past the end of its device allocation.
 *
 * Streams and synchronization: work runs on the given HIP stream (a hipStream_t passed as void *, NULL = the null
 * stream) on the device selected by nttm_init.  These entries return with the stream drained (they read flags or
 * lengths back): nttm_mul, nttm_mul_add, nttm_mul_cut, nttm_mul_batch, nttm_mul_each, nttm_mul_each_async,
 * nttm_bs_tree, nttm_e_fraction, nttm_e_write_packed, nttm_limbs_mod, nttm_crt, nttm_ntt_fwd_mfma,
 * nttm_batch_plan_create, nttm_batch_plan_ncn.  These only enqueue work and return: nttm_ntt_fwd, nttm_ntt_inv,
 * nttm_ntt_inv_pw, nttm_ntt_fwd3, nttm_ntt_inv3, nttm_ntt_inv3_pw, nttm_pw, nttm_load, nttm_pack60,
 * nttm_batch_plan_run.  (The reference's NTT calls are asynchronous on the caller's stream, E/ntt.h:1-2; its products
 * return complete, E/rns_mul.c:792-1028.)  Growing the library's workspace synchronizes the device.  One host thread
 * per library instance; calls are not reentrant.
 *
 * Errors: every entry returns 0 on success or a negative NTTM_E* code, and nttm_last_error() describes the calling
 * thread's last failure.  The reference instead ends the process through ec_fatal (E/fatal.h:19-23); its limits are kept as
 * errors here: transform length above 2^32 (E/ntt.c:517-520), a product above the prime set's term bound
 * (E/crt.c:91-99), limbs >= 10^18 are the caller's contract (not checked on the device).
 *
 * Reference interfaces each entry replaces (paths relative to the reference repository, E/ = ecalc/):
 *   nttm_init / nttm_shutdown     rns_init / rns_shutdown                         E/rns_mul.h:21,40
 *   nttm_mul, nttm_mul_add        rns_mul_dist_db on one device = rns_local_core   E/rns_mul.h:54,76; E/rns_mul.c:1057
 *   nttm_mul_batch                rns_mul_batch(rns_prod *, N)                     E/rns_mul.h:50-53; E/rns_mul.c:792
 *   nttm_mul_cut                  rns_mul_low_db / _high_db / _band_db, rns_mul_local  E/rns_mul.h:66-73
 *   nttm_prod                     rns_prod                                         E/rns_mul.h:50-51
 *   nttm_ntt_fwd / nttm_ntt_inv   ntt_fwd / ntt_inv                                E/ntt.h:45-46
 *   nttm_pw                       ntt_pw                                           E/ntt.h:47
 *   nttm_ntt_inv_pw               ntt_inv_pw_y (FULL / BCAST / PAIR)               E/ntt.h:52
 *   nttm_ntt_fwd3 / _inv3         ntt_fwd3 / ntt_inv3 (3 2^k lengths)              E/ntt.h:53-54; E/ntt.c:671-779
 *   nttm_ntt_inv3_pw              ntt_inv3_pw_y (FULL / BCAST / PAIR)              E/ntt.h:55-57
 *   nttm_root3                    ec_root3                                         E/modarith.h:206-214
 *   nttm_load                     ntt_load                                         E/ntt.h:53
 *   nttm_crt                      k_crt_batch + spill merge for one product        E/rns_int.h:50-52; E/rns_mul.c:336
 *   nttm_max_terms                ec_np3_max_terms / ec_np_check                   E/modarith.h:40-43; E/crt.c:28-44
 */
#ifndef NTTM_H
#define NTTM_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    NTTM_OK = 0,
    NTTM_EINVAL = -1,      /* bad argument (length, pointer, mode) */
    NTTM_ENOMEM = -2,      /* device allocation failed */
    NTTM_EHIP = -3,        /* a HIP runtime call failed */
    NTTM_EBOUND = -4,      /* the product exceeds what the prime set reconstructs, or the length exceeds 2^32 */
    NTTM_ESTATE = -5,      /* not initialised */
    NTTM_ECHECK = -6,      /* a post-condition failed (a carry beyond the top limb, or the residue check) */
    NTTM_EIO = -7          /* a file could not be written */
};

#define NTTM_MAX_PRIMES 4

/* y layouts of the fused inverse (E/ntt.h:31-35) */
enum { NTTM_Y_FULL = 0, NTTM_Y_BCAST = 1, NTTM_Y_PAIR = 2 };

/* transform lengths of the products (nttm_config.radix3): 2^k always, or 3 2^k where the product fits it */
enum { NTTM_R3_AUTO = 0,       /* 3 2^k where it fits and moves fewer bytes than the 2^k transform */
       NTTM_R3_NEVER = 1,      /* 2^k only */
       NTTM_R3_FIT = 2 };      /* 3 2^k wherever it fits (the reference's rule, E/rns_mul.c:270-277) */

typedef struct nttm_config {
    int nprimes;                /* 0: choose per product (3 while the bound allows, else 4); 3 or 4: fixed */
    size_t workspace_bytes;     /* 0: grow on demand; else the cap of the library's device workspace */
    int tile_log;               /* points per tile of the strided transform passes, log2: 12 or 13 (0: default) */
    int pts;                    /* transform points per thread: 8 or 16 (0: default) */
    int radix3;                 /* NTTM_R3_*: when products use 3 2^k lengths (0: NTTM_R3_AUTO) */
    int check;                  /* 1: every whole product (single, batch, each) is checked afterwards: c = a b + x modulo
                                 * the prime 4611686018427388039, NTTM_ECHECK on a mismatch (the reference's residue
                                 * identity of "detect and retry", E/dr.h:1-8); costs one read of a, b, x and c.  Not
                                 * applied by cut products, plan replays or nttm_mul_noalloc.  0: off (default) */
} nttm_config;

/* one product of a batch: c = a b (+ x).  c has na + nb limbs, or na + nb + 1 when x is given (nx <= na + nb),
 * whose top limb is the final carry.  ncn receives the normalized length of c (the reference's k_norm). */
typedef struct nttm_prod {
    const uint64_t *a; size_t na;
    const uint64_t *b; size_t nb;
    uint64_t *c;
    const uint64_t *x; size_t nx;
    size_t ncn;
} nttm_prod;

int  nttm_init(int device, const nttm_config *cfg);
void nttm_shutdown(void);
const char *nttm_last_error(void);

/* prime i of the set (0 <= i < NTTM_MAX_PRIMES), its 2-adic order, and the primitive 2^logn-th root the transforms
 * use (g^((p-1)/2^logn) for the smallest generator g) */
uint64_t nttm_prime(int i);
uint64_t nttm_root(int i, int logn);
/* largest min(na, nb) a product on np primes reconstructs exactly: floor((p_0 ... p_{np-1} - 1) / (10^18 - 1)^2) */
uint64_t nttm_max_terms(int np);

/* c[0 .. na+nb) = a b; device or host memory (see the top).  na or nb == 0 gives c = 0.  A product that no single transform holds
 * (more than 2^32 points, more terms than the prime set reconstructs, or more workspace than nttm_config caps) is
 * formed as a grid of piece products added into c (the reference's grid, E/rns_dist.c:610-700). */
int nttm_mul(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, void *stream);
/* c[0 .. na+nb) = a b with no allocation, for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  Device memory only; one transform
 * of at most 2^20 points on three primes (e.g. 2^19 x 2^19 limbs; larger products, four primes or a grid give
 * NTTM_EINVAL).  The planes and CRT scratch live in a 40 MiB file-scope device array, one per GPU, held until the process
 * ends; calls that overlap on one GPU share it and race.  Synchronizes the stream. */
int nttm_mul_noalloc(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, void *stream);
/* c[0 .. na+nb+1) = a b + x, nx <= na + nb; the top limb is the final carry */
int nttm_mul_add(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
                 const uint64_t *x, size_t nx, void *stream);
/* the cut product: c[0 .. min(na+nb, w)) = a b mod 10^(18 w) (w = SIZE_MAX: the whole product).  The operands are cut
 * to their low w limbs first.  The caller reads only limbs at or above lowcut (lowcut <= min(na+nb, w)).  A product
 * that fits one transform is formed whole, so every limb written is exact; a product formed as a grid of pieces
 * (see nttm_mul) leaves out the pieces ending at or below lowcut, so its limbs from lowcut on may fall short of the
 * exact product's by what those pieces carry into them (at most one unit of 10^(18 lowcut) per piece left out).
 * No carry beyond limb w-1 is checked. */
int nttm_mul_cut(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, size_t lowcut, size_t w,
                 void *stream);
/* the residue post-condition on demand: NTTM_ECHECK unless c (nc limbs) = a b + x modulo 4611686018427388039 (the
 * first prime above 2^62; x may be null); device memory; synchronizes.  No counterpart in the reference's API (its
 * checks run inside the product, E/dr.h:1-8). */
int nttm_check_product(const uint64_t *c, size_t nc, const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
                       const uint64_t *x, size_t nx, void *stream);
/* N independent products; P is HOST memory; its a, b, x, c may each be device or host memory (host arrays staged, see
 * the top; a batch with more than 1 GiB of host arrays runs as consecutive sub-batches).  Products are grouped by
 * transform length and run batched; ncn is filled on return (the call synchronizes the stream). */
int nttm_mul_batch(nttm_prod *P, size_t N, void *stream);
/* a node of a binary-splitting tree: P and Q as device limb arrays */
typedef struct nttm_bs_node {
    const uint64_t *p; size_t pn;
    const uint64_t *q; size_t qn;
} nttm_bs_node;
/* the product tree of binary splitting (E/binsplit.c:1583-1700): the n leaves are combined pairwise, level by level,
 * into (P_L Q_R + P_R, Q_L Q_R) -- each level one nttm_mul_batch call with the fused + x and Q_R shared by the pair's
 * two products; an odd last node moves up unchanged -- until one node remains.  Its P and Q are written to p and q
 * (capacities p_cap and q_cap limbs, NTTM_EINVAL when too small) with their normalized lengths in *pn and *qn.  The
 * library holds the inner levels in its own device memory.  Every pair needs |P_R| <= |P_L| + |Q_R| limbs (the product's
 * x bound), which holds when every leaf has P < Q, as in a convergent series: then every node does.  Synchronizes the
 * stream. */
int nttm_bs_tree(const nttm_bs_node *leaves, size_t n, uint64_t *p, size_t p_cap, size_t *pn, uint64_t *q, size_t q_cap,
                 size_t *qn, void *stream);
/* e's fraction to m limbs: out[0 .. m) = floor((e - 2) 10^(18 m)), least significant limb first, HOST memory (the
 * limbs read from the top are e's decimal digits after "2.", 18 per limb; NTTM_EINVAL for device memory, before any
 * work).  The reference's pipeline on one GPU
 * (E/ALGORITHM.md:12-20,59-94): binary splitting of sum 1/k! (leaves of up to 64 terms folded on the device, then
 * nttm_bs_tree), a Newton reciprocal of Q with doubling precision and a corrected quotient, every product on the
 * device.  N terms with N! > 10^(18 (m + 2)) must stay below 10^9 - 2 (m up to about 4.7 10^8 limbs, 8.5 10^9 digits).  Synchronous. */
int nttm_e_fraction(uint64_t *out, size_t m, void *stream);
/* wall time of the phases of the last nttm_e_fraction (ms): leaf folding, the product tree (with the copies of P and
 * Q to the host), the reciprocal, the quotient (the reference's run prints its phase split, E/ecalc.c:237) */
int nttm_e_last_phase_ms(double *seed_ms, double *tree_ms, double *recip_ms, double *quot_ms);

/* ---- packed digit output (the reference's ECPACK18 version-2 file, E/packed_fmt.h:1-20,50-55) ---- */
/* bytes of n limbs in the 60-bit form: 15 (n / 2) + 8 (n & 1)  (E/packed_fmt.h:56 ecp60_bytes) */
uint64_t nttm_pack60_bytes(size_t n);
/* the limbs [lo, hi) of x (device) in file order, most significant first, as the 60-bit little-endian bit stream
 * (E/packed_fmt.h:58-78 ecp60_put_pair / ecp60_pack); out: device, 4-byte aligned, nttm_pack60_bytes(hi - lo) bytes.
 * Asynchronous on the stream. */
int nttm_pack60(uint8_t *out, const uint64_t *x, size_t lo, size_t hi, void *stream);
/* res[j] = (sum_k x[k] 10^(18 k)) mod q[j] for nq <= 8 odd moduli in (10^18, 2^63); x on the device, res on the host
 * (E/verify.c:30 vf_limbs_mod, the residues in the packed header).  Synchronous. */
int nttm_limbs_mod(uint64_t *res, const uint64_t *x, size_t n, const uint64_t *q, int nq, void *stream);
/* writes the single-part packed file of e's digits (E/mn_out.c:299-313 and E/packed_fmt.h:32-48 for the header):
 * X = 2 10^(18 m) + frac (frac: nttm_e_fraction's m limbs, host memory), nl = m + 1 limbs, pad 17, d = 18 m digits,
 * d_out of them in the ASCII form ("2." + digits 1..d_out + "\n"), the residues of X mod the eight primes above 2^62
 * (E/verify.c:18).  Synchronous. */
int nttm_e_write_packed(const char *path, const uint64_t *frac, size_t m, uint64_t d_out, void *stream);
/* a batch fixed once and run many times (the binary-splitting leaves' repeated rns_mul_batch calls, E/rns_mul.c:792):
 * the products are grouped by transform length; with autotune != 0 each group's schedule is chosen by timing the
 * candidates on these operands -- per length class, 3 2^k points for the products that fit them or 2^(k+2); per group,
 * a strided-pass tile of 2^12 or 2^13 points and 8 or 16 points per thread -- else the library's defaults are taken.
 * The choices are remembered until nttm_shutdown, per prime count, transform length and group size; autotune == 2
 * takes the remembered choice of a group of the same size, else of the nearest size in the same power-of-two class,
 * without timing it (autotune == 1 times every group and refreshes what is remembered).
 * Every descriptor is uploaded once and all launches are captured into one hipGraph.  P is HOST memory holding device
 * pointers; the operands and outputs it names must stay allocated until the plan is destroyed (their contents may
 * change between runs).  Creating an autotuned plan runs the products several times.  Destroy plans before
 * nttm_shutdown.  No counterpart in the reference (it plans per call). */
typedef struct nttm_batch_plan nttm_batch_plan;
int nttm_batch_plan_create(const nttm_prod *P, size_t N, int autotune, nttm_batch_plan **plan, void *stream);
/* all N products, one graph launch on the stream (asynchronous).  NTTM_ESTATE for a plan made under an earlier
 * nttm_init (the library's tables it captured are gone); device memory only (a plan refuses host operands). */
int nttm_batch_plan_run(nttm_batch_plan *plan, void *stream);
/* after a run: the normalized lengths (ncn[i] for product i; HOST memory, N entries); synchronizes the stream and
 * checks the CRT's carry flags */
int nttm_batch_plan_ncn(nttm_batch_plan *plan, size_t *ncn, void *stream);
/* the chosen schedule as text, one line per group */
int nttm_batch_plan_info(const nttm_batch_plan *plan, char *buf, size_t cap);
void nttm_batch_plan_destroy(nttm_batch_plan *plan);
/* the same N products, one product call each (the unbatched baseline of the batch tier): every call synchronizes and
 * the normalized lengths are read back limb by limb.  No counterpart in the reference. */
int nttm_mul_each(nttm_prod *P, size_t N, void *stream);
/* the same N products, one product call each, launched back to back with no synchronization between them: one at the
 * end for the CRT flags, and the normalized lengths by one kernel over all N.  Device memory only.  No counterpart in
 * the reference (the one-launch-per-product baseline without host round trips). */
int nttm_mul_each_async(nttm_prod *P, size_t N, void *stream);

/* ---- transform primitives on canonical residues in [0, p) (uint64_t), batch transforms of 2^logn contiguous ---- */
/* natural order in -> bit-reversed out */
int nttm_ntt_fwd(int prime, uint64_t *x, int logn, size_t batch, void *stream);
/* the same transform (identical output) with its first four levels as one split-precision radix-16 stage on INT8
 * matrix cores: residues split into 7-bit digits, the 16-point DFT as 15 shifted digit products of exact INT32 sums
 * (v_mfma_i32_16x16x64_i8), recombined mod p (the north star's split-precision stage; no counterpart in the
 * reference).  logn >= 8.  It is slower than the FP64 stage it replaces, so no product uses it by default.
 * Synchronous. */
int nttm_ntt_fwd_mfma(int prime, uint64_t *x, int logn, size_t batch, void *stream);
/* bit-reversed in -> natural out, times n^-1 */
int nttm_ntt_inv(int prime, uint64_t *x, int logn, size_t batch, void *stream);
/* x[i] = x[i] y[i] mod p */
int nttm_pw(int prime, uint64_t *x, const uint64_t *y, size_t count, void *stream);
/* the pointwise product with y laid out per ymode, then the inverse: FULL y[t][k]; BCAST y[k]; PAIR y[t/2][k] */
int nttm_ntt_inv_pw(int prime, uint64_t *x, const uint64_t *y, int ymode, int logn, size_t batch, void *stream);
/* ---- transforms of 3 2^logk points (1 <= logk <= 32), batch transforms of 3 2^logk contiguous residues ----
 * forward: natural order in; third r (points [r 2^logk, (r+1) 2^logk)), position brev_logk(t), holds X[3t + r] for
 * the root nttm_root3(prime, logk) of order 3 2^logk.  Pointwise products need both operands in this order only. */
int nttm_ntt_fwd3(int prime, uint64_t *x, int logk, size_t batch, void *stream);
/* that order in -> natural out, times (3 2^logk)^-1 */
int nttm_ntt_inv3(int prime, uint64_t *x, int logk, size_t batch, void *stream);
/* the pointwise product with y (whole transforms of 3 2^logk points laid out per ymode), then the inverse */
int nttm_ntt_inv3_pw(int prime, uint64_t *x, const uint64_t *y, int ymode, int logk, size_t batch, void *stream);
/* the primitive 3 2^logk-th root: g^((p-1) / (3 2^logk)) for the smallest generator g */
uint64_t nttm_root3(int i, int logk);

/* dst[0 .. npoints) = src limbs mod p, zero-extended (nlimbs <= npoints) */
int nttm_load(int prime, uint64_t *dst, const uint64_t *src, size_t nlimbs, size_t npoints, void *stream);
/* limbs of sum_k CRT(planes[0..np)[k]) 10^(18 k) + x: planes hold ncoef canonical residues each (ncoef <= nc); c gets
 * nc limbs.  Returns NTTM_ECHECK when the value needs more than nc limbs. */
int nttm_crt(uint64_t *c, size_t nc, const uint64_t *const *planes, int np, size_t ncoef,
             const uint64_t *x, size_t nx, void *stream);

/* device-side time of the phases of the last nttm_mul / nttm_mul_add (ms): the forward transforms of b (limb load
 * fused), those of a with the fused pointwise product and the inverse transforms, the CRT */
int nttm_last_phase_ms(double *b_fwd_ms, double *a_fwd_pw_inv_ms, double *crt_ms);

/* ---- live kernel timing (no reference counterpart; the reference times phases with events, E/rns_mul.c:548-587) ---- */
typedef struct nttm_kstat {
    char name[40];              /* kernel family */
    uint64_t launches;
    double total_ms;            /* summed device time between events recorded around each launch */
    double alg_bytes;           /* summed algorithmic HBM bytes of those launches (the library's byte model) */
} nttm_kstat;
/* 1: from now on bracket every kernel launch with HIP events on its stream (clears earlier records); 0: stop */
int nttm_profile(int enable);
/* synchronizes the device, aggregates the records per kernel family into out[0 .. cap); returns the family count */
int nttm_profile_read(nttm_kstat *out, int cap);

/* ---- host-only introspection (no device calls; the reference has no counterpart) ---- */
/* the pass plan of a 2^logn transform batch: 4 ints (lo, m, logw, logh) per pass into out[0 .. cap); returns passes.
 * tile_log / last_log: log2 of points per tile of the strided passes / of the last (contiguous) pass (0: defaults;
 * the product path uses the configured tile_log and 12) */
int nttm_debug_plan(int logn, uint64_t batch, int tile_log, int last_log, int *out, int cap);
/* the pass plan a product transform of 2^logn points (batch of them) takes: 4 ints (lo, m, logw, logh) per pass, tile_log
 * the configured strided tile (0: default); the last pass may take fewer than 12 bits in 4096-point tiles */
int nttm_debug_product_plan(int logn, uint64_t batch, int tile_log, int *out, int cap);
/* the transform length the products take for ncoef = na + nb - 1 coefficients under a radix3 rule (NTTM_R3_*) and
 * tile_log (0: default): 2^logn (r3 = 0) or 3 2^logn (r3 = 1) points */
int nttm_debug_product_len(size_t ncoef, int radix3, int tile_log, int *logn, int *r3);
/* the e pipeline's device addition / subtraction (k_addsub with the CRT's carry scan): r[0 .. nout) = a 10^(18 sa)
 * +- b 10^(18 sb) mod 10^(18 nout), device pointers; *ge = the carry out of limb nout - 1 (for a subtraction: 1 when
 * the difference is >= 0).  Synchronous. */
int nttm_debug_addsub(uint64_t *r, size_t nout, const uint64_t *a, size_t na, size_t sa, const uint64_t *b, size_t nb,
                      size_t sb, int sub, int *ge, void *stream);
/* synthetic operand limbs on the device: x[j] = splitmix64 value j + 1 of the stream seeded `seed`, mod 10^18, a zero
 * top limb set to 1 (the reference's generator, E/tests/harness.h:55-81).  Asynchronous. */
int nttm_debug_gen_limbs(uint64_t *x, size_t n, uint64_t seed, void *stream);
/* the global element indices of tile `tile` of a pass of that geometry; returns the tile's element count */
int nttm_debug_tile_indices(int lo, int m, int logw, int logh, uint64_t tile, uint64_t *out);
/* the normalized length of x[0 .. n) (one more than the index of its top non-zero limb, 0 if none) into the device word
 * *len_dev, by the scan the e pipeline and the planned tree normalize with.  Asynchronous. */
int nttm_debug_top_nonzero(const uint64_t *x, size_t n, uint64_t *len_dev, void *stream);
/* the cyclic product behind the e pipeline's wrapped Newton product: v[0 .. L + 3) = sum_k c_k 10^(18 k) for
 * c_k = sum over i + j = k mod L of a_i b_j (a b mod (x^L - 1) at x = 10^18, carries not wrapped), L = 2^k or 3 2^k
 * (k >= 4), na, nb <= L, device pointers.  Synchronous. */
int nttm_debug_mul_cyclic(uint64_t *v, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, size_t L,
                          void *stream);

#ifdef __cplusplus
}
#endif
#endif
