/* api.hip - the C ABI of libnttm.so: context and tables, the product schedule (one large product, batches of leaf
 * products) and the transform primitives.  See include/nttm.h for the contract and the reference entry points each
 * function replaces. */
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <stdlib.h>
#include <vector>
#include <map>
#include <array>
#include <algorithm>
#include <chrono>
#include <thread>
#include "internal.h"
#include "plan.h"

int nttm_run_pass(bool inv, const PassArgs &a, const PassGeom &g, hipStream_t s);
int nttm_run_fused(const PassArgs &a, const PassGeom &g, hipStream_t s);
int nttm_run_fused2(const PassArgs &a, const PassArgs &bq, const PassGeom &g, hipStream_t s);
bool nttm_fwd_pass_streams(const PassArgs &a, const PassGeom &g);
void nttm_set_pts(int pts);
int nttm_get_pts(void);
void nttm_set_fixed_shapes(int on);
void nttm_set_stream(int on, int ncu);
#define NTTM_FUSED_LOG 12          /* tile of the fused forward + pointwise + inverse pass (k_fused) */
#define NTTM_TILE_LOG_DEFAULT 13   /* tile of the strided passes: 2^13 points (64 KiB of LDS), runs of 32 columns */
#define NTTM_PRODUCT_LOGN_MIN 4    /* products use at least 2^4 points, so every fused tile has >= 16 */
#define NTTM_STAGE_CHUNK ((size_t)1 << 30)   /* staged bytes per sub-batch of a batch with host operands (default) */
struct LoadArgs;
int nttm_run_u64_to_f64(double *x, uint64_t count, hipStream_t s);
int nttm_run_cvt_u64_f64(double *dst, const uint64_t *src, uint64_t count, hipStream_t s);
int nttm_run_f64_to_canon(double *x, uint64_t count, const PrimeK &pk, hipStream_t s);
int nttm_run_pw_u64(uint64_t *x, const uint64_t *y, uint64_t count, const PrimeK &pk, hipStream_t s);
int nttm_run_load_planes(double *dst, uint64_t plane, const uint64_t *src, uint64_t nlimbs, uint64_t npoints, int np,
                         const PrimeK *pk, const double2 *scale, int out_u64, hipStream_t s);
int nttm_run_load_batch(double *dst, uint64_t plane, const void *desc, int which, uint64_t count, int logn, int np,
                        const PrimeK *pk, const double2 *scale, hipStream_t s);
int nttm_crt_launch_batch(const void *desc_dev, uint64_t nprod, uint64_t nblocks, int np, const CrtConst &cc,
                          uint32_t *scratch, hipStream_t s, double alg_bytes, int per = 8);
uint64_t nttm_crt_blocks(uint64_t nc, int per = 8);      /* per: CRT limbs per thread, 8 (2048-limb blocks) or 1 (256) */
uint64_t nttm_crt_scratch_words(uint64_t nblocks);
int nttm_run_norm(const void *desc_dev, uint64_t nprod, uint64_t *ncn_dev, hipStream_t s);
int nttm_run_fill_x(const void *desc_dev, uint64_t nprod, hipStream_t s);
int nttm_run_dft16_mfma(const uint64_t *x, double *y, int logn, uint64_t batch, const void *etab, const double *tlo,
                        const double *thi, const double *consts, double p, hipStream_t s);
int nttm_run_add_carry(uint64_t *c, uint64_t n, const uint64_t *carry, hipStream_t s);
int nttm_run_r3(bool inv, double *x, uint64_t batch, int logk, const PrimeK &pk, const double2 *b3,
                const uint64_t *limbs, uint64_t nlimbs, const LoadDesc *ldesc, int which, const double2 *scale,
                hipStream_t s);

/* ---- errors ----------------------------------------------------------------------------------------------------- */
/* per host thread: the library's own helper threads (the e pipeline's download preparation) may fail too */
static thread_local char g_err[512] = "";
int nttm_fail(int code, const char *fmt, ...)
{
    va_list ap; va_start(ap, fmt); vsnprintf(g_err, sizeof g_err, fmt, ap); va_end(ap);
    return code;
}
int nttm_hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    return nttm_fail(NTTM_EHIP, "%s: %s", what, hipGetErrorString(e));
}
extern "C" const char *nttm_last_error(void) { return g_err; }

/* ---- host modular arithmetic (exact, __int128) -------------------------------------------------------------------- */
typedef unsigned __int128 u128;
static uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)((u128)a * b % p); }
static uint64_t h_powmod(uint64_t a, uint64_t e, uint64_t p)
{
    uint64_t r = 1 % p; a %= p;
    while (e) { if (e & 1) r = h_mulmod(r, a, p); a = h_mulmod(a, a, p); e >>= 1; }
    return r;
}
static uint64_t h_inv(uint64_t a, uint64_t p) { return h_powmod(a, p - 2, p); }
static uint64_t h_generator(uint64_t p)
{
    uint64_t f[64]; int nf = 0, t2 = 0; uint64_t t = p - 1;
    while (!(t & 1)) { t >>= 1; t2 = 1; }
    if (t2) f[nf++] = 2;
    for (uint64_t q = 3; q * q <= t; q += 2) if (t % q == 0) { f[nf++] = q; while (t % q == 0) t /= q; }
    if (t > 1) f[nf++] = t;
    for (uint64_t g = 2; g < p; g++) {
        int ok = 1;
        for (int i = 0; i < nf && ok; i++) if (h_powmod(g, (p - 1) / f[i], p) == 1) ok = 0;
        if (ok) return g;
    }
    return 0;
}
static double h_bal(uint64_t v, uint64_t p) { return v > p / 2 ? -(double)(p - v) : (double)v; }
static double2 h_shoup(uint64_t v, uint64_t p) { double w = h_bal(v, p); return make_double2(w, w / (double)p); }

/* ---- the context ------------------------------------------------------------------------------------------------------ */
struct Ctx {
    int init, dev;
    nttm_config cfg;
    PrimeK pk[NTTM_MAX_PRIMES];
    uint64_t root32[NTTM_MAX_PRIMES];        /* root of order 2^32 */
    uint64_t root3x32[NTTM_MAX_PRIMES];      /* root of order 3 2^32 */
    TwTables twf[NTTM_MAX_PRIMES], twi[NTTM_MAX_PRIMES];
    double2 *tables;
    CrtConst cc;
    char *ws; size_t ws_bytes;               /* device workspace */
    uint64_t *tmp; size_t tmp_limbs;         /* a piece product of a grid (above one transform) */
    uint64_t *bsp[2]; size_t bsp_limbs[2];   /* the inner levels of a binary-splitting tree (even / odd levels) */
    char *tdesc; size_t tdesc_bytes;         /* the planned tree's descriptors, every group of every level */
    char *tpin; size_t tpin_bytes;           /* their pinned host image, and the groups' flags */
    hipStream_t tcs;                         /* the tree's descriptor copies (ahead of the launches on the caller's stream) */
    void *dfree[8]; int ndfree;              /* buffers the planned tree outgrew, freed after its final synchronization */
    void *hfree[4]; int nhfree;
    hipEvent_t tev[8];                       /* the last copies on tcs, in turn */
    hipStream_t ps[NTTM_MAX_PRIMES - 1];     /* primes 1.. of a small product, beside the caller's stream (prime 0) */
    hipEvent_t pfev, pjev[NTTM_MAX_PRIMES - 1];   /* their fork from and join to the caller's stream */
    uint64_t *hpin; size_t hpin_bytes;       /* pinned host staging (descriptors, flags) */
    char *stage; size_t stage_bytes;         /* device copies of host-resident operands and results */
    size_t stage_chunk;                      /* staged bytes per sub-batch of a batch */
    char *chk; size_t chk_bytes;             /* device scratch of the residue post-condition */
    char *bounce; size_t bounce_half;        /* pinned bounce buffer of the host copies: two halves */
    char *dl; hipEvent_t dlev[2];            /* pinned halves of the large downloads (nttm_copy_to_host) */
    hipEvent_t bev[2]; int bev_live[2];      /* the last copy through each half */
    uint64_t gen;                            /* which nttm_init this context is (plans record it) */
    hipEvent_t ev[4];
    hipEvent_t pev[2 * NTTM_MAX_PRIMES];      /* around B's transform of each prime (phase times of nttm_mul) */
    double last_ms[3];
};
static Ctx g;
static uint64_t g_init_count;
static int g_prof_on;                        /* live kernel timing on (nttm_profile) */
static int g_no_defer;                       /* NTTM_NO_DEFER=1: every pass applies its own boundary twiddles */
static int g_tree_levelwise;                 /* NTTM_TREE_LEVELWISE=1: the tree level by level (for comparison) */
static int g_fused_bits_env;                 /* NTTM_FUSED_BITS=10|11|12: that last-pass width for every product
                                              * transform above it (for comparison measurements) */

/* the workspace of nttm_mul_noalloc: a file-scope device array, one copy per GPU, held until the process ends.  Host
 * code never uses its name as an address (on the host it is a stand-in, not the device address): nttm_init looks up
 * the address on its device with hipGetSymbolAddress and keeps it in g_scratch_dev[device]. */
#define NTTM_NOALLOC_BYTES ((size_t)40 << 20)
#define NTTM_NOALLOC_MAX_DEV 64
static __device__ __align__(256) char scratch[NTTM_NOALLOC_BYTES];
static char *g_scratch_dev[NTTM_NOALLOC_MAX_DEV];                /* nttm_init calls so far: a plan made under an earlier one is stale */

static uint64_t max_terms_np(int np)
{
    /* floor((prod p - 1) / (B - 1)^2) by long division of the product (< 2^200) */
    uint64_t M[4] = {1, 0, 0, 0};
    for (int i = 0; i < np; i++) {
        uint64_t carry = 0;
        for (int w = 0; w < 4; w++) { u128 s = (u128)M[w] * NTTM_PRIMES[i] + carry; M[w] = (uint64_t)s; carry = (uint64_t)(s >> 64); }
    }
    for (int w = 0; w < 4; w++) { if (M[w]--) break; }
    const u128 D = (u128)(NTTM_B - 1) * (NTTM_B - 1);
    u128 rem = 0; uint64_t q[4] = {0, 0, 0, 0};
    for (int bit = 255; bit >= 0; bit--) {
        int top = (int)(rem >> 127);
        rem = (rem << 1) | ((M[bit / 64] >> (bit % 64)) & 1);
        if (top || rem >= D) { rem -= D; q[bit / 64] |= 1ULL << (bit % 64); }
    }
    if (q[1] || q[2] || q[3]) return UINT64_MAX;
    return q[0];
}

extern "C" uint64_t nttm_prime(int i) { return (i >= 0 && i < NTTM_MAX_PRIMES) ? NTTM_PRIMES[i] : 0; }
extern "C" uint64_t nttm_max_terms(int np) { return (np >= 1 && np <= NTTM_MAX_PRIMES) ? max_terms_np(np) : 0; }
extern "C" uint64_t nttm_root(int i, int logn)
{
    if (i < 0 || i >= NTTM_MAX_PRIMES || logn < 0 || logn > NTTM_LOGN_MAX) return 0;
    const uint64_t p = NTTM_PRIMES[i];
    return h_powmod(h_generator(p), (p - 1) >> logn, p);
}
extern "C" uint64_t nttm_root3(int i, int logk)
{
    if (i < 0 || i >= NTTM_MAX_PRIMES || logk < 0 || logk > NTTM_LOGN_MAX) return 0;
    const uint64_t p = NTTM_PRIMES[i];
    return h_powmod(h_generator(p), ((p - 1) >> logk) / 3, p);
}

static void build_crt_const(CrtConst &cc)
{
    memset(&cc, 0, sizeof cc);
    for (int i = 0; i < NTTM_MAX_PRIMES; i++) {
        const uint64_t p = NTTM_PRIMES[i];
        cc.pk[i].p = (double)p; cc.pk[i].pinv = 1.0 / (double)p; cc.pk[i].pu = p;
        /* prefix product p_0 ... p_{i-1} in words, and its residue mod p_i */
        uint64_t M[NTTM_MAX_PRIMES] = {1, 0, 0, 0};
        for (int j = 0; j < i; j++) {
            uint64_t carry = 0;
            for (int w = 0; w < NTTM_MAX_PRIMES; w++) { u128 s = (u128)M[w] * NTTM_PRIMES[j] + carry; M[w] = (uint64_t)s; carry = (uint64_t)(s >> 64); }
        }
        for (int w = 0; w < NTTM_MAX_PRIMES; w++) cc.mprod[i][w] = M[w];
        {   /* the same in base 10^18 (digits j < i; the product is below 2^(50 i) < 10^(18 i)) */
            uint64_t W[NTTM_MAX_PRIMES];
            for (int w = 0; w < NTTM_MAX_PRIMES; w++) W[w] = M[w];
            for (int j = 0; j < NTTM_MAX_PRIMES; j++) {
                u128 rem = 0;
                for (int w = NTTM_MAX_PRIMES - 1; w >= 0; w--) {
                    const u128 cur = (rem << 64) | W[w];
                    W[w] = (uint64_t)(cur / NTTM_B); rem = cur % NTTM_B;
                }
                cc.mb[i][j] = (uint64_t)rem;
                cc.mbp[i][j] = (uint64_t)(((u128)rem << 64) / NTTM_B);
            }
        }
        for (int j = 1; j <= i; j++) {
            uint64_t m = 1;
            for (int k = 0; k < j; k++) m = h_mulmod(m, NTTM_PRIMES[k] % p, p);
            if (j < i) cc.mmod[i][j] = h_bal(m, p);
            if (j == i) cc.cinv[i] = h_bal(h_inv(m, p), p);
        }
    }
}

static int build_tables(void)
{
    const size_t nst = (size_t)1 << NTTM_STAGE_LOG;
    const size_t nbq = (size_t)4 << NTTM_BQ_LOG, nb3 = nbq + 1;
    const size_t per_dir = nst + nbq + nb3, total = (size_t)NTTM_MAX_PRIMES * 2 * per_dir;
    std::vector<double2> h(total);
    for (int i = 0; i < NTTM_MAX_PRIMES; i++) {
        const uint64_t p = NTTM_PRIMES[i];
        const uint64_t gen = h_generator(p);
        g.root32[i] = h_powmod(gen, (p - 1) >> 32, p);
        g.root3x32[i] = h_powmod(gen, ((p - 1) >> 32) / 3, p);
        for (int dir = 0; dir < 2; dir++) {
            const uint64_t w32 = dir ? h_inv(g.root32[i], p) : g.root32[i];
            double2 *st = &h[((size_t)i * 2 + dir) * per_dir], *bq = st + nst;
            double2 *b3 = bq + nbq;
            const uint64_t W3 = dir ? h_inv(g.root3x32[i], p) : g.root3x32[i];
            for (int k = 0; k < 4; k++) {                                /* b3[k 2^8 + x] = W3^(x 2^(8k)) */
                const uint64_t wk = h_powmod(W3, (uint64_t)1 << (NTTM_BQ_LOG * k), p);
                uint64_t v = 1;
                for (size_t x = 0; x < ((size_t)1 << NTTM_BQ_LOG); x++) {
                    b3[((size_t)k << NTTM_BQ_LOG) + x] = h_shoup(v, p); v = h_mulmod(v, wk, p);
                }
            }
            b3[nbq] = h_shoup(h_powmod(g.root3x32[i], (uint64_t)1 << 32, p), p);   /* the forward cube root */
            st[0] = h_shoup(1, p);
            for (int s = 0; s + 1 < NTTM_STAGE_LOG + 1 && ((size_t)1 << s) < nst; s++) {
                const uint64_t ws = h_powmod(w32, (uint64_t)1 << (32 - (s + 1)), p);   /* order 2^(s+1) */
                uint64_t v = 1;
                for (size_t j = 0; j < ((size_t)1 << s); j++) { st[((size_t)1 << s) + j] = h_shoup(v, p); v = h_mulmod(v, ws, p); }
            }
            for (int k = 0; k < 4; k++) {                                /* bq[k 2^8 + x] = w32^(x 2^(8k)) */
                const uint64_t wk = h_powmod(w32, (uint64_t)1 << (NTTM_BQ_LOG * k), p);
                uint64_t v = 1;
                for (size_t x = 0; x < ((size_t)1 << NTTM_BQ_LOG); x++) {
                    bq[((size_t)k << NTTM_BQ_LOG) + x] = h_shoup(v, p); v = h_mulmod(v, wk, p);
                }
            }
        }
    }
    HIPCK(hipMalloc(&g.tables, total * sizeof(double2)));
    HIPCK(hipMemcpy(g.tables, h.data(), total * sizeof(double2), hipMemcpyHostToDevice));
    for (int i = 0; i < NTTM_MAX_PRIMES; i++)
        for (int dir = 0; dir < 2; dir++) {
            const double2 *st = g.tables + ((size_t)i * 2 + dir) * per_dir;
            TwTables t; t.stage = st; t.bq = st + nst;
            t.b3 = t.bq + nbq;
            if (dir) g.twi[i] = t; else g.twf[i] = t;
        }
    return 0;
}

/* an outgrown buffer: freed at once after a device synchronization, or (defer: work queued before it may still read
 * it, the caller synchronizes later) listed for free_deferred, so that growing it does not wait for that work */
static int drop_dev(void *p, bool defer)
{
    if (!p) return 0;
    if (defer && g.ndfree < 8) { g.dfree[g.ndfree++] = p; return 0; }
    HIPCK(hipDeviceSynchronize());
    HIPCK(hipFree(p));
    return 0;
}
static int drop_host(void *p, bool defer)
{
    if (!p) return 0;
    if (defer && g.nhfree < 4) { g.hfree[g.nhfree++] = p; return 0; }
    HIPCK(hipDeviceSynchronize());
    HIPCK(hipHostFree(p));
    return 0;
}
/* after the synchronization that ends every use of the listed buffers */
static void free_deferred(void)
{
    for (int i = 0; i < g.ndfree; i++) (void)hipFree(g.dfree[i]);
    for (int i = 0; i < g.nhfree; i++) (void)hipHostFree(g.hfree[i]);
    g.ndfree = g.nhfree = 0;
}
static int ws_need(size_t bytes, bool defer = false)
{
    if (g.cfg.workspace_bytes && bytes > g.cfg.workspace_bytes)
        return nttm_fail(NTTM_ENOMEM, "workspace of %zu bytes exceeds the configured cap %zu", bytes, g.cfg.workspace_bytes);
    if (bytes <= g.ws_bytes) return 0;
    if (g.ws) { RCK(drop_dev(g.ws, defer)); g.ws = nullptr; g.ws_bytes = 0; }
    if (hipMalloc(&g.ws, bytes) != hipSuccess) { (void)hipGetLastError(); return nttm_fail(NTTM_ENOMEM, "hipMalloc of %zu bytes failed", bytes); }
    g.ws_bytes = bytes;
    return 0;
}
static int hpin_need(size_t bytes)
{
    if (bytes <= g.hpin_bytes) return 0;
    if (g.hpin) { HIPCK(hipDeviceSynchronize()); HIPCK(hipHostFree(g.hpin)); g.hpin = nullptr; g.hpin_bytes = 0; }
    HIPCK(hipHostMalloc((void **)&g.hpin, bytes, hipHostMallocDefault));
    g.hpin_bytes = bytes;
    return 0;
}

/* ---- where an operand lives -------------------------------------------------------------------------------------------
 * Kernels read and write HBM only (XNACK is off: a kernel touching host memory faults the GPU).  Every pointer an entry
 * point takes is classed before any launch: device memory of this device is used in place; host memory -- pageable,
 * pinned (hipHostMalloc) or managed -- is copied through library-owned HBM by the entry points that stage (the
 * products, as the reference's batch stages products outside registered memory, E/rns_mul.h:42-47,
 * E/rns_mul.c:807-817) and refused with NTTM_EINVAL by the others.  One MemMap lives for one call: it remembers the
 * device allocations it has seen, so the operands of a large batch carved from a few allocations cost one runtime
 * query per allocation. */
enum { MEM_DEV = 0, MEM_HOST = 1, MEM_PINNED = 2, MEM_BAD = 3 };
struct MemMap {
    std::map<uintptr_t, uintptr_t> dev;                          /* base -> end of the device allocations seen */
    uintptr_t last_b = 0, last_e = 0;                            /* the last one hit */
    const char *why = "";
    int kind(const void *p, size_t bytes)
    {
        if (!p || !bytes) return MEM_DEV;                        /* nothing is read or written */
        const uintptr_t a = (uintptr_t)p, e = a + bytes;
        if (e < a) { why = "wraps the address space"; return MEM_BAD; }
        if (a >= last_b && a < last_e) return e <= last_e ? MEM_DEV : (why = "runs past the end of its device allocation", MEM_BAD);
        auto it = dev.upper_bound(a);
        if (it != dev.begin()) {
            --it;
            if (a < it->second) {
                last_b = it->first; last_e = it->second;
                return e <= it->second ? MEM_DEV : (why = "runs past the end of its device allocation", MEM_BAD);
            }
        }
        hipPointerAttribute_t at;
        memset(&at, 0, sizeof at);
        if (hipPointerGetAttributes(&at, p) != hipSuccess) { (void)hipGetLastError(); return MEM_HOST; }   /* pageable */
        if (at.type == hipMemoryTypeDevice) {
            if (at.device != g.dev) { why = "is device memory of another device"; return MEM_BAD; }
            hipDeviceptr_t base = nullptr; size_t sz = 0;
            if (hipMemGetAddressRange(&base, &sz, (hipDeviceptr_t)p) == hipSuccess && base && sz) {
                const uintptr_t b = (uintptr_t)base;
                dev[b] = b + sz; last_b = b; last_e = b + sz;
                if (e > b + sz) { why = "runs past the end of its device allocation"; return MEM_BAD; }
            } else (void)hipGetLastError();
            return MEM_DEV;
        }
        if (at.type == hipMemoryTypeHost) return MEM_PINNED;
        return MEM_HOST;                                         /* managed, unified or unregistered */
    }
};
/* NTTM_EINVAL unless [p, p + bytes) is device memory of the library's device (entry points that do not stage) */
int nttm_require_device(const void *p, size_t bytes, const char *what)
{
    MemMap mm;
    const int k = mm.kind(p, bytes);
    if (k == MEM_DEV) return 0;
    return nttm_fail(NTTM_EINVAL, "%s %s", what, k == MEM_BAD ? mm.why : "is host memory (this entry point takes device memory)");
}
#define DEVCK(p, bytes, what) RCK(nttm_require_device((p), (bytes), (what)))
/* NTTM_EINVAL unless [p, p + bytes) is host memory (pageable, pinned or managed): entry points whose output the host
 * writes or first-touches */
int nttm_require_host(const void *p, size_t bytes, const char *what)
{
    MemMap mm;
    const int k = mm.kind(p, bytes);
    if (k == MEM_HOST || k == MEM_PINNED) return 0;
    return nttm_fail(NTTM_EINVAL, "%s %s", what, k == MEM_BAD ? mm.why : "is device memory (this entry point takes host memory)");
}

extern "C" int nttm_init(int device, const nttm_config *cfg)
{
    if (g.init) return nttm_fail(NTTM_ESTATE, "nttm_init: already initialised");
    memset(&g, 0, sizeof g);
    if (cfg) g.cfg = *cfg;
    if (g.cfg.nprimes && g.cfg.nprimes != 3 && g.cfg.nprimes != 4)
        return nttm_fail(NTTM_EINVAL, "nttm_init: nprimes must be 0, 3 or 4");
    if (!g.cfg.tile_log) g.cfg.tile_log = NTTM_TILE_LOG_DEFAULT;
    if (g.cfg.tile_log != 12 && g.cfg.tile_log != 13) return nttm_fail(NTTM_EINVAL, "nttm_init: tile_log must be 12 or 13");
    if (g.cfg.pts && g.cfg.pts != 8 && g.cfg.pts != 16) return nttm_fail(NTTM_EINVAL, "nttm_init: pts must be 0, 8 or 16");
    if (g.cfg.radix3 < NTTM_R3_AUTO || g.cfg.radix3 > NTTM_R3_FIT)
        return nttm_fail(NTTM_EINVAL, "nttm_init: radix3 must be 0, 1 or 2");
    if (g.cfg.check != 0 && g.cfg.check != 1) return nttm_fail(NTTM_EINVAL, "nttm_init: check must be 0 or 1");
    nttm_set_pts(g.cfg.pts ? g.cfg.pts : 16);
    {   /* NTTM_GENERIC_PASSES=1: runtime-geometry pass kernels only (for comparison) */
        const char *e = getenv("NTTM_GENERIC_PASSES");
        nttm_set_fixed_shapes(!(e && e[0] == '1'));
    }
    int nd = 0;
    HIPCK(hipGetDeviceCount(&nd));
    if (device < 0 || device >= nd) return nttm_fail(NTTM_EINVAL, "nttm_init: device %d of %d", device, nd);
    HIPCK(hipSetDevice(device));
    g.dev = device;
    if (device < NTTM_NOALLOC_MAX_DEV && !g_scratch_dev[device]) {   /* once per device, with it current */
        void *ptr = nullptr;
        HIPCK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return nttm_fail(NTTM_EHIP, "nttm_init: no device address for the no-allocation workspace");
        g_scratch_dev[device] = (char *)ptr;
    }
    {   /* NTTM_STREAM=0: strided passes without the LDS-streaming kernel (for comparison); one streaming block per CU */
        const char *st = getenv("NTTM_STREAM");
        int ncu = 0;
        HIPCK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
        nttm_set_stream(!(st && st[0] == '0'), ncu);
    }
    for (int i = 0; i < NTTM_MAX_PRIMES; i++) { g.pk[i].p = (double)NTTM_PRIMES[i]; g.pk[i].pinv = 1.0 / (double)NTTM_PRIMES[i]; g.pk[i].pu = NTTM_PRIMES[i]; }
    build_crt_const(g.cc);
    RCK(build_tables());
    for (int i = 0; i < 4; i++) HIPCK(hipEventCreate(&g.ev[i]));
    for (int i = 0; i < 2 * NTTM_MAX_PRIMES; i++) HIPCK(hipEventCreate(&g.pev[i]));
    RCK(hpin_need(1 << 16));
    {
        const char *e = getenv("NTTM_NO_DEFER");
        g_no_defer = e && e[0] == '1';
    }
    {
        const char *e = getenv("NTTM_TREE_LEVELWISE");
        g_tree_levelwise = e && e[0] == '1';
    }
    {
        const char *e = getenv("NTTM_FUSED_BITS");
        g_fused_bits_env = e ? atoi(e) : 0;
    }
    {   /* NTTM_STAGE_CHUNK=<bytes>: staged bytes per sub-batch (default 1 GiB; smaller values exercise the split) */
        const char *e = getenv("NTTM_STAGE_CHUNK");
        g.stage_chunk = e && atoll(e) > 0 ? (size_t)atoll(e) : NTTM_STAGE_CHUNK;
    }
    g.gen = ++g_init_count;
    g.init = 1;
    return 0;
}

static void prof_release(void);
extern "C" void nttm_shutdown(void)
{
    if (!g.init) return;
    (void)hipDeviceSynchronize();
    prof_release();
    if (g.ws) (void)hipFree(g.ws);
    if (g.tmp) (void)hipFree(g.tmp);
    for (int k = 0; k < 2; k++) if (g.bsp[k]) (void)hipFree(g.bsp[k]);
    if (g.tdesc) (void)hipFree(g.tdesc);
    if (g.tpin) (void)hipHostFree(g.tpin);
    if (g.tables) (void)hipFree(g.tables);
    if (g.hpin) (void)hipHostFree(g.hpin);
    if (g.stage) (void)hipFree(g.stage);
    if (g.chk) (void)hipFree(g.chk);
    if (g.bounce) (void)hipHostFree(g.bounce);
    if (g.dl) (void)hipHostFree(g.dl);
    for (int k = 0; k < 2; k++) if (g.dlev[k]) (void)hipEventDestroy(g.dlev[k]);
    for (int k = 0; k < 2; k++) if (g.bev[k]) (void)hipEventDestroy(g.bev[k]);
    for (int k = 0; k < 8; k++) if (g.tev[k]) (void)hipEventDestroy(g.tev[k]);
    if (g.tcs) (void)hipStreamDestroy(g.tcs);
    for (int k = 0; k < NTTM_MAX_PRIMES - 1; k++) {
        if (g.ps[k]) (void)hipStreamDestroy(g.ps[k]);
        if (g.pjev[k]) (void)hipEventDestroy(g.pjev[k]);
    }
    if (g.pfev) (void)hipEventDestroy(g.pfev);
    free_deferred();
    for (int i = 0; i < 4; i++) (void)hipEventDestroy(g.ev[i]);
    for (int i = 0; i < 2 * NTTM_MAX_PRIMES; i++) if (g.pev[i]) (void)hipEventDestroy(g.pev[i]);
    memset(&g, 0, sizeof g);
}

/* ---- transforms on FP64 planes ------------------------------------------------------------------------------------------ */
static int ceil_log2(uint64_t n) { int l = 0; while (((uint64_t)1 << l) < n) l++; return l; }

static PassArgs base_args(int prime, double *x, int logn, uint64_t batch, bool inv)
{
    PassArgs a; memset(&a, 0, sizeof a);
    a.x = x; a.y = nullptr; a.logn = logn; a.batch = batch; a.ymode = NTTM_Y_FULL;
    a.pk = g.pk[prime]; a.tw = inv ? g.twi[prime] : g.twf[prime]; a.twi = g.twi[prime];
    return a;
}
static int fwd_planes(int prime, double *x, int logn, uint64_t batch, hipStream_t s)
{
    Plan pl; plan_make(&pl, logn, batch, g.cfg.tile_log);
    PassArgs a = base_args(prime, x, logn, batch, false);
    for (int i = 0; i < pl.npass; i++) RCK(nttm_run_pass(false, a, pl.p[i], s));
    return 0;
}
/* inverse with the pointwise product by y (ymode layout) fused into the first pass; scale by n^-1 in the last pass.
 * r3y: the batch is the thirds of transforms of 3 2^logn points, y holds whole such transforms */
static int inv_planes(int prime, double *x, const double *y, int ymode, int logn, uint64_t batch, int scale, hipStream_t s,
                      int r3y = 0)
{
    Plan pl; plan_make(&pl, logn, batch, g.cfg.tile_log);
    PassArgs a = base_args(prime, x, logn, batch, true);
    for (int i = pl.npass - 1; i >= 0; i--) {
        PassArgs b = a;
        if (i == pl.npass - 1) { b.y = y; b.ymode = ymode; b.r3y = r3y; }
        if (i == 0 && scale) {
            const uint64_t p = NTTM_PRIMES[prime];
            b.scale = h_shoup(h_inv(((uint64_t)1 << logn) % p, p), p); b.do_scale = 1;
        }
        RCK(nttm_run_pass(true, b, pl.p[i], s));
    }
    return 0;
}

/* ---- the product ------------------------------------------------------------------------------------------------------- */
struct CrtDesc {                 /* must match crt.hip */
    const double *pl; uint64_t plane;
    uint64_t ncoef;
    const uint64_t *x; uint64_t nx;
    uint64_t *c; uint64_t nc;
    uint64_t first_block;
    uint64_t trunc;
};
static_assert(sizeof(CrtDesc) <= 96, "a product's pinned slot holds its CRT descriptor below byte 96");

static int choose_np(size_t mn, int *np)
{
    int want = g.cfg.nprimes;
    if (!want) want = (uint64_t)mn <= max_terms_np(3) ? 3 : 4;
    if ((uint64_t)mn > max_terms_np(want))
        return nttm_fail(NTTM_EBOUND, "a product of %zu terms exceeds the %d-prime bound of %llu terms", mn, want,
                         (unsigned long long)max_terms_np(want));
    *np = want;
    return 0;
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

/* the transforms of one length group of products (cnt transforms of 2^logn points per prime):
 *   B: forward, the first pass reading b's limbs (times n^-1, folded here once);
 *   A: forward up to its last pass, the first pass reading a's limbs;
 *   A: the fused last forward pass + pointwise product with B + first inverse pass (k_fused);
 *   A: the remaining inverse passes.
 * Operand limbs come from (a, na, b, nb) when cnt == 1 and ldesc is null, else from the device descriptors ldesc.
 * Result: plane i of A holds the cyclic convolution mod p_i. */
/* the strided-pass tile for a 2^logn product transform: the configured one, or 2^13 where that saves a pass */
static int product_tile_log(int logn, int tl = 0)
{
    if (!tl) tl = g.cfg.tile_log;
    const int R = logn > NTTM_FUSED_LOG ? logn - NTTM_FUSED_LOG : 0;
    const int k_cfg = (R + tl - 5) / (tl - 4), k_13 = (R + 8) / 9;
    return k_13 < k_cfg ? 13 : tl;
}

/* bits of the double-fused last pass of a 2^logn product transform (its tiles stay 2^12 points): 12, or 10 and 11 for
 * 2^26 and 2^27 points, so that their two strided passes take 8 bits each (8+8+10, 8+8+11, 8+8+12 up to 2^28) and run
 * as the 8 x 32 streamed shape -- a level moved out of the fused pass (instruction-bound: both forward transforms and
 * the inverse run there) costs less in an HBM-bound strided pass.  Measured per pass at 2^28 points on one MI355X:
 * 8-bit streamed passes 0.78-0.85 ms, 9-bit (16-column rows) 0.96-1.14 ms, the fused pass 2.08 ms at 12 bits and 1.68
 * at 10, so 9+9+10 loses to 8+8+12 there, and 2^29 / 2^30 points keep 9+8+12 / 9+9+12. */
static int fused_bits(int logn, int tile_log)
{
    if (tile_log == 13 && g_fused_bits_env >= 10 && g_fused_bits_env <= 12 && logn > g_fused_bits_env &&
        logn - g_fused_bits_env <= 18)
        return g_fused_bits_env;
    if (tile_log != 13 || logn < 26 || logn > 30) return NTTM_FUSED_LOG;
    return logn <= 28 ? logn - 16 : NTTM_FUSED_LOG;
}
static void product_plan(Plan *pl, int logn, uint64_t batch, int tile_log)
{
    plan_make3(pl, logn, batch, tile_log, fused_bits(logn, tile_log), NTTM_FUSED_LOG);
}

/* the transforms of cnt products of one length: 2^logn points (r3 = 0), or 3 2^logn points (r3 = 1: a radix-3 layer,
 * reading the limbs, and the 2^logn engine on the 3 cnt thirds).  Plane i of A holds the cnt transforms of prime i back
 * to back.  After it, A holds the cyclic convolutions mod each prime, natural order.  B has one plane, reused by every
 * prime: B's transform of prime i is made just before A's (np + 1 planes in all instead of 2 np).
 * Shared B operands (bmap non-null): B's plane holds the cnt_b distinct operands of ldesc_b; A's engine transform t
 * pairs with B's engine transform bmap[t] in the double-fused pass.
 * pev (or null): 2 np events recorded around B's transform of each prime. */
/* the first forward pass (from limbs) leaves its boundary twiddles to the second when both are streamed strided
 * passes (a transform of three or more passes): the limb pass has no VALU to spare, the second pass, bound by HBM, has.
 * q: the first pass's arguments (limbs set, in_scale as the caller wants it applied) */
static bool defer_ok(const PassArgs &q, const Plan &pl, int last)
{
    if (last < 2 || g_no_defer || !(q.limbs || q.ldesc)) return false;
    PassArgs nx = q; nx.limbs = nullptr; nx.ldesc = nullptr; nx.do_in_scale = 0;
    return nttm_fwd_pass_streams(q, pl.p[0]) && nttm_fwd_pass_streams(nx, pl.p[1]) && pl.p[1].m >= 4;
}
/* the deferral's flags for pass j of the plan (q carries pass 0's in_scale until pass 0 has run) */
static void defer_set(PassArgs &q, const Plan &pl, int j, bool on)
{
    q.dfr_out = on && j == 0;
    if (on && j == 1) {
        q.dfr_in = 1; q.dfr_lom = pl.p[0].lo + pl.p[0].m; q.dfr_gm = pl.p[0].m;
    } else q.dfr_in = 0;
    if (on && j == 0) { q.dfr_do_scale = q.do_in_scale; q.dfr_scale = q.in_scale; }
    if (j > 1) q.dfr_do_scale = 0;
}
/* ps: prime i's work on stream ps[i] with its own B plane at Bp + i bstride (the caller forks and joins them); null:
 * every prime on s, one after another, through the one B plane at Bp */
static int product_transforms(int np, int logn, int r3, uint64_t cnt, double *A, double *Bp, const uint64_t *a,
                              uint64_t na, const uint64_t *b, uint64_t nb, const LoadDesc *ldesc, hipStream_t s0,
                              hipEvent_t *pev, uint64_t cnt_b = 0, const LoadDesc *ldesc_b = nullptr,
                              const uint32_t *bmap = nullptr, const hipStream_t *ps = nullptr, uint64_t bstride = 0)
{
    double *const Bp0 = Bp;
    if (!bmap) { cnt_b = cnt; ldesc_b = ldesc; }
    const uint64_t tb = r3 ? 3 * cnt : cnt, tbb = r3 ? 3 * cnt_b : cnt_b;   /* engine transforms of 2^logn points */
    Plan pl; product_plan(&pl, logn, tb, product_tile_log(logn));
    Plan plb; product_plan(&plb, logn, tbb, product_tile_log(logn));
    const uint64_t n = (uint64_t)1 << logn, plane = tb * n, ntot = r3 ? 3 * n : n;
    const int last = pl.npass - 1;
    if (bmap && (plb.npass != pl.npass || pl.p[last].m + pl.p[last].logh > logn || (last == 0 && !r3)))
        return nttm_fail(NTTM_EINVAL, "internal: shared B operands on this length");
    for (int i = 0; i < np; i++) {
        double *Ai = A + (size_t)i * plane;
        const uint64_t p = NTTM_PRIMES[i];
        const hipStream_t s = ps ? ps[i] : s0;
        Bp = ps ? Bp0 + (size_t)i * bstride : Bp0;
        /* B, prime i, all passes but the last */
        if (pev) HIPCK(hipEventRecord(pev[2 * i], s));
        const double2 ninv = h_shoup(h_inv(ntot % p, p), p);     /* the inverse's 1/length, folded into B */
        PassArgs fb = base_args(i, Bp, logn, tbb, false);
        if (r3) RCK(nttm_run_r3(false, Bp, cnt_b, logn, g.pk[i], g.twf[i].b3, ldesc_b ? nullptr : b, nb, ldesc_b, 1, &ninv, s));
        else { fb.limbs = ldesc_b ? nullptr : b; fb.nlimbs = nb; fb.ldesc = ldesc_b; fb.which = 1; fb.in_scale = ninv; fb.do_in_scale = 1; }
        const bool dfb = !r3 && defer_ok(fb, plb, last);
        for (int j = 0; j < last; j++) {
            defer_set(fb, plb, j, dfb);
            RCK(nttm_run_pass(false, fb, plb.p[j], s));
            fb.limbs = nullptr; fb.ldesc = nullptr; fb.do_in_scale = 0;
        }
        if (pev) HIPCK(hipEventRecord(pev[2 * i + 1], s));
        /* A, prime i */
        PassArgs fa = base_args(i, Ai, logn, tb, false);
        if (r3) RCK(nttm_run_r3(false, Ai, cnt, logn, g.pk[i], g.twf[i].b3, ldesc ? nullptr : a, na, ldesc, 0, nullptr, s));
        else { fa.limbs = ldesc ? nullptr : a; fa.nlimbs = na; fa.ldesc = ldesc; fa.which = 0; }
        const bool dfa = !r3 && defer_ok(fa, pl, last);
        for (int j = 0; j < last; j++) {
            defer_set(fa, pl, j, dfa);
            RCK(nttm_run_pass(false, fa, pl.p[j], s));
            fa.limbs = nullptr; fa.ldesc = nullptr;
        }
        fa.dfr_in = 0;
        /* B's last forward pass, A's last forward pass, the pointwise product and A's first inverse pass */
        PassArgs fbl = base_args(i, Bp, logn, tb, false);
        fbl.bmap = bmap;
        if (last == 0 && !r3) {                                  /* a single-pass transform reads B's limbs here */
            fbl.limbs = ldesc ? nullptr : b; fbl.nlimbs = nb; fbl.ldesc = ldesc; fbl.which = 1;
            fbl.in_scale = h_shoup(h_inv(n % p, p), p); fbl.do_in_scale = 1;
        }
        RCK(nttm_run_fused2(fa, fbl, pl.p[last], s));
        PassArgs ia = base_args(i, Ai, logn, tb, true);
        for (int j = last - 1; j >= 0; j--) RCK(nttm_run_pass(true, ia, pl.p[j], s));
        if (r3) RCK(nttm_run_r3(true, Ai, cnt, logn, g.pk[i], g.twi[i].b3, nullptr, 0, nullptr, 0, nullptr, s));
    }
    return 0;
}

/* HBM bytes per point of a 2^logn transform of one prime in the product schedule (both operands' forward
 * transforms, the double-fused pass, the inverse; limb reads and the CRT excluded: equal for both lengths) */
static double sched_bytes(int logn, int r3, int tile_log)
{
    static double memo[NTTM_LOGN_MAX + 1][2][2];                 /* per length, radix, tile 12 / 13 (0: not yet) */
    const bool cacheable = logn >= 0 && logn <= NTTM_LOGN_MAX && (tile_log == 12 || tile_log == 13);
    double *m = cacheable ? &memo[logn][r3 ? 1 : 0][tile_log - 12] : nullptr;
    if (m && *m > 0) return *m;
    Plan pl; product_plan(&pl, logn, 1, product_tile_log(logn, tile_log));
    const int P = pl.npass;
    const double v = r3 ? 3.0 * (2 * (8 + 16 * (P - 1)) + 24 + 16 * (P - 1) + 16)
                        : (P >= 2 ? 2 * (8 + 16 * (P - 2)) + 24 + 16 * (P - 1) : 8);
    if (m) *m = v;
    return v;
}
/* the transform length of a product with ncoef coefficients: 2^logn, or 3 2^(logn - 2) when it fits and the rule
 * (NTTM_R3_*) takes it (the reference takes it whenever it fits, E/rns_mul.c:270-277) */
static void choose_len(uint64_t ncoef, int mode, int tile_log, int *logn_out, int *r3_out)
{
    int logn = ceil_log2(ncoef); if (logn < NTTM_PRODUCT_LOGN_MIN) logn = NTTM_PRODUCT_LOGN_MIN;
    int r3 = 0;
    if (mode != NTTM_R3_NEVER && logn - 2 >= NTTM_PRODUCT_LOGN_MIN && ncoef <= ((uint64_t)3 << (logn - 2)))
        r3 = mode == NTTM_R3_FIT ||
             sched_bytes(logn - 2, 1, tile_log) * ldexp(1.0, logn - 2) < sched_bytes(logn, 0, tile_log) * ldexp(1.0, logn);
    *logn_out = r3 ? logn - 2 : logn; *r3_out = r3;
}
static void choose_len(uint64_t ncoef, int *logn_out, int *r3_out)
{
    choose_len(ncoef, g.cfg.radix3, g.cfg.tile_log, logn_out, r3_out);
}

/* c = a b (+ x), device pointers; synchronous (returns after the stream has drained) */
/* the one-transform schedule of an na x nb product writing nc limbs: primes, length and workspace bytes; nonzero when
 * no transform holds it (the prime set's bound, or more than 2^32 points) */
struct OnePlan { int np, logn, r3; uint64_t n; size_t planes_bytes, crt_words, ws_bytes; };
static int one_plan(size_t na, size_t nb, size_t nc, OnePlan *o)
{
    const int want = g.cfg.nprimes;
    const size_t mn = na < nb ? na : nb;
    o->np = want ? want : (uint64_t)mn <= max_terms_np(3) ? 3 : 4;
    if ((uint64_t)mn > max_terms_np(o->np)) return 1;
    choose_len(na + nb - 1, &o->logn, &o->r3);
    if (o->logn > NTTM_LOGN_MAX) return 1;
    o->n = (uint64_t)(o->r3 ? 3 : 1) << o->logn;
    o->planes_bytes = (size_t)(o->np + 1) * o->n * 8;
    o->crt_words = nttm_crt_scratch_words(nttm_crt_blocks(nc));
    o->ws_bytes = align_up(o->planes_bytes, 256) + 256 + o->crt_words * 4 + 256;
    return 0;
}
static bool fits_one(size_t na, size_t nb, size_t nc)
{
    OnePlan o;
    return !one_plan(na, nb, nc, &o) && (!g.cfg.workspace_bytes || o.ws_bytes <= g.cfg.workspace_bytes);
}

/* the workspace of a one-transform product: the planes, its CRT descriptor and the CRT's scratch (bytes from ws) */
struct CoreLayout { int np, logn, r3, par; uint64_t n, ncoef; size_t nc, nc_full, off_desc, off_crt, bytes; };
/* transforms of at most this many points run their primes concurrently (a launch of one prime's pass has too few tiles
 * to fill the device); NTTM_PAR_POINTS overrides it, 0 turns it off */
static uint64_t par_points(void)
{
    static uint64_t v = 0;
    static bool set = false;
    if (!set) {
        const char *e = getenv("NTTM_PAR_POINTS");
        v = e ? strtoull(e, nullptr, 0) : ((uint64_t)1 << 22);
        set = true;
    }
    return v;
}
/* allow_par: the primes may run concurrently, each with its own B plane (the library's workspace, not captured) */
static int core_layout(size_t na, size_t nb, bool has_x, size_t wcut, CoreLayout *L, bool allow_par = false)
{
    L->nc_full = na + nb + (has_x ? 1 : 0); L->nc = L->nc_full < wcut ? L->nc_full : wcut;
    RCK(choose_np(na < nb ? na : nb, &L->np));
    const uint64_t ncoef_full = na + nb - 1;
    L->ncoef = ncoef_full < L->nc ? ncoef_full : L->nc;           /* coefficients read */
    choose_len(ncoef_full, &L->logn, &L->r3);
    if (L->logn > NTTM_LOGN_MAX) return nttm_fail(NTTM_EBOUND, "transform length 2^%d above 2^%d", L->logn, NTTM_LOGN_MAX);
    L->n = (uint64_t)(L->r3 ? 3 : 1) << L->logn;                  /* points per plane */
    /* np planes of A and one of B: B's transform of each prime is made just before A's of that prime; concurrent
     * primes take a B plane each */
    const size_t crt_words = nttm_crt_scratch_words(nttm_crt_blocks(L->nc));
    for (int par = allow_par && L->np > 1 && L->n <= par_points(); par >= 0; par--) {
        const size_t planes_bytes = (size_t)(par ? 2 * L->np : L->np + 1) * L->n * 8;
        L->par = par;
        L->off_desc = align_up(planes_bytes, 256); L->off_crt = L->off_desc + 256;
        L->bytes = L->off_crt + crt_words * 4 + 256;
        if (!g.cfg.workspace_bytes || L->bytes <= g.cfg.workspace_bytes) break;   /* concurrent only within the cap */
    }
    return 0;
}

/* c = a b (+ x), device pointers.  wcut: write only the low wcut limbs of the result (the value mod B^wcut; no check of
 * what lies above).  ws: the workspace (null: the library's, grown to fit; else ws_cap bytes of the caller's).
 * sync = 1: returns after the stream has drained and the CRT's flags are checked; sync = 0: returns after the launches,
 * the flags copied to pinned slot `slot` (g.hpin + 128 slot + 64, read by the caller after a synchronization). */
static int mul_core(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *x, size_t nx,
                    hipStream_t s, size_t wcut = SIZE_MAX, char *ws = nullptr, size_t ws_cap = 0, int sync = 1,
                    size_t slot = 0)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    const size_t nc_full = na + nb + (x ? 1 : 0), nc = nc_full < wcut ? nc_full : wcut;
    if (x && nx > na + nb) return nttm_fail(NTTM_EINVAL, "nx = %zu exceeds na + nb = %zu", nx, na + nb);
    if (!c && nc) return nttm_fail(NTTM_EINVAL, "null result pointer");
    if ((na && !a) || (nb && !b) || (nx && !x)) return nttm_fail(NTTM_EINVAL, "null operand pointer");
    if (!na || !nb) {
        if (nc) HIPCK(hipMemsetAsync(c, 0, nc * 8, s));
        if (x && nx && nc) HIPCK(hipMemcpyAsync(c, x, (nx < nc ? nx : nc) * 8, hipMemcpyDeviceToDevice, s));
        if (sync) HIPCK(hipStreamSynchronize(s));
        else ((uint32_t *)((char *)g.hpin + 128 * slot + 96))[1] = 0;
        return 0;
    }
    if (!nc) return 0;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (!ws && hipStreamIsCapturing(s, &cap) != hipSuccess) { (void)hipGetLastError(); cap = hipStreamCaptureStatusActive; }
    CoreLayout L; RCK(core_layout(na, nb, x != nullptr, wcut, &L, !ws && cap == hipStreamCaptureStatusNone));
    if (ws) {
        if (L.bytes > ws_cap) return nttm_fail(NTTM_EINVAL, "the product needs %zu bytes of workspace, above the %zu given", L.bytes, ws_cap);
    } else {
        RCK(ws_need(L.bytes));
        ws = g.ws;
    }
    double *A = (double *)ws, *Bp = A + (size_t)L.np * L.n;
    CrtDesc *ddesc = (CrtDesc *)(ws + L.off_desc);
    uint32_t *crt_scr = (uint32_t *)(ws + L.off_crt);
    char *hs = (char *)g.hpin + 128 * slot;                       /* pinned: the descriptor, then (at 96) the flags */

    if (sync) HIPCK(hipEventRecord(g.ev[0], s));
    hipStream_t ps[NTTM_MAX_PRIMES];
    if (L.par) {                                                 /* prime 0 on s, the others forked from it */
        if (!g.pfev) {
            HIPCK(hipEventCreateWithFlags(&g.pfev, hipEventDisableTiming));
            for (int k = 0; k < NTTM_MAX_PRIMES - 1; k++) {
                HIPCK(hipStreamCreateWithFlags(&g.ps[k], hipStreamNonBlocking));
                HIPCK(hipEventCreateWithFlags(&g.pjev[k], hipEventDisableTiming));
            }
        }
        HIPCK(hipEventRecord(g.pfev, s));
        ps[0] = s;
        for (int i = 1; i < L.np; i++) { ps[i] = g.ps[i - 1]; HIPCK(hipStreamWaitEvent(ps[i], g.pfev, 0)); }
    }
    RCK(product_transforms(L.np, L.logn, L.r3, 1, A, Bp, a, na, b, nb, nullptr, s, sync ? g.pev : nullptr, 0, nullptr,
                           nullptr, L.par ? ps : nullptr, L.n));
    if (L.par)
        for (int i = 1; i < L.np; i++) { HIPCK(hipEventRecord(g.pjev[i - 1], ps[i])); HIPCK(hipStreamWaitEvent(s, g.pjev[i - 1], 0)); }
    if (sync) HIPCK(hipEventRecord(g.ev[2], s));
    CrtDesc hd; memset(&hd, 0, sizeof hd);
    hd.pl = A; hd.plane = L.n; hd.ncoef = L.ncoef; hd.x = x; hd.nx = x ? nx : 0; hd.c = c; hd.nc = nc; hd.first_block = 0;
    hd.trunc = nc < nc_full;
    memcpy(hs, &hd, sizeof hd);
    HIPCK(hipMemcpyAsync(ddesc, hs, sizeof hd, hipMemcpyHostToDevice, s));
    RCK(nttm_crt_launch_batch(ddesc, 1, nttm_crt_blocks(nc), L.np, g.cc, crt_scr, s,
                              8.0 * L.np * (double)L.ncoef + 8.0 * (double)nc + 8.0 * (double)(x ? (nx < nc ? nx : nc) : 0)));
    if (sync) HIPCK(hipEventRecord(g.ev[3], s));
    HIPCK(hipMemcpyAsync(hs + 96, crt_scr, 8, hipMemcpyDeviceToHost, s));
    if (!sync) return 0;
    HIPCK(hipStreamSynchronize(s));
    float ms, tb = 0, t;
    for (int i = 0; i < L.np; i++) { HIPCK(hipEventElapsedTime(&t, g.pev[2 * i], g.pev[2 * i + 1])); tb += t; }
    HIPCK(hipEventElapsedTime(&ms, g.ev[0], g.ev[2]));
    g.last_ms[0] = tb; g.last_ms[1] = ms - tb;
    HIPCK(hipEventElapsedTime(&ms, g.ev[2], g.ev[3])); g.last_ms[2] = ms;
    const uint32_t flags = ((const uint32_t *)(hs + 96))[1];
    if (flags & 1) return nttm_fail(NTTM_ECHECK, "crt: non-zero carry beyond the top limb (wrong residues)");
    return 0;
}

/* ---- the residue post-condition ------------------------------------------------------------------------------------
 * c = a b + x modulo the prime NTTM_CHECK_Q = 4611686018427388039 (the first prime above 2^62).  The reference's
 * "detect and retry" checks the same identity on residues it tracks through the run (level 2, E/dr.h:1-8,
 * ALGORITHM.md:441-458, modulo the same prime DR_Q); here the residues are computed from the arrays themselves on the
 * device after the product, every array of a call in 64Ki-limb chunks by one kernel (k_res_chunks), the chunks of an
 * array weighted by powers of B^65536 on the host.  Costs one read of a, b, x and c. */
#define NTTM_CHECK_Q 4611686018427388039ULL
struct ResChunkH { const uint64_t *x; uint64_t n; };          /* must match digits.hip's ResChunk */
/* the cyclic product of length L = 2^logn or 3 2^logn (at least 16 points): v[0 .. L + 3) = sum_k c_k B^k over the
 * cyclic convolution c_k = sum_(i + j = k mod L) a_i b_j, i.e. a b mod (x^L - 1) at x = B before its carries wrap (the
 * caller folds v modulo B^L - 1).  The transforms are DFTs of length L, so the pointwise product gives the cyclic
 * convolution directly; each c_k has at most min(na, nb) terms, the bound of an ordinary product.  na, nb <= L, device
 * pointers.  flag null: synchronous; else the launches only, the CRT's flag word copied to the pinned word *flag (bit
 * 0 set: a carry beyond the top limb), read by the caller after it synchronizes the stream.  The Newton reciprocal's
 * wrapped product (GMP's mpn_mulmod_bnm1 plays this part there). */
int nttm_mul_cyclic(uint64_t *v, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, size_t L, hipStream_t s,
                    const uint32_t **flag)
{
    if (flag) *flag = nullptr;
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    int logn = -1, r3 = 0;
    if (L >= 16 && !(L & (L - 1))) logn = __builtin_ctzll(L);
    else if (L % 3 == 0 && L / 3 >= 16 && !((L / 3) & (L / 3 - 1))) { logn = __builtin_ctzll(L / 3); r3 = 1; }
    if (logn < NTTM_PRODUCT_LOGN_MIN || logn > NTTM_LOGN_MAX)
        return nttm_fail(NTTM_EINVAL, "cyclic product: length %zu is not 2^k or 3 2^k (k in [%d, %d])", L,
                         NTTM_PRODUCT_LOGN_MIN, NTTM_LOGN_MAX);
    if (na > L || nb > L) return nttm_fail(NTTM_EINVAL, "cyclic product: operands of %zu and %zu limbs above %zu", na, nb, L);
    if (!v || (na && !a) || (nb && !b)) return nttm_fail(NTTM_EINVAL, "cyclic product: null pointer");
    const size_t nc = L + 3;
    if (!na || !nb) { HIPCK(hipMemsetAsync(v, 0, nc * 8, s)); return flag ? 0 : nttm_hip_check(hipStreamSynchronize(s), "cyclic product"); }
    int np; RCK(choose_np(na < nb ? na : nb, &np));
    const size_t planes_bytes = (size_t)(np + 1) * L * 8;
    const size_t off_desc = align_up(planes_bytes, 256), off_crt = off_desc + 256;
    RCK(ws_need(off_crt + nttm_crt_scratch_words(nttm_crt_blocks(nc)) * 4 + 256));
    double *A = (double *)g.ws, *Bp = A + (size_t)np * L;
    RCK(product_transforms(np, logn, r3, 1, A, Bp, a, na, b, nb, nullptr, s, nullptr));
    CrtDesc hd; memset(&hd, 0, sizeof hd);
    hd.pl = A; hd.plane = L; hd.ncoef = L; hd.c = v; hd.nc = nc;
    char *hs = (char *)g.hpin;
    memcpy(hs, &hd, sizeof hd);
    HIPCK(hipMemcpyAsync(g.ws + off_desc, hs, sizeof hd, hipMemcpyHostToDevice, s));
    uint32_t *crt_scr = (uint32_t *)(g.ws + off_crt);
    RCK(nttm_crt_launch_batch(g.ws + off_desc, 1, nttm_crt_blocks(nc), np, g.cc, crt_scr, s, 8.0 * np * L + 8.0 * nc));
    HIPCK(hipMemcpyAsync(hs + 96, crt_scr, 8, hipMemcpyDeviceToHost, s));
    if (flag) { *flag = (const uint32_t *)(hs + 96) + 1; return 0; }
    HIPCK(hipStreamSynchronize(s));
    if (((const uint32_t *)(hs + 96))[1] & 1) return nttm_fail(NTTM_ECHECK, "cyclic product: crt: non-zero carry beyond the top limb");
    return 0;
}

extern "C" int nttm_debug_mul_cyclic(uint64_t *v, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, size_t L,
                                     void *stream)
{
    if ((na && nttm_require_device(a, na * 8, "a")) || (nb && nttm_require_device(b, nb * 8, "b")) ||
        nttm_require_device(v, (L + 3) * 8, "v"))
        return nttm_fail(NTTM_EINVAL, "nttm_debug_mul_cyclic: device pointers only");
    return nttm_mul_cyclic(v, a, na, b, nb, L, (hipStream_t)stream, nullptr);
}
/* the transform length (points) the product rule gives `need` coefficients: a supported length of the cyclic product */
size_t nttm_wrap_len(size_t need)
{
    int logn, r3;
    choose_len(need ? need : 1, &logn, &r3);
    return (size_t)(r3 ? 3 : 1) << logn;
}

int nttm_run_res_chunks(const void *chunks_dev, uint64_t nchunks, uint64_t q, uint64_t *out_dev, hipStream_t s);
uint64_t nttm_res_chunk_limbs(void);
static int chk_need(size_t bytes)
{
    if (bytes <= g.chk_bytes) return 0;
    if (g.chk) { HIPCK(hipDeviceSynchronize()); HIPCK(hipFree(g.chk)); g.chk = nullptr; g.chk_bytes = 0; }
    if (hipMalloc(&g.chk, bytes) != hipSuccess) { (void)hipGetLastError(); return nttm_fail(NTTM_ENOMEM, "residue check: %zu bytes", bytes); }
    g.chk_bytes = bytes;
    return 0;
}
/* res[i] = (sum_k arr[i][k] B^k) mod NTTM_CHECK_Q for device arrays; synchronizes */
static int residues_q(uint64_t *res, const uint64_t *const *arr, const size_t *len, size_t count, hipStream_t s)
{
    const uint64_t CH = nttm_res_chunk_limbs(), q = NTTM_CHECK_Q;
    std::vector<ResChunkH> ch;
    std::vector<size_t> owner;
    for (size_t i = 0; i < count; i++) {
        res[i] = 0;
        for (size_t off = 0; off < len[i]; off += CH) {
            ch.push_back({arr[i] + off, std::min<uint64_t>(CH, len[i] - off)});
            owner.push_back(i);
        }
    }
    if (ch.empty()) return 0;
    const size_t off_out = align_up(ch.size() * sizeof(ResChunkH), 256);
    RCK(chk_need(off_out + ch.size() * 8));
    std::vector<uint64_t> out(ch.size());
    HIPCK(hipMemcpyAsync(g.chk, ch.data(), ch.size() * sizeof(ResChunkH), hipMemcpyHostToDevice, s));
    RCK(nttm_run_res_chunks(g.chk, ch.size(), q, (uint64_t *)(g.chk + off_out), s));
    HIPCK(hipMemcpyAsync(out.data(), g.chk + off_out, ch.size() * 8, hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    const uint64_t bch = h_powmod(NTTM_B % q, CH, q);
    uint64_t w = 1;
    for (size_t k = 0; k < ch.size(); k++) {
        if (k == 0 || owner[k] != owner[k - 1]) w = 1;
        res[owner[k]] = (uint64_t)(((u128)out[k] * w + res[owner[k]]) % q);
        w = h_mulmod(w, bch, q);
    }
    return 0;
}
/* the post-condition over N products on device memory: NTTM_ECHECK naming the first that fails */
static int check_products(const nttm_prod *P, size_t N, hipStream_t s)
{
    std::vector<const uint64_t *> arr(4 * N);
    std::vector<size_t> len(4 * N);
    for (size_t i = 0; i < N; i++) {
        const nttm_prod &q = P[i];
        arr[4 * i] = q.a; len[4 * i] = q.na;
        arr[4 * i + 1] = q.b; len[4 * i + 1] = q.nb;
        arr[4 * i + 2] = q.x; len[4 * i + 2] = q.x ? q.nx : 0;
        arr[4 * i + 3] = q.c; len[4 * i + 3] = q.na + q.nb + (q.x ? 1 : 0);
    }
    std::vector<uint64_t> r(4 * N);
    RCK(residues_q(r.data(), arr.data(), len.data(), 4 * N, s));
    for (size_t i = 0; i < N; i++) {
        const uint64_t want = (uint64_t)(((u128)r[4 * i] * r[4 * i + 1] + r[4 * i + 2]) % NTTM_CHECK_Q);
        if (want != r[4 * i + 3])
            return nttm_fail(NTTM_ECHECK, "product %zu: c differs from a b + x modulo %llu (the residue post-condition)", i,
                             (unsigned long long)NTTM_CHECK_Q);
    }
    return 0;
}

/* ---- products above one transform: the grid (E/rns_dist.c:610-700, our own form) ---------------------------------------
 * A product that no single transform holds (more than 2^32 points, the prime set's term bound, or the configured
 * workspace cap) is formed from ka x kb piece products a_i b_j (pieces of ceil(na/ka) and ceil(nb/kb) limbs).  The grid
 * takes the fewest plane points in total, then the fewest pieces (i, j <= 64).  Each piece is added into c at limb
 * o = oa + ob by one product with the window c[o .. o + la + lb) as its x (c = a_i b_j + window, written back), and its
 * top limb (0 or 1) carried into c above the window on the device.  The cuts skip whole pieces: a piece at or above
 * limb w only adds a multiple of B^w, and one ending at or below lowcut is left out (the caller reads the limbs above
 * lowcut and tolerates what those pieces would have carried into them, E/rns_dist.c:678-679). */
static int grid_shape(size_t na, size_t nb, int *ka, int *kb)
{
    double best = 0; *ka = *kb = 0;
    for (int i = 1; i <= 64; i++)
        for (int j = 1; j <= 64; j++) {
            const size_t pa = (na + i - 1) / i, pb = (nb + j - 1) / j;
            if (!fits_one(pa, pb, pa + pb + 1)) continue;
            OnePlan o; one_plan(pa, pb, pa + pb + 1, &o);
            const double cost = (double)i * j * (double)o.n;
            if (!*ka || cost < best || (cost == best && i * j < *ka * *kb)) { best = cost; *ka = i; *kb = j; }
        }
    if (!*ka) return nttm_fail(NTTM_EBOUND, "no grid of at most 64 x 64 pieces forms a %zu x %zu-limb product", na, nb);
    return 0;
}
static int tmp_need(size_t limbs)
{
    if (limbs <= g.tmp_limbs) return 0;
    if (g.tmp) { HIPCK(hipDeviceSynchronize()); HIPCK(hipFree(g.tmp)); g.tmp = nullptr; g.tmp_limbs = 0; }
    if (hipMalloc(&g.tmp, limbs * 8) != hipSuccess) { (void)hipGetLastError(); return nttm_fail(NTTM_ENOMEM, "hipMalloc of %zu limbs failed", limbs); }
    g.tmp_limbs = limbs;
    return 0;
}
static int mul_grid(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *x,
                    size_t nx, hipStream_t s, size_t lowcut, size_t w)
{
    const size_t nc_full = na + nb + (x ? 1 : 0), nout = nc_full < w ? nc_full : w;
    int ka, kb; RCK(grid_shape(na, nb, &ka, &kb));
    const size_t pa = (na + ka - 1) / ka, pb = (nb + kb - 1) / kb;
    RCK(tmp_need(pa + pb + 1));
    HIPCK(hipMemsetAsync(c, 0, nout * 8, s));
    if (x && nx) HIPCK(hipMemcpyAsync(c, x, (nx < nout ? nx : nout) * 8, hipMemcpyDeviceToDevice, s));
    for (int j = 0; j < kb; j++)
        for (int i = 0; i < ka; i++) {
            const size_t oa = (size_t)i * pa, ob = (size_t)j * pb;
            if (oa >= na || ob >= nb) continue;
            const size_t la = na - oa < pa ? na - oa : pa, lb = nb - ob < pb ? nb - ob : pb, o = oa + ob, len = la + lb;
            if (o >= nout || o + len <= lowcut) continue;             /* at or above w; below the low cut */
            const size_t wp = nout - o;                               /* limbs of c from o on */
            if (wp <= len) {                                          /* the piece reaches the top: mod B^wp */
                RCK(mul_core(g.tmp, a + oa, la, b + ob, lb, c + o, wp, s, wp));
                HIPCK(hipMemcpyAsync(c + o, g.tmp, wp * 8, hipMemcpyDeviceToDevice, s));
            } else {
                RCK(mul_core(g.tmp, a + oa, la, b + ob, lb, c + o, len, s));
                HIPCK(hipMemcpyAsync(c + o, g.tmp, len * 8, hipMemcpyDeviceToDevice, s));
                RCK(nttm_run_add_carry(c + o + len, wp - len, g.tmp + len, s));
            }
        }
    HIPCK(hipStreamSynchronize(s));
    return 0;
}
/* one transform when it holds the product, else the grid */
static int mul_any(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *x, size_t nx,
                   hipStream_t s, size_t lowcut = 0, size_t w = SIZE_MAX)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    const size_t nc_full = na + nb + (x ? 1 : 0), nout = nc_full < w ? nc_full : w;
    if (!na || !nb || !nout || fits_one(na, nb, nout)) return mul_core(c, a, na, b, nb, x, nx, s, w);
    if (x && nx > na + nb) return nttm_fail(NTTM_EINVAL, "nx = %zu exceeds na + nb = %zu", nx, na + nb);
    if (!c || !a || !b || (nx && !x)) return nttm_fail(NTTM_EINVAL, "null pointer");
    return mul_grid(c, a, na, b, nb, x, nx, s, lowcut, w);
}
/* mul_any and, with nttm_config.check, the residue post-condition on a whole product (w = SIZE_MAX) */
static int mul_full(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *x, size_t nx,
                    hipStream_t s, size_t lowcut = 0, size_t w = SIZE_MAX)
{
    RCK(mul_any(c, a, na, b, nb, x, nx, s, lowcut, w));
    if (!g.cfg.check || w != SIZE_MAX) return 0;
    nttm_prod q; memset(&q, 0, sizeof q);
    q.a = a; q.na = na; q.b = b; q.nb = nb; q.c = c; q.x = x; q.nx = x ? nx : 0;
    return check_products(&q, 1, s);
}

/* c = a b on device memory without a synchronization where it fits one transform (no residue check configured): the
 * launches only, the CRT's flag word copied to the pinned word *flag (bit 0 set: a carry beyond the top limb), which
 * the caller reads after synchronizing the stream; *flag = null when the product ran synchronously (a grid) */
int nttm_mul_launch(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, hipStream_t s,
                    const uint32_t **flag)
{
    *flag = nullptr;
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (!na || !nb || g.cfg.check || !fits_one(na, nb, na + nb)) return mul_full(c, a, na, b, nb, nullptr, 0, s);
    RCK(mul_core(c, a, na, b, nb, nullptr, 0, s, SIZE_MAX, nullptr, 0, 0, 0));
    *flag = (const uint32_t *)((char *)g.hpin + 96) + 1;
    return 0;
}

extern "C" int nttm_last_phase_ms(double *f, double *i, double *c)
{
    if (f) *f = g.last_ms[0];
    if (i) *i = g.last_ms[1];
    if (c) *c = g.last_ms[2];
    return 0;
}

/* ---- batches ---------------------------------------------------------------------------------------------------------- */
/* one length group of a batch: its products' descriptors and their layout in device memory (offsets from the group's
 * descriptor base), for a given schedule (length, strided-pass tile, points per thread) */
struct GroupBuild {
    int r3 = 0, logn = 0, tile = 0, pts = 0;
    int crt_per = 8;                                             /* CRT limbs per thread: 1 when every product is small */
    uint64_t n = 0, cnt = 0, cnt_b = 0, nblk = 0;
    bool share = false;
    std::vector<CrtDesc> cd;
    std::vector<LoadDesc> ld, ldb;
    std::vector<uint32_t> bmap;
    size_t planes_bytes = 0;                                     /* (np cnt + cnt_b) n doubles */
    size_t off_cd = 0, off_ld = 0, off_ldb = 0, off_bmap = 0, off_ncn = 0, off_crt = 0, desc_bytes = 0;
    double crt_bytes = 0;
};
static void group_build(GroupBuild &gb, const nttm_prod *P, const std::vector<size_t> &G, int np, int r3, int logn,
                        int tile, int pts)
{
    gb.r3 = r3; gb.logn = logn; gb.tile = tile; gb.pts = pts;
    const uint64_t n = (uint64_t)(r3 ? 3 : 1) << logn, cnt = G.size();          /* points per product plane */
    gb.n = n; gb.cnt = cnt; gb.nblk = 0; gb.crt_bytes = 0;
    gb.crt_per = n <= 1024 ? 1 : 8;                              /* products of at most n + 1 limbs: 256-limb blocks */
    gb.cd.assign(cnt, CrtDesc()); gb.ld.assign(cnt, LoadDesc()); gb.ldb.clear(); gb.bmap.clear();
    for (uint64_t j = 0; j < cnt; j++) {
        const nttm_prod &q = P[G[j]];
        LoadDesc &l = gb.ld[j];
        CrtDesc &c = gb.cd[j];
        l.a = q.a; l.na = q.na; l.b = q.b; l.nb = q.nb;
        c.pl = nullptr; c.plane = cnt * n; c.ncoef = q.na + q.nb - 1; c.x = q.x; c.nx = q.x ? q.nx : 0;
        c.c = q.c; c.nc = q.na + q.nb + (q.x ? 1 : 0); c.first_block = gb.nblk; c.trunc = 0;
        gb.nblk += nttm_crt_blocks(c.nc, gb.crt_per);
        gb.crt_bytes += 8.0 * np * c.ncoef + 8.0 * c.nc + 8.0 * c.nx;
    }
    /* products that share their B operand (the same b and nb: the binary-splitting tree's P1 Q2 + P2 and Q1 Q2,
     * E/rns_mul.c:476-495) transform it once, where B has planes (a transform of more than one pass, or a radix-3
     * length) and the double-fused tiles lie inside one transform */
    std::vector<uint32_t> slot(cnt);
    Plan pq; product_plan(&pq, logn, r3 ? 3 * cnt : cnt, product_tile_log(logn, tile));
    const bool can = logn >= NTTM_FUSED_LOG && (pq.npass > 1 || r3);
    std::map<std::pair<const uint64_t *, uint64_t>, uint32_t> seen;
    for (uint64_t j = 0; j < cnt && can; j++) {
        auto key = std::make_pair(gb.ld[j].b, gb.ld[j].nb);
        auto it = seen.find(key);
        if (it == seen.end()) {
            it = seen.emplace(key, (uint32_t)gb.ldb.size()).first;
            gb.ldb.push_back(gb.ld[j]);
        }
        slot[j] = it->second;
    }
    gb.cnt_b = can ? gb.ldb.size() : cnt;                       /* no sharing: one B per product */
    gb.share = gb.cnt_b < cnt;
    if (!gb.share) gb.ldb.clear();
    const uint64_t tb = r3 ? 3 * cnt : cnt;
    if (gb.share) {
        gb.bmap.resize(tb);
        for (uint64_t t = 0; t < tb; t++) gb.bmap[t] = r3 ? 3 * slot[t / 3] + (uint32_t)(t % 3) : slot[t];
    }
    gb.planes_bytes = ((size_t)np * cnt + gb.cnt_b) * n * 8;
    gb.off_cd = 0;
    gb.off_ld = align_up(gb.off_cd + cnt * sizeof(CrtDesc), 256);
    gb.off_ldb = align_up(gb.off_ld + cnt * sizeof(LoadDesc), 256);
    gb.off_bmap = align_up(gb.off_ldb + gb.ldb.size() * sizeof(LoadDesc), 256);
    gb.off_ncn = align_up(gb.off_bmap + gb.bmap.size() * 4, 256);
    gb.off_crt = align_up(gb.off_ncn + cnt * 8, 256);
    gb.desc_bytes = align_up(gb.off_crt + nttm_crt_scratch_words(gb.nblk) * 4, 256);
}
/* points per thread of a batch group's passes: the configured value, else 8 up to 2^14 points and 16 above (what the
 * batch plans' autotuner chose for every such group of the 4096-leaf workload) */
static int batch_pts(int logn) { return g.cfg.pts ? g.cfg.pts : logn <= 14 ? 8 : 16; }
/* the group's descriptors as one host image (planes at A), to copy to the group's descriptor base */
static void group_image(GroupBuild &gb, double *A, char *img)
{
    for (uint64_t j = 0; j < gb.cnt; j++) gb.cd[j].pl = A + j * gb.n;          /* (the alignment gaps stay unread) */
    memcpy(img + gb.off_cd, gb.cd.data(), gb.cnt * sizeof(CrtDesc));
    memcpy(img + gb.off_ld, gb.ld.data(), gb.cnt * sizeof(LoadDesc));
    if (gb.share) {
        memcpy(img + gb.off_ldb, gb.ldb.data(), gb.ldb.size() * sizeof(LoadDesc));
        memcpy(img + gb.off_bmap, gb.bmap.data(), gb.bmap.size() * 4);
    }
}
/* the group's launches: transforms, CRT with the carry scan, normalized lengths; under the group's tile and points
 * per thread (the library's settings are restored afterwards) */
static int group_launch(const GroupBuild &gb, int np, double *A, char *dbase, hipStream_t s, bool norm = true)
{
    const int tile0 = g.cfg.tile_log, pts0 = nttm_get_pts();
    g.cfg.tile_log = gb.tile; nttm_set_pts(gb.pts);
    double *Bp = A + (size_t)np * gb.cnt * gb.n;
    int rc = product_transforms(np, gb.logn, gb.r3, gb.cnt, A, Bp, nullptr, 0, nullptr, 0, (const LoadDesc *)(dbase + gb.off_ld),
                                s, nullptr, gb.cnt_b, gb.share ? (const LoadDesc *)(dbase + gb.off_ldb) : nullptr,
                                gb.share ? (const uint32_t *)(dbase + gb.off_bmap) : nullptr);
    if (!rc) rc = nttm_crt_launch_batch(dbase + gb.off_cd, gb.cnt, gb.nblk, np, g.cc, (uint32_t *)(dbase + gb.off_crt), s,
                                        gb.crt_bytes, gb.crt_per);
    if (!rc && norm) rc = nttm_run_norm(dbase + gb.off_cd, gb.cnt, (uint64_t *)(dbase + gb.off_ncn), s);
    g.cfg.tile_log = tile0; nttm_set_pts(pts0);
    return rc;
}

/* the batch on device memory (every pointer of P classed MEM_DEV) */
static int mul_batch_dev(nttm_prod *P, size_t N, hipStream_t s)
{
    if (!N) return 0;
    /* validate; zero-operand products are written directly */
    size_t mn_max = 0;
    /* groups by transform length: 2^logn at [logn], 3 2^logn at [NTTM_LOGN_MAX + 1 + logn] */
    std::vector<std::vector<size_t>> groups(2 * (NTTM_LOGN_MAX + 1));
    for (size_t i = 0; i < N; i++) {
        const nttm_prod &q = P[i];
        if (q.x && q.nx > q.na + q.nb) return nttm_fail(NTTM_EINVAL, "product %zu: nx exceeds na + nb", i);
        if (!q.c || (q.na && !q.a) || (q.nb && !q.b)) return nttm_fail(NTTM_EINVAL, "product %zu: null pointer", i);
        if (!q.na || !q.nb) continue;
        size_t mn = q.na < q.nb ? q.na : q.nb; if (mn > mn_max) mn_max = mn;
        int logn, r3;
        choose_len(q.na + q.nb - 1, &logn, &r3);
        if (logn > NTTM_LOGN_MAX) return nttm_fail(NTTM_EBOUND, "product %zu: transform above 2^%d", i, NTTM_LOGN_MAX);
        groups[r3 * (NTTM_LOGN_MAX + 1) + logn].push_back(i);
    }
    int np; RCK(choose_np(mn_max ? mn_max : 1, &np));
    for (size_t i = 0; i < N; i++) {
        const nttm_prod &q = P[i];
        if (q.na && q.nb) continue;
        const size_t nc = q.na + q.nb + (q.x ? 1 : 0);
        if (nc) HIPCK(hipMemsetAsync(q.c, 0, nc * 8, s));
        if (q.x && q.nx) HIPCK(hipMemcpyAsync(q.c, q.x, q.nx * 8, hipMemcpyDeviceToDevice, s));
    }
    /* one group per transform length: batched load, transforms, CRT */
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const std::vector<size_t> &G = groups[gi];
        if (G.empty()) continue;
        const int r3 = gi > NTTM_LOGN_MAX, logn = (int)(gi % (NTTM_LOGN_MAX + 1));
        GroupBuild gb;
        group_build(gb, P, G, np, r3, logn, g.cfg.tile_log, batch_pts(logn));
        const size_t off_d = align_up(gb.planes_bytes, 256);
        RCK(ws_need(off_d + gb.desc_bytes + 256));
        RCK(hpin_need(gb.desc_bytes + 4096));
        char *hp = (char *)g.hpin;
        HIPCK(hipStreamSynchronize(s));                      /* the pinned staging is reused per group */
        group_image(gb, (double *)g.ws, hp);
        HIPCK(hipMemcpyAsync(g.ws + off_d, hp, gb.off_ncn, hipMemcpyHostToDevice, s));   /* descriptors up to ncn */
        RCK(group_launch(gb, np, (double *)g.ws, g.ws + off_d, s));
        HIPCK(hipMemcpyAsync(hp, g.ws + off_d + gb.off_ncn, gb.cnt * 8, hipMemcpyDeviceToHost, s));
        HIPCK(hipMemcpyAsync(hp + gb.cnt * 8, g.ws + off_d + gb.off_crt, 8, hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        const uint32_t flags = ((const uint32_t *)(hp + gb.cnt * 8))[1];
        if (flags & 1) return nttm_fail(NTTM_ECHECK, "crt: non-zero carry beyond a top limb (wrong residues)");
        for (uint64_t j = 0; j < gb.cnt; j++) P[G[j]].ncn = ((const uint64_t *)hp)[j];
    }
    /* zero-operand products: normalized length of x */
    for (size_t i = 0; i < N; i++) {
        nttm_prod &q = P[i];
        if (q.na && q.nb) continue;
        size_t n = q.x ? q.nx : 0;
        HIPCK(hipStreamSynchronize(s));
        while (n) { uint64_t v; HIPCK(hipMemcpy(&v, q.x + n - 1, 8, hipMemcpyDeviceToHost)); if (v) break; n--; }
        q.ncn = n;
    }
    return g.cfg.check ? check_products(P, N, s) : 0;
}

/* ---- host-resident operands: staging through HBM -----------------------------------------------------------------------
 * The products accept operands and results in host memory, as the reference's batch does: it reads registered host
 * memory in place and stages anything else through device buffers (E/rns_mul.h:42-47, E/rns_mul.c:807-817).  Kernels
 * here read HBM only, so every host input is copied into a library-owned device region (g.stage) before the products
 * and every host result copied back after them.  Copies go through a pinned bounce buffer of two halves, so that the
 * host copy into one half overlaps the DMA out of the other; pinned inputs and results of at least a half move by
 * DMA directly.  A host operand named by several products of one call (the shared B of a binary-splitting pair) is
 * copied once, so those products still share its transform.  A batch whose staged bytes exceed g.stage_chunk (1 GiB)
 * runs as consecutive sub-batches (a product larger than that is staged alone). */
#define NTTM_BOUNCE_HALF ((size_t)4 << 20)       /* bytes per half of the pinned bounce buffer */
#define NTTM_STAGE_ALIGN 256                     /* every staged array starts on a 256-byte boundary */

static int stage_need(size_t bytes)
{
    if (bytes <= g.stage_bytes) return 0;
    if (g.stage) { HIPCK(hipDeviceSynchronize()); HIPCK(hipFree(g.stage)); g.stage = nullptr; g.stage_bytes = 0; }
    if (hipMalloc(&g.stage, bytes) != hipSuccess) { (void)hipGetLastError(); return nttm_fail(NTTM_ENOMEM, "staging: hipMalloc of %zu bytes failed", bytes); }
    g.stage_bytes = bytes;
    return 0;
}
static int bounce_need(void)
{
    if (g.bounce) return 0;
    HIPCK(hipHostMalloc((void **)&g.bounce, 2 * NTTM_BOUNCE_HALF, hipHostMallocDefault));
    g.bounce_half = NTTM_BOUNCE_HALF;
    for (int k = 0; k < 2; k++) { HIPCK(hipEventCreateWithFlags(&g.bev[k], hipEventDisableTiming)); g.bev_live[k] = 0; }
    return 0;
}
static int bounce_wait(int h)                    /* the last copy through half h has completed */
{
    if (g.bev_live[h]) { HIPCK(hipEventSynchronize(g.bev[h])); g.bev_live[h] = 0; }
    return 0;
}

/* the host arrays of one staged call and their offsets in g.stage: the inputs first, then the results */
struct StageSet {
    struct Item { const void *p; size_t bytes; int kind; size_t off; };
    std::vector<Item> in, out;
    std::map<std::pair<const void *, size_t>, size_t> seen;
    size_t bytes = 0;
    size_t add_in(const void *p, size_t n, int kind)
    {
        const auto key = std::make_pair(p, n);
        const auto it = seen.find(key);
        if (it != seen.end()) return it->second;
        in.push_back({p, n, kind, 0});
        seen.emplace(key, in.size() - 1);
        return in.size() - 1;
    }
    size_t add_out(void *p, size_t n, int kind) { out.push_back({p, n, kind, 0}); return out.size() - 1; }
    void layout()
    {
        size_t o = 0;
        for (Item &i : in) { i.off = o; o = align_up(o + i.bytes, NTTM_STAGE_ALIGN); }
        for (Item &i : out) { i.off = o; o = align_up(o + i.bytes, NTTM_STAGE_ALIGN); }
        bytes = o ? o : NTTM_STAGE_ALIGN;
    }
    const uint64_t *din(size_t k) const { return (const uint64_t *)(g.stage + in[k].off); }
    uint64_t *dout(size_t k) const { return (uint64_t *)(g.stage + out[k].off); }
};

/* the inputs into g.stage: consecutive inputs are packed into a bounce half (the alignment gaps zero-filled), and a
 * full half goes to the device by one DMA while the host fills the other half */
static int stage_upload(const StageSet &st, hipStream_t s)
{
    if (st.in.empty()) return 0;
    RCK(bounce_need());
    const size_t half = g.bounce_half;
    int h = 0;
    size_t fill = 0, base = 0;                                   /* half h holds fill bytes for g.stage + base */
    RCK(bounce_wait(0));
    auto flush = [&]() -> int {
        if (!fill) return 0;
        HIPCK(hipMemcpyAsync(g.stage + base, g.bounce + h * half, fill, hipMemcpyHostToDevice, s));
        HIPCK(hipEventRecord(g.bev[h], s)); g.bev_live[h] = 1;
        h ^= 1; fill = 0;
        return bounce_wait(h);
    };
    for (const StageSet::Item &i : st.in) {
        if (!i.bytes) continue;
        if (i.kind == MEM_PINNED && i.bytes >= half) {           /* pinned and large: one direct DMA */
            RCK(flush());
            HIPCK(hipMemcpyAsync(g.stage + i.off, i.p, i.bytes, hipMemcpyHostToDevice, s));
            continue;
        }
        const char *src = (const char *)i.p;
        size_t left = i.bytes, off = i.off;
        while (left) {
            if (!fill) base = off;
            else if (base + fill != off) {                       /* an alignment gap: zero-filled while it fits */
                const size_t pad = off - (base + fill);
                if (fill + pad < half) { memset(g.bounce + h * half + fill, 0, pad); fill += pad; }
                else { RCK(flush()); base = off; }
            }
            const size_t take = std::min(left, half - fill);
            memcpy(g.bounce + h * half + fill, src, take);
            fill += take; src += take; off += take; left -= take;
            if (fill == half) RCK(flush());
        }
    }
    return flush();
}

/* the results out of g.stage: windows of at most a half DMA'd into alternating halves, each landed window copied to
 * its host arrays while the next one is in flight; pinned results of at least a half by DMA directly.  Returns after
 * the stream has drained. */
static int stage_download(const StageSet &st, hipStream_t s)
{
    if (st.out.empty()) return nttm_hip_check(hipStreamSynchronize(s), "staging");
    RCK(bounce_need());
    const size_t half = g.bounce_half;
    struct Piece { char *dst; size_t off, bytes; };
    struct Win { size_t w0, w1, p0, p1; };                       /* device range, pieces [p0, p1) */
    std::vector<Piece> pc;
    for (const StageSet::Item &o : st.out) {
        if (!o.bytes) continue;
        if (o.kind == MEM_PINNED && o.bytes >= half) {
            HIPCK(hipMemcpyAsync((void *)o.p, g.stage + o.off, o.bytes, hipMemcpyDeviceToHost, s));
            continue;
        }
        for (size_t k = 0; k < o.bytes; k += half) pc.push_back({(char *)o.p + k, o.off + k, std::min(half, o.bytes - k)});
    }
    std::vector<Win> win;
    for (size_t k = 0; k < pc.size(); k++) {
        if (!win.empty() && pc[k].off + pc[k].bytes - win.back().w0 <= half) { win.back().w1 = pc[k].off + pc[k].bytes; win.back().p1 = k + 1; }
        else win.push_back({pc[k].off, pc[k].off + pc[k].bytes, k, k + 1});
    }
    RCK(bounce_wait(0)); RCK(bounce_wait(1));
    auto scatter = [&](size_t w) {
        const Win &x = win[w];
        const char *hb = g.bounce + (w & 1) * half;
        for (size_t k = x.p0; k < x.p1; k++) memcpy(pc[k].dst, hb + (pc[k].off - x.w0), pc[k].bytes);
    };
    for (size_t w = 0; w < win.size(); w++) {
        const int h = (int)(w & 1);
        HIPCK(hipMemcpyAsync(g.bounce + h * half, g.stage + win[w].w0, win[w].w1 - win[w].w0, hipMemcpyDeviceToHost, s));
        HIPCK(hipEventRecord(g.bev[h], s)); g.bev_live[h] = 1;
        if (w) { RCK(bounce_wait(h ^ 1)); scatter(w - 1); }
    }
    if (!win.empty()) { RCK(bounce_wait((int)((win.size() - 1) & 1))); scatter(win.size() - 1); }
    return nttm_hip_check(hipStreamSynchronize(s), "staging");
}

/* one large device array to host memory.  Pinned destinations take one DMA.  Pageable ones go through two pinned
 * 64 MiB halves: the DMA of chunk k overlaps the host copy of chunk k - 1, which is split over up to 8 threads (the
 * destination's pages are first touched there, so their faults are taken in parallel).  Synchronizes. */
#define NTTM_DL_HALF ((size_t)64 << 20)
/* the pinned halves of nttm_copy_to_host, allocated ahead of the copy (pinning 128 MiB takes milliseconds; a caller can
 * run this on another host thread while the device works) */
int nttm_copy_to_host_prepare(void)
{
    if (g.dl) return 0;
    char *dl = nullptr;
    HIPCK(hipHostMalloc((void **)&dl, 2 * NTTM_DL_HALF, hipHostMallocDefault));
    memset(dl, 0, 2 * NTTM_DL_HALF);                             /* its pages touched here, not during the copy */
    for (int k = 0; k < 2; k++) HIPCK(hipEventCreateWithFlags(&g.dlev[k], hipEventDisableTiming));
    g.dl = dl;
    return 0;
}
int nttm_copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    if (!bytes) return nttm_hip_check(hipStreamSynchronize(s), "copy to host");
    MemMap mm;
    if (mm.kind(dst, bytes) == MEM_PINNED || bytes <= NTTM_DL_HALF) {
        HIPCK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        return nttm_hip_check(hipStreamSynchronize(s), "copy to host");
    }
    RCK(nttm_copy_to_host_prepare());
    const size_t H = NTTM_DL_HALF, nch = (bytes + H - 1) / H;
    const unsigned hw = std::thread::hardware_concurrency(), T = hw >= 8 ? 8 : hw ? hw : 1;
    auto host_copy = [&](size_t c) {
        const size_t off = c * H, len = std::min(H, bytes - off), part = (len + T - 1) / T;
        const char *from = g.dl + (c & 1) * H;
        char *to = (char *)dst + off;
        std::vector<std::thread> th;
        for (unsigned i = 1; i < T; i++)
            if ((size_t)i * part < len)
                th.emplace_back([=]() { memcpy(to + (size_t)i * part, from + (size_t)i * part, std::min(part, len - (size_t)i * part)); });
        memcpy(to, from, std::min(part, len));
        for (std::thread &t : th) t.join();
    };
    for (size_t c = 0; c < nch; c++) {
        const size_t off = c * H, len = std::min(H, bytes - off);
        HIPCK(hipMemcpyAsync(g.dl + (c & 1) * H, (const char *)src + off, len, hipMemcpyDeviceToHost, s));
        HIPCK(hipEventRecord(g.dlev[c & 1], s));
        if (c) { HIPCK(hipEventSynchronize(g.dlev[(c - 1) & 1])); host_copy(c - 1); }
    }
    HIPCK(hipEventSynchronize(g.dlev[(nch - 1) & 1]));
    host_copy(nch - 1);
    return nttm_hip_check(hipStreamSynchronize(s), "copy to host");
}

/* the four pointers of a product classed (a, b, x, c with their byte counts); NTTM_EINVAL for memory no entry can use */
static int prod_kinds(MemMap &mm, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *x, size_t nx,
                      const uint64_t *c, size_t nout, int *k, size_t idx)
{
    static const char *const nm[4] = {"a", "b", "x", "c"};
    k[0] = mm.kind(a, na * 8); k[1] = mm.kind(b, nb * 8); k[2] = x ? mm.kind(x, nx * 8) : MEM_DEV; k[3] = mm.kind(c, nout * 8);
    for (int j = 0; j < 4; j++)
        if (k[j] == MEM_BAD) return nttm_fail(NTTM_EINVAL, "product %zu: operand %s %s", idx, nm[j], mm.why);
    return 0;
}
static bool any_host(const int *k) { return k[0] != MEM_DEV || k[1] != MEM_DEV || k[2] != MEM_DEV || k[3] != MEM_DEV; }

/* one product with host memory among its operands (kinds k): staged, then formed on the device copies */
static int mul_staged(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *x,
                      size_t nx, hipStream_t s, size_t lowcut, size_t w, const int *k)
{
    const size_t nc_full = na + nb + (x ? 1 : 0), nout = nc_full < w ? nc_full : w;
    StageSet st;
    const size_t ia = k[0] != MEM_DEV ? st.add_in(a, na * 8, k[0]) : SIZE_MAX;
    const size_t ib = k[1] != MEM_DEV ? st.add_in(b, nb * 8, k[1]) : SIZE_MAX;
    const size_t ix = k[2] != MEM_DEV ? st.add_in(x, nx * 8, k[2]) : SIZE_MAX;
    const size_t ic = k[3] != MEM_DEV ? st.add_out(c, nout * 8, k[3]) : SIZE_MAX;
    st.layout();
    RCK(stage_need(st.bytes));
    RCK(stage_upload(st, s));
    RCK(mul_full(ic == SIZE_MAX ? c : st.dout(ic), ia == SIZE_MAX ? a : st.din(ia), na, ib == SIZE_MAX ? b : st.din(ib), nb,
                 ix == SIZE_MAX ? x : st.din(ix), nx, s, lowcut, w));
    return stage_download(st, s);
}

/* the entry of the single products: device operands in place, host ones staged */
static int mul_entry(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *x,
                     size_t nx, hipStream_t s, size_t lowcut = 0, size_t w = SIZE_MAX)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (x && nx > na + nb) return nttm_fail(NTTM_EINVAL, "nx = %zu exceeds na + nb = %zu", nx, na + nb);
    const size_t nc_full = na + nb + (x ? 1 : 0), nout = nc_full < w ? nc_full : w;
    MemMap mm;
    int k[4];
    RCK(prod_kinds(mm, a, na, b, nb, x, nx, c, nout, k, 0));
    if (!any_host(k)) return mul_full(c, a, na, b, nb, x, nx, s, lowcut, w);
    return mul_staged(c, a, na, b, nb, x, nx, s, lowcut, w, k);
}

extern "C" int nttm_mul(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, void *stream)
{
    return mul_entry(c, a, na, b, nb, nullptr, 0, (hipStream_t)stream);
}
extern "C" int nttm_mul_add(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *x,
                            size_t nx, void *stream)
{
    if (!x) return nttm_fail(NTTM_EINVAL, "nttm_mul_add: null x");
    return mul_entry(c, a, na, b, nb, x, nx, (hipStream_t)stream);
}
/* the cut product (E/rns_dist.c:581-589,684-693; E/rns_mul.c rns_mul_local): the operands cut to their low w limbs,
 * the low min(na + nb, w) limbs of the product written.  A product that fits one transform is formed whole (the
 * reference skips no piece of it either); a grid skips the pieces ending at or below lowcut. */
extern "C" int nttm_mul_cut(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, size_t lowcut,
                            size_t w, void *stream)
{
    const size_t nout = na + nb < w ? na + nb : w;
    if (lowcut > nout) return nttm_fail(NTTM_EINVAL, "nttm_mul_cut: lowcut %zu above the %zu limbs written", lowcut, nout);
    return mul_entry(c, a, na < w ? na : w, b, nb < w ? nb : w, nullptr, 0, (hipStream_t)stream, lowcut, w);
}

/* c = a b with no allocation: the same product as nttm_mul on device memory, its planes, CRT descriptor and CRT scratch
 * in the file-scope array `scratch` of this GPU (NTTM_NOALLOC_BYTES = 40 MiB: one transform of at most 2^20 points on
 * three primes, e.g. 2^19 x 2^19 limbs).  No hipMalloc, hipFree or other allocating call: the operands are classed by
 * pointer queries without the per-call cache, live kernel timing is paused for the call, and a product that would
 * need more (a larger transform, four primes beyond the array, a grid, host operands) is refused with NTTM_EINVAL.
 * Nothing in the array must start at zero: the limb passes write every plane point before any read, and every CRT
 * block writes its status word and the carry scan its control words.  Calls that overlap on one GPU share the array
 * and race; calls in one stream or one after another are safe.  Synchronizes the stream. */
static int noalloc_dev_ok(const void *p, size_t bytes)
{
    if (!p || !bytes) return 1;
    hipPointerAttribute_t at;
    memset(&at, 0, sizeof at);
    if (hipPointerGetAttributes(&at, p) != hipSuccess) { (void)hipGetLastError(); return 0; }
    if (at.type != hipMemoryTypeDevice || at.device != g.dev) return 0;
    hipDeviceptr_t base = nullptr; size_t sz = 0;
    if (hipMemGetAddressRange(&base, &sz, (hipDeviceptr_t)p) != hipSuccess) { (void)hipGetLastError(); return 1; }
    return (uintptr_t)p + bytes <= (uintptr_t)base + sz;
}
extern "C" int nttm_mul_noalloc(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, void *stream)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    hipStream_t s = (hipStream_t)stream;
    if (g.dev >= NTTM_NOALLOC_MAX_DEV || !g_scratch_dev[g.dev]) return nttm_fail(NTTM_ESTATE, "nttm_mul_noalloc: no workspace on device %d", g.dev);
    if ((na && !a) || (nb && !b) || (na + nb && !c)) return nttm_fail(NTTM_EINVAL, "nttm_mul_noalloc: null pointer");
    if (!noalloc_dev_ok(a, na * 8) || !noalloc_dev_ok(b, nb * 8) || !noalloc_dev_ok(c, (na + nb) * 8))
        return nttm_fail(NTTM_EINVAL, "nttm_mul_noalloc: operands must be device memory of this device, inside their allocations");
    if (na && nb) {
        if (!fits_one(na, nb, na + nb)) return nttm_fail(NTTM_EINVAL, "nttm_mul_noalloc: a %zu x %zu-limb product needs a grid", na, nb);
        CoreLayout L; RCK(core_layout(na, nb, false, SIZE_MAX, &L));
        if (L.bytes > NTTM_NOALLOC_BYTES)
            return nttm_fail(NTTM_EINVAL, "nttm_mul_noalloc: a %zu x %zu-limb product needs %zu bytes of workspace, above the "
                             "%zu-byte maximum (one transform of at most 2^20 points on three primes)", na, nb, L.bytes,
                             (size_t)NTTM_NOALLOC_BYTES);
    }
    const int prof0 = g_prof_on;
    g_prof_on = 0;                                               /* event records could create events */
    const int rc = mul_core(c, a, na, b, nb, nullptr, 0, s, SIZE_MAX, g_scratch_dev[g.dev], NTTM_NOALLOC_BYTES);
    g_prof_on = prof0;
    return rc;
}

/* the residue post-condition on demand: c (nc limbs) = a b + x modulo NTTM_CHECK_Q; device memory; synchronizes */
extern "C" int nttm_check_product(const uint64_t *c, size_t nc, const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
                                  const uint64_t *x, size_t nx, void *stream)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if ((nc && !c) || (na && !a) || (nb && !b) || (nx && !x)) return nttm_fail(NTTM_EINVAL, "nttm_check_product: null pointer");
    DEVCK(c, nc * 8, "c"); DEVCK(a, na * 8, "a"); DEVCK(b, nb * 8, "b"); DEVCK(x, nx * 8, "x");
    const uint64_t *arr[4] = {a, b, x, c};
    const size_t len[4] = {na, nb, x ? nx : 0, nc};
    uint64_t r[4];
    RCK(residues_q(r, arr, len, 4, (hipStream_t)stream));
    if ((uint64_t)(((u128)r[0] * r[1] + r[2]) % NTTM_CHECK_Q) != r[3])
        return nttm_fail(NTTM_ECHECK, "c differs from a b + x modulo %llu (the residue post-condition)", (unsigned long long)NTTM_CHECK_Q);
    return 0;
}

extern "C" int nttm_mul_batch(nttm_prod *P, size_t N, void *stream)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (!N) return 0;
    if (!P) return nttm_fail(NTTM_EINVAL, "nttm_mul_batch: null P");
    hipStream_t s = (hipStream_t)stream;
    MemMap mm;
    std::vector<int> k(4 * N);
    std::vector<size_t> est(N, 0);                               /* staged bytes per product */
    bool host = false;
    for (size_t i = 0; i < N; i++) {
        const nttm_prod &q = P[i];
        if (q.x && q.nx > q.na + q.nb) return nttm_fail(NTTM_EINVAL, "product %zu: nx exceeds na + nb", i);
        const size_t nc = q.na + q.nb + (q.x ? 1 : 0);
        int *ki = &k[4 * i];
        RCK(prod_kinds(mm, q.a, q.na, q.b, q.nb, q.x, q.nx, q.c, nc, ki, i));
        if (!any_host(ki)) continue;
        host = true;
        const size_t by[4] = {q.na * 8, q.nb * 8, q.x ? q.nx * 8 : 0, nc * 8};
        for (int j = 0; j < 4; j++) if (ki[j] != MEM_DEV) est[i] += align_up(by[j], NTTM_STAGE_ALIGN);
    }
    if (!host) return mul_batch_dev(P, N, s);
    /* sub-batches of at most NTTM_STAGE_CHUNK staged bytes */
    std::vector<size_t> idx;
    auto run = [&]() -> int {
        StageSet st;
        std::vector<size_t> ia(idx.size()), ib(idx.size()), ix(idx.size()), ic(idx.size());
        for (size_t j = 0; j < idx.size(); j++) {
            const nttm_prod &q = P[idx[j]];
            const int *ki = &k[4 * idx[j]];
            const size_t nc = q.na + q.nb + (q.x ? 1 : 0);
            ia[j] = ki[0] != MEM_DEV ? st.add_in(q.a, q.na * 8, ki[0]) : SIZE_MAX;
            ib[j] = ki[1] != MEM_DEV ? st.add_in(q.b, q.nb * 8, ki[1]) : SIZE_MAX;
            ix[j] = ki[2] != MEM_DEV ? st.add_in(q.x, q.nx * 8, ki[2]) : SIZE_MAX;
            ic[j] = ki[3] != MEM_DEV ? st.add_out(q.c, nc * 8, ki[3]) : SIZE_MAX;
        }
        st.layout();
        RCK(stage_need(st.bytes));
        RCK(stage_upload(st, s));
        std::vector<nttm_prod> D(idx.size());
        for (size_t j = 0; j < idx.size(); j++) {
            D[j] = P[idx[j]];
            if (ia[j] != SIZE_MAX) D[j].a = st.din(ia[j]);
            if (ib[j] != SIZE_MAX) D[j].b = st.din(ib[j]);
            if (ix[j] != SIZE_MAX) D[j].x = st.din(ix[j]);
            if (ic[j] != SIZE_MAX) D[j].c = st.dout(ic[j]);
        }
        RCK(mul_batch_dev(D.data(), D.size(), s));
        RCK(stage_download(st, s));
        for (size_t j = 0; j < idx.size(); j++) P[idx[j]].ncn = D[j].ncn;
        idx.clear();
        return 0;
    };
    size_t bytes = 0;
    for (size_t i = 0; i < N; i++) {
        if (!idx.empty() && bytes + est[i] > g.stage_chunk) { RCK(run()); bytes = 0; }
        idx.push_back(i); bytes += est[i];
    }
    return idx.empty() ? 0 : run();
}

/* the same N products, one product call each (the unbatched baseline of the batch tier); host memory staged per
 * product.  Every call synchronizes, and the normalized lengths are read back limb by limb. */
extern "C" int nttm_mul_each(nttm_prod *P, size_t N, void *stream)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (N && !P) return nttm_fail(NTTM_EINVAL, "nttm_mul_each: null P");
    hipStream_t s = (hipStream_t)stream;
    MemMap mm;
    std::vector<int> k(4 * N);
    for (size_t i = 0; i < N; i++) {
        const nttm_prod &q = P[i];
        if (q.x && q.nx > q.na + q.nb) return nttm_fail(NTTM_EINVAL, "product %zu: nx exceeds na + nb", i);
        const size_t nc = q.na + q.nb + (q.x ? 1 : 0);
        int *ki = &k[4 * i];
        RCK(prod_kinds(mm, q.a, q.na, q.b, q.nb, q.x, q.nx, q.c, nc, ki, i));
        if (any_host(ki)) RCK(mul_staged(q.c, q.a, q.na, q.b, q.nb, q.x, q.nx, s, 0, SIZE_MAX, ki));
        else RCK(mul_full(q.c, q.a, q.na, q.b, q.nb, q.x, q.nx, s));
        P[i].ncn = nc;
    }
    /* normalized lengths: the top limbs read back */
    for (size_t i = 0; i < N; i++) {
        size_t n = P[i].ncn;
        if (k[4 * i + 3] != MEM_DEV) { while (n && !P[i].c[n - 1]) n--; }
        else while (n) {
            uint64_t v; HIPCK(hipMemcpy(&v, P[i].c + n - 1, 8, hipMemcpyDeviceToHost));
            if (v) break;
            n--;
        }
        P[i].ncn = n;
    }
    return 0;
}

/* the same N products, one product call each, launched back to back: no synchronization between products, one at the
 * end for every product's CRT flags, and the normalized lengths by one k_norm launch over all N.  Device memory only;
 * a product above one transform (a grid) runs synchronously inside the sequence. */
extern "C" int nttm_mul_each_async(nttm_prod *P, size_t N, void *stream)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (!N) return 0;
    if (!P) return nttm_fail(NTTM_EINVAL, "nttm_mul_each_async: null P");
    hipStream_t s = (hipStream_t)stream;
    MemMap mm;
    size_t ws_max = 0;
    for (size_t i = 0; i < N; i++) {
        const nttm_prod &q = P[i];
        if (q.x && q.nx > q.na + q.nb) return nttm_fail(NTTM_EINVAL, "product %zu: nx exceeds na + nb", i);
        const size_t nc = q.na + q.nb + (q.x ? 1 : 0);
        int ki[4];
        RCK(prod_kinds(mm, q.a, q.na, q.b, q.nb, q.x, q.nx, q.c, nc, ki, i));
        if (any_host(ki)) return nttm_fail(NTTM_EINVAL, "nttm_mul_each_async: product %zu: host memory (device memory only)", i);
        if (q.na && q.nb && fits_one(q.na, q.nb, nc)) {                /* (with the concurrent primes' B planes) */
            CoreLayout L; RCK(core_layout(q.na, q.nb, q.x != nullptr, SIZE_MAX, &L, true));
            ws_max = std::max(ws_max, L.bytes);
        }
    }
    const size_t off_ncn = align_up(N * sizeof(CrtDesc), 256);
    RCK(ws_need(ws_max));                                        /* sized once: no reallocation between launches */
    RCK(hpin_need(128 * N + 8 * N + 256));
    RCK(stage_need(off_ncn + N * 8));
    std::vector<char> fl(N, 0);                                  /* 1: flags in pinned slot i */
    for (size_t i = 0; i < N; i++) {
        const nttm_prod &q = P[i];
        const size_t nc = q.na + q.nb + (q.x ? 1 : 0);
        if (q.na && q.nb && !fits_one(q.na, q.nb, nc)) RCK(mul_any(q.c, q.a, q.na, q.b, q.nb, q.x, q.nx, s));
        else { RCK(mul_core(q.c, q.a, q.na, q.b, q.nb, q.x, q.nx, s, SIZE_MAX, nullptr, 0, 0, i)); fl[i] = 1; }
    }
    std::vector<CrtDesc> cd(N);
    for (size_t i = 0; i < N; i++) { memset(&cd[i], 0, sizeof(CrtDesc)); cd[i].c = P[i].c; cd[i].nc = P[i].na + P[i].nb + (P[i].x ? 1 : 0); }
    HIPCK(hipMemcpyAsync(g.stage, cd.data(), N * sizeof(CrtDesc), hipMemcpyHostToDevice, s));
    RCK(nttm_run_norm(g.stage, N, (uint64_t *)(g.stage + off_ncn), s));
    uint64_t *hn = (uint64_t *)((char *)g.hpin + 128 * N);
    HIPCK(hipMemcpyAsync(hn, g.stage + off_ncn, N * 8, hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    for (size_t i = 0; i < N; i++) {
        if (fl[i] && (((const uint32_t *)((char *)g.hpin + 128 * i + 96))[1] & 1))
            return nttm_fail(NTTM_ECHECK, "product %zu: crt: non-zero carry beyond the top limb (wrong residues)", i);
        P[i].ncn = hn[i];
    }
    return g.cfg.check ? check_products(P, N, s) : 0;
}

/* ---- the binary-splitting product tree ----------------------------------------------------------------------------------- */
static int bsp_need(int k, size_t limbs, bool defer = false)
{
    if (limbs <= g.bsp_limbs[k]) return 0;
    if (g.bsp[k]) { RCK(drop_dev(g.bsp[k], defer)); g.bsp[k] = nullptr; g.bsp_limbs[k] = 0; }
    if (hipMalloc(&g.bsp[k], limbs * 8) != hipSuccess) { (void)hipGetLastError(); return nttm_fail(NTTM_ENOMEM, "hipMalloc of %zu limbs failed", limbs); }
    g.bsp_limbs[k] = limbs;
    return 0;
}
/* ---- the binary-splitting tree planned on the host ---------------------------------------------------------------------
 * Every level's products take their operands' unnormalized lengths: a parent's P has |P_L| + |Q_R| + 1 limbs and its Q
 * |Q_L| + |Q_R| (top limbs possibly 0, which changes no value).  So every level's shape, grouping and pool layout is
 * known before any product runs, and the levels follow each other on the stream with no synchronization: the host
 * builds level k's descriptors while the device runs level k - 1 (a level's large groups in chunks of TREE_CHUNK
 * products), and copies them from a pinned image that no later group reuses.  One synchronization at the end reads every group's CRT flags and the root's top limbs (normalized
 * there).  The reference's level loop reads each level's lengths back before the next (E/binsplit.c:1583-1651).
 * Returns 1, having launched nothing, when a pair's |P_R| exceeds |P_L| + |Q_R| on the planned lengths or the residue
 * post-condition is configured (the level-by-level path handles both). */
int nttm_run_top_nonzero(const uint64_t *x, uint64_t n, uint64_t *out_dev, hipStream_t s);
static int tdesc_need(size_t bytes, bool defer = false)
{
    if (bytes <= g.tdesc_bytes) return 0;
    if (g.tdesc) { RCK(drop_dev(g.tdesc, defer)); g.tdesc = nullptr; g.tdesc_bytes = 0; }
    if (hipMalloc(&g.tdesc, bytes) != hipSuccess) { (void)hipGetLastError(); return nttm_fail(NTTM_ENOMEM, "tree descriptors: %zu bytes", bytes); }
    g.tdesc_bytes = bytes;
    return 0;
}
static int tpin_need(size_t bytes, bool defer = false)
{
    if (bytes <= g.tpin_bytes) return 0;
    if (g.tpin) { RCK(drop_host(g.tpin, defer)); g.tpin = nullptr; g.tpin_bytes = 0; }
    HIPCK(hipHostMalloc((void **)&g.tpin, bytes, hipHostMallocDefault));
    g.tpin_bytes = bytes;
    return 0;
}
/* products per launched group: a level's large length group is launched in chunks, so that the device starts on the
 * first chunk while the host builds the next (the lowest levels have 10^5 products) */
#define TREE_CHUNK 16384
#define TREE_NKEY (2 * (NTTM_LOGN_MAX + 1))       /* transform-length classes: 2^logn at logn, 3 2^logn at 33 + logn */
/* the transform-length class of a product of ncoef coefficients.  choose_len depends on ncoef only through
 * lg = max(ceil_log2(ncoef), NTTM_PRODUCT_LOGN_MIN) and whether ncoef <= 3 2^(lg - 2), so each of those pairs is
 * resolved once, on a representative ncoef of the pair */
struct LenClass {
    int8_t key[65][2];
    LenClass() { memset(key, -1, sizeof key); }
    int operator()(uint64_t ncoef)
    {
        int lg = 63 - __builtin_clzll(ncoef | 1); if (((uint64_t)1 << lg) < ncoef) lg++;
        if (lg < NTTM_PRODUCT_LOGN_MIN) lg = NTTM_PRODUCT_LOGN_MIN;
        const int f = ncoef <= ((uint64_t)3 << (lg - 2));
        if (key[lg][f] < 0) {
            int logn, r3;
            choose_len(f ? (uint64_t)3 << (lg - 2) : (uint64_t)1 << lg, &logn, &r3);
            key[lg][f] = (int8_t)(r3 * (NTTM_LOGN_MAX + 1) + logn);
        }
        return key[lg][f];
    }
};
static_assert(TREE_NKEY <= 127, "length classes fit int8_t");
static int bs_tree_planned(const nttm_bs_node *leaves, size_t n, uint64_t *p, size_t p_cap, size_t *pn, uint64_t *q,
                           size_t q_cap, size_t *qn, hipStream_t s)
{
    if (g.cfg.check || n < 2) return 1;
    struct Len { size_t pn, qn; };
    std::vector<std::vector<Len>> len(1, std::vector<Len>(n));
    for (size_t i = 0; i < n; i++) len[0][i] = {leaves[i].pn, leaves[i].qn};
    while (len.back().size() > 1) {
        const std::vector<Len> &c = len.back();
        const size_t npairs = c.size() / 2;
        std::vector<Len> nx(npairs + (c.size() & 1));
        for (size_t i = 0; i < npairs; i++) {
            const Len &l = c[2 * i], &r = c[2 * i + 1];
            if (r.pn > l.pn + r.qn) return 1;
            nx[i] = {l.pn + r.qn + 1, l.qn + r.qn};
        }
        if (c.size() & 1) nx.back() = c.back();
        len.push_back(std::move(nx));
    }
    const int nlev = (int)len.size() - 1;                        /* levels of products */
    /* resources for every level, sized before the first launch: pools by level parity, planes, descriptors; each
     * product's length class kept for its level's grouping (-1: an empty operand) */
    size_t pool_max[2] = {1, 1}, planes_max = 0, desc_total = 0, img_total = 0, ngroups = 0;
    std::vector<int> lev_np(nlev + 1, 3);
    std::vector<std::vector<int8_t>> keys(nlev + 1);
    std::vector<std::array<uint64_t, TREE_NKEY>> lev_cnt(nlev + 1);
    LenClass lclass;
    for (int k = 1; k <= nlev; k++) {
        size_t tot = 0;
        for (const Len &x : len[k]) tot += x.pn + x.qn;
        pool_max[k & 1] = std::max(pool_max[k & 1], tot);
        const std::vector<Len> &c = len[k - 1];
        const size_t npairs = c.size() / 2;
        size_t mn_max = 1;
        uint64_t *cnt = lev_cnt[k].data(), blk[TREE_NKEY];     /* per class: products, CRT blocks */
        for (int j = 0; j < TREE_NKEY; j++) cnt[j] = blk[j] = 0;
        std::vector<int8_t> &kk = keys[k];
        kk.resize(2 * npairs);
        for (size_t i = 0; i < npairs; i++) {
            const Len &l = c[2 * i], &r = c[2 * i + 1];
            const size_t pr[2][2] = {{l.pn, r.qn}, {l.qn, r.qn}};
            for (int t = 0; t < 2; t++) {
                const size_t na = pr[t][0], nb = pr[t][1];
                if (!na || !nb) { kk[2 * i + t] = -1; continue; }
                mn_max = std::max(mn_max, std::min(na, nb));
                const int key = lclass(na + nb - 1);
                const int logn = key % (NTTM_LOGN_MAX + 1), r3 = key > NTTM_LOGN_MAX;
                const uint64_t npts = (uint64_t)(r3 ? 3 : 1) << logn;
                kk[2 * i + t] = (int8_t)key;
                cnt[key]++;
                blk[key] += nttm_crt_blocks(na + nb + (t == 0 ? 1 : 0), npts <= 1024 ? 1 : 8);
            }
        }
        RCK(choose_np(mn_max, &lev_np[k]));
        for (int key = 0; key < TREE_NKEY; key++) {
            if (!cnt[key]) continue;
            const int logn = key % (NTTM_LOGN_MAX + 1), r3 = key > NTTM_LOGN_MAX;
            const uint64_t npts = (uint64_t)(r3 ? 3 : 1) << logn, cnt_all = cnt[key];
            if (logn > NTTM_LOGN_MAX) return 1;
            for (uint64_t c0 = 0; c0 < cnt_all; c0 += TREE_CHUNK) {   /* the group's chunks (below) */
                const uint64_t cn = std::min<uint64_t>(TREE_CHUNK, cnt_all - c0), tb = r3 ? 3 * cn : cn;
                planes_max = std::max(planes_max, (size_t)(lev_np[k] + 1) * cn * npts * 8);
                const size_t img = align_up(cn * sizeof(CrtDesc), 256) + 2 * align_up(cn * sizeof(LoadDesc), 256) +
                                   align_up(tb * 4, 256) + 256;
                img_total += img;
                desc_total += img + align_up(cn * 8, 256) + align_up(nttm_crt_scratch_words(blk[key]) * 4, 256) + 256;
                ngroups++;
            }
        }
    }
    /* the root's P and Q go straight to p and q when their planned lengths fit there and no operand of the last level
     * lies in them (its operands are pool slots above the leaves) */
    const Len &rl = len[nlev][0];
    bool root_direct = rl.pn <= p_cap && rl.qn <= q_cap;
    if (root_direct && nlev == 1) {
        for (size_t i = 0; i < 2 && root_direct; i++) {
            const nttm_bs_node &L = leaves[i];
            const uint64_t *arr[2] = {L.p, L.q};
            const size_t nn[2] = {L.pn, L.qn};
            for (int t = 0; t < 2; t++)
                for (int o = 0; o < 2; o++) {
                    const uint64_t *d = o ? q : p;
                    const size_t dn = o ? rl.qn : rl.pn;
                    if (nn[t] && dn && arr[t] < d + dn && d < arr[t] + nn[t]) root_direct = false;
                }
        }
    }
    /* grown without waiting for the work queued before the tree (the caller's leaves): the outgrown buffers are freed
     * after the tree's final synchronization (a failed call's, at the next call's start) */
    if (g.ndfree || g.nhfree) { HIPCK(hipDeviceSynchronize()); free_deferred(); }
    RCK(bsp_need(0, pool_max[0], true)); RCK(bsp_need(1, pool_max[1], true));
    RCK(ws_need(planes_max ? planes_max : 256, true));
    RCK(tdesc_need(desc_total + 512, true));
    RCK(tpin_need(img_total + 8 * ngroups + 2048, true));
    if (!g.tcs) {
        HIPCK(hipStreamCreateWithFlags(&g.tcs, hipStreamNonBlocking));
        for (int j = 0; j < 8; j++) HIPCK(hipEventCreateWithFlags(&g.tev[j], hipEventDisableTiming));
    }
    HIPCK(hipStreamSynchronize(g.tcs));                           /* (a failed call's copies) before the image is reused */
    /* the levels */
    std::vector<nttm_bs_node> cur(leaves, leaves + n), nxt;
    std::vector<nttm_prod> prods;
    std::vector<const char *> flag_at;                          /* each group's CRT flag word in device memory */
    size_t dpos = 0, ipos = 0, ncopy = 0;
    for (int k = 1; k <= nlev; k++) {
        const size_t npairs = cur.size() / 2, odd = cur.size() & 1;
        const bool top = k == nlev && root_direct;
        uint64_t *pool = g.bsp[k & 1];
        nxt.assign(npairs + odd, nttm_bs_node());
        prods.resize(2 * npairs);
        size_t off = 0;
        for (size_t i = 0; i < npairs; i++) {                    /* P = P_L Q_R + P_R, Q = Q_L Q_R: Q_R shared */
            const nttm_bs_node &L = cur[2 * i], &R = cur[2 * i + 1];
            uint64_t *pc = top ? p : pool + off; off += L.pn + R.qn + 1;
            uint64_t *qc = top ? q : pool + off; off += L.qn + R.qn;
            nttm_prod &a = prods[2 * i], &b = prods[2 * i + 1];
            a.a = L.p; a.na = L.pn; a.b = R.q; a.nb = R.qn; a.c = pc; a.x = R.pn ? R.p : pc; a.nx = R.pn; a.ncn = 0;
            b.a = L.q; b.na = L.qn; b.b = R.q; b.nb = R.qn; b.c = qc; b.x = nullptr; b.nx = 0; b.ncn = 0;
            nxt[i].p = pc; nxt[i].pn = L.pn + R.qn + 1; nxt[i].q = qc; nxt[i].qn = L.qn + R.qn;
        }
        if (odd) {                                               /* the last node moves up (into this level's pool) */
            const nttm_bs_node &L = cur.back();
            uint64_t *pc = pool + off; off += L.pn;
            uint64_t *qc = pool + off; off += L.qn;
            if (L.pn) HIPCK(hipMemcpyAsync(pc, L.p, L.pn * 8, hipMemcpyDeviceToDevice, s));
            if (L.qn) HIPCK(hipMemcpyAsync(qc, L.q, L.qn * 8, hipMemcpyDeviceToDevice, s));
            nxt[npairs] = L; nxt[npairs].p = pc; nxt[npairs].q = qc;
        }
        /* zero-operand products (c = x, zero-extended); the others by length class */
        const std::vector<int8_t> &kk = keys[k];
        std::vector<std::vector<size_t>> groups(TREE_NKEY);
        for (int key = 0; key < TREE_NKEY; key++) groups[key].reserve(lev_cnt[k][key]);
        for (size_t i = 0; i < prods.size(); i++) {
            if (kk[i] >= 0) { groups[kk[i]].push_back(i); continue; }
            const nttm_prod &pq = prods[i];
            const size_t nc = pq.na + pq.nb + (pq.x ? 1 : 0);
            if (nc) HIPCK(hipMemsetAsync(pq.c, 0, nc * 8, s));
            if (pq.x && pq.nx) HIPCK(hipMemcpyAsync(pq.c, pq.x, pq.nx * 8, hipMemcpyDeviceToDevice, s));
        }
        for (int gi = 0; gi < TREE_NKEY; gi++) {
            const std::vector<size_t> &Gall = groups[gi];
            for (size_t c0 = 0; c0 < Gall.size(); c0 += TREE_CHUNK) {
                const std::vector<size_t> G(Gall.begin() + c0, Gall.begin() + std::min(Gall.size(), c0 + TREE_CHUNK));
                const int r3 = gi > NTTM_LOGN_MAX, logn = gi % (NTTM_LOGN_MAX + 1);
                GroupBuild gb;
                group_build(gb, prods.data(), G, lev_np[k], r3, logn, g.cfg.tile_log, batch_pts(logn));
                if (dpos + gb.desc_bytes > g.tdesc_bytes || ipos + gb.off_ncn > img_total || gb.planes_bytes > g.ws_bytes)
                    return nttm_fail(NTTM_EINVAL, "internal: tree plan exceeds its sizing (level %d)", k);
                char *img = g.tpin + ipos, *dbase = g.tdesc + dpos;
                group_image(gb, (double *)g.ws, img);
                /* the descriptors copied on the side stream, which runs ahead of the launches (their region of the
                 * image and of tdesc is this group's alone): the caller's stream waits on the copy, not behind it */
                hipEvent_t ev = g.tev[ncopy++ % 8];
                HIPCK(hipMemcpyAsync(dbase, img, gb.off_ncn, hipMemcpyHostToDevice, g.tcs));
                HIPCK(hipEventRecord(ev, g.tcs));
                HIPCK(hipStreamWaitEvent(s, ev, 0));
                RCK(group_launch(gb, lev_np[k], (double *)g.ws, dbase, s, false));
                flag_at.push_back(dbase + gb.off_crt + 4);
                ipos = align_up(ipos + gb.off_ncn, 256);
                dpos = align_up(dpos + gb.desc_bytes, 256);
            }
        }
        cur.swap(nxt);
    }
    /* every group's flags and the root's normalized lengths (its leading zero limbs found on the device), then one
     * synchronization */
    const nttm_bs_node root = cur[0];
    char *hf = g.tpin + align_up(ipos, 256);
    uint64_t *hlen = (uint64_t *)(hf + align_up(flag_at.size() * 4, 64));
    if ((size_t)((char *)(hlen + 2) - g.tpin) > g.tpin_bytes || dpos + 16 > g.tdesc_bytes)
        return nttm_fail(NTTM_EINVAL, "internal: tree flags exceed their staging");
    uint64_t *dlen = (uint64_t *)(g.tdesc + dpos);
    RCK(nttm_run_top_nonzero(root.p, root.pn, dlen, s));
    RCK(nttm_run_top_nonzero(root.q, root.qn, dlen + 1, s));
    HIPCK(hipMemcpyAsync(hlen, dlen, 16, hipMemcpyDeviceToHost, s));
    for (size_t j = 0; j < flag_at.size(); j++) HIPCK(hipMemcpyAsync(hf + 4 * j, flag_at[j], 4, hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    free_deferred();
    for (size_t j = 0; j < flag_at.size(); j++)
        if (((const uint32_t *)hf)[j] & 1) return nttm_fail(NTTM_ECHECK, "crt: non-zero carry beyond a top limb (wrong residues)");
    const size_t rp = hlen[0], rq = hlen[1];
    if (rp > p_cap || rq > q_cap)
        return nttm_fail(NTTM_EINVAL, "nttm_bs_tree: the root needs %zu and %zu limbs (capacities %zu, %zu)", rp, rq,
                         p_cap, q_cap);
    if (!root_direct) {
        if (rp) HIPCK(hipMemcpyAsync(p, root.p, rp * 8, hipMemcpyDeviceToDevice, s));
        if (rq) HIPCK(hipMemcpyAsync(q, root.q, rq * 8, hipMemcpyDeviceToDevice, s));
        HIPCK(hipStreamSynchronize(s));
    }
    *pn = rp; *qn = rq;
    return 0;
}
extern "C" int nttm_bs_tree(const nttm_bs_node *leaves, size_t n, uint64_t *p, size_t p_cap, size_t *pn, uint64_t *q,
                            size_t q_cap, size_t *qn, void *stream)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (!leaves || !n || !pn || !qn || (p_cap && !p) || (q_cap && !q)) return nttm_fail(NTTM_EINVAL, "nttm_bs_tree arguments");
    hipStream_t s = (hipStream_t)stream;
    MemMap mm;
    for (size_t i = 0; i < n; i++) {
        if ((leaves[i].pn && !leaves[i].p) || (leaves[i].qn && !leaves[i].q)) return nttm_fail(NTTM_EINVAL, "leaf %zu: null limbs", i);
        if (mm.kind(leaves[i].p, leaves[i].pn * 8) != MEM_DEV || mm.kind(leaves[i].q, leaves[i].qn * 8) != MEM_DEV)
            return nttm_fail(NTTM_EINVAL, "nttm_bs_tree: leaf %zu is not device memory of this device", i);
    }
    if (mm.kind(p, p_cap * 8) != MEM_DEV || mm.kind(q, q_cap * 8) != MEM_DEV)
        return nttm_fail(NTTM_EINVAL, "nttm_bs_tree: p or q is not device memory of this device");
    if (!g_tree_levelwise) {
        const int rc = bs_tree_planned(leaves, n, p, p_cap, pn, q, q_cap, qn, s);
        if (rc <= 0) return rc;
    }
    std::vector<nttm_bs_node> cur(leaves, leaves + n), nxt;
    std::vector<nttm_prod> prods;
    for (int level = 0; cur.size() > 1; level++) {
        const size_t npairs = cur.size() / 2, odd = cur.size() & 1;
        size_t total = 0;                                        /* the parents' limbs (the odd node copied) */
        for (size_t i = 0; i < npairs; i++) total += cur[2 * i].pn + cur[2 * i + 1].qn + 1 + cur[2 * i].qn + cur[2 * i + 1].qn;
        if (odd) total += cur.back().pn + cur.back().qn;
        RCK(bsp_need(level & 1, total ? total : 1));
        uint64_t *pool = g.bsp[level & 1];
        nxt.assign(npairs + odd, nttm_bs_node());
        prods.assign(2 * npairs, nttm_prod());
        size_t off = 0;
        for (size_t i = 0; i < npairs; i++) {                    /* P = P_L Q_R + P_R, Q = Q_L Q_R: Q_R shared */
            const nttm_bs_node &L = cur[2 * i], &R = cur[2 * i + 1];
            if (R.pn > L.pn + R.qn)
                return nttm_fail(NTTM_EINVAL, "nttm_bs_tree: level %d pair %zu: |P_R| = %zu exceeds |P_L| + |Q_R| = %zu",
                                 level, i, R.pn, L.pn + R.qn);
            uint64_t *pc = pool + off; off += L.pn + R.qn + 1;
            uint64_t *qc = pool + off; off += L.qn + R.qn;
            nttm_prod &a = prods[2 * i], &b = prods[2 * i + 1];
            a.a = L.p; a.na = L.pn; a.b = R.q; a.nb = R.qn; a.c = pc; a.x = R.pn ? R.p : nullptr; a.nx = R.pn;
            b.a = L.q; b.na = L.qn; b.b = R.q; b.nb = R.qn; b.c = qc; b.x = nullptr; b.nx = 0;
            if (!a.x) { a.x = pc; a.nx = 0; }                     /* P_R = 0: an empty x keeps the carry limb */
            nxt[i].p = pc; nxt[i].q = qc;
        }
        if (odd) {                                               /* the last node moves up (into this level's pool) */
            const nttm_bs_node &L = cur.back();
            uint64_t *pc = pool + off; off += L.pn;
            uint64_t *qc = pool + off; off += L.qn;
            if (L.pn) HIPCK(hipMemcpyAsync(pc, L.p, L.pn * 8, hipMemcpyDeviceToDevice, s));
            if (L.qn) HIPCK(hipMemcpyAsync(qc, L.q, L.qn * 8, hipMemcpyDeviceToDevice, s));
            nxt[npairs].p = pc; nxt[npairs].pn = L.pn; nxt[npairs].q = qc; nxt[npairs].qn = L.qn;
        }
        RCK(nttm_mul_batch(prods.data(), prods.size(), stream));
        for (size_t i = 0; i < npairs; i++) { nxt[i].pn = prods[2 * i].ncn; nxt[i].qn = prods[2 * i + 1].ncn; }
        cur.swap(nxt);
    }
    const nttm_bs_node &root = cur[0];
    if (root.pn > p_cap || root.qn > q_cap)
        return nttm_fail(NTTM_EINVAL, "nttm_bs_tree: the root needs %zu and %zu limbs (capacities %zu, %zu)", root.pn,
                         root.qn, p_cap, q_cap);
    if (root.pn) HIPCK(hipMemcpyAsync(p, root.p, root.pn * 8, hipMemcpyDeviceToDevice, s));
    if (root.qn) HIPCK(hipMemcpyAsync(q, root.q, root.qn * 8, hipMemcpyDeviceToDevice, s));
    HIPCK(hipStreamSynchronize(s));
    *pn = root.pn; *qn = root.qn;
    return 0;
}

/* ---- transform primitives on canonical uint64 residues ---------------------------------------------------------------- */
/* the arguments of a transform primitive: x holds batch transforms of (3 if r3) 2^logn residues in device memory */
static int prim_check(int prime, const void *x, int logn, size_t batch, int r3 = 0)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (prime < 0 || prime >= NTTM_MAX_PRIMES) return nttm_fail(NTTM_EINVAL, "prime %d", prime);
    if (logn < 1 || logn > NTTM_LOGN_MAX) return nttm_fail(NTTM_EINVAL, "logn %d outside [1, %d]", logn, NTTM_LOGN_MAX);
    if (!x || !batch) return nttm_fail(NTTM_EINVAL, "empty transform");
    DEVCK(x, ((uint64_t)batch << logn) * (r3 ? 3 : 1) * 8, "x");
    return 0;
}
extern "C" int nttm_ntt_fwd(int prime, uint64_t *x, int logn, size_t batch, void *stream)
{
    RCK(prim_check(prime, x, logn, batch));
    hipStream_t s = (hipStream_t)stream;
    const uint64_t cnt = (uint64_t)batch << logn;
    RCK(nttm_run_u64_to_f64((double *)x, cnt, s));
    RCK(fwd_planes(prime, (double *)x, logn, batch, s));
    return nttm_run_f64_to_canon((double *)x, cnt, g.pk[prime], s);
}
extern "C" int nttm_ntt_inv(int prime, uint64_t *x, int logn, size_t batch, void *stream)
{
    RCK(prim_check(prime, x, logn, batch));
    hipStream_t s = (hipStream_t)stream;
    const uint64_t cnt = (uint64_t)batch << logn;
    RCK(nttm_run_u64_to_f64((double *)x, cnt, s));
    RCK(inv_planes(prime, (double *)x, nullptr, NTTM_Y_FULL, logn, batch, 1, s));
    return nttm_run_f64_to_canon((double *)x, cnt, g.pk[prime], s);
}
extern "C" int nttm_ntt_inv_pw(int prime, uint64_t *x, const uint64_t *y, int ymode, int logn, size_t batch, void *stream)
{
    RCK(prim_check(prime, x, logn, batch));
    if (!y) return nttm_fail(NTTM_EINVAL, "null y");
    if (ymode < NTTM_Y_FULL || ymode > NTTM_Y_PAIR) return nttm_fail(NTTM_EINVAL, "ymode %d", ymode);
    hipStream_t s = (hipStream_t)stream;
    const uint64_t cnt = (uint64_t)batch << logn;
    const uint64_t ycnt = ymode == NTTM_Y_FULL ? cnt : ymode == NTTM_Y_BCAST ? ((uint64_t)1 << logn)
                                                     : ((uint64_t)((batch + 1) / 2) << logn);
    DEVCK(y, ycnt * 8, "y");
    RCK(ws_need(ycnt * 8));
    RCK(nttm_run_cvt_u64_f64((double *)g.ws, y, ycnt, s));
    RCK(nttm_run_u64_to_f64((double *)x, cnt, s));
    RCK(inv_planes(prime, (double *)x, (const double *)g.ws, ymode, logn, batch, 1, s));
    return nttm_run_f64_to_canon((double *)x, cnt, g.pk[prime], s);
}

/* the forward transform with its first four levels as one radix-16 stage on INT8 matrix cores (mfma.hip), the
 * remaining levels on the FP64 engine as 16 transforms of n/16 points each: the same canonical output as nttm_ntt_fwd */
extern "C" int nttm_ntt_fwd_mfma(int prime, uint64_t *x, int logn, size_t batch, void *stream)
{
    RCK(prim_check(prime, x, logn, batch));
    if (logn < 8) return nttm_fail(NTTM_EINVAL, "nttm_ntt_fwd_mfma: 2^%d points (at least 2^8)", logn);
    hipStream_t s = (hipStream_t)stream;
    const uint64_t p = NTTM_PRIMES[prime], n = (uint64_t)1 << logn, cnt = (uint64_t)batch << logn;
    const uint64_t w = nttm_root(prime, logn), W = nttm_root(prime, 4);
    /* host tables: the 15 x 2 A operands (30720 bytes), w^i and w^(4096 i), the recombination constants */
    const uint64_t nhi = ((n >> 4) + 4095) / 4096;
    std::vector<uint8_t> et(15 * 2 * 64 * 16, 0);
    uint64_t Wp[16];
    for (int i = 0; i < 16; i++) Wp[i] = h_powmod(W, i, p);
    for (int sh = 0; sh < 15; sh++)
        for (int h = 0; h < 2; h++)
            for (int l = 0; l < 64; l++)
                for (int by = 0; by < 16; by++) {
                    const int k = l & 15, t = 2 * (l >> 4) + by / 8 + 8 * h, b = by % 8, c = sh - b;
                    if (c < 0 || c > 7) continue;
                    et[((sh * 2 + h) * 64 + l) * 16 + by] = (uint8_t)((Wp[(t * k) & 15] >> (7 * c)) & 127);
                }
    std::vector<double> tw(4096 + nhi + 4);
    const uint64_t w4096 = h_powmod(w, 4096, p);
    for (uint64_t i = 0, v = 1; i < 4096; i++, v = h_mulmod(v, w, p)) tw[i] = h_bal(v, p);
    for (uint64_t i = 0, v = 1; i < nhi; i++, v = h_mulmod(v, w4096, p)) tw[4096 + i] = h_bal(v, p);
    const double2 c35 = h_shoup(h_powmod(2, 35, p), p), c70 = h_shoup(h_powmod(2, 70, p), p);
    double *cs = tw.data() + 4096 + nhi;
    cs[0] = c35.x; cs[1] = c35.y; cs[2] = c70.x; cs[3] = c70.y;
    const size_t off_e = align_up(cnt * 8, 256), off_t = align_up(off_e + et.size(), 256);
    RCK(ws_need(off_t + tw.size() * 8 + 256));
    HIPCK(hipMemcpyAsync(g.ws + off_e, et.data(), et.size(), hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(g.ws + off_t, tw.data(), tw.size() * 8, hipMemcpyHostToDevice, s));
    double *y = (double *)g.ws;
    const double *dt = (const double *)(g.ws + off_t);
    RCK(nttm_run_dft16_mfma(x, y, logn, batch, g.ws + off_e, dt, dt + 4096, cs, (double)p, s));   /* cs: host copy */
    RCK(fwd_planes(prime, y, logn - 4, 16 * (uint64_t)batch, s));
    HIPCK(hipMemcpyAsync(x, y, cnt * 8, hipMemcpyDeviceToDevice, s));
    RCK(nttm_run_f64_to_canon((double *)x, cnt, g.pk[prime], s));
    return nttm_hip_check(hipStreamSynchronize(s), "nttm_ntt_fwd_mfma");                 /* the host tables */
}

/* transforms of 3 2^logk points: the radix-3 layer (nttm_run_r3) and the 2^logk engine on the thirds (batch 3) */
extern "C" int nttm_ntt_fwd3(int prime, uint64_t *x, int logk, size_t batch, void *stream)
{
    RCK(prim_check(prime, x, logk, batch, 1));
    hipStream_t s = (hipStream_t)stream;
    const uint64_t cnt = 3 * ((uint64_t)batch << logk);
    RCK(nttm_run_u64_to_f64((double *)x, cnt, s));
    RCK(nttm_run_r3(false, (double *)x, batch, logk, g.pk[prime], g.twf[prime].b3, nullptr, 0, nullptr, 0, nullptr, s));
    RCK(fwd_planes(prime, (double *)x, logk, 3 * (uint64_t)batch, s));
    return nttm_run_f64_to_canon((double *)x, cnt, g.pk[prime], s);
}
/* the inverse: the engine's inverses scale by 2^-logk, the radix-3 layer by 3^-1 */
static int inv3_planes(int prime, double *x, const double *y, int ymode, int logk, uint64_t batch, hipStream_t s)
{
    const uint64_t p = NTTM_PRIMES[prime];
    RCK(inv_planes(prime, x, y, ymode, logk, 3 * batch, 1, s, 1));
    const double2 third = h_shoup(h_inv(3, p), p);
    return nttm_run_r3(true, x, batch, logk, g.pk[prime], g.twi[prime].b3, nullptr, 0, nullptr, 0, &third, s);
}
extern "C" int nttm_ntt_inv3(int prime, uint64_t *x, int logk, size_t batch, void *stream)
{
    RCK(prim_check(prime, x, logk, batch, 1));
    hipStream_t s = (hipStream_t)stream;
    const uint64_t cnt = 3 * ((uint64_t)batch << logk);
    RCK(nttm_run_u64_to_f64((double *)x, cnt, s));
    RCK(inv3_planes(prime, (double *)x, nullptr, NTTM_Y_FULL, logk, batch, s));
    return nttm_run_f64_to_canon((double *)x, cnt, g.pk[prime], s);
}
extern "C" int nttm_ntt_inv3_pw(int prime, uint64_t *x, const uint64_t *y, int ymode, int logk, size_t batch, void *stream)
{
    RCK(prim_check(prime, x, logk, batch, 1));
    if (!y) return nttm_fail(NTTM_EINVAL, "null y");
    if (ymode < NTTM_Y_FULL || ymode > NTTM_Y_PAIR) return nttm_fail(NTTM_EINVAL, "ymode %d", ymode);
    hipStream_t s = (hipStream_t)stream;
    const uint64_t n3 = (uint64_t)3 << logk, cnt = n3 * batch;
    const uint64_t ycnt = ymode == NTTM_Y_FULL ? cnt : ymode == NTTM_Y_BCAST ? n3 : (uint64_t)((batch + 1) / 2) * n3;
    DEVCK(y, ycnt * 8, "y");
    RCK(ws_need(ycnt * 8));
    RCK(nttm_run_cvt_u64_f64((double *)g.ws, y, ycnt, s));
    RCK(nttm_run_u64_to_f64((double *)x, cnt, s));
    RCK(inv3_planes(prime, (double *)x, (const double *)g.ws, ymode, logk, batch, s));
    return nttm_run_f64_to_canon((double *)x, cnt, g.pk[prime], s);
}
extern "C" int nttm_pw(int prime, uint64_t *x, const uint64_t *y, size_t count, void *stream)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (prime < 0 || prime >= NTTM_MAX_PRIMES || !x || !y) return nttm_fail(NTTM_EINVAL, "nttm_pw arguments");
    if (!count) return 0;
    DEVCK(x, count * 8, "x");
    DEVCK(y, count * 8, "y");
    return nttm_run_pw_u64(x, y, count, g.pk[prime], (hipStream_t)stream);
}
extern "C" int nttm_load(int prime, uint64_t *dst, const uint64_t *src, size_t nlimbs, size_t npoints, void *stream)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (prime < 0 || prime >= NTTM_MAX_PRIMES || !dst || (nlimbs && !src) || nlimbs > npoints)
        return nttm_fail(NTTM_EINVAL, "nttm_load arguments");
    if (!npoints) return 0;
    DEVCK(dst, npoints * 8, "dst");
    DEVCK(src, nlimbs * 8, "src");
    return nttm_run_load_planes((double *)dst, npoints, src, nlimbs, npoints, 1, &g.pk[prime], nullptr, 1, (hipStream_t)stream);
}
extern "C" int nttm_crt(uint64_t *c, size_t nc, const uint64_t *const *planes, int np, size_t ncoef, const uint64_t *x,
                        size_t nx, void *stream)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (np < 3 || np > NTTM_MAX_PRIMES || !planes || (nc && !c) || (nx && !x)) return nttm_fail(NTTM_EINVAL, "nttm_crt arguments");
    if (ncoef > nc) return nttm_fail(NTTM_EINVAL, "nttm_crt: ncoef %zu exceeds nc %zu", ncoef, nc);
    if (!nc) return 0;
    DEVCK(c, nc * 8, "c");
    DEVCK(x, nx * 8, "x");
    for (int i = 0; i < np; i++) DEVCK(planes[i], ncoef * 8, "plane");
    hipStream_t s = (hipStream_t)stream;
    const size_t pbytes = align_up((size_t)np * ncoef * 8, 256);
    const size_t crt_words = nttm_crt_scratch_words(nttm_crt_blocks(nc));
    RCK(ws_need(pbytes + 256 + crt_words * 4 + 256));
    double *pl = (double *)g.ws;
    for (int i = 0; i < np; i++) {
        if (!planes[i] && ncoef) return nttm_fail(NTTM_EINVAL, "null plane %d", i);
        if (ncoef) RCK(nttm_run_cvt_u64_f64(pl + (size_t)i * ncoef, planes[i], ncoef, s));
    }
    CrtDesc hd; memset(&hd, 0, sizeof hd);
    hd.pl = pl; hd.plane = ncoef; hd.ncoef = ncoef; hd.x = x; hd.nx = x ? nx : 0; hd.c = c; hd.nc = nc; hd.first_block = 0;
    HIPCK(hipStreamSynchronize(s));
    memcpy(g.hpin, &hd, sizeof hd);
    CrtDesc *dd = (CrtDesc *)(g.ws + pbytes);
    uint32_t *crt_scr = (uint32_t *)(g.ws + pbytes + 256);
    HIPCK(hipMemcpyAsync(dd, g.hpin, sizeof hd, hipMemcpyHostToDevice, s));
    RCK(nttm_crt_launch_batch(dd, 1, nttm_crt_blocks(nc), np, g.cc, crt_scr, s,
                              8.0 * np * (double)ncoef + 8.0 * (double)nc + 8.0 * (double)(x ? nx : 0)));
    HIPCK(hipMemcpyAsync(g.hpin + 64, crt_scr, 8, hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    const uint32_t flags = ((const uint32_t *)(g.hpin + 64))[1];
    if (flags & 1) return nttm_fail(NTTM_ECHECK, "crt: the value needs more than nc limbs");
    return 0;
}

/* ---- live kernel timing ------------------------------------------------------------------------------------------------ */
struct ProfRec { const char *name; double bytes; hipEvent_t a, b; };
static std::vector<ProfRec> g_prof;
static std::vector<hipEvent_t> g_evpool;
static hipEvent_t ev_get(void)
{
    if (!g_evpool.empty()) { hipEvent_t e = g_evpool.back(); g_evpool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
    return e;
}
void prof_begin(const char *name, double alg_bytes, hipStream_t s)
{
    if (!g_prof_on) return;
    ProfRec r; r.name = name; r.bytes = alg_bytes; r.a = ev_get(); r.b = ev_get();
    if (!r.a || !r.b) return;
    (void)hipEventRecord(r.a, s);
    g_prof.push_back(r);
}
void prof_end(hipStream_t s)
{
    if (!g_prof_on || g_prof.empty()) return;
    (void)hipEventRecord(g_prof.back().b, s);
}
static void prof_clear(void)
{
    for (auto &r : g_prof) { g_evpool.push_back(r.a); g_evpool.push_back(r.b); }
    g_prof.clear();
}
static void prof_release(void)
{
    prof_clear();
    for (hipEvent_t e : g_evpool) (void)hipEventDestroy(e);
    g_evpool.clear();
    g_prof_on = 0;
}
extern "C" int nttm_profile(int enable)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    HIPCK(hipDeviceSynchronize());
    prof_clear();
    g_prof_on = enable != 0;
    return 0;
}
extern "C" int nttm_profile_read(nttm_kstat *out, int cap)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    HIPCK(hipDeviceSynchronize());
    int n = 0;
    for (auto &r : g_prof) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) { (void)hipGetLastError(); continue; }
        int k = 0;
        while (k < n && strcmp(out[k].name, r.name)) k++;
        if (k == n) {
            if (n >= cap) continue;
            memset(&out[n], 0, sizeof out[n]);
            snprintf(out[n].name, sizeof out[n].name, "%s", r.name);
            n++;
        }
        out[k].launches++; out[k].total_ms += ms; out[k].alg_bytes += r.bytes;
    }
    return n;
}

/* ---- batch plans: a batch's whole schedule as one hipGraph, each group's schedule chosen by measurement -------------
 * The binary-splitting leaves of the reference go through rns_mul_batch again and again with the same shapes
 * (E/rns_mul.c:792-1028).  A plan fixes one batch: it groups the products, chooses per length class whether the
 * products that fit 3 2^k points take that length or 2^(k+2), and per group the strided-pass tile (2^12 or 2^13
 * points) and the points per thread (8 or 16) by timing every candidate on the batch's own operands; uploads every
 * descriptor once; and captures all groups' launches (transforms, CRT, carry scan, lengths) as one graph that a run
 * launches with one call. */
struct nttm_batch_plan {
    size_t N = 0;
    uint64_t gen = 0;                                            /* the nttm_init it was made under */
    int np = 3;
    std::vector<GroupBuild> gb;
    std::vector<std::vector<size_t>> gidx;
    std::vector<double> gms;                                     /* each group's measured time (ms; 0 untuned) */
    std::vector<size_t> goff;                                    /* descriptor bases in blob */
    std::vector<size_t> zidx;                                    /* products with an empty operand */
    size_t off_z = 0, off_zncn = 0;
    char *blob = nullptr;
    size_t blob_bytes = 0;
    uint64_t *hstage = nullptr;                                  /* pinned: lengths and flags read back */
    size_t hstage_words = 0;
    hipStream_t cs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double tune_ms = 0;
    int candidates = 0;
    int reused = 0;                                              /* groups that took a remembered choice */
};

static void plan_free(nttm_batch_plan *pl)
{
    if (!pl) return;
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->cs) (void)hipStreamDestroy(pl->cs);
    if (pl->blob) (void)hipFree(pl->blob);
    if (pl->hstage) (void)hipHostFree(pl->hstage);
    delete pl;
}

/* a device buffer that grows (the autotuner's trial runs) */
struct TuneBuf {
    char *p = nullptr; size_t bytes = 0;
    ~TuneBuf() { release(); }
    void release() { if (p) { (void)hipDeviceSynchronize(); (void)hipFree(p); } p = nullptr; bytes = 0; }
    int need(size_t b)
    {
        if (b <= bytes) return 0;
        if (p) { (void)hipDeviceSynchronize(); (void)hipFree(p); p = nullptr; bytes = 0; }
        if (hipMalloc(&p, b) != hipSuccess) { (void)hipGetLastError(); return nttm_fail(NTTM_ENOMEM, "plan: %zu bytes", b); }
        bytes = b;
        return 0;
    }
};
/* the autotuner's choices, kept across plans: per prime count, transform length (2^logn, or 3 2^logn with r3) and group
 * size (its products), the (tile, points per thread) that won and its time per product.  Each entry belongs to the
 * nttm_init it was timed under.  Plans created with autotune == 2 take the entry of the same size, else the nearest
 * size of the same power-of-two class, instead of timing the candidates again */
struct TuneMemo { uint64_t gen; int np, r3, logn; uint64_t cnt; int tile, pts; double ms_per; };
#define TUNE_MEMO_MAX 4096
static TuneMemo g_memo[TUNE_MEMO_MAX];
static int g_nmemo = 0;
static const TuneMemo *memo_find(int np, int r3, int logn, uint64_t cnt)
{
    const TuneMemo *best = nullptr;
    uint64_t dbest = 0;
    for (int i = 0; i < g_nmemo; i++) {
        const TuneMemo &m = g_memo[i];
        if (m.gen != g.gen || m.np != np || m.r3 != r3 || m.logn != logn || ceil_log2(m.cnt) != ceil_log2(cnt)) continue;
        const uint64_t d = m.cnt > cnt ? m.cnt - cnt : cnt - m.cnt;
        if (!best || d < dbest) { best = &m; dbest = d; }
    }
    return best;
}
static void memo_put(const TuneMemo &e)
{
    TuneMemo *slot = nullptr;
    for (int i = 0; i < g_nmemo && !slot; i++) {                 /* the same key, or an entry of an earlier nttm_init */
        TuneMemo &m = g_memo[i];
        if (m.gen != g.gen || (m.np == e.np && m.r3 == e.r3 && m.logn == e.logn && m.cnt == e.cnt)) slot = &m;
    }
    if (!slot && g_nmemo < TUNE_MEMO_MAX) slot = &g_memo[g_nmemo++];
    if (slot) *slot = e;                                         /* (a full table keeps its entries) */
}
/* the fastest (tile, points per thread) for one group of products, timed on the stream; *ms its time.  reuse: a
 * remembered choice for the same class is taken untimed (*ms scaled from its time per product; *nreused counts it) */
static int tune_group(const nttm_prod *P, const std::vector<size_t> &G, int np, int r3, int logn, bool tune, TuneBuf &tb,
                      std::vector<char> &img, hipStream_t s, GroupBuild &best, double *ms, int *ntried, bool reuse = false,
                      int *nreused = nullptr)
{
    static const int cand[4][2] = {{12, 16}, {13, 16}, {12, 8}, {13, 8}};
    *ms = 0;
    if (!tune) { group_build(best, P, G, np, r3, logn, g.cfg.tile_log, batch_pts(logn)); return 0; }
    if (reuse) {
        if (const TuneMemo *m = memo_find(np, r3, logn, G.size())) {
            group_build(best, P, G, np, r3, logn, m->tile, m->pts);
            *ms = m->ms_per * (double)G.size();
            if (nreused) (*nreused)++;
            return 0;
        }
    }
    double best_ms = 1e30;
    hipEvent_t e0 = g.ev[0], e1 = g.ev[1];
    for (int c = 0; c < 4; c++) {
        GroupBuild gb;
        group_build(gb, P, G, np, r3, logn, cand[c][0], cand[c][1]);
        const size_t off_d = align_up(gb.planes_bytes, 256);
        RCK(tb.need(off_d + gb.desc_bytes + 256));
        img.resize(gb.desc_bytes);
        group_image(gb, (double *)tb.p, img.data());
        HIPCK(hipMemcpyAsync(tb.p + off_d, img.data(), gb.desc_bytes, hipMemcpyHostToDevice, s));
        RCK(group_launch(gb, np, (double *)tb.p, tb.p + off_d, s));       /* warm */
        double t = 1e30;
        for (int r = 0; r < 2; r++) {
            HIPCK(hipEventRecord(e0, s));
            RCK(group_launch(gb, np, (double *)tb.p, tb.p + off_d, s));
            HIPCK(hipEventRecord(e1, s));
            HIPCK(hipEventSynchronize(e1));
            float f = 0; HIPCK(hipEventElapsedTime(&f, e0, e1));
            t = std::min(t, (double)f);
        }
        (*ntried)++;
        if (t < best_ms) { best_ms = t; best = std::move(gb); }
    }
    *ms = best_ms;
    memo_put(TuneMemo{g.gen, np, r3, logn, (uint64_t)G.size(), best.tile, best.pts, best_ms / (double)G.size()});
    return 0;
}

extern "C" int nttm_batch_plan_create(const nttm_prod *P, size_t N, int autotune, nttm_batch_plan **out, void *stream)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (!out || (N && !P)) return nttm_fail(NTTM_EINVAL, "nttm_batch_plan_create arguments");
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    /* the graph has the operands' addresses built in: device memory only */
    {
        MemMap mm;
        for (size_t i = 0; i < N; i++) {
            const nttm_prod &q = P[i];
            int k[4];
            RCK(prod_kinds(mm, q.a, q.na, q.b, q.nb, q.x, q.nx, q.c, q.na + q.nb + (q.x ? 1 : 0), k, i));
            if (any_host(k)) return nttm_fail(NTTM_EINVAL, "plan: product %zu: host memory (a plan takes device memory)", i);
        }
    }
    nttm_batch_plan *pl = new nttm_batch_plan();
    pl->N = N;
    pl->gen = g.gen;
    auto fail = [&](int rc) { plan_free(pl); return rc; };
    /* validate and class the products by their 2^k length */
    size_t mn_max = 0;
    std::vector<std::vector<size_t>> cls(NTTM_LOGN_MAX + 1);
    for (size_t i = 0; i < N; i++) {
        const nttm_prod &q = P[i];
        if (q.x && q.nx > q.na + q.nb) return fail(nttm_fail(NTTM_EINVAL, "product %zu: nx exceeds na + nb", i));
        if (!q.c || (q.na && !q.a) || (q.nb && !q.b) || (q.nx && !q.x)) return fail(nttm_fail(NTTM_EINVAL, "product %zu: null pointer", i));
        if (!q.na || !q.nb) { pl->zidx.push_back(i); continue; }
        mn_max = std::max(mn_max, (size_t)std::min(q.na, q.nb));
        int logn, r3;
        choose_len(q.na + q.nb - 1, NTTM_R3_NEVER, g.cfg.tile_log, &logn, &r3);
        if (logn > NTTM_LOGN_MAX) return fail(nttm_fail(NTTM_EBOUND, "product %zu: transform above 2^%d", i, NTTM_LOGN_MAX));
        cls[logn].push_back(i);
    }
    int rc = choose_np(mn_max ? mn_max : 1, &pl->np);
    if (rc) return fail(rc);
    const int np = pl->np;
    /* the schedule: per class k the radix-3 subset (products of at most 3 2^(k-2) coefficients) either takes 3 2^(k-2)
     * points or stays at 2^k; per group the tile and points per thread */
    TuneBuf tb;
    std::vector<char> img;
    const bool tune = autotune != 0, reuse = autotune == 2;
    auto t_start = std::chrono::steady_clock::now();
    for (int k = 0; k <= NTTM_LOGN_MAX; k++) {
        const std::vector<size_t> &C = cls[k];
        if (C.empty()) continue;
        std::vector<size_t> S3, S2;
        for (size_t i : C) {
            int logn, r3;
            choose_len(P[i].na + P[i].nb - 1, tune ? NTTM_R3_FIT : g.cfg.radix3, g.cfg.tile_log, &logn, &r3);
            (r3 ? S3 : S2).push_back(i);
        }
        GroupBuild a, b3, b2;
        double ta = 0, t3 = 0, t2 = 0;
        if (tune && !S3.empty()) {                               /* both variants timed */
            if ((rc = tune_group(P, C, np, 0, k, true, tb, img, s, a, &ta, &pl->candidates, reuse, &pl->reused))) return fail(rc);
            if ((rc = tune_group(P, S3, np, 1, k - 2, true, tb, img, s, b3, &t3, &pl->candidates, reuse, &pl->reused)))
                return fail(rc);
            if (!S2.empty() && (rc = tune_group(P, S2, np, 0, k, true, tb, img, s, b2, &t2, &pl->candidates, reuse, &pl->reused)))
                return fail(rc);
            if (ta <= t3 + t2) { pl->gb.push_back(std::move(a)); pl->gidx.push_back(C); pl->gms.push_back(ta); continue; }
        } else if (S3.empty()) {
            if ((rc = tune_group(P, C, np, 0, k, tune, tb, img, s, a, &ta, &pl->candidates, reuse, &pl->reused))) return fail(rc);
            pl->gb.push_back(std::move(a)); pl->gidx.push_back(C); pl->gms.push_back(ta);
            continue;
        } else {                                                 /* untuned: the library's rule */
            if ((rc = tune_group(P, S3, np, 1, k - 2, false, tb, img, s, b3, &t3, &pl->candidates))) return fail(rc);
            if (!S2.empty() && (rc = tune_group(P, S2, np, 0, k, false, tb, img, s, b2, &t2, &pl->candidates))) return fail(rc);
        }
        pl->gb.push_back(std::move(b3)); pl->gidx.push_back(S3); pl->gms.push_back(t3);
        if (!S2.empty()) { pl->gb.push_back(std::move(b2)); pl->gidx.push_back(S2); pl->gms.push_back(t2); }
    }
    pl->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    /* one device blob: the planes (the largest group's; groups run one after another), then every descriptor */
    size_t planes = 0;
    for (const GroupBuild &x : pl->gb) planes = std::max(planes, x.planes_bytes);
    size_t off = align_up(std::max(planes, (size_t)256), 256);
    for (const GroupBuild &x : pl->gb) { pl->goff.push_back(off); off = align_up(off + x.desc_bytes, 256); }
    pl->off_z = off; off = align_up(off + pl->zidx.size() * sizeof(CrtDesc), 256);
    pl->off_zncn = off; off = align_up(off + pl->zidx.size() * 8, 256);
    pl->blob_bytes = off;
    tb.release();
    if (hipMalloc(&pl->blob, pl->blob_bytes) != hipSuccess) {
        (void)hipGetLastError(); pl->blob = nullptr;
        return fail(nttm_fail(NTTM_ENOMEM, "plan: %zu bytes", pl->blob_bytes));
    }
    const size_t first = pl->goff.empty() ? pl->off_z : pl->goff[0];         /* the descriptors, after the planes */
    std::vector<char> himg(pl->blob_bytes - first, 0);
    double *A = (double *)pl->blob;
    for (size_t i = 0; i < pl->gb.size(); i++) group_image(pl->gb[i], A, himg.data() + (pl->goff[i] - first));
    for (size_t z = 0; z < pl->zidx.size(); z++) {
        const nttm_prod &q = P[pl->zidx[z]];
        CrtDesc c{}; c.c = q.c; c.nc = q.na + q.nb + (q.x ? 1 : 0); c.x = q.x; c.nx = q.x ? q.nx : 0;
        memcpy(himg.data() + (pl->off_z - first) + z * sizeof(CrtDesc), &c, sizeof c);
    }
    if ((rc = nttm_hip_check(hipMemcpy(pl->blob + first, himg.data(), himg.size(), hipMemcpyHostToDevice),
                             "plan descriptors"))) return fail(rc);
    size_t words = pl->zidx.size() + 2 * pl->gb.size();
    for (const GroupBuild &x : pl->gb) words += x.cnt;
    pl->hstage_words = words + 1;
    if (hipHostMalloc((void **)&pl->hstage, pl->hstage_words * 8, 0) != hipSuccess) {
        (void)hipGetLastError(); pl->hstage = nullptr; return fail(nttm_fail(NTTM_ENOMEM, "plan: pinned staging"));
    }
    /* capture every group's launches on a stream of the plan's own (default priority) */
    if ((rc = nttm_hip_check(hipStreamCreate(&pl->cs), "plan stream"))) return fail(rc);
    if ((rc = nttm_hip_check(hipStreamSynchronize(s), "plan tuning"))) return fail(rc);
    const int prof0 = g_prof_on;
    g_prof_on = 0;                                               /* no event records inside the graph */
    if ((rc = nttm_hip_check(hipStreamBeginCapture(pl->cs, hipStreamCaptureModeRelaxed), "begin capture"))) { g_prof_on = prof0; return fail(rc); }
    for (size_t i = 0; i < pl->gb.size() && !rc; i++) rc = group_launch(pl->gb[i], np, A, pl->blob + pl->goff[i], pl->cs);
    if (!rc) rc = nttm_run_fill_x(pl->blob + pl->off_z, pl->zidx.size(), pl->cs);
    if (!rc) rc = nttm_run_norm(pl->blob + pl->off_z, pl->zidx.size(), (uint64_t *)(pl->blob + pl->off_zncn), pl->cs);
    hipGraph_t gr = nullptr;
    const int rc_end = nttm_hip_check(hipStreamEndCapture(pl->cs, &gr), "end capture");
    g_prof_on = prof0;
    pl->graph = gr;
    if (rc || rc_end) return fail(rc ? rc : rc_end);
    if ((rc = nttm_hip_check(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0), "graph instantiate"))) return fail(rc);
    *out = pl;
    return 0;
}

extern "C" int nttm_batch_plan_run(nttm_batch_plan *pl, void *stream)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (!pl) return nttm_fail(NTTM_EINVAL, "null plan");
    if (pl->gen != g.gen) return nttm_fail(NTTM_ESTATE, "plan made before the last nttm_shutdown: its tables are gone");
    if (!pl->exec) return 0;
    return nttm_hip_check(hipGraphLaunch(pl->exec, (hipStream_t)stream), "graph launch");
}

extern "C" int nttm_batch_plan_ncn(nttm_batch_plan *pl, size_t *ncn, void *stream)
{
    if (!g.init) return nttm_fail(NTTM_ESTATE, "not initialised");
    if (!pl || !ncn) return nttm_fail(NTTM_EINVAL, "nttm_batch_plan_ncn arguments");
    if (pl->gen != g.gen) return nttm_fail(NTTM_ESTATE, "plan made before the last nttm_shutdown: its tables are gone");
    hipStream_t s = (hipStream_t)stream;
    uint64_t *h = pl->hstage;
    size_t w = 0;
    std::vector<size_t> at(pl->gb.size());
    for (size_t i = 0; i < pl->gb.size(); i++) {
        const GroupBuild &x = pl->gb[i];
        at[i] = w;
        HIPCK(hipMemcpyAsync(h + w, pl->blob + pl->goff[i] + x.off_ncn, x.cnt * 8, hipMemcpyDeviceToHost, s));
        w += x.cnt;
        HIPCK(hipMemcpyAsync(h + w, pl->blob + pl->goff[i] + x.off_crt, 8, hipMemcpyDeviceToHost, s));
        w += 2;
    }
    if (!pl->zidx.empty()) HIPCK(hipMemcpyAsync(h + w, pl->blob + pl->off_zncn, pl->zidx.size() * 8, hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    for (size_t i = 0; i < pl->gb.size(); i++) {
        const GroupBuild &x = pl->gb[i];
        const uint32_t flags = ((const uint32_t *)(h + at[i] + x.cnt))[1];
        if (flags & 1) return nttm_fail(NTTM_ECHECK, "crt: non-zero carry beyond a top limb (wrong residues)");
        for (uint64_t j = 0; j < x.cnt; j++) ncn[pl->gidx[i][j]] = h[at[i] + j];
    }
    for (size_t z = 0; z < pl->zidx.size(); z++) ncn[pl->zidx[z]] = h[w + z];
    return 0;
}

/* the chosen schedule as text: one line per group, then the tuning summary */
extern "C" int nttm_batch_plan_info(const nttm_batch_plan *pl, char *buf, size_t cap)
{
    if (!pl || !buf || !cap) return nttm_fail(NTTM_EINVAL, "nttm_batch_plan_info arguments");
    size_t k = 0;
    buf[0] = 0;
    for (size_t i = 0; i < pl->gb.size() && k < cap; i++) {
        const GroupBuild &x = pl->gb[i];
        k += snprintf(buf + k, cap - k, "group points=%llu r3=%d logn=%d products=%llu shared_b=%llu tile=%d pts=%d ms=%.4f\n",
                      (unsigned long long)x.n, x.r3, x.logn, (unsigned long long)x.cnt,
                      (unsigned long long)(x.share ? x.cnt - x.cnt_b : 0), x.tile, x.pts, pl->gms[i]);
    }
    if (k < cap) k += snprintf(buf + k, cap - k, "plan products=%zu empty=%zu primes=%d candidates=%d reused=%d tune_ms=%.1f "
                               "blob_bytes=%zu\n", pl->N, pl->zidx.size(), pl->np, pl->candidates, pl->reused, pl->tune_ms,
                               pl->blob_bytes);
    return k < cap ? 0 : nttm_fail(NTTM_EINVAL, "nttm_batch_plan_info: buffer too small");
}

extern "C" void nttm_batch_plan_destroy(nttm_batch_plan *pl)
{
    if (pl && pl->cs) (void)hipStreamSynchronize(pl->cs);
    (void)hipDeviceSynchronize();                                /* no replay still reading the plan's memory */
    plan_free(pl);
}

/* ---- host-side introspection for tests: the pass plan and the tile -> element map ---------------------------------- */
extern "C" int nttm_debug_plan(int logn, uint64_t batch, int tile_log, int last_log, int *out, int cap)
{
    if (tile_log <= 0) tile_log = NTTM_TILE_LOG_DEFAULT;
    if (tile_log > NTTM_TILE_LOG || last_log > NTTM_TILE_LOG) return nttm_fail(NTTM_EINVAL, "tile logs above %d", NTTM_TILE_LOG);
    Plan pl; plan_make2(&pl, logn, batch, tile_log, last_log > 0 ? last_log : tile_log);
    int k = 0;
    for (int i = 0; i < pl.npass && k + 4 <= cap; i++) { out[k++] = pl.p[i].lo; out[k++] = pl.p[i].m; out[k++] = pl.p[i].logw; out[k++] = pl.p[i].logh; }
    return pl.npass;
}
extern "C" int nttm_debug_product_plan(int logn, uint64_t batch, int tile_log, int *out, int cap)
{
    if (tile_log <= 0) tile_log = NTTM_TILE_LOG_DEFAULT;
    if (tile_log != 12 && tile_log != 13) return nttm_fail(NTTM_EINVAL, "tile_log %d", tile_log);
    if (logn < 1 || logn > NTTM_LOGN_MAX || !batch) return nttm_fail(NTTM_EINVAL, "nttm_debug_product_plan arguments");
    Plan pl; product_plan(&pl, logn, batch, product_tile_log(logn, tile_log));
    int k = 0;
    for (int i = 0; i < pl.npass && k + 4 <= cap; i++) { out[k++] = pl.p[i].lo; out[k++] = pl.p[i].m; out[k++] = pl.p[i].logw; out[k++] = pl.p[i].logh; }
    return pl.npass;
}
extern "C" int nttm_debug_product_len(size_t ncoef, int radix3, int tile_log, int *logn, int *r3)
{
    if (!ncoef || !logn || !r3) return nttm_fail(NTTM_EINVAL, "nttm_debug_product_len arguments");
    if (radix3 < NTTM_R3_AUTO || radix3 > NTTM_R3_FIT) return nttm_fail(NTTM_EINVAL, "radix3 mode %d", radix3);
    if (tile_log <= 0) tile_log = NTTM_TILE_LOG_DEFAULT;
    if (tile_log != 12 && tile_log != 13) return nttm_fail(NTTM_EINVAL, "tile_log %d", tile_log);
    choose_len(ncoef, radix3, tile_log, logn, r3);
    return 0;
}
extern "C" int nttm_debug_tile_indices(int lo, int m, int logw, int logh, uint64_t tile, uint64_t *out)
{
    PassGeom gg; gg.lo = lo; gg.m = m; gg.logw = logw; gg.logh = logh;
    const int T = 1 << (m + logw + logh);
    for (int e = 0; e < T; e++) out[e] = plan_gidx(gg, tile, e);
    return T;
}
