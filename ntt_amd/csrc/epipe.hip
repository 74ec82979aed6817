/* epipe.hip - digits of e on one GPU through the product library: the caller side of the hot path (binary splitting
 * with batched products, then a Newton reciprocal and a corrected quotient, every product on the device).
 *
 * The reference's pipeline (E/ALGORITHM.md:12-20,59-94) sums 1/k! by binary splitting into P/Q, takes a Newton
 * reciprocal of Q with cut products (E/newton_db.c) and divides.  Here, in base B = 10^18 (the quotient's limbs are the
 * decimal digits, 18 per limb, with no radix conversion):
 *   1. N terms with N! > B^(m+2).  The term k has (P_k, Q_k) = (1, k + 1).  k_e_seed folds runs of G consecutive terms
 *      into one leaf per thread (P <- P f + 1, Q <- Q f for f = k + 1, base-10^9 digits in registers); nttm_bs_tree
 *      combines the leaves into P = sum_{k=1..N} N!/k!, Q = N!;
 *   2. e - 2 = P/Q - 1 + tail, so the fraction is A / Q with A = P - Q < Q;
 *   3. x ~ B^(n+m) / Q by Newton's iteration with doubling precision (ebig.h), n = the limbs of Q;
 *   4. Y = floor(A x / B^n), corrected by the remainder A B^m - Y Q to floor(A B^m / Q).
 * The series tail after N terms is below 2/(N+1)! < B^-(m+2), so Y is e's fraction to m limbs unless the m+1-th and
 * m+2-th limbs of the exact fraction are all B - 1 (which the caller's check windows would catch).
 * Linear host work (additions, subtractions, shifts) runs on the host; every product runs on the device (nttm_mul). */
#include <math.h>
#include <string.h>
#include <chrono>
#include <vector>
#include "internal.h"
#include "ebig.h"

#define E_SEED_WD 64              /* base-10^9 digits one thread holds per number */
#define E_SEED_GMAX 64            /* terms per leaf at most */
#define E_D9 1000000000ULL

/* leaf i folds the terms [i G, min((i + 1) G, nterms)): P = sum_j prod_{k > j} (k + 1), Q = prod (k + 1); written as
 * wl limbs each at pq + 2 wl i (P) and pq + 2 wl i + wl (Q), normalized lengths at lens[2 i], lens[2 i + 1].  The host
 * bounds the digits (G log10(nterms + 1) + 3 < E_SEED_WD) and nterms < 10^9 - 1, so one new digit per term suffices. */
__global__ __launch_bounds__(256) void k_e_seed(uint64_t nterms, uint32_t G, uint64_t nleaf, uint32_t wl, uint64_t *pq,
                                                uint32_t *lens)
{
    const uint64_t leaf = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= nleaf) return;
    const uint64_t k0 = leaf * G, k1 = k0 + G < nterms ? k0 + G : nterms;
    uint32_t p[E_SEED_WD], q[E_SEED_WD];
#pragma unroll
    for (int j = 0; j < E_SEED_WD; j++) { p[j] = 0; q[j] = 0; }
    p[0] = 1;
    q[0] = (uint32_t)(k0 + 1);
    int len = 1;                                                   /* digits in use */
    for (uint64_t k = k0 + 1; k < k1; k++) {
        const uint64_t f = k + 1;
        uint64_t cp = 1, cq = 0;
        int grow = 0;
#pragma unroll
        for (int j = 0; j < E_SEED_WD; j++) {
            if (j <= len) {                                        /* digit len takes the carry out of the top */
                const uint64_t tp = (uint64_t)p[j] * f + cp, tq = (uint64_t)q[j] * f + cq;
                cp = tp / E_D9; cq = tq / E_D9;
                p[j] = (uint32_t)(tp - cp * E_D9); q[j] = (uint32_t)(tq - cq * E_D9);
                if (j == len && (p[j] | q[j])) grow = 1;
            }
        }
        len += grow;
    }
    uint64_t *po = pq + 2 * (uint64_t)wl * leaf, *qo = po + wl;
    uint32_t pn = 0, qn = 0;
#pragma unroll
    for (int i = 0; i < E_SEED_WD / 2; i++) {
        if ((uint32_t)i < wl) {
            const uint64_t lp = p[2 * i] + (uint64_t)p[2 * i + 1] * E_D9, lq = q[2 * i] + (uint64_t)q[2 * i + 1] * E_D9;
            po[i] = lp; qo[i] = lq;
            if (lp) pn = i + 1;
            if (lq) qn = i + 1;
        }
    }
    lens[2 * leaf] = pn;
    lens[2 * leaf + 1] = qn;
}

/* a b on the device, the operands copied from and the product back to the host */
struct DevMul {
    uint64_t *da = nullptr, *db = nullptr, *dc = nullptr;
    size_t cap = 0;
    hipStream_t s;
    ~DevMul() { release(); }
    void release()
    {
        if (da) (void)hipFree(da);
        if (db) (void)hipFree(db);
        if (dc) (void)hipFree(dc);
        da = db = dc = nullptr; cap = 0;
    }
    int grow(size_t limbs)
    {
        if (limbs <= cap) return 0;
        release();
        const size_t c = limbs + limbs / 4 + 64;
        if (hipMalloc(&da, c * 8) != hipSuccess || hipMalloc(&db, c * 8) != hipSuccess || hipMalloc(&dc, 2 * c * 8) != hipSuccess) {
            (void)hipGetLastError(); release();
            return nttm_fail(NTTM_ENOMEM, "e pipeline: device buffers of %zu limbs", c);
        }
        cap = c;
        return 0;
    }
    int operator()(const Big &a, const Big &b, Big &c)
    {
        if (a.empty() || b.empty()) { c.clear(); return 0; }
        int rc = grow(std::max(a.size(), b.size()));
        if (rc) return rc;
        if ((rc = nttm_hip_check(hipMemcpyAsync(da, a.data(), a.size() * 8, hipMemcpyHostToDevice, s), "copy"))) return rc;
        if ((rc = nttm_hip_check(hipMemcpyAsync(db, b.data(), b.size() * 8, hipMemcpyHostToDevice, s), "copy"))) return rc;
        if ((rc = nttm_mul(dc, da, a.size(), db, b.size(), s))) return rc;
        c.resize(a.size() + b.size());
        if ((rc = nttm_hip_check(hipMemcpyAsync(c.data(), dc, c.size() * 8, hipMemcpyDeviceToHost, s), "copy"))) return rc;
        if ((rc = nttm_hip_check(hipStreamSynchronize(s), "sync"))) return rc;
        big_norm(c);
        return 0;
    }
};

static double g_e_ms[4];                                          /* seed, tree, reciprocal, quotient (wall ms) */
typedef std::chrono::steady_clock eclk;
static double ms_since(eclk::time_point t0) { return std::chrono::duration<double, std::milli>(eclk::now() - t0).count(); }

/* terms per leaf: the largest power of two up to E_SEED_GMAX whose digits fit a thread's registers */
static int seed_shape(size_t N, uint32_t *G, uint32_t *wl)
{
    if ((double)N + 2 >= (double)E_D9) return nttm_fail(NTTM_EBOUND, "e pipeline: %zu terms (factors must stay below 10^9)", N);
    const double lg = log10((double)N + 1.0);
    for (uint32_t g = E_SEED_GMAX; g >= 1; g /= 2) {
        const uint32_t digits = (uint32_t)ceil(((double)g * lg + 1.0) / 9.0) + 2;      /* Q's digits, P's extra, one spare */
        if (digits + 1 < E_SEED_WD) { *G = g; *wl = (digits + 1) / 2; return 0; }
    }
    return nttm_fail(NTTM_EBOUND, "e pipeline: no leaf shape for %zu terms", N);
}

/* P, Q of the N-term series on the host */
static int e_series(size_t N, Big &P, Big &Q, hipStream_t s)
{
    uint32_t G, wl;
    RCK(seed_shape(N, &G, &wl));
    const uint64_t nleaf = (N + G - 1) / G;
    uint64_t *pq = nullptr, *dp = nullptr, *dq = nullptr;
    uint32_t *lens = nullptr;
    int rc = 0;
    auto t0 = eclk::now();
    if (hipMalloc(&pq, nleaf * 2 * wl * 8) != hipSuccess || hipMalloc(&lens, nleaf * 2 * 4) != hipSuccess) {
        (void)hipGetLastError(); rc = nttm_fail(NTTM_ENOMEM, "e pipeline: %llu leaves", (unsigned long long)nleaf);
    }
    std::vector<uint32_t> hl;
    if (!rc) {
        hipLaunchKernelGGL(k_e_seed, dim3((unsigned)((nleaf + 255) / 256)), dim3(256), 0, s, (uint64_t)N, G, nleaf, wl, pq, lens);
        rc = nttm_hip_check(hipGetLastError(), "k_e_seed launch");
    }
    if (!rc) {
        hl.resize(2 * nleaf);
        rc = nttm_hip_check(hipMemcpyAsync(hl.data(), lens, nleaf * 2 * 4, hipMemcpyDeviceToHost, s), "copy leaf lengths");
        if (!rc) rc = nttm_hip_check(hipStreamSynchronize(s), "k_e_seed");
    }
    g_e_ms[0] = ms_since(t0);
    t0 = eclk::now();
    size_t cap = 2;
    std::vector<nttm_bs_node> leaves;
    if (!rc) {
        leaves.resize(nleaf);
        for (uint64_t i = 0; i < nleaf; i++) {
            leaves[i].p = pq + 2 * wl * i; leaves[i].pn = hl[2 * i];
            leaves[i].q = pq + 2 * wl * i + wl; leaves[i].qn = hl[2 * i + 1];
            cap += hl[2 * i + 1];                                  /* |Q| <= sum of the leaves' |Q|; P < 2 Q */
        }
        if (hipMalloc(&dp, cap * 8) != hipSuccess || hipMalloc(&dq, cap * 8) != hipSuccess) {
            (void)hipGetLastError(); rc = nttm_fail(NTTM_ENOMEM, "e pipeline: P and Q of %zu limbs", cap);
        }
    }
    if (!rc) {
        size_t pn = 0, qn = 0;
        rc = nttm_bs_tree(leaves.data(), nleaf, dp, cap, &pn, dq, cap, &qn, s);
        if (!rc) {
            P.resize(pn); Q.resize(qn);
            rc = nttm_hip_check(hipMemcpy(P.data(), dp, pn * 8, hipMemcpyDeviceToHost), "copy P");
            if (!rc) rc = nttm_hip_check(hipMemcpy(Q.data(), dq, qn * 8, hipMemcpyDeviceToHost), "copy Q");
        }
    }
    if (pq) (void)hipFree(pq);
    if (lens) (void)hipFree(lens);
    if (dp) (void)hipFree(dp);
    if (dq) (void)hipFree(dq);
    g_e_ms[1] = ms_since(t0);
    big_norm(P); big_norm(Q);
    return rc;
}

extern "C" int nttm_e_fraction(uint64_t *out, size_t m, void *stream)
{
    if (!out || !m) return nttm_fail(NTTM_EINVAL, "nttm_e_fraction arguments");
    hipStream_t s = (hipStream_t)stream;
    for (double &v : g_e_ms) v = 0;
    /* N: log(N!) > 18 (m + 2) log 10 */
    const double need = 18.0 * (double)(m + 2) * log(10.0);
    size_t N = 2;
    while (lgamma((double)N + 1.0) < need) N += N / 2 + 1;
    { size_t lo = N / 2, hi = N; while (lo + 1 < hi) { size_t md = (lo + hi) / 2; if (lgamma((double)md + 1.0) < need) lo = md; else hi = md; } N = hi; }
    Big P, Q;
    RCK(e_series(N, P, Q, s));
    if (big_cmp(P, Q) <= 0) return nttm_fail(NTTM_ECHECK, "e pipeline: P <= Q");
    const Big A = big_sub(P, Q);                                  /* e - 2 ~ A / Q, A < Q */
    P.clear(); P.shrink_to_fit();
    DevMul dm; dm.s = s;
    Big x, Y;
    auto t0 = eclk::now();
    int rc = eb_recip(dm, Q, m, x);                               /* x ~ B^(n+m) / Q */
    if (rc) return rc > 0 ? rc : nttm_fail(NTTM_ECHECK, "e pipeline: reciprocal invariant broken");
    g_e_ms[2] = ms_since(t0);
    t0 = eclk::now();
    rc = eb_quotient(dm, A, Q, m, x, Y);                          /* floor(A B^m / Q) */
    if (rc) return rc > 0 ? rc : nttm_fail(NTTM_ECHECK, "e pipeline: quotient correction did not converge");
    g_e_ms[3] = ms_since(t0);
    if (Y.size() > m) return nttm_fail(NTTM_ECHECK, "e pipeline: fraction of %zu limbs above %zu", Y.size(), m);
    memset(out, 0, m * 8);
    memcpy(out, Y.data(), Y.size() * 8);
    return 0;
}

extern "C" int nttm_e_last_phase_ms(double *seed_ms, double *tree_ms, double *recip_ms, double *quot_ms)
{
    if (seed_ms) *seed_ms = g_e_ms[0];
    if (tree_ms) *tree_ms = g_e_ms[1];
    if (recip_ms) *recip_ms = g_e_ms[2];
    if (quot_ms) *quot_ms = g_e_ms[3];
    return 0;
}
