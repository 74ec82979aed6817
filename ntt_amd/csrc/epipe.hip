/* epipe.hip - digits of e on one GPU through the product library: the caller side of the hot path (binary splitting
 * with batched products, then a Newton reciprocal and a corrected quotient, all products on the device).
 *
 * The reference's pipeline (E/ALGORITHM.md:12-20,59-94) sums 1/k! by binary splitting into P/Q, takes a Newton
 * reciprocal of Q with cut products (E/newton_db.c) and divides.  Here, in base B = 10^18 (so the quotient's limbs are
 * the decimal digits, 18 per limb, with no radix conversion):
 *   1. N terms with N! > B^(m+2): leaves (P_k, Q_k) = (1, k + 1), nttm_bs_tree -> P = sum_{k=1..N} N!/k!, Q = N!;
 *   2. e - 2 = P/Q - 1 + tail, so the fraction is A / Q with A = P - Q < Q;
 *   3. x = floor(B^K / Q), K = n + m (n = the limbs of Q), by Newton's iteration x <- x + x (B^K - Q x) / B^K at full
 *      length (each step two device products; the number of correct limbs doubles), then made exact by remainders;
 *   4. Y = floor(A x / B^n), corrected by the remainder A B^m - Y Q to floor(A B^m / Q).
 * The series tail after N terms is below 1/N! < B^-(m+2), so Y is e's fraction to m limbs unless the m+1-th and m+2-th
 * limbs of the exact fraction are all B - 1 (which the caller's check windows would catch).
 * Host code: additions, subtractions and comparisons of limb vectors are linear and run on the host; every product
 * runs on the device (nttm_mul). */
#include <math.h>
#include <string.h>
#include <vector>
#include "internal.h"

typedef std::vector<uint64_t> Big;                 /* base-10^18 limbs, least significant first, normalized */

static void big_norm(Big &a) { while (!a.empty() && !a.back()) a.pop_back(); }
static int big_cmp(const Big &a, const Big &b)
{
    if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
    for (size_t i = a.size(); i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
static Big big_add(const Big &a, const Big &b)
{
    Big r(std::max(a.size(), b.size()) + 1, 0);
    uint64_t c = 0;
    for (size_t i = 0; i < r.size(); i++) {
        uint64_t s = c + (i < a.size() ? a[i] : 0) + (i < b.size() ? b[i] : 0);
        c = s >= NTTM_B; r[i] = c ? s - NTTM_B : s;
    }
    big_norm(r);
    return r;
}
static Big big_sub(const Big &a, const Big &b)               /* a >= b */
{
    Big r(a.size(), 0);
    uint64_t br = 0;
    for (size_t i = 0; i < a.size(); i++) {
        const uint64_t s = (i < b.size() ? b[i] : 0) + br;
        if (a[i] >= s) { r[i] = a[i] - s; br = 0; } else { r[i] = a[i] + NTTM_B - s; br = 1; }
    }
    big_norm(r);
    return r;
}
static Big big_shr(const Big &a, size_t k) { return a.size() > k ? Big(a.begin() + k, a.end()) : Big(); }
static Big big_shl(const Big &a, size_t k)
{
    if (a.empty()) return a;
    Big r(k, 0);
    r.insert(r.end(), a.begin(), a.end());
    return r;
}
static Big big_small(uint64_t v) { Big r; if (v) r.push_back(v); return r; }

/* a b on the device */
struct DevMul {
    uint64_t *da = nullptr, *db = nullptr, *dc = nullptr;
    size_t cap = 0;
    hipStream_t s;
    ~DevMul() { if (da) (void)hipFree(da); if (db) (void)hipFree(db); if (dc) (void)hipFree(dc); }
    int grow(size_t limbs)
    {
        if (limbs <= cap) return 0;
        if (da) { (void)hipFree(da); (void)hipFree(db); (void)hipFree(dc); da = db = dc = nullptr; }
        cap = limbs + limbs / 4 + 64;
        if (hipMalloc(&da, cap * 8) != hipSuccess || hipMalloc(&db, cap * 8) != hipSuccess || hipMalloc(&dc, 2 * cap * 8) != hipSuccess) {
            (void)hipGetLastError(); return nttm_fail(NTTM_ENOMEM, "e pipeline: device buffers of %zu limbs", cap);
        }
        return 0;
    }
    int mul(const Big &a, const Big &b, Big &c)
    {
        if (a.empty() || b.empty()) { c.clear(); return 0; }
        int rc = grow(std::max(a.size(), b.size()));
        if (rc) return rc;
        if ((rc = nttm_hip_check(hipMemcpyAsync(da, a.data(), a.size() * 8, hipMemcpyHostToDevice, s), "copy"))) return rc;
        if ((rc = nttm_hip_check(hipMemcpyAsync(db, b.data(), b.size() * 8, hipMemcpyHostToDevice, s), "copy"))) return rc;
        if ((rc = nttm_mul(dc, da, a.size(), db, b.size(), s))) return rc;
        c.assign(a.size() + b.size(), 0);
        if ((rc = nttm_hip_check(hipMemcpyAsync(c.data(), dc, c.size() * 8, hipMemcpyDeviceToHost, s), "copy"))) return rc;
        if ((rc = nttm_hip_check(hipStreamSynchronize(s), "sync"))) return rc;
        big_norm(c);
        return 0;
    }
};

/* floor(B^K / Q) for Q > 0 of n limbs, K >= n */
static int recip(DevMul &dm, const Big &Q, size_t K, Big &x)
{
    const size_t n = Q.size();
    /* the first limb from the top two of Q: B^2 / (Q_top + 1) <= B^K / Q scaled by B^(n - 2 - K) */
    const long double qt = (long double)Q[n - 1] + (n >= 2 ? (long double)Q[n - 2] / 1e18L : 0.0L);
    const long double t = 1e18L / (qt + 1e-18L);                 /* B^(n-1) / Q to about 18 digits, in (1/B, 1] x B */
    uint64_t hi = (uint64_t)(t * 0.999999L);                     /* below the true value */
    if (!hi) hi = 1;
    x = big_shl(big_small(hi), K - n);                           /* ~ B^(K-n) B^(n-1) / Q ... within a factor 1 +- 1e-15 */
    Big BK = big_shl(big_small(1), K), T, E, D;
    for (int it = 0; it < 80; it++) {                            /* Newton: x += x (B^K - Q x) / B^K */
        int rc = dm.mul(Q, x, T);
        if (rc) return rc;
        const int sg = big_cmp(BK, T);
        E = sg >= 0 ? big_sub(BK, T) : big_sub(T, BK);
        if (big_cmp(E, Q) < 0) break;                            /* |B^K - Q x| < Q: x is within 1 of the floor */
        if ((rc = dm.mul(x, E, D))) return rc;
        Big step = big_shr(D, K);
        if (step.empty()) step = big_small(1);
        x = sg >= 0 ? big_add(x, step) : big_sub(x, step);
    }
    /* exact: 0 <= B^K - Q x < Q */
    Big T2;
    int rc = dm.mul(Q, x, T2);
    if (rc) return rc;
    while (big_cmp(T2, BK) > 0) { x = big_sub(x, big_small(1)); T2 = big_sub(T2, Q); }
    Big R = big_sub(BK, T2);
    while (big_cmp(R, Q) >= 0) { x = big_add(x, big_small(1)); R = big_sub(R, Q); }
    return 0;
}

extern "C" int nttm_e_fraction(uint64_t *out, size_t m, void *stream)
{
    if (!out || !m) return nttm_fail(NTTM_EINVAL, "nttm_e_fraction arguments");
    hipStream_t s = (hipStream_t)stream;
    /* N: log10(N!) > 18 (m + 2) */
    const double need = 18.0 * (double)(m + 2) * log(10.0);
    size_t N = 2;
    while (lgamma((double)N + 1.0) < need) N += N / 2 + 1;
    { size_t lo = N / 2, hi = N; while (lo + 1 < hi) { size_t md = (lo + hi) / 2; if (lgamma((double)md + 1.0) < need) lo = md; else hi = md; } N = hi; }
    /* leaves on the device: P_k = 1, Q_k = k + 1 */
    std::vector<uint64_t> hl(2 * N);
    for (size_t k = 0; k < N; k++) { hl[2 * k] = 1; hl[2 * k + 1] = k + 1; }
    uint64_t *dl = nullptr, *dp = nullptr, *dq = nullptr;
    const size_t cap = N + 64;                                    /* P, Q < N! N: at most N limbs (each factor < B) */
    int rc = 0;
    if (hipMalloc(&dl, 2 * N * 8) != hipSuccess || hipMalloc(&dp, cap * 8) != hipSuccess || hipMalloc(&dq, cap * 8) != hipSuccess) {
        (void)hipGetLastError(); rc = nttm_fail(NTTM_ENOMEM, "e pipeline: leaves of %zu terms", N);
    }
    Big P, Q;
    if (!rc) rc = nttm_hip_check(hipMemcpy(dl, hl.data(), 2 * N * 8, hipMemcpyHostToDevice), "copy leaves");
    if (!rc) {
        std::vector<nttm_bs_node> leaves(N);
        for (size_t k = 0; k < N; k++) { leaves[k].p = dl + 2 * k; leaves[k].pn = 1; leaves[k].q = dl + 2 * k + 1; leaves[k].qn = 1; }
        size_t pn = 0, qn = 0;
        rc = nttm_bs_tree(leaves.data(), N, dp, cap, &pn, dq, cap, &qn, stream);
        if (!rc) {
            P.resize(pn); Q.resize(qn);
            rc = nttm_hip_check(hipMemcpy(P.data(), dp, pn * 8, hipMemcpyDeviceToHost), "copy P");
            if (!rc) rc = nttm_hip_check(hipMemcpy(Q.data(), dq, qn * 8, hipMemcpyDeviceToHost), "copy Q");
        }
    }
    if (dl) (void)hipFree(dl);
    if (dp) (void)hipFree(dp);
    if (dq) (void)hipFree(dq);
    if (rc) return rc;
    big_norm(P); big_norm(Q);
    if (big_cmp(P, Q) <= 0) return nttm_fail(NTTM_ECHECK, "e pipeline: P <= Q");
    const Big A = big_sub(P, Q);                                  /* e - 2 ~ A / Q, A < Q */
    const size_t n = Q.size();
    DevMul dm; dm.s = s;
    Big x;
    if ((rc = recip(dm, Q, n + m, x))) return rc;                /* floor(B^(n+m) / Q) */
    Big Y, T;
    if ((rc = dm.mul(A, x, T))) return rc;
    Y = big_shr(T, n);                                            /* within a few units of floor(A B^m / Q) */
    const Big ABm = big_shl(A, m);
    if ((rc = dm.mul(Y, Q, T))) return rc;
    while (big_cmp(T, ABm) > 0) { Y = big_sub(Y, big_small(1)); T = big_sub(T, Q); }
    Big R = big_sub(ABm, T);
    while (big_cmp(R, Q) >= 0) { Y = big_add(Y, big_small(1)); R = big_sub(R, Q); }
    if (Y.size() > m) return nttm_fail(NTTM_ECHECK, "e pipeline: fraction of %zu limbs above %zu", Y.size(), m);
    memset(out, 0, m * 8);
    memcpy(out, Y.data(), Y.size() * 8);
    return 0;
}
