/* epipe.hip - digits of e on one GPU through the product library: the caller side of the hot path (binary splitting
 * with batched products, then a Newton reciprocal and a corrected quotient, every number in device memory).
 *
 * The reference's pipeline (E/ALGORITHM.md:12-20,59-94) sums 1/k! by binary splitting into P/Q, takes a Newton
 * reciprocal of Q with cut products (E/newton_db.c) and divides.  Here, in base B = 10^18 (the quotient's limbs are the
 * decimal digits, 18 per limb, with no radix conversion):
 *   1. N terms with N! > B^(m+2).  The term k has (P_k, Q_k) = (1, k + 1).  k_e_seed folds runs of G consecutive terms
 *      into one leaf per thread (P <- P f + 1, Q <- Q f for f = k + 1, base-10^9 digits in registers); nttm_bs_tree
 *      combines the leaves into P = sum_{k=1..N} N!/k!, Q = N!;
 *   2. e - 2 = P/Q - 1 + tail, so the fraction is A / Q with A = P - Q < Q;
 *   3. x ~ B^(n+m+1) / Q by Newton's iteration with doubling precision (ebig.h), n = the limbs of Q, one guard limb;
 *   4. Y = floor(A x / B^(n+1)), exact unless limb n of A x lies within 3 of a limb boundary (then corrected by the
 *      remainder A B^m - Y Q) = floor(A B^m / Q).
 * The series tail after N terms is below 2/(N+1)! < B^-(m+2), so Y is e's fraction to m limbs unless the m+1-th and
 * m+2-th limbs of the exact fraction are all B - 1 (which the caller's check windows would catch).
 * DevOps is ebig.h's backend on the device: products by the library's product (wrapped products by its cyclic
 * convolution), additions and subtractions by k_addsub with the CRT's block carry maps (crt.hip), each result's
 * normalized length (and carry) read back behind one synchronization; only the few-limb base of the Newton iteration
 * and the final fraction cross to the host. */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <chrono>
#include <memory>
#include <thread>
#include <vector>
#include "internal.h"
#include "ebig.h"

#define E_SEED_WD 64              /* base-10^9 digits one thread holds per number */
#define E_SEED_GMAX 64            /* terms per leaf at most */
#define E_D9 1000000000ULL

int nttm_run_addsub(uint64_t *r, uint64_t nout, const uint64_t *a, uint64_t na, uint64_t sa, int64_t apow, const uint64_t *b,
                    uint64_t nb, uint64_t sb, int64_t bpow, int sub, uint32_t *scratch, hipStream_t s);
int nttm_run_top_nonzero(const uint64_t *x, uint64_t n, uint64_t *out_dev, hipStream_t s);
int nttm_copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t s);
int nttm_copy_to_host_prepare(void);
int nttm_mul_cyclic(uint64_t *v, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, size_t L, hipStream_t s,
                    const uint32_t **flag);
int nttm_mul_launch(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, hipStream_t s,
                    const uint32_t **flag);
size_t nttm_wrap_len(size_t need);
uint64_t nttm_crt_blocks(uint64_t nc, int per = 8);
uint64_t nttm_crt_scratch_words(uint64_t nblocks);

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

/* ---- the lowest tree levels by schoolbook products ------------------------------------------------------------------
 * One wave per pair of adjacent nodes (fixed strides: node i of the input holds P at in + 2 w i and Q at in + 2 w i + w,
 * lengths lens_in[2 i], lens_in[2 i + 1]): P = P_L Q_R + P_R, Q = Q_L Q_R into node i of the output (stride wo = 2 w + 1);
 * an odd last node is copied.  Column k of a product, sum_i a_i b_(k-i) (+ P_R's limb k), is below w B^2 + B < 2^128
 * for w <= 256; it splits into three base-10^18 digits (div_b), and limb k = d0[k] + d1[k-1] + d2[k-2] + carry (one
 * lane, sequentially; the carry is at most 2). */
#define BS_MAXW 256
#define E_SCHOOL_W 128            /* the widest schoolbook input of the e pipeline's leaves: the 247 -> 495-limb level runs
                                   * as 2^9-point transforms (k_fused2<8, 9, 3, true>), 5 ms faster at 10^9 digits */
__device__ __forceinline__ void split3(unsigned __int128 v, uint64_t *d0, uint64_t *d1, uint64_t *d2)
{
    const uint64_t hi = (uint64_t)(v >> 64), lo = (uint64_t)v;
    const uint64_t h1 = hi / NTTM_B, h0 = hi - h1 * NTTM_B;
    uint64_t q, r;
    div_b(h0, lo, &q, &r);                                         /* v = (h1 2^64 + q) B + r */
    *d0 = r;
    div_b(h1, q, d2, d1);                                          /* h1 2^64 + q = d2 B + d1 */
}
__device__ void school_one(const uint64_t *a, uint32_t na, const uint64_t *b, uint32_t nb, const uint64_t *x, uint32_t nx,
                           uint64_t *s0, uint64_t *s1, uint64_t *s2, uint64_t *out, uint32_t wo, uint32_t *len)
{
    const int lane = threadIdx.x;
    const uint32_t ncol = na && nb ? na + nb - 1 : 0, nout = (na && nb ? na + nb : 0) > nx ? na + nb : nx;
    const uint32_t ntot = nout + 1 < wo ? nout + 1 : wo;            /* one limb for the carry of + x */
    for (uint32_t k = lane; k < ntot + 2; k += 64) {
        unsigned __int128 acc = k < nx ? x[k] : 0;
        if (k < ncol) {
            const uint32_t i0 = k + 1 > nb ? k + 1 - nb : 0, i1 = k < na - 1 ? k : na - 1;
            for (uint32_t i = i0; i <= i1; i++) acc += (unsigned __int128)a[i] * b[k - i];
        }
        split3(acc, &s0[k], &s1[k], &s2[k]);
    }
    __syncthreads();
    /* v_k = d0[k] + d1[k-1] + d2[k-2] < 2B + 351: q_k = floor(v_k / B) <= 2 goes to limb k+1 directly, and with
     * w_k = v_k - q_k B + q_(k-1) <= B + 1 the rest is a one-bit carry o_k = [w_k + o_(k-1) >= B]: a 64-lane adder of
     * the ballots X = [w >= B], Y = X | [w = B - 1] per word of 64 limbs, words in order (as in the CRT) */
    uint32_t cin = 0, qprev = 0, n = 0;
    for (uint32_t base = 0; base < ntot; base += 64) {
        const uint32_t k = base + lane;
        const uint64_t v = k < ntot ? s0[k] + (k >= 1 ? s1[k - 1] : 0) + (k >= 2 ? s2[k - 2] : 0) : 0;
        const uint32_t q = v >= 2 * NTTM_B ? 2u : v >= NTTM_B ? 1u : 0u;
        uint32_t qp = __shfl_up(q, 1, 64);
        if (lane == 0) qp = qprev;
        qprev = __builtin_amdgcn_readlane(q, 63);
        const uint64_t wv = v - (uint64_t)q * NTTM_B + qp;
        const uint64_t X = __ballot(k < ntot && wv >= NTTM_B), Y = X | __ballot(k < ntot && wv == NTTM_B - 1);
        const uint64_t XY = X + Y, S = XY + cin;
        const uint32_t cout = (uint32_t)(XY < X) | (uint32_t)(S < XY);            /* carry out of X + Y + cin */
        const uint64_t CI = S ^ X ^ Y;                                               /* bit b: the carry into bit b */
        const uint32_t oin = (uint32_t)(CI >> lane) & 1u;
        const uint32_t oout = lane == 63 ? cout : (uint32_t)(CI >> (lane + 1)) & 1u;
        const uint64_t limb = wv + oin - (oout ? NTTM_B : 0);
        if (k < ntot) out[k] = limb;
        const uint64_t nz = __ballot(k < ntot && limb != 0);
        if (nz) n = base + 64 - __clzll((long long)nz);
        cin = cout;
    }
    if (lane == 0) *len = n;
    __syncthreads();
    for (uint32_t k = ntot + lane; k < wo; k += 64) out[k] = 0;
}
/* MAXW: the largest input node width the instance holds in LDS (the small instance keeps more waves per CU) */
template <int MAXW>
__global__ __launch_bounds__(64) void k_bs_school(const uint64_t *in, uint32_t w, const uint32_t *lens_in, uint64_t nin,
                                                  uint64_t *out, uint32_t wo, uint32_t *lens_out)
{
    __shared__ uint64_t sPL[MAXW], sQL[MAXW], sPR[MAXW], sQR[MAXW];
    __shared__ uint64_t s0[2 * MAXW + 4], s1[2 * MAXW + 4], s2[2 * MAXW + 4];
    const uint64_t pair = blockIdx.x, L = 2 * pair, R = L + 1;
    const int lane = threadIdx.x;
    uint64_t *po = out + 2 * (uint64_t)wo * pair, *qo = po + wo;
    const uint64_t *pl = in + 2 * (uint64_t)w * L, *ql = pl + w;
    const uint32_t pnL = lens_in[2 * L], qnL = lens_in[2 * L + 1];
    if (R >= nin) {                                                /* the odd last node moves up */
        for (uint32_t k = lane; k < wo; k += 64) { po[k] = k < pnL ? pl[k] : 0; qo[k] = k < qnL ? ql[k] : 0; }
        if (lane == 0) { lens_out[2 * pair] = pnL; lens_out[2 * pair + 1] = qnL; }
        return;
    }
    const uint64_t *pr = in + 2 * (uint64_t)w * R, *qr = pr + w;
    const uint32_t pnR = lens_in[2 * R], qnR = lens_in[2 * R + 1];
    for (uint32_t k = lane; k < w; k += 64) {
        sPL[k] = k < pnL ? pl[k] : 0; sQL[k] = k < qnL ? ql[k] : 0;
        sPR[k] = k < pnR ? pr[k] : 0; sQR[k] = k < qnR ? qr[k] : 0;
    }
    __syncthreads();
    school_one(sQL, qnL, sQR, qnR, nullptr, 0, s0, s1, s2, qo, wo, &lens_out[2 * pair + 1]);
    school_one(sPL, pnL, sQR, qnR, sPR, pnR, s0, s1, s2, po, wo, &lens_out[2 * pair]);
}

/* ---- the device backend of ebig.h ------------------------------------------------------------------------------------ */
struct DBuf {
    uint64_t *p = nullptr;
    bool own = true;                                              /* false: a slice of the pipeline's arena */
    ~DBuf() { if (p && own) (void)hipFree(p); }
};
/* a normalized number: n limbs at buf + off (views share the buffer) */
struct DNum {
    std::shared_ptr<DBuf> buf;
    size_t off = 0, n = 0;
    const uint64_t *ptr() const { return buf ? buf->p + off : nullptr; }
    uint64_t *wptr() const { return buf ? buf->p + off : nullptr; }
};

struct DevOps {
    typedef DNum Num;
    hipStream_t s = nullptr;
    /* the numbers' memory: slices of one arena, never reused within the pipeline (hipMalloc and hipFree per number
     * would each stall the stream); hipMalloc only when the arena is exhausted */
    char *arena = nullptr; size_t arena_bytes = 0, arena_used = 0;
    uint32_t *scratch = nullptr; size_t scratch_words = 0;        /* carry status words of k_addsub */
    uint64_t *dword = nullptr;                                    /* one device word (lengths) */
    uint64_t *hword = nullptr;                                    /* pinned host words for read-back */
    int err = 0;
    ~DevOps()
    {
        if (arena) (void)hipFree(arena);
        if (scratch) (void)hipFree(scratch);
        if (dword) (void)hipFree(dword);
        if (hword) (void)hipHostFree(hword);
    }
    int init(hipStream_t st, size_t arena_limbs = 0)
    {
        s = st;
        if (arena_limbs) {
            if (hipMalloc(&arena, arena_limbs * 8) == hipSuccess) arena_bytes = arena_limbs * 8;
            else { (void)hipGetLastError(); arena = nullptr; }    /* without: hipMalloc per number */
        }
        if (hipMalloc(&dword, 64) != hipSuccess || hipHostMalloc((void **)&hword, 64, 0) != hipSuccess) {
            (void)hipGetLastError(); return nttm_fail(NTTM_ENOMEM, "e pipeline: device words");
        }
        return 0;
    }
    int alloc(size_t n, DNum &r)
    {
        r = DNum();
        std::shared_ptr<DBuf> b = std::make_shared<DBuf>();
        const size_t bytes = ((n ? n : 1) * 8 + 255) & ~(size_t)255;
        if (arena && arena_used + bytes <= arena_bytes) {
            b->p = (uint64_t *)(arena + arena_used); b->own = false; arena_used += bytes;
            r.buf = b; r.n = n;
            return 0;
        }
        if (hipMalloc(&b->p, (n ? n : 1) * 8) != hipSuccess) {
            (void)hipGetLastError(); b->p = nullptr;
            return nttm_fail(NTTM_ENOMEM, "e pipeline: %zu limbs", n);
        }
        r.buf = b; r.n = n;
        return 0;
    }
    int read_word(const void *src, size_t bytes, uint64_t *v)
    {
        *v = 0;
        RCK(nttm_hip_check(hipMemcpyAsync(hword, src, bytes, hipMemcpyDeviceToHost, s), "read back"));
        RCK(nttm_hip_check(hipStreamSynchronize(s), "read back"));
        memcpy(v, hword, bytes);
        return 0;
    }
    /* drop leading zero limbs: the device scan from the top (a few limbs' work when there are few) and one read */
    int norm(DNum &a)
    {
        if (!a.n) { a = DNum(); return 0; }
        uint64_t v;
        RCK(nttm_run_top_nonzero(a.ptr(), a.n, dword, s));
        RCK(read_word(dword, 8, &v));
        a.n = v;
        if (!a.n) a = DNum();
        return 0;
    }
    size_t size(const DNum &a) const { return a.n; }
    DNum top(const DNum &a, size_t t) const { if (a.n <= t) return a; DNum r = a; r.off += a.n - t; r.n = t; return r; }
    DNum shr(const DNum &a, size_t k) const { if (a.n <= k) return DNum(); DNum r = a; r.off += k; r.n -= k; return r; }
    uint64_t limb(const DNum &a, size_t i)
    {
        uint64_t v = ~(uint64_t)0;
        if (i >= a.n) return 0;
        if ((err = read_word(a.ptr() + i, 8, &v))) return ~(uint64_t)0;
        return v;
    }
    int low(const DNum &a, size_t k, DNum &r) { r = a; r.n = std::min(k, a.n); return norm(r); }
    /* the product and its normalized length (normalized operands: at most one zero limb on top) behind one
     * synchronization */
    int mul(const DNum &a, const DNum &b, DNum &c)
    {
        if (!a.n || !b.n) { c = DNum(); return 0; }
        DNum t;
        RCK(alloc(a.n + b.n, t));
        const uint32_t *flag = nullptr;
        RCK(nttm_mul_launch(t.wptr(), a.ptr(), a.n, b.ptr(), b.n, s, &flag));
        uint64_t v;
        RCK(read_word(t.ptr() + t.n - 1, 8, &v));
        if (flag && (*flag & 1)) return nttm_fail(NTTM_ECHECK, "e pipeline: crt: non-zero carry beyond the top limb");
        if (!v) t.n--;
        c = t;
        return 0;
    }
    /* r = a B^sa +- (b B^sb, or B^bpow); for a subtraction *ge = (the difference >= 0), r empty otherwise */
    int addsub_raw(DNum &r, const DNum *a, size_t sa, int64_t apow, const DNum *b, size_t sb, int64_t bpow, bool sub, bool *ge)
    {
        const size_t ea = apow >= 0 ? (size_t)apow + 1 : (a && a->n ? a->n + sa : 0);
        const size_t eb = bpow >= 0 ? (size_t)bpow + 1 : (b && b->n ? b->n + sb : 0);
        const size_t nout = std::max(ea, eb) + (sub ? 0 : 1);
        if (ge) *ge = true;
        if (!nout) { r = DNum(); return 0; }
        const size_t words = nttm_crt_scratch_words(nttm_crt_blocks(nout));
        if (words > scratch_words) {
            if (scratch) { (void)hipStreamSynchronize(s); (void)hipFree(scratch); scratch = nullptr; scratch_words = 0; }
            if (hipMalloc(&scratch, words * 4) != hipSuccess) { (void)hipGetLastError(); return nttm_fail(NTTM_ENOMEM, "e pipeline: carry words"); }
            scratch_words = words;
        }
        DNum t;
        RCK(alloc(nout, t));
        RCK(nttm_run_addsub(t.wptr(), nout, a ? a->ptr() : nullptr, a ? a->n : 0, sa, apow, b ? b->ptr() : nullptr,
                            b ? b->n : 0, sb, bpow, sub ? 1 : 0, scratch, s));
        /* the normalized length and (a subtraction) the final carry out, ctl[1], behind one synchronization */
        RCK(nttm_run_top_nonzero(t.ptr(), nout, dword, s));
        RCK(nttm_hip_check(hipMemcpyAsync(hword, dword, 8, hipMemcpyDeviceToHost, s), "read back"));
        if (sub) RCK(nttm_hip_check(hipMemcpyAsync(hword + 1, scratch, 8, hipMemcpyDeviceToHost, s), "read back"));
        RCK(nttm_hip_check(hipStreamSynchronize(s), "read back"));
        if (sub) {
            const bool g = (hword[1] >> 32) != 0;
            if (ge) *ge = g;
            if (!g) { r = DNum(); return 0; }
        }
        t.n = hword[0];
        if (!t.n) t = DNum();
        r = t;
        return 0;
    }
    int addsub(DNum &r, const DNum &a, size_t sa, const DNum &b, size_t sb, bool sub, bool *ge)
    {
        return addsub_raw(r, &a, sa, -1, &b, sb, -1, sub, ge);
    }
    int pow_sub(size_t K, const DNum &a, DNum &r)
    {
        bool ge;
        RCK(addsub_raw(r, nullptr, 0, (int64_t)K, &a, 0, -1, true, &ge));
        return ge ? 0 : nttm_fail(NTTM_ECHECK, "e pipeline: B^%zu below its subtrahend", K);
    }
    int add_one(DNum &a) { DNum r; RCK(addsub_raw(r, &a, 0, -1, nullptr, 0, 0, false, nullptr)); a = r; return 0; }
    int sub_pow(DNum &r, const DNum &a, size_t K, bool *ge) { return addsub_raw(r, &a, 0, -1, nullptr, 0, (int64_t)K, true, ge); }
    /* the wrapped product: r = a b mod (B^L - 1) (B^L - 1 itself may stand for 0), through the cyclic convolution of
     * length L and a fold of its top limbs onto the bottom; used from h = wrap_h limbs of reciprocal up (below, its
     * extra additions cost more than the shorter transform saves; NTTM_E_NO_WRAP=1: never) */
    size_t wrap_h = (size_t)1 << 15;
    size_t wrap_min() const { return wrap_h; }
    size_t wrap_len(size_t need) const { return nttm_wrap_len(need); }
    int mul_wrap(const DNum &a, const DNum &b, size_t L, DNum &r)
    {
        if (!a.n || !b.n) { r = DNum(); return 0; }
        DNum v, lo, hi, t;
        RCK(alloc(L + 3, v));
        const uint32_t *flag = nullptr;
        RCK(nttm_mul_cyclic(v.wptr(), a.ptr(), a.n, b.ptr(), b.n, L, s, &flag));
        RCK(low(v, L, lo));                                       /* synchronizes: the CRT's flags are in */
        if (flag && (*flag & 1)) return nttm_fail(NTTM_ECHECK, "e pipeline: cyclic product: carry beyond the top limb");
        hi = shr(v, L);                                           /* 3 limbs, top ones possibly 0 */
        RCK(addsub(t, lo, 0, hi, 0, false, nullptr));             /* v mod B^L + v div B^L < B^L + B^3 */
        if (t.n > L) {                                            /* B^L = 1 (mod B^L - 1) */
            RCK(low(t, L, lo));
            RCK(add_one(lo));
            t = lo;
        }
        r = t;
        return 0;
    }
    int sub_one(DNum &a)
    {
        bool ge; DNum r;
        RCK(addsub_raw(r, &a, 0, -1, nullptr, 0, 0, true, &ge));
        if (!ge) return nttm_fail(NTTM_ECHECK, "e pipeline: 0 - 1");
        a = r;
        return 0;
    }
    int to_host(const DNum &a, Big &h)
    {
        h.assign(a.n, 0);
        if (a.n) {
            RCK(nttm_hip_check(hipMemcpyAsync(h.data(), a.ptr(), a.n * 8, hipMemcpyDeviceToHost, s), "copy to host"));
            RCK(nttm_hip_check(hipStreamSynchronize(s), "copy to host"));
        }
        return 0;
    }
    int from_host(const Big &h, DNum &a)
    {
        if (h.empty()) { a = DNum(); return 0; }
        RCK(alloc(h.size(), a));
        RCK(nttm_hip_check(hipMemcpyAsync(a.wptr(), h.data(), h.size() * 8, hipMemcpyHostToDevice, s), "copy to device"));
        return nttm_hip_check(hipStreamSynchronize(s), "copy to device");
    }
    int div_pow(size_t e, const DNum &d, DNum &r)                  /* a few limbs: on the host */
    {
        Big hd;
        RCK(to_host(d, hd));
        return from_host(big_div_pow(e, hd), r);
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

/* upper bounds of the leaves' normalized lengths from the series itself, so that the tree is planned on the host while
 * the device still forms the leaves: leaf i covers the terms [k0, k1) = [i s, min((i + 1) s, N)), Q_i = k1! / k0!, and
 * P_i / Q_i = sum_(j < k1) 1 / ((k0 + 1) ... (j + 1)) < 1 / k0 (< 2 for k0 = 0).  With L = log10 Q_i from the
 * difference of FP64 lgamma values (a few ulp each: below 10^-6 absolute up to 10^9 terms, under 10^-7 limbs),
 * |Q_i| <= floor(L / 18 + 10^-5) + 1 and |P_i| <= floor((L - log10 k0) / 18 + 10^-5) + 1 (L + log10 2 for k0 = 0):
 * at most one limb above the true length, clamped to the slot width w (every slot holds zeros above its number) */
static void leaf_bounds(uint64_t N, uint64_t span, uint64_t nodes, uint32_t w, std::vector<uint32_t> &hl)
{
    hl.assign(2 * nodes, 0);
    const double ln10 = log(10.0);
    double lg0 = 0.0;                                              /* lgamma(k0 + 1), k0 = 0 */
    for (uint64_t i = 0; i < nodes; i++) {
        const uint64_t k0 = i * span, k1 = std::min<uint64_t>(k0 + span, N);
        const double lg1 = lgamma((double)k1 + 1.0);
        const double L = (lg1 - lg0) / ln10, Lp = k0 ? L - log10((double)k0) : L + log10(2.0);
        const uint64_t qn = (uint64_t)floor(L / 18.0 + 1e-5) + 1;
        const uint64_t pn = Lp > 0 ? (uint64_t)floor(Lp / 18.0 + 1e-5) + 1 : 1;
        hl[2 * i] = (uint32_t)std::min<uint64_t>(pn, w);
        hl[2 * i + 1] = (uint32_t)std::min<uint64_t>(qn, w);
        lg0 = lg1;
    }
}

/* P, Q of the N-term series, in device memory.  The leaves (k_e_seed and the schoolbook levels) are launched without
 * a synchronization, the tree planned from the leaves' length bounds meanwhile and queued behind them; leaf and tree
 * times come from events around the leaves. */
static int e_series(DevOps &o, size_t N, DNum &P, DNum &Q)
{
    hipStream_t s = o.s;
    uint32_t G, wl;
    RCK(seed_shape(N, &G, &wl));
    const uint64_t nleaf = (N + G - 1) / G;
    const char *sw = getenv("NTTM_E_SCHOOL_W");                  /* the widest schoolbook input (for comparisons) */
    const uint32_t school_w = sw && atoi(sw) > 0 && atoi(sw) <= BS_MAXW ? (uint32_t)atoi(sw) : E_SCHOOL_W;
    const char *lc = getenv("NTTM_E_LEAF_CHECK");                /* 1: check the length bounds against the leaves */
    const bool leaf_check = lc && lc[0] == '1';
    hipEvent_t ev[2] = {nullptr, nullptr};
    for (int k = 0; k < 2; k++) RCK(nttm_hip_check(hipEventCreate(&ev[k]), "event"));
    struct EvFree { hipEvent_t *e; ~EvFree() { for (int k = 0; k < 2; k++) if (e[k]) (void)hipEventDestroy(e[k]); } } evf{ev};
    auto t0 = eclk::now();
    DNum pq, lw;
    RCK(o.alloc(nleaf * 2 * wl, pq));
    RCK(o.alloc(nleaf + 1, lw));                                   /* the seed's lengths, 2 nleaf words */
    uint32_t *lens = (uint32_t *)lw.wptr();
    RCK(nttm_hip_check(hipEventRecord(ev[0], s), "event"));
    hipLaunchKernelGGL(k_e_seed, dim3((unsigned)((nleaf + 255) / 256)), dim3(256), 0, s, (uint64_t)N, G, nleaf, wl, pq.wptr(), lens);
    RCK(nttm_hip_check(hipGetLastError(), "k_e_seed launch"));
    /* the lowest tree levels by schoolbook products up to inputs of school_w limbs (30 -> 61 -> 123 -> 247 limbs) */
    uint64_t nodes = nleaf, span = G;
    uint32_t w = wl;
    while (nodes > 1 && w <= school_w) {
        const uint64_t nout = (nodes + 1) / 2;
        const uint32_t wo = 2 * w + 1;
        DNum next, lw2;
        RCK(o.alloc(nout * 2 * wo, next));
        RCK(o.alloc(nout + 1, lw2));
        uint32_t *lens2 = (uint32_t *)lw2.wptr();
        if (w <= 64)
            hipLaunchKernelGGL(k_bs_school<64>, dim3((unsigned)nout), dim3(64), 0, s, pq.ptr(), w, lens, nodes, next.wptr(), wo, lens2);
        else
            hipLaunchKernelGGL(k_bs_school<BS_MAXW>, dim3((unsigned)nout), dim3(64), 0, s, pq.ptr(), w, lens, nodes, next.wptr(), wo, lens2);
        RCK(nttm_hip_check(hipGetLastError(), "k_bs_school launch"));
        lens = lens2;
        pq = next;
        nodes = nout;
        w = wo;
        span *= 2;
    }
    RCK(nttm_hip_check(hipEventRecord(ev[1], s), "event"));
    std::vector<uint32_t> hl;
    leaf_bounds(N, span, nodes, w, hl);
    if (leaf_check) {
        std::vector<uint32_t> tl(2 * nodes);
        RCK(nttm_hip_check(hipMemcpyAsync(tl.data(), lens, nodes * 2 * 4, hipMemcpyDeviceToHost, s), "copy leaf lengths"));
        RCK(nttm_hip_check(hipStreamSynchronize(s), "e leaves"));
        for (uint64_t i = 0; i < 2 * nodes; i++)
            if (hl[i] < tl[i] || hl[i] > tl[i] + 1)
                return nttm_fail(NTTM_ECHECK, "e pipeline: leaf %llu: length bound %u against %u limbs",
                                 (unsigned long long)(i / 2), hl[i], tl[i]);
    }
    size_t cap = 2;
    std::vector<nttm_bs_node> leaves(nodes);
    for (uint64_t i = 0; i < nodes; i++) {
        leaves[i].p = pq.ptr() + 2 * (uint64_t)w * i; leaves[i].pn = hl[2 * i];
        leaves[i].q = pq.ptr() + 2 * (uint64_t)w * i + w; leaves[i].qn = hl[2 * i + 1];
        cap += hl[2 * i + 1];                                      /* |Q| <= sum of the leaves' |Q|; P < 2 Q */
    }
    /* P's buffer also holds the root's unnormalized length (|P_0| <= |Q_0| + 1, one limb per level), so the tree
     * writes the root in place */
    const size_t pcap = cap + 72;
    RCK(o.alloc(pcap, P));
    RCK(o.alloc(cap, Q));
    size_t pn = 0, qn = 0;
    RCK(nttm_bs_tree(leaves.data(), nodes, P.wptr(), pcap, &pn, Q.wptr(), cap, &qn, s));   /* synchronizes */
    P.n = pn; Q.n = qn;                                            /* normalized by the tree */
    float leaf_ms = 0;
    RCK(nttm_hip_check(hipEventElapsedTime(&leaf_ms, ev[0], ev[1]), "event"));
    const double total = ms_since(t0);
    g_e_ms[0] = leaf_ms;
    g_e_ms[1] = total - leaf_ms;
    return 0;
}

extern "C" int nttm_e_fraction(uint64_t *out, size_t m, void *stream)
{
    if (!out || !m) return nttm_fail(NTTM_EINVAL, "nttm_e_fraction arguments");
    RCK(nttm_require_host(out, m * 8, "out"));                    /* host threads write its pages */
    for (double &v : g_e_ms) v = 0;
    /* N: log(N!) > 18 (m + 2) log 10 */
    const double need = 18.0 * (double)(m + 2) * log(10.0);
    size_t N = 2;
    while (lgamma((double)N + 1.0) < need) N += N / 2 + 1;
    { size_t lo = N / 2, hi = N; while (lo + 1 < hi) { size_t md = (lo + hi) / 2; if (lgamma((double)md + 1.0) < need) lo = md; else hi = md; } N = hi; }
    /* the download's pinned buffers (a fraction above 64 MiB goes through them), allocated and first touched on
     * another host thread while the device forms the series (inside the copy, their first use cost about 7 ms at
     * 10^9 digits) */
    struct Prep {
        std::thread t;
        explicit Prep(bool on)
        {
            int dev = 0;
            if (on && hipGetDevice(&dev) == hipSuccess)         /* a new thread starts on device 0: the caller's device */
                t = std::thread([dev]() { if (hipSetDevice(dev) == hipSuccess) (void)nttm_copy_to_host_prepare(); });
        }
        ~Prep() { if (t.joinable()) t.join(); }
    } prep(m * 8 > ((size_t)64 << 20));
    DevOps o;
    /* the arena: the series' leaves (about 2 m limbs per schoolbook level) and root, the Newton levels' numbers
     * (their products, wrapped products of L + 3 limbs and differences: about 12 m limbs over all levels) and the
     * quotient's (about 5 m), with room to spare (14 GB at 10^9 digits); past it, a number is allocated on its own;
     * NTTM_E_ARENA=0 allocates every number on its own */
    const char *ea = getenv("NTTM_E_ARENA");
    RCK(o.init((hipStream_t)stream, ea && ea[0] == '0' ? 0 : 32 * (m + 1024)));
    const char *nw = getenv("NTTM_E_NO_WRAP");
    if (nw && nw[0] == '1') o.wrap_h = SIZE_MAX;
    DNum P, Q, A, x, Y;
    RCK(e_series(o, N, P, Q));
    /* the first touch of the caller's output pages (zeros) on 4 host threads while the device runs the reciprocal:
     * inside the copy, their page faults cost about 4 ms at 10^9 digits (started after the tree, whose planning
     * runs on this thread while the leaves are formed) */
    struct Touch {
        std::thread t[4];
        Touch(bool on, char *p, size_t bytes)
        {
            const size_t part = (bytes / 4 + 4095) & ~(size_t)4095;
            for (int i = 0; on && i < 4; i++) {
                const size_t off = i * part, len = off < bytes ? std::min(part, bytes - off) : 0;
                if (len) t[i] = std::thread([=]() { memset(p + off, 0, len); });
            }
        }
        void join() { for (std::thread &x : t) if (x.joinable()) x.join(); }
        ~Touch() { join(); }
    } touch(m * 8 > ((size_t)64 << 20), (char *)out, m * 8);
    bool ge;
    RCK(o.addsub(A, P, 0, Q, 0, true, &ge));                      /* e - 2 ~ A / Q, 0 <= A < Q */
    if (!ge || !A.n) return nttm_fail(NTTM_ECHECK, "e pipeline: P <= Q");
    P = DNum();
    auto t0 = eclk::now();
    int rc = eb_recip(o, Q, m + 1, x);                            /* x ~ B^(n+m+1) / Q (one guard limb) */
    if (rc) return rc > 0 ? rc : nttm_fail(NTTM_ECHECK, "e pipeline: reciprocal invariant broken%s", o.err ? " (read-back failed)" : "");
    g_e_ms[2] = ms_since(t0);
    t0 = eclk::now();
    const char *fc = getenv("NTTM_E_FULL_CHECK");                 /* 1: always form the remainder (for tests) */
    rc = eb_quotient(o, A, Q, m, x, Y, fc && fc[0] == '1');       /* floor(A B^m / Q) */
    if (rc) return rc > 0 ? rc : nttm_fail(NTTM_ECHECK, "e pipeline: quotient correction did not converge");
    if (Y.n > m) return nttm_fail(NTTM_ECHECK, "e pipeline: fraction of %zu limbs above %zu", Y.n, m);
    if (prep.t.joinable()) prep.t.join();
    touch.join();
    if (Y.n) RCK(nttm_copy_to_host(out, Y.ptr(), Y.n * 8, o.s));
    if (Y.n < m) memset(out + Y.n, 0, (m - Y.n) * 8);
    g_e_ms[3] = ms_since(t0);
    return 0;
}

/* the device addition / subtraction, for tests: r[0 .. nout) = a B^sa +- b B^sb (device pointers), *ge = the final
 * carry out (for a subtraction: the difference is >= 0) */
extern "C" int nttm_debug_addsub(uint64_t *r, size_t nout, const uint64_t *a, size_t na, size_t sa, const uint64_t *b,
                                 size_t nb, size_t sb, int sub, int *ge, void *stream)
{
    if (!r || !ge || (na && !a) || (nb && !b)) return nttm_fail(NTTM_EINVAL, "nttm_debug_addsub arguments");
    RCK(nttm_require_device(r, nout * 8, "r"));
    RCK(nttm_require_device(a, na * 8, "a"));
    RCK(nttm_require_device(b, nb * 8, "b"));
    hipStream_t s = (hipStream_t)stream;
    uint32_t *scratch = nullptr;
    const size_t words = nttm_crt_scratch_words(nttm_crt_blocks(nout ? nout : 1));
    if (hipMalloc(&scratch, words * 4) != hipSuccess) { (void)hipGetLastError(); return nttm_fail(NTTM_ENOMEM, "addsub scratch"); }
    int rc = nttm_run_addsub(r, nout, a, na, sa, -1, b, nb, sb, -1, sub, scratch, s);
    uint32_t ctl[2] = {0, 0};
    if (!rc && nout) rc = nttm_hip_check(hipMemcpyAsync(ctl, scratch, 8, hipMemcpyDeviceToHost, s), "read flag");
    if (!rc) rc = nttm_hip_check(hipStreamSynchronize(s), "addsub");
    (void)hipFree(scratch);
    *ge = ctl[1] != 0;
    return rc;
}

extern "C" int nttm_e_last_phase_ms(double *seed_ms, double *tree_ms, double *recip_ms, double *quot_ms)
{
    if (seed_ms) *seed_ms = g_e_ms[0];
    if (tree_ms) *tree_ms = g_e_ms[1];
    if (recip_ms) *recip_ms = g_e_ms[2];
    if (quot_ms) *quot_ms = g_e_ms[3];
    return 0;
}
