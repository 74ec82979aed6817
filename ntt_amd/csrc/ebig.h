/* ebig.h - the Newton reciprocal and the corrected quotient of the e pipeline, written over a number backend.
 *
 * Numbers are base-10^18 limb vectors (least significant first, normalized: no leading zero limb; 0 is empty).  The
 * algorithms below use only the backend's operations (the device backend in epipe.hip, HostOps here for the host):
 *   size(a), top(a, t) (the top t limbs), shr(a, k) (drop the low k limbs), limb(a, i), low(a, k, r) (a mod B^k),
 *   mul(a, b, c), pow_sub(K, a, r) (B^K - a, for a < B^K), sub_pow(r, a, K, &ge) (a - B^K), addsub(r, a, sa, b, sb, sub,
 *   &ge) (a B^sa +- b B^sb; for a subtraction ge tells whether it was >= 0, and r is only meaningful then), add_one /
 *   sub_one, div_pow(e, d, r) (floor(B^e / d) for a few limbs), and the wrapped product mul_wrap(a, b, L, r)
 *   (a b mod (B^L - 1)) with wrap_len(need) (a length L >= need it supports) and wrap_min() (the reciprocal size from
 *   which eb_recip takes it).
 *
 * Reciprocal (the reference's Newton correction form, E/ALGORITHM.md:197-246, E/newton_db.c:218): x_l approximates
 * B^(n+l) / Q (Q of n limbs) to within 3 units.  From x_h to x_l with l <= 2h - 2, using only the top t = l + 2 limbs
 * Q_t of Q:  E = B^(t+h) - Q_t x_h,  x_l = x_h B^(l-h) + floor(x_h E / B^(t+2h-l)).
 * The relative error after the step is eps_h^2 + B^(-l-1) (the truncation of Q) + one unit of rounding + 1/B of a
 * unit (x_h E formed from E's top limbs only, below); with eps_h <= 3 B^-h and l <= 2h - 2 the squared term is below
 * 9/B units, so the error stays within 3 units at every level.  Each step costs two products, of about 1.5 l and
 * l limbs; the levels halve down to a base of at most 8 limbs, which is exact long division.  Q_t x_h is B^K - E with |E| <= B^(h+1) + 3 B^t <= 4 B^t, so above wrap_min() it is formed
 * modulo B^L - 1 only (L >= t + 2, about 2/3 of the whole product's length), and E read off the residue: the
 * wrap-around of the Newton product (as GMP's mpn_invertappr does with mpn_mulmod_bnm1).
 * Quotient, from x = x_(m+1) (one guard limb): for A < Q, A B^m / Q = (T + A d) / B^(n+1) with T = A x and |d| <= 3,
 * so with Y = floor(T / B^(n+1)) and T_lo = T mod B^(n+1), A B^m / Q = Y + (T_lo + A d) / B^(n+1), where
 * |A d| < 3 B^n.  When limb n of T (the top limb of T_lo) lies in [3, B - 4], 0 < T_lo + A d < B^(n+1) and Y is
 * exact: one product and one limb read.  Otherwise (a fraction A B^m / Q within about 3 B^-1 of an integer, e.g. an
 * exact quotient) Y lies within 1 of the quotient, and the remainder R = A B^m - Y Q (one more product) with a few
 * additions or subtractions of Q makes it exact. */
#ifndef NTTM_EBIG_H
#define NTTM_EBIG_H
#include <stdint.h>
#include <algorithm>
#include <vector>

typedef std::vector<uint64_t> Big;
static const uint64_t EB_BASE = 1000000000000000000ULL;

static inline void big_norm(Big &a) { while (!a.empty() && !a.back()) a.pop_back(); }
static inline int big_cmp(const Big &a, const Big &b)
{
    if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
    for (size_t i = a.size(); i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
static inline Big big_add(const Big &a, const Big &b)
{
    const Big &l = a.size() >= b.size() ? a : b, &s = a.size() >= b.size() ? b : a;
    Big r(l.size() + 1);
    uint64_t c = 0;
    size_t i = 0;
    for (; i < s.size(); i++) { const uint64_t v = l[i] + s[i] + c; c = v >= EB_BASE; r[i] = c ? v - EB_BASE : v; }
    for (; i < l.size(); i++) { const uint64_t v = l[i] + c; c = v >= EB_BASE; r[i] = c ? v - EB_BASE : v; }
    r[i] = c;
    big_norm(r);
    return r;
}
/* a - b for a >= b */
static inline Big big_sub(const Big &a, const Big &b)
{
    Big r(a.size());
    uint64_t br = 0;
    size_t i = 0;
    for (; i < b.size(); i++) {
        const uint64_t s = b[i] + br;
        if (a[i] >= s) { r[i] = a[i] - s; br = 0; } else { r[i] = a[i] + EB_BASE - s; br = 1; }
    }
    for (; i < a.size(); i++) {
        if (a[i] >= br) { r[i] = a[i] - br; br = 0; } else { r[i] = EB_BASE - 1; }
    }
    big_norm(r);
    return r;
}
static inline Big big_shr(const Big &a, size_t k) { return a.size() > k ? Big(a.begin() + k, a.end()) : Big(); }
static inline Big big_shl(const Big &a, size_t k)
{
    if (a.empty()) return a;
    Big r(k + a.size(), 0);
    std::copy(a.begin(), a.end(), r.begin() + k);
    return r;
}
static inline Big big_small(uint64_t v) { Big r; if (v) r.push_back(v); return r; }
static inline Big big_pow(size_t k) { Big r(k + 1, 0); r[k] = 1; return r; }            /* B^k */
static inline Big big_top(const Big &a, size_t t) { return a.size() > t ? Big(a.end() - t, a.end()) : a; }
static inline Big big_mul_small(const Big &a, uint64_t d)                              /* d < B */
{
    Big r(a.size() + 1);
    unsigned __int128 c = 0;
    for (size_t i = 0; i < a.size(); i++) {
        const unsigned __int128 t = (unsigned __int128)a[i] * d + c;
        c = t / EB_BASE;
        r[i] = (uint64_t)(t - c * EB_BASE);
    }
    r[a.size()] = (uint64_t)c;
    big_norm(r);
    return r;
}
/* floor(B^e / D) by long division, each quotient limb found by bisection (small sizes only) */
static inline Big big_div_pow(size_t e, const Big &D)
{
    Big q(e + 1, 0), rem;
    for (size_t pos = e + 1; pos-- > 0;) {
        rem = big_shl(rem, 1);
        if (pos == e) rem = big_add(rem, big_small(1));
        uint64_t lo = 0, hi = EB_BASE - 1;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo + 1) / 2;
            if (big_cmp(big_mul_small(D, mid), rem) <= 0) lo = mid; else hi = mid - 1;
        }
        if (lo) rem = big_sub(rem, big_mul_small(D, lo));
        q[pos] = lo;
    }
    big_norm(q);
    return q;
}

/* the host backend: limb vectors, products by the caller's multiplier int mul(const Big &, const Big &, Big &) */
template <class Mul> struct HostOps {
    typedef Big Num;
    Mul &m;
    explicit HostOps(Mul &mm) : m(mm) {}
    size_t size(const Big &a) const { return a.size(); }
    Big top(const Big &a, size_t t) const { return big_top(a, t); }
    Big shr(const Big &a, size_t k) const { return big_shr(a, k); }
    uint64_t limb(const Big &a, size_t i) const { return i < a.size() ? a[i] : 0; }
    int low(const Big &a, size_t k, Big &r) const { r.assign(a.begin(), a.begin() + std::min(k, a.size())); big_norm(r); return 0; }
    int mul(const Big &a, const Big &b, Big &c) { return m(a, b, c); }
    int pow_sub(size_t K, const Big &a, Big &r) const { r = big_sub(big_pow(K), a); return 0; }
    int addsub(Big &r, const Big &a, size_t sa, const Big &b, size_t sb, bool sub, bool *ge) const
    {
        const Big A = big_shl(a, sa), Bv = big_shl(b, sb);
        if (!sub) { r = big_add(A, Bv); if (ge) *ge = true; return 0; }
        const bool g = big_cmp(A, Bv) >= 0;
        if (ge) *ge = g;
        r = g ? big_sub(A, Bv) : Big();
        return 0;
    }
    int add_one(Big &a) const { a = big_add(a, big_small(1)); return 0; }
    int sub_pow(Big &r, const Big &a, size_t K, bool *ge) const
    {
        const Big p = big_pow(K);
        *ge = big_cmp(a, p) >= 0;
        r = *ge ? big_sub(a, p) : Big();
        return 0;
    }
    /* the wrapped product a b mod (B^L - 1): the whole product folded in L-limb pieces (wrap_h: from which reciprocal
     * size on eb_recip takes it; any length is a valid L here, the lengths below mirror the device's 2^k and 3 2^k) */
    size_t wrap_h = 0;
    long wraps = 0;                                               /* wrapped products made */
    size_t wrap_min() const { return wrap_h; }
    size_t wrap_len(size_t need) const
    {
        size_t p2 = 16, p3 = 48;
        while (p2 < need) p2 *= 2;
        while (p3 < need) p3 *= 2;
        return std::min(p2, p3);
    }
    int mul_wrap(const Big &a, const Big &b, size_t L, Big &r)
    {
        Big c;
        int rc = m(a, b, c);
        if (rc) return rc;
        wraps++;
        r.clear();
        for (size_t i = 0; i < c.size(); i += L)
            r = big_add(r, Big(c.begin() + i, c.begin() + std::min(c.size(), i + L)));
        while (r.size() > L) r = big_add(Big(r.begin(), r.begin() + L), Big(r.begin() + L, r.end()));
        big_norm(r);
        return 0;
    }
    int sub_one(Big &a) const { if (a.empty()) return -1; a = big_sub(a, big_small(1)); return 0; }
    int div_pow(size_t e, const Big &d, Big &r) const { r = big_div_pow(e, d); return 0; }
};

/* x ~ B^(n+m) / Q within 3 units (Q normalized, n = |Q| >= 1); returns 0, the backend's error, or -1 on a broken
 * invariant */
template <class O> int eb_recip(O &o, const typename O::Num &Q, size_t m, typename O::Num &x)
{
    typedef typename O::Num N;
    const size_t n = o.size(Q);
    if (!n) return -1;
    std::vector<size_t> lv;
    for (size_t l = m;;) { lv.push_back(l); if (l <= 8) break; l = (l + 3) / 2; }    /* h = ceil((l + 2) / 2) */
    size_t h = lv.back();
    const size_t t0 = std::min(n, h + 2);
    int rc = o.div_pow(t0 + h, o.top(Q, t0), x);
    if (rc) return rc;
    for (size_t i = lv.size() - 1; i-- > 0;) {
        const size_t l = lv[i], t = std::min(n, l + 2), K = t + h;
        N T, E, D, xn;
        bool neg;
        const N Qt = o.top(Q, t);
        const size_t L = o.wrap_len(t + 2);
        if (h >= o.wrap_min() && L < o.wrap_len(o.size(Qt) + o.size(x) - 1)) {
            /* T modulo B^L - 1 only: E = B^K - T has |E| < 3 Q_t < 3 B^t <= B^(L-2), so its residue U = B^(K mod L) - T
             * mod (B^L - 1) is E itself when small (size <= L - 1) and B^L - 1 - |E| when E < 0 (size L) */
            const size_t Kw = K % L;
            if ((rc = o.mul_wrap(Qt, x, L, T))) return rc;
            if (o.size(T) <= Kw) {                                  /* T < B^Kw: U = B^Kw - T, small */
                if ((rc = o.pow_sub(Kw, T, E))) return rc;
                neg = false;
            } else {                                                /* V = T - B^Kw = -U */
                bool ge;
                N V;
                if ((rc = o.sub_pow(V, T, Kw, &ge))) return rc;
                if (!ge) return -1;
                if (o.size(V) < L) { E = V; neg = true; }           /* E = -V */
                else {                                              /* E = B^L - 1 - V */
                    if ((rc = o.pow_sub(L, V, E)) || (rc = o.sub_one(E))) return rc;
                    neg = false;
                }
            }
            if (o.size(E) > t + 1) return -1;
        } else {
            if ((rc = o.mul(Qt, x, T))) return rc;
            if (o.size(T) > K) {                                    /* T >= B^K: E = T - B^K, T < 2 B^K */
                if (o.size(T) != K + 1 || o.limb(T, K) != 1) return -1;
                if ((rc = o.low(T, K, E))) return rc;
                neg = true;
            } else {
                if ((rc = o.pow_sub(K, T, E))) return rc;
                neg = false;
            }
        }
        T = N();
        /* only floor(x E / B^sh) is used: dropping E's low k = sh - h - 2 limbs moves x E by less than
         * x B^k < (B^(h+1) + 3) B^k, under 1/B of a unit after the shift, so the product takes E's top limbs */
        const size_t sh = t + 2 * h - l, k = sh > h + 2 ? sh - h - 2 : 0;
        if ((rc = o.mul(x, o.shr(E, k), D))) return rc;
        E = N();
        bool ge = true;
        if ((rc = o.addsub(xn, x, l - h, o.shr(D, sh - k), 0, neg, &ge))) return rc;   /* x B^(l-h) +- step */
        if (!ge) return -1;
        x = xn;
        h = l;
    }
    return 0;
}

/* Y = floor(A B^m / Q) for 0 <= A < Q, from x ~ B^(n+m+1) / Q within 3 units; full_check: always form the remainder
 * (the path the rare undecided case takes) */
template <class O> int eb_quotient(O &o, const typename O::Num &A, const typename O::Num &Q, size_t m,
                                   const typename O::Num &x, typename O::Num &Y, bool full_check = false)
{
    typedef typename O::Num N;
    const size_t n = o.size(Q);
    N T, R, R2;
    int rc = o.mul(A, x, T);
    if (rc) return rc;
    Y = o.shr(T, n + 1);
    const uint64_t tn = o.limb(T, n);
    if (tn == ~(uint64_t)0) return -1;                               /* (a failed read) */
    if (!full_check && tn >= 3 && tn <= EB_BASE - 4) return 0;
    T = N();
    if ((rc = o.mul(Y, Q, T))) return rc;
    bool ge, neg = false;
    if ((rc = o.addsub(R, A, m, T, 0, true, &ge))) return rc;      /* R = A B^m - Y Q */
    if (!ge) {                                                       /* negative: R = Y Q - A B^m */
        if ((rc = o.addsub(R, T, 0, A, m, true, &ge))) return rc;
        neg = true;
    }
    T = N();
    int k = 0;
    for (; k < 16 && neg; k++) {                                     /* R < 0: Y -= 1, R += Q */
        if ((rc = o.sub_one(Y))) return rc;
        if ((rc = o.addsub(R2, Q, 0, R, 0, true, &ge))) return rc;  /* Q - |R| */
        if (ge) { R = R2; neg = false; }
        else { if ((rc = o.addsub(R2, R, 0, Q, 0, true, &ge))) return rc; R = R2; }
    }
    for (; k < 16; k++) {                                            /* R >= Q: Y += 1, R -= Q */
        if ((rc = o.addsub(R2, R, 0, Q, 0, true, &ge))) return rc;
        if (!ge) break;
        R = R2;
        if ((rc = o.add_one(Y))) return rc;
    }
    return k < 16 ? 0 : -1;
}
#endif
