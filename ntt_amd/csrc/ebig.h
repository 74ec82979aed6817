/* ebig.h - host-side limb vectors and the Newton reciprocal / corrected quotient of the e pipeline.
 *
 * Numbers are base-10^18 limb vectors (least significant first, normalized: no leading zero limb; 0 is empty).  The
 * linear work (additions, subtractions, shifts by whole limbs, comparisons) is done here; every product goes through a
 * caller-supplied multiplier `int mul(const Big &a, const Big &b, Big &c)` (the device product in the library).
 *
 * Reciprocal (the reference's Newton correction form, E/ALGORITHM.md:197-246, E/newton_db.c:218): x_l approximates
 * B^(n+l) / Q (Q of n limbs) to within 3 units.  From x_h to x_l with l <= 2h - 2, using only the top t = l + 2 limbs
 * Q_t of Q:  E = B^(t+h) - Q_t x_h,  x_l = x_h B^(l-h) + floor(x_h E / B^(t+2h-l)).
 * The relative error after the step is eps_h^2 + B^(-l-1) (the truncation of Q) + one unit of rounding; with
 * eps_h <= 3 B^-h and l <= 2h - 2 the squared term is below 9/B units, so the error stays within 3 units at every
 * level.  Each step costs two products of about 1.5 l limbs; the levels halve down to a base of at most 8 limbs, which
 * is exact long division on the host.
 * Quotient: Y = floor(A x_m / B^n) lies within 4 of floor(A B^m / Q) for A < Q; one product Y Q and a few
 * subtractions make it exact. */
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

/* x ~ B^(n+m) / Q within 3 units (Q normalized, n = |Q| >= 1); returns 0 or the multiplier's error, -1 on a broken
 * invariant */
template <class Mul> int eb_recip(Mul &mul, const Big &Q, size_t m, Big &x)
{
    const size_t n = Q.size();
    if (!n) return -1;
    std::vector<size_t> lv;
    for (size_t l = m;;) { lv.push_back(l); if (l <= 8) break; l = (l + 3) / 2; }    /* h = ceil((l + 2) / 2) */
    size_t h = lv.back();
    {
        const size_t t0 = std::min(n, h + 2);
        x = big_div_pow(t0 + h, big_top(Q, t0));
    }
    Big T, E, D;
    for (size_t i = lv.size() - 1; i-- > 0;) {
        const size_t l = lv[i], t = std::min(n, l + 2), K = t + h;
        int rc = mul(big_top(Q, t), x, T);
        if (rc) return rc;
        bool neg;
        if (T.size() > K) {                                        /* T >= B^K: E = T - B^K, T < 2 B^K */
            if (T.size() != K + 1 || T[K] != 1) return -1;
            T.pop_back(); big_norm(T); E.swap(T); neg = true;
        } else {
            E = big_sub(big_pow(K), T); neg = false;
        }
        if ((rc = mul(x, E, D))) return rc;
        const Big st = big_shr(D, t + 2 * h - l), y = big_shl(x, l - h);
        if (neg && big_cmp(y, st) < 0) return -1;
        x = neg ? big_sub(y, st) : big_add(y, st);
        h = l;
    }
    return 0;
}

/* Y = floor(A B^m / Q) for 0 <= A < Q, from x ~ B^(n+m) / Q within 3 units */
template <class Mul> int eb_quotient(Mul &mul, const Big &A, const Big &Q, size_t m, const Big &x, Big &Y)
{
    const size_t n = Q.size();
    Big T;
    int rc = mul(A, x, T);
    if (rc) return rc;
    Y = big_shr(T, n);
    if ((rc = mul(Y, Q, T))) return rc;
    const Big ABm = big_shl(A, m);
    int k = 0;
    for (; k < 16 && big_cmp(T, ABm) > 0; k++) { Y = big_sub(Y, big_small(1)); T = big_sub(T, Q); }
    Big R = big_sub(ABm, T);
    for (; k < 16 && big_cmp(R, Q) >= 0; k++) { Y = big_add(Y, big_small(1)); R = big_sub(R, Q); }
    return k < 16 ? 0 : -1;
}
#endif
