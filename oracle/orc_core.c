/* orc_core.c - CPU restatement of the reference product path (test infrastructure only; see oracle.h).
 * Exact arithmetic throughout: unsigned __int128 for every modular and multi-word step. */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef unsigned __int128 u128;

/* ---- generator (E/tests/harness.h:55-81) ------------------------------------------------------------------------ */
uint64_t orc_splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void orc_gen_limbs(uint64_t *a, size_t n, int kind, uint64_t *st)
{
    if (!n) return;
    switch (kind) {
    case ORC_GEN_UNIFORM:
        for (size_t i = 0; i < n; i++) a[i] = orc_splitmix64(st);
        break;
    case ORC_GEN_ONES:
        for (size_t i = 0; i < n; i++) a[i] = ORC_B - 1;
        break;
    case ORC_GEN_ZEROS:
        memset(a, 0, n * sizeof *a); a[n - 1] = 1;
        break;
    case ORC_GEN_BIT:
        memset(a, 0, n * sizeof *a); a[n - 1] = 1ULL << (orc_splitmix64(st) & 63);
        break;
    case ORC_GEN_SPARSE:
        memset(a, 0, n * sizeof *a); a[n - 1] = 1;
        for (size_t i = 0; i < n / 16 + 1; i++) {
            /* the reference's statement draws the bit position before the index (gcc's evaluation order,
             * pinned against the reference's own generator) */
            uint64_t bit = orc_splitmix64(st) & 63;
            uint64_t at = orc_splitmix64(st) % n;
            a[at] |= 1ULL << bit;
        }
        break;
    }
    if (kind != ORC_GEN_ONES)
        for (size_t i = 0; i < n; i++) a[i] %= ORC_B;
    if (a[n - 1] == 0) a[n - 1] = 1;
}

/* ---- base-10^18 division helpers ---------------------------------------------------------------------------------- */
void orc_div1e18(uint64_t hi, uint64_t lo, uint64_t *q, uint64_t *r)
{
    u128 x = ((u128)hi << 64) | lo;
    *q = (uint64_t)(x / ORC_B);
    *r = (uint64_t)(x % ORC_B);
}

/* w (nw words, little-endian) /= B in place; returns the remainder */
static uint64_t words_divB(uint64_t *w, int nw)
{
    uint64_t rem = 0;
    for (int i = nw - 1; i >= 0; i--) { uint64_t q; orc_div1e18(rem, w[i], &q, &rem); w[i] = q; }
    return rem;
}

void orc_words_to_dec(const uint64_t *w, int nw, uint64_t *d, int nd)
{
    uint64_t t[8];
    for (int i = 0; i < nw; i++) t[i] = w[i];
    for (int k = 0; k < nd; k++) d[k] = words_divB(t, nw);
}

/* ---- schoolbook (E/bigint.c:77-110): column sums in three words, then one carry ripple ---------------------------- */
void orc_mul_add_school(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
                        const uint64_t *x, size_t nx)
{
    size_t n = na + nb;
    if (!na || !nb) {
        for (size_t k = 0; k < n + (x ? 1 : 0); k++) c[k] = (x && k < nx) ? x[k] : 0;
        return;
    }
    uint64_t carry[3] = {0, 0, 0};
    for (size_t k = 0; k < n; k++) {
        uint64_t acc[3] = {carry[0], carry[1], carry[2]};
        if (x && k < nx) { u128 s = (u128)acc[0] + x[k]; acc[0] = (uint64_t)s; s = (u128)acc[1] + (uint64_t)(s >> 64); acc[1] = (uint64_t)s; acc[2] += (uint64_t)(s >> 64); }
        size_t i0 = k >= nb - 1 ? k - (nb - 1) : 0, i1 = k < na ? k : na - 1;
        for (size_t i = i0; i <= i1; i++) {
            u128 p = (u128)a[i] * b[k - i];
            u128 s = (u128)acc[0] + (uint64_t)p; acc[0] = (uint64_t)s;
            s = (u128)acc[1] + (uint64_t)(p >> 64) + (uint64_t)(s >> 64); acc[1] = (uint64_t)s;
            acc[2] += (uint64_t)(s >> 64);
        }
        c[k] = words_divB(acc, 3);
        carry[0] = acc[0]; carry[1] = acc[1]; carry[2] = acc[2];
    }
    if (x) c[n] = carry[0];          /* a b + x < 2 B^n: the top limb is the final carry (0 or 1) */
}

void orc_mul_school(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb)
{
    orc_mul_add_school(c, a, na, b, nb, NULL, 0);
}

/* ---- exact modular helpers ----------------------------------------------------------------------------------------- */
uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)((u128)a * b % p); }
uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t p)
{
    uint64_t r = 1 % p; a %= p;
    while (e) { if (e & 1) r = orc_mulmod(r, a, p); a = orc_mulmod(a, a, p); e >>= 1; }
    return r;
}
uint64_t orc_invmod(uint64_t a, uint64_t p) { return orc_powmod(a, p - 2, p); }

int orc_is_prime(uint64_t n)
{
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; i++) { if (n % small[i] == 0) return n == small[i]; }
    uint64_t d = n - 1; int s = 0;
    while (!(d & 1)) { d >>= 1; s++; }
    for (int i = 0; i < 12; i++) {        /* deterministic for n < 3.3e24 with these bases */
        uint64_t x = orc_powmod(small[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int r = 1; r < s; r++) { x = orc_mulmod(x, x, n); if (x == n - 1) { comp = 0; break; } }
        if (comp) return 0;
    }
    return 1;
}

uint64_t orc_generator(uint64_t p)
{
    uint64_t f[64]; int nf = 0;
    uint64_t t = p - 1;
    for (uint64_t q = 2; q * q <= t; q++)            /* trial division: p - 1 = c 2^k with small c for NTT primes */
        if (t % q == 0) { f[nf++] = q; while (t % q == 0) t /= q; }
    if (t > 1) f[nf++] = t;
    for (uint64_t g = 2; g < p; g++) {
        int ok = 1;
        for (int i = 0; i < nf && ok; i++) if (orc_powmod(g, (p - 1) / f[i], p) == 1) ok = 0;
        if (ok) return g;
    }
    return 0;
}

uint64_t orc_root2k(uint64_t p, int logn)
{
    if (logn < 0 || logn > 63 || ((p - 1) & ((1ULL << logn) - 1))) return 0;
    return orc_powmod(orc_generator(p), (p - 1) >> logn, p);
}

/* ---- NTT definitions (E/ntt.h:1-8, E/ntt.c:630-669) ------------------------------------------------------------------ */
void orc_ntt_fwd(uint64_t *x, int logn, uint64_t p, uint64_t w)
{
    size_t n = (size_t)1 << logn;
    for (size_t h = n >> 1; h >= 1; h >>= 1) {               /* DIF: half-size h, root of order 2h */
        uint64_t wh = orc_powmod(w, n / (2 * h), p);
        for (size_t s = 0; s < n; s += 2 * h) {
            uint64_t wj = 1;
            for (size_t j = 0; j < h; j++) {
                uint64_t u = x[s + j], v = x[s + j + h];
                x[s + j] = u + v >= p ? u + v - p : u + v;
                x[s + j + h] = orc_mulmod(u >= v ? u - v : u + p - v, wj, p);
                wj = orc_mulmod(wj, wh, p);
            }
        }
        if (h == 1) break;
    }
}

void orc_ntt_inv(uint64_t *x, int logn, uint64_t p, uint64_t w)
{
    size_t n = (size_t)1 << logn;
    uint64_t wi = orc_invmod(w, p);
    for (size_t h = 1; h < n; h <<= 1) {                     /* DIT on bit-reversed input */
        uint64_t wh = orc_powmod(wi, n / (2 * h), p);
        for (size_t s = 0; s < n; s += 2 * h) {
            uint64_t wj = 1;
            for (size_t j = 0; j < h; j++) {
                uint64_t u = x[s + j], v = orc_mulmod(x[s + j + h], wj, p);
                x[s + j] = u + v >= p ? u + v - p : u + v;
                x[s + j + h] = u >= v ? u - v : u + p - v;
                wj = orc_mulmod(wj, wh, p);
            }
        }
    }
    uint64_t ninv = orc_invmod((uint64_t)(n % p), p);
    for (size_t i = 0; i < n; i++) x[i] = orc_mulmod(x[i], ninv, p);
}

static size_t bitrev(size_t k, int logn)
{
    size_t r = 0;
    for (int i = 0; i < logn; i++) { r = (r << 1) | (k & 1); k >>= 1; }
    return r;
}

void orc_dft_bitrev(const uint64_t *x, uint64_t *X, int logn, uint64_t p, uint64_t w)
{
    size_t n = (size_t)1 << logn;
    for (size_t k = 0; k < n; k++) {
        uint64_t wk = orc_powmod(w, k, p), t = 1, acc = 0;
        for (size_t i = 0; i < n; i++) { acc = (uint64_t)(((u128)x[i] * t + acc) % p); t = orc_mulmod(t, wk, p); }
        X[bitrev(k, logn)] = acc;
    }
}

/* ---- transforms of 3 2^k points (E/ntt.c:671-779; roots E/modarith.h:206-214) ------------------------------------------ */
uint64_t orc_root3k(uint64_t p, int logk)
{
    if (logk < 0 || logk > 63 || ((p - 1) & ((1ULL << logk) - 1)) || ((p - 1) >> logk) % 3) return 0;
    return orc_powmod(orc_generator(p), ((p - 1) >> logk) / 3, p);
}

/* the definition: X[k] = sum_i x[i] w^(ik) over n = 3 2^logk points, X[3t + r] written at r 2^logk + bitrev(t) */
void orc_dft3(const uint64_t *x, uint64_t *X, int logk, uint64_t p, uint64_t w)
{
    size_t m = (size_t)1 << logk, n = 3 * m;
    for (size_t k = 0; k < n; k++) {
        uint64_t wk = orc_powmod(w, k, p), t = 1, acc = 0;
        for (size_t i = 0; i < n; i++) { acc = (uint64_t)(((u128)x[i] * t + acc) % p); t = orc_mulmod(t, wk, p); }
        X[(k % 3) * m + bitrev(k / 3, logk)] = acc;
    }
}

/* the reference's decomposition: one DIF radix-3 layer over the thirds with the twiddles w^(r j), then the 2^logk
 * DIF transforms of the thirds with the root w^3 (E/ntt.c:671-681, k_r3_fwd) */
void orc_ntt3_fwd(uint64_t *x, int logk, uint64_t p, uint64_t w)
{
    size_t m = (size_t)1 << logk;
    uint64_t w3 = orc_powmod(w, m, p), w3s = orc_mulmod(w3, w3, p), wj = 1;
    for (size_t j = 0; j < m; j++) {
        uint64_t a0 = x[j], a1 = x[j + m], a2 = x[j + 2 * m];
        uint64_t y0 = (uint64_t)(((u128)a0 + a1 + a2) % p);
        uint64_t y1 = (uint64_t)(((u128)a0 + orc_mulmod(a1, w3, p) + orc_mulmod(a2, w3s, p)) % p);
        uint64_t y2 = (uint64_t)(((u128)a0 + orc_mulmod(a1, w3s, p) + orc_mulmod(a2, w3, p)) % p);
        x[j] = y0; x[j + m] = orc_mulmod(y1, wj, p); x[j + 2 * m] = orc_mulmod(y2, orc_mulmod(wj, wj, p), p);
        wj = orc_mulmod(wj, w, p);
    }
    uint64_t wm = orc_powmod(w, 3, p);
    for (int r = 0; r < 3; r++) orc_ntt_fwd(x + r * m, logk, p, wm);
}

/* the inverse: the 2^logk inverses of the thirds (x 2^-logk), then the mirrored layer with w^(-r j), w3^(-r i) and
 * 3^-1 (E/ntt.c:683-694, k_r3_inv) */
void orc_ntt3_inv(uint64_t *x, int logk, uint64_t p, uint64_t w)
{
    size_t m = (size_t)1 << logk;
    uint64_t wm = orc_powmod(w, 3, p);
    for (int r = 0; r < 3; r++) orc_ntt_inv(x + r * m, logk, p, wm);
    uint64_t wi = orc_invmod(w, p), w3i = orc_powmod(wi, m, p), w3is = orc_mulmod(w3i, w3i, p), inv3 = orc_invmod(3, p);
    uint64_t wj = 1;
    for (size_t j = 0; j < m; j++) {
        uint64_t a0 = x[j], a1 = orc_mulmod(x[j + m], wj, p), a2 = orc_mulmod(x[j + 2 * m], orc_mulmod(wj, wj, p), p);
        uint64_t x0 = (uint64_t)(((u128)a0 + a1 + a2) % p);
        uint64_t x1 = (uint64_t)(((u128)a0 + orc_mulmod(a1, w3i, p) + orc_mulmod(a2, w3is, p)) % p);
        uint64_t x2 = (uint64_t)(((u128)a0 + orc_mulmod(a1, w3is, p) + orc_mulmod(a2, w3i, p)) % p);
        x[j] = orc_mulmod(x0, inv3, p); x[j + m] = orc_mulmod(x1, inv3, p); x[j + 2 * m] = orc_mulmod(x2, inv3, p);
        wj = orc_mulmod(wj, wi, p);
    }
}

/* ---- Garner (E/crt.c:114-144, any np) ---------------------------------------------------------------------------------- */
/* w (nw words) mod p */
static uint64_t words_mod(const uint64_t *w, int nw, uint64_t p)
{
    u128 t = 0;
    for (int i = nw - 1; i >= 0; i--) t = ((t << 64) | w[i]) % p;
    return (uint64_t)t;
}
/* w += m * t (m has nm words, t one word); w has room for nw words */
static void words_addmul(uint64_t *w, int nw, const uint64_t *m, int nm, uint64_t t)
{
    uint64_t carry = 0;
    for (int i = 0; i < nw; i++) {
        u128 s = (u128)w[i] + carry + (i < nm ? (u128)m[i] * t : 0);
        w[i] = (uint64_t)s; carry = (uint64_t)(s >> 64);
    }
}

void orc_garner(const uint64_t *r, const uint64_t *P, int np, uint64_t *out)
{
    uint64_t x[8] = {0}, M[8] = {0};
    x[0] = r[0]; M[0] = P[0];
    for (int i = 1; i < np; i++) {
        uint64_t xm = words_mod(x, np, P[i]);
        uint64_t c = orc_invmod(words_mod(M, np, P[i]), P[i]);
        uint64_t d = r[i] >= xm ? r[i] - xm : r[i] + P[i] - xm;
        uint64_t t = orc_mulmod(d, c, P[i]);
        words_addmul(x, np, M, np, t);
        uint64_t z[8] = {0};
        words_addmul(z, np, M, np, P[i]);
        memcpy(M, z, sizeof M);
    }
    for (int i = 0; i < np; i++) out[i] = x[i];
}

/* ---- CRT with the windowed base-10^18 limb sum (E/crt.c:146-199, E/rns_mul.c:332-405) ---------------------------------- */
uint64_t orc_crt_carry(const uint64_t *const *planes, const uint64_t *P, int np, size_t ncoef,
                       const uint64_t *x, size_t nx, uint64_t *c, size_t nc)
{
    const int nd = np;                                        /* every P[i] < B, so prod P < B^np */
    uint64_t ring[8][8];                                      /* digits of coefficients k-nd+1 .. k */
    memset(ring, 0, sizeof ring);
    u128 carry = 0;
    size_t kend = nc + nd;
    uint64_t overflow = 0;
    for (size_t k = 0; k < kend; k++) {
        uint64_t *d = ring[k % 8];
        if (k < ncoef) {
            uint64_t r[8], w[8];
            for (int i = 0; i < np; i++) r[i] = planes[i][k];
            orc_garner(r, P, np, w);
            orc_words_to_dec(w, np, d, nd);
        } else {
            memset(d, 0, sizeof ring[0]);
        }
        u128 s = carry + ((x && k < nx) ? x[k] : 0);
        for (int j = 0; j < nd; j++) if (k >= (size_t)j) s += ring[(k - j) % 8][j];
        if (k < nc) c[k] = (uint64_t)(s % ORC_B);
        else overflow |= (uint64_t)(s % ORC_B);
        carry = s / ORC_B;
    }
    return overflow | (uint64_t)carry;
}

/* ---- the three-prime bound (E/crt.c:28-44) for any np: floor((prod P - 1) / (B-1)^2) --------------------------------- */
uint64_t orc_max_terms(const uint64_t *P, int np)
{
    uint64_t M[8] = {1, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < np; i++) { uint64_t z[8] = {0}; words_addmul(z, 8, M, 8, P[i]); memcpy(M, z, sizeof M); }
    /* M - 1 */
    for (int i = 0; i < 8; i++) { if (M[i]--) break; }
    u128 D = (u128)(ORC_B - 1) * (ORC_B - 1);
    /* binary long division of M (512 bits) by D (< 2^120); the quotient must fit 64 bits for the caller's use */
    u128 rem = 0; uint64_t q[8] = {0};
    for (int bit = 511; bit >= 0; bit--) {
        int top = (int)(rem >> 127);
        rem = (rem << 1) | ((M[bit / 64] >> (bit % 64)) & 1);
        if (top || rem >= D) { rem -= D; q[bit / 64] |= 1ULL << (bit % 64); }
    }
    for (int i = 1; i < 8; i++) if (q[i]) return UINT64_MAX;
    return q[0];
}

/* ---- the product path on the CPU --------------------------------------------------------------------------------------- */
int orc_mul_ntt(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *x, size_t nx,
                const uint64_t *P, int np)
{
    size_t nc = na + nb + (x ? 1 : 0);
    if (!na || !nb) { for (size_t k = 0; k < nc; k++) c[k] = (x && k < nx) ? x[k] : 0; return 0; }
    size_t ncoef = na + nb - 1, mn = na < nb ? na : nb;
    int logn = 1;
    while (((size_t)1 << logn) < ncoef) logn++;
    size_t n = (size_t)1 << logn;
    if (mn > orc_max_terms(P, np)) return -1;
    uint64_t *buf = (uint64_t *)calloc((size_t)2 * np * n, sizeof *buf);
    if (!buf) return -1;
    const uint64_t *planes[8];
    for (int i = 0; i < np; i++) {
        uint64_t w = orc_root2k(P[i], logn);
        if (!w) { free(buf); return -1; }
        uint64_t *A = buf + (size_t)2 * i * n, *Bp = A + n;
        for (size_t k = 0; k < na; k++) A[k] = a[k] % P[i];
        for (size_t k = 0; k < nb; k++) Bp[k] = b[k] % P[i];
        orc_ntt_fwd(A, logn, P[i], w);
        orc_ntt_fwd(Bp, logn, P[i], w);
        for (size_t k = 0; k < n; k++) A[k] = orc_mulmod(A[k], Bp[k], P[i]);
        orc_ntt_inv(A, logn, P[i], w);
        planes[i] = A;
    }
    uint64_t ov = orc_crt_carry(planes, P, np, ncoef, x, nx, c, nc);
    free(buf);
    return ov ? -1 : 0;
}

/* ---- residue checks (E/tests/harness.h:259-264, E/tests/tref.h:46-79) --------------------------------------------------- */
/* Horner over B from the top; above 2^16 limbs the array is cut into chunks whose residues and B^len are combined
 * (chunks in parallel under OpenMP) */
static uint64_t limbs_mod_serial(const uint64_t *a, size_t n, uint64_t q)
{
    uint64_t r = 0, bm = ORC_B % q;
    for (size_t i = n; i-- > 0;) r = (uint64_t)(((u128)r * bm + a[i] % q) % q);
    return r;
}
uint64_t orc_limbs_mod(const uint64_t *a, size_t n, uint64_t q)
{
    if (n <= ((size_t)1 << 16)) return limbs_mod_serial(a, n, q);
    enum { T = 64 };
    uint64_t res[T], pw[T];
    const uint64_t bm = ORC_B % q;
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < T; t++) {
        size_t lo = n * t / T, hi = n * (t + 1) / T;
        res[t] = limbs_mod_serial(a + lo, hi - lo, q);
        pw[t] = orc_powmod(bm, hi - lo, q);                   /* B^(chunk length) mod q */
    }
    uint64_t r = 0;
    for (int t = T - 1; t >= 0; t--) r = (uint64_t)(((u128)r * pw[t] + res[t]) % q);
    return r;
}

static size_t norm_len(const uint64_t *a, size_t n) { while (n && a[n - 1] == 0) n--; return n; }

/* t1_q[0..2] of the reference (E/verify.c:18): the first primes above 2^62 */
static const uint64_t T1Q[3] = {4611686018427388039ULL, 4611686018427388073ULL, 4611686018427388081ULL};

int orc_tref_res(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *c, size_t nc)
{
    na = norm_len(a, na); nb = norm_len(b, nb); nc = norm_len(c, nc);
    if (!na || !nb) return nc == 0;
    for (size_t i = 0; i < nc; i++) if (c[i] >= ORC_B) return 0;
    if (nc != na + nb && nc + 1 != na + nb) return 0;
    for (int j = 0; j < 3; j++) {
        uint64_t ra = orc_limbs_mod(a, na, T1Q[j]), rb = orc_limbs_mod(b, nb, T1Q[j]), rc = orc_limbs_mod(c, nc, T1Q[j]);
        if (orc_mulmod(ra, rb, T1Q[j]) != rc) return 0;
    }
    return 1;
}

/* ---- the 60-bit packed digit stream (E/packed_fmt.h:50-88) ---------------------------------------------------------
 * orc_pack60 writes the definition bit by bit: file limb j occupies bits [60 j, 60 j + 60) of a little-endian stream,
 * padded with zero bits to 15 (n / 2) + 8 (n & 1) bytes.  orc_unpack60 reads it back pairwise: 15 bytes hold two limbs
 * (E/packed_fmt.h:64-70), a last odd limb 8 bytes. */
uint64_t orc_pack60_bytes(size_t n) { return 15 * (uint64_t)(n / 2) + 8 * (uint64_t)(n & 1); }
void orc_pack60(const uint64_t *l, size_t n, uint8_t *out)
{
    memset(out, 0, orc_pack60_bytes(n));
    for (size_t j = 0; j < n; j++)
        for (int b = 0; b < 60; b++)
            if ((l[j] >> b) & 1) { const size_t bit = 60 * j + (size_t)b; out[bit >> 3] |= (uint8_t)(1u << (bit & 7)); }
}
void orc_unpack60(const uint8_t *in, size_t n, uint64_t *l)
{
    const uint64_t m60 = ((uint64_t)1 << 60) - 1;
    for (size_t i = 0; 2 * i + 1 < n; i++) {
        const uint8_t *p = in + 15 * i;
        uint64_t w0 = 0, w1 = 0;
        for (int k = 0; k < 8; k++) w0 |= (uint64_t)p[k] << (8 * k);
        for (int k = 0; k < 7; k++) w1 |= (uint64_t)p[8 + k] << (8 * k);
        l[2 * i] = w0 & m60;
        l[2 * i + 1] = (w0 >> 60) | (w1 << 4);
    }
    if (n & 1) {
        const uint8_t *p = in + 15 * (n / 2);
        uint64_t a = 0;
        for (int k = 0; k < 8; k++) a |= (uint64_t)p[k] << (8 * k);
        l[n - 1] = a;
    }
}
