/* orc_gmp.c - GMP bridge of the oracle (test infrastructure only; see oracle.h).
 * Conversions between base-10^18 limbs and mpz by a product tree: value = lo + hi B^h with h = 2^k, n/2 <= h < n, the
 * powers B^(2^k) cached (the scheme of E/tests/harness.h:104-199); halves run as OpenMP tasks above a size threshold.
 * Built without GMP (no gmp.h), every entry reports its absence. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "oracle.h"

#ifdef ORC_HAVE_GMP
#include <gmp.h>
#include <pthread.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LEAF 16
#define TASK_MIN 8192
static mpz_t g_pw[48];
static int g_npw = 0;
static pthread_mutex_t g_pw_mx = PTHREAD_MUTEX_INITIALIZER;

static const mpz_t *pow2k(int k)                              /* B^(2^k) */
{
    if (__atomic_load_n(&g_npw, __ATOMIC_ACQUIRE) <= k) {
        pthread_mutex_lock(&g_pw_mx);
        for (int n = g_npw; n <= k; n++) {
            mpz_init(g_pw[n]);
            if (!n) mpz_set_ui(g_pw[0], ORC_B); else mpz_mul(g_pw[n], g_pw[n - 1], g_pw[n - 1]);
            __atomic_store_n(&g_npw, n + 1, __ATOMIC_RELEASE);
        }
        pthread_mutex_unlock(&g_pw_mx);
    }
    return (const mpz_t *)&g_pw[k];
}
static int split_k(size_t n) { int k = 0; while (((size_t)2 << k) < n) k++; return k; }

static int g_progress;                                        /* 1: report the top combines on stderr */
static double now_s(void);
static void from_rec(mpz_t z, const uint64_t *a, size_t n, int par)
{
    if (n <= LEAF) {
        mpz_set_ui(z, 0);
        for (size_t i = n; i-- > 0;) { mpz_mul_ui(z, z, ORC_B); mpz_add_ui(z, z, a[i]); }
        return;
    }
    int k = split_k(n); size_t h = (size_t)1 << k;
    mpz_t t; mpz_init(t);
#pragma omp task shared(t) if (par && n >= TASK_MIN)
    from_rec(t, a, h, par);
    from_rec(z, a + h, n - h, par);
#pragma omp taskwait
    if (g_progress && n >= ((size_t)1 << 25)) { fprintf(stderr, "[oracle] combine %zu limbs at %.1f s\n", n, now_s()); fflush(stderr); }
    mpz_mul(z, z, *pow2k(k)); mpz_add(z, z, t);
    mpz_clear(t);
}
static void to_rec(uint64_t *a, size_t n, const mpz_t z, int par)    /* the n limbs of z mod B^n */
{
    if (n <= LEAF) {
        mpz_t t; mpz_init_set(t, z);
        for (size_t i = 0; i < n; i++) a[i] = mpz_fdiv_q_ui(t, t, ORC_B);
        mpz_clear(t); return;
    }
    int k = split_k(n); size_t h = (size_t)1 << k;
    mpz_t q, r; mpz_inits(q, r, NULL);
    mpz_tdiv_qr(q, r, z, *pow2k(k));
#pragma omp task shared(r) if (par && n >= TASK_MIN)
    to_rec(a, h, r, par);
    to_rec(a + h, n - h, q, par);
#pragma omp taskwait
    mpz_clears(q, r, NULL);
}
static int use_par(size_t n)
{
    if (n > LEAF) pow2k(split_k(n));
#ifdef _OPENMP
    return n >= TASK_MIN && !omp_in_parallel();
#else
    return 0;
#endif
}
static void mpz_from_limbs(mpz_t z, const uint64_t *a, size_t n)
{
    if (!n) { mpz_set_ui(z, 0); return; }
    if (use_par(n)) {
#pragma omp parallel
#pragma omp single
        from_rec(z, a, n, 1);
    } else from_rec(z, a, n, 0);
}
static void mpz_to_limbs_n(uint64_t *a, size_t n, const mpz_t z)
{
    if (!n) return;
    if (use_par(n)) {
#pragma omp parallel
#pragma omp single
        to_rec(a, n, z, 1);
    } else to_rec(a, n, z, 0);
}
static double now_s(void) { struct timespec t; clock_gettime(CLOCK_MONOTONIC, &t); return t.tv_sec + 1e-9 * t.tv_nsec; }

/* the full-size CPU baseline and exact check of one large product, reporting each step on stderr: a and b converted
 * (t[0] seconds, all host threads), one mpz_mul (t[1], one thread), c converted (t[2]); returns 1 when c = a b exactly,
 * 0 when not (limbs >= B included) */
int orc_gmp_full_check(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *c, size_t nc, double *t)
{
    for (size_t i = 0; i < nc; i++) if (c[i] >= ORC_B) return 0;
    g_progress = 1;
    mpz_t x, y, z, w; mpz_inits(x, y, z, w, NULL);
    double t0 = now_s();
    fprintf(stderr, "[oracle] converting a and b (%zu, %zu limbs)\n", na, nb); fflush(stderr);
    mpz_from_limbs(x, a, na); mpz_from_limbs(y, b, nb);
    t[0] = now_s() - t0;
    fprintf(stderr, "[oracle] converted in %.1f s; mpz_mul\n", t[0]); fflush(stderr);
    t0 = now_s(); mpz_mul(z, x, y); t[1] = now_s() - t0;
    fprintf(stderr, "[oracle] mpz_mul %.1f s; converting c (%zu limbs)\n", t[1], nc); fflush(stderr);
    mpz_clears(x, y, NULL);
    t0 = now_s(); mpz_from_limbs(w, c, nc); t[2] = now_s() - t0;
    const int ok = mpz_cmp(z, w) == 0;
    fprintf(stderr, "[oracle] c converted in %.1f s; equal %d\n", t[2], ok); fflush(stderr);
    mpz_clears(z, w, NULL);
    g_progress = 0;
    return ok;
}
static int cmp_d(const void *x, const void *y) { double a = *(const double *)x, b = *(const double *)y; return (a > b) - (a < b); }

int orc_have_gmp(void) { return 1; }

int orc_gmp_mul(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb)
{
    mpz_t x, y, z; mpz_inits(x, y, z, NULL);
    mpz_from_limbs(x, a, na); mpz_from_limbs(y, b, nb); mpz_mul(z, x, y);
    mpz_to_limbs_n(c, na + nb, z);
    mpz_clears(x, y, z, NULL);
    return 0;
}

double orc_gmp_time_mul(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int reps)
{
    if (reps < 1) reps = 1;
    if (reps > 64) reps = 64;
    double t[64];
    mpz_t x, y, z; mpz_inits(x, y, z, NULL);
    mpz_from_limbs(x, a, na); mpz_from_limbs(y, b, nb);
    mpz_mul(z, x, y);                                           /* warm: z allocated at full size */
    for (int r = 0; r < reps; r++) { double t0 = now_s(); mpz_mul(z, x, y); t[r] = now_s() - t0; }
    mpz_clears(x, y, z, NULL);
    qsort(t, reps, sizeof t[0], cmp_d);
    return t[reps / 2];
}

/* n products a[i] b[i]: seconds of the mpz_mul loop alone (operands converted and results sized beforehand), on
 * `threads` OpenMP threads (dynamic schedule), median of reps */
double orc_gmp_time_batch(const uint64_t *const *a, const size_t *na, const uint64_t *const *b, const size_t *nb, size_t n,
                          int threads, int reps)
{
    if (reps < 1) reps = 1;
    if (reps > 64) reps = 64;
    if (threads < 1) threads = 1;
    if (!n) return 0.0;
    mpz_t *x = (mpz_t *)malloc(3 * n * sizeof(mpz_t));
    if (!x) return -1.0;
    mpz_t *y = x + n, *z = y + n;
#pragma omp parallel for schedule(dynamic, 1)    /* conversions on every host thread (the power cache is locked) */
    for (long i = 0; i < (long)n; i++) {
        mpz_inits(x[i], y[i], z[i], NULL);
        mpz_from_limbs(x[i], a[i], na[i]);
        mpz_from_limbs(y[i], b[i], nb[i]);
        mpz_mul(z[i], x[i], y[i]);                              /* warm: every result allocated at full size */
    }
    double t[64];
    for (int r = 0; r < reps; r++) {
        const double t0 = now_s();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
        for (long i = 0; i < (long)n; i++) mpz_mul(z[i], x[i], y[i]);
        t[r] = now_s() - t0;
    }
    for (size_t i = 0; i < n; i++) mpz_clears(x[i], y[i], z[i], NULL);
    free(x);
    qsort(t, reps, sizeof t[0], cmp_d);
    return t[reps / 2];
}

int orc_gmp_check(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *xx, size_t nx,
                  const uint64_t *c, size_t nc)
{
    for (size_t i = 0; i < nc; i++) if (c[i] >= ORC_B) return 0;
    mpz_t x, y, z, w; mpz_inits(x, y, z, w, NULL);
    mpz_from_limbs(x, a, na); mpz_from_limbs(y, b, nb); mpz_mul(z, x, y);
    if (xx && nx) { mpz_from_limbs(w, xx, nx); mpz_add(z, z, w); }
    mpz_from_limbs(w, c, nc);
    int ok = mpz_cmp(z, w) == 0;
    mpz_clears(x, y, z, w, NULL);
    return ok;
}

/* n products checked on every host thread (one product per thread at a time): ok[i] = c_i == a_i b_i exactly and every
 * limb of c_i below B; returns the number that fail */
size_t orc_gmp_check_batch(const uint64_t *const *a, const size_t *na, const uint64_t *const *b, const size_t *nb,
                           const uint64_t *const *c, const size_t *nc, size_t n, uint8_t *ok)
{
    size_t bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
    for (long i = 0; i < (long)n; i++) {
        int good = 1;
        for (size_t k = 0; k < nc[i] && good; k++) if (c[i][k] >= ORC_B) good = 0;
        if (good) {
            mpz_t x, y, z, w; mpz_inits(x, y, z, w, NULL);
            from_rec(x, a[i], na[i], 0); from_rec(y, b[i], nb[i], 0); mpz_mul(z, x, y);
            from_rec(w, c[i], nc[i], 0);
            good = mpz_cmp(z, w) == 0;
            mpz_clears(x, y, z, w, NULL);
        }
        ok[i] = (uint8_t)good;
        bad += !good;
    }
    return bad;
}

static size_t norm_len(const uint64_t *a, size_t n) { while (n && a[n - 1] == 0) n--; return n; }

/* E/tests/tref.h:59-78: the low 64 limbs exact; floor(c / B^(ka+kb)) within [Ah Bh, Ah Bh + Ah + Bh] */
int orc_gmp_tref_edges(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *c, size_t nc)
{
    enum { W = 64 };
    na = norm_len(a, na); nb = norm_len(b, nb); nc = norm_len(c, nc);
    if (!na || !nb) return nc == 0;
    mpz_t x, y, z, m; mpz_inits(x, y, z, m, NULL);
    int ok = 1;
    {
        size_t la = na < W ? na : W, lb = nb < W ? nb : W;
        uint64_t want[W];
        mpz_from_limbs(x, a, la); mpz_from_limbs(y, b, lb); mpz_mul(z, x, y);
        mpz_fdiv_r(z, z, *pow2k(6));
        mpz_to_limbs_n(want, W, z);
        for (size_t i = 0; i < W && ok; i++) if ((i < nc ? c[i] : 0) != want[i]) ok = 0;
    }
    if (ok) {
        size_t ka = na > W ? na - W : 0, kb = nb > W ? nb - W : 0;
        mpz_from_limbs(x, a + ka, na - ka); mpz_from_limbs(y, b + kb, nb - kb); mpz_mul(z, x, y);
        mpz_from_limbs(m, c + ka + kb, nc > ka + kb ? nc - ka - kb : 0);
        if (mpz_cmp(m, z) < 0) ok = 0;
        else { mpz_add(z, z, x); mpz_add(z, z, y); if (mpz_cmp(m, z) > 0) ok = 0; }
    }
    mpz_clears(x, y, z, m, NULL);
    return ok;
}

#else  /* no GMP */
int orc_have_gmp(void) { return 0; }
int orc_gmp_mul(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb)
{ (void)c; (void)a; (void)na; (void)b; (void)nb; return -1; }
double orc_gmp_time_mul(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int reps)
{ (void)a; (void)na; (void)b; (void)nb; (void)reps; return -1.0; }
double orc_gmp_time_batch(const uint64_t *const *a, const size_t *na, const uint64_t *const *b, const size_t *nb, size_t n,
                          int threads, int reps)
{ (void)a; (void)na; (void)b; (void)nb; (void)n; (void)threads; (void)reps; return -1.0; }
int orc_gmp_check(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *x, size_t nx,
                  const uint64_t *c, size_t nc)
{ (void)a; (void)na; (void)b; (void)nb; (void)x; (void)nx; (void)c; (void)nc; return -1; }
int orc_gmp_tref_edges(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *c, size_t nc)
{ (void)a; (void)na; (void)b; (void)nb; (void)c; (void)nc; return -1; }
int orc_gmp_full_check(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *c, size_t nc, double *t)
{ (void)a; (void)na; (void)b; (void)nb; (void)c; (void)nc; (void)t; return -1; }
size_t orc_gmp_check_batch(const uint64_t *const *a, const size_t *na, const uint64_t *const *b, const size_t *nb,
                           const uint64_t *const *c, const size_t *nc, size_t n, uint8_t *ok)
{ (void)a; (void)na; (void)b; (void)nb; (void)c; (void)nc; (void)n; (void)ok; return (size_t)-1; }
#endif
