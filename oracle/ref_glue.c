/* ref_glue.c - exports the reference's own header-only helpers (static inline functions of E/modarith.h, the operand
 * generator and conversions of E/tests/harness.h, the product checker of E/tests/tref.h) from the reference build under
 * oracle/_ref, so the tests can run the reference's code beside the oracle.  Compiled together with the reference's
 * sources where they lie; nothing of the reference is copied. */
#define TREF_NO_DEVICE
#include "tref.h"
#include "modarith.h"
#include "crt.h"
#include "packed_fmt.h"
#include "verify.h"

void ref_gen_limbs(uint64_t *a, size_t n, int kind, uint64_t *seed)
{
    rng_t r; r.s = *seed; gen_limbs(a, n, kind, &r); *seed = r.s;
}
void ref_div1e18(uint64_t hi, uint64_t lo, uint64_t *q, uint64_t *r) { ec_div1e18(hi, lo, q, r); }
void ref_words_to_dec3(const uint64_t *c, uint64_t *d) { ec_words_to_dec3(c, d); }
void ref_words_to_dec4(const uint64_t *c, uint64_t *d) { ec_words_to_dec4(c, d); }
void ref_primes(uint64_t *p) { for (int i = 0; i < EC_NP; i++) p[i] = ec_P[i]; }
uint64_t ref_np3_max_terms(void) { crt_init(); return ec_np3_max_terms; }
void ref_set_np(int np) { crt_init(); ec_np = np; }
int ref_tref_ok(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, const uint64_t *c, size_t nc, int mode)
{
    char why[160]; return tref_ok(a, na, b, nb, c, nc, mode, why, sizeof why);
}
uint64_t ref_limbs_mod_u64(const uint64_t *a, size_t n, uint64_t q) { return limbs_mod_u64(a, n, q); }
void ref_mpz_mul_limbs(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb)
{
    mpz_t x, y, z; mpz_inits(x, y, z, NULL);
    mpz_from_limbs(x, a, na); mpz_from_limbs(y, b, nb); mpz_mul(z, x, y);
    hv_to_limbs_n(c, na + nb, z);
    mpz_clears(x, y, z, NULL);
}
void ref_tables(uint64_t *gen, uint64_t *w33, uint64_t *w3x33, uint64_t *mu115)
{
    for (int i = 0; i < EC_NP; i++) { gen[i] = ec_G[i]; w33[i] = ec_W33[i]; w3x33[i] = ec_W3X33[i]; mu115[i] = ec_MU115[i]; }
}
uint64_t ref_root(int i, int logn) { return ec_root(i, logn); }
uint64_t ref_root3(int i, int logk) { return ec_root3(i, logk); }
/* the packed digit stream (E/packed_fmt.h) and the T1 primes of its header residues (E/verify.c:18) */
void ref_ecp60_pack(const uint64_t *l, size_t n, uint8_t *out) { ecp60_pack(l, n, out); }
uint64_t ref_ecp60_bytes(uint64_t n) { return ecp60_bytes(n); }
void ref_t1_q(uint64_t *q) { for (int i = 0; i < T1_NQ; i++) q[i] = t1_q[i]; }
