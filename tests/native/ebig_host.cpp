/* ebig_host.cpp - the e pipeline's host-side Newton reciprocal and corrected quotient (ntt_amd/csrc/ebig.h), built for
 * the host with a schoolbook multiplier in place of the device product, so the tests can check them against exact
 * integers without a GPU. */
#include <string.h>
#include "../../ntt_amd/csrc/ebig.h"

static long g_products;
static size_t g_max_operand;

struct SchoolMul {
    int operator()(const Big &a, const Big &b, Big &c)
    {
        g_products++;
        g_max_operand = std::max(g_max_operand, std::max(a.size(), b.size()));
        if (a.empty() || b.empty()) { c.clear(); return 0; }
        c.assign(a.size() + b.size(), 0);
        for (size_t i = 0; i < a.size(); i++) {
            unsigned __int128 carry = 0;
            for (size_t j = 0; j < b.size(); j++) {
                const unsigned __int128 t = (unsigned __int128)a[i] * b[j] + c[i + j] + carry;
                carry = t / EB_BASE;
                c[i + j] = (uint64_t)(t - carry * EB_BASE);
            }
            size_t k = i + b.size();
            while (carry) {
                const unsigned __int128 t = (unsigned __int128)c[k] + carry;
                carry = t / EB_BASE;
                c[k++] = (uint64_t)(t - carry * EB_BASE);
            }
        }
        big_norm(c);
        return 0;
    }
};

static Big load(const uint64_t *v, size_t n) { Big r(v, v + n); big_norm(r); return r; }
static size_t store(const Big &b, uint64_t *out, size_t cap)
{
    if (b.size() > cap) return (size_t)-1;
    if (!b.empty()) memcpy(out, b.data(), b.size() * 8);    /* an empty Big has no data pointer to copy from */
    return b.size();
}

extern "C" {
long eb_t_products(void) { return g_products; }
size_t eb_t_max_operand(void) { return g_max_operand; }
void eb_t_reset(void) { g_products = 0; g_max_operand = 0; }
static long g_wraps;
long eb_t_wraps(void) { return g_wraps; }
/* x ~ B^(n+m) / Q; returns the limbs of x written, or -1.  wrap: the Newton product modulo B^L - 1 wherever it is
 * shorter (0: always the whole product) */
long eb_t_recip(const uint64_t *q, size_t n, size_t m, uint64_t *x, size_t cap, int wrap)
{
    SchoolMul mul;
    HostOps<SchoolMul> o(mul);
    if (!wrap) o.wrap_h = (size_t)-1;
    Big X;
    const int rc = eb_recip(o, load(q, n), m, X);
    g_wraps = o.wraps;
    if (rc) return -1;
    return (long)store(X, x, cap);
}
/* floor(A B^m / Q) through eb_recip (one guard limb) and eb_quotient (full: always through the remainder); returns
 * the limbs written, or -1 */
long eb_t_quotient(const uint64_t *a, size_t na, const uint64_t *q, size_t n, size_t m, uint64_t *y, size_t cap, int full)
{
    SchoolMul mul;
    HostOps<SchoolMul> o(mul);
    const Big Q = load(q, n);
    Big X, Y;
    if (eb_recip(o, Q, m + 1, X) || eb_quotient(o, load(a, na), Q, m, X, Y, full != 0)) return -1;
    return (long)store(Y, y, cap);
}
/* floor(B^e / D) by the host long division */
long eb_t_div_pow(size_t e, const uint64_t *d, size_t nd, uint64_t *out, size_t cap)
{
    return (long)store(big_div_pow(e, load(d, nd)), out, cap);
}
}
