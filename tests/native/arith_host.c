/* arith_host.c - host build of the product library's arithmetic header (ntt_amd/csrc/arith.h) for the CPU tests:
 * the same IEEE FP64 operations as the device (compiled with -ffp-contract=off), exported for ctypes. */
#include "../../ntt_amd/csrc/arith.h"

double t_mul_shoup(double a, double w, double wq, double p) { return mul_shoup(a, w, wq, p); }
double t_mul_barrett(double a, double b, double p, double pinv) { return mul_barrett(a, b, p, pinv); }
double t_reduce(double s, double p, double pinv) { return reduce(s, p, pinv); }
double t_limb_to_res(uint64_t x, double p, double pinv) { return limb_to_res(x, p, pinv); }
uint64_t t_res_canon(double v, double p, double pinv) { return res_canon(v, p, pinv); }
void t_div_b(uint64_t hi, uint64_t lo, uint64_t *q, uint64_t *r) { div_b(hi, lo, q, r); }
void t_mulmod_b(uint64_t t, uint64_t m, uint64_t mp, uint64_t *a, uint64_t *b) { mulmod_b(t, m, mp, a, b); }
uint64_t t_f64_to_u52(double v) { return f64_to_u52(v); }
uint64_t t_prime(int i) { return NTTM_PRIMES[i]; }
uint64_t t_divb_v(void) { return NTTM_DIVB_V; }
