/* rns_nttm.c - the ecalc side of the drop-in: the reference's one-device product tiers implemented over libnttm.so.
 *
 * A maintainer adds this file to ecalc's build in place of the one-device paths of rns_mul.c (E/Makefile:2-5 compiles
 * C with hipcc -x hip) and links -lnttm.  It includes the reference's own headers (rns_mul.h, dbig.h, fatal.h) and
 * include/nttm.h, so it compiles only where both are present; the layout checks below make the compiler prove that
 * rns_prod (E/rns_mul.h:50-51) and nttm_prod are the same record, so the batch passes the caller's array through
 * unchanged.  Errors end the process through ec_fatal, the reference's convention (E/fatal.h:19-23).
 *
 * Residence: binsplit's batch tier hands over host pool memory (E/binsplit.c:1622,1636-1644) and the tests device
 * memory; nttm_mul_batch takes both and stages host arrays through HBM, as rns_mul_batch does (E/rns_mul.c:807-817).
 * A dbig of the one-device layout (ECALC_PE_PER_GPU=1, E/dbig.h:14-16) keeps its limbs at q[0] + off on this device. */
#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>
#include "rns_mul.h"
#include "dbig.h"
#include "fatal.h"
#include "nttm.h"

_Static_assert(sizeof(rns_prod) == sizeof(nttm_prod), "rns_prod and nttm_prod differ in size");
_Static_assert(offsetof(rns_prod, a) == offsetof(nttm_prod, a), "field a");
_Static_assert(offsetof(rns_prod, na) == offsetof(nttm_prod, na), "field na");
_Static_assert(offsetof(rns_prod, b) == offsetof(nttm_prod, b), "field b");
_Static_assert(offsetof(rns_prod, nb) == offsetof(nttm_prod, nb), "field nb");
_Static_assert(offsetof(rns_prod, c) == offsetof(nttm_prod, c), "field c");
_Static_assert(offsetof(rns_prod, x) == offsetof(nttm_prod, x), "field x");
_Static_assert(offsetof(rns_prod, nx) == offsetof(nttm_prod, nx), "field nx");
_Static_assert(offsetof(rns_prod, ncn) == offsetof(nttm_prod, ncn), "field ncn");

static void nttm_ck(int rc, const char *what)
{
    if (rc) ec_fatal(EC_RC_FATAL, "%s: nttm error %d: %s\n", what, rc, nttm_last_error());
}

/* the limbs of a one-device dbig */
static uint64_t *dev_limbs(const dbig *x) { return x->q[0] ? x->q[0] + x->off : NULL; }

int rns_init(int pool_log)                                   /* E/rns_mul.h:40 */
{
    (void)pool_log;                                          /* nttm sizes its workspace per product */
    nttm_config cfg = {0};                                   /* nprimes 0: three primes while the bound allows, else four */
    nttm_ck(nttm_init(0, &cfg), "rns_init");
    return 1;                                                /* one device */
}

void rns_shutdown(void) { nttm_shutdown(); }                 /* E/rns_mul.h:21 */

/* E/rns_mul.h:50-53: N products c = a b (+ x), operands in host or device memory; ncn filled */
void rns_mul_batch(rns_prod *P, size_t N)
{
    nttm_ck(nttm_mul_batch((nttm_prod *)P, N, NULL), "rns_mul_batch");
}

/* E/rns_mul.h:76: C = A B with nc >= na + nb limbs of room reserved, nc limbs written (the limbs above na + nb zero) */
void rns_local_core(dbig *C, const dbig *A, const dbig *B, size_t nc)
{
    const size_t na = A->n, nb = B->n;
    if (nc < na + nb) ec_fatal(EC_RC_FATAL, "rns_local_core: nc %zu below na + nb = %zu\n", nc, na + nb);
    db_reserve(C, nc);
    nttm_ck(nttm_mul(dev_limbs(C), dev_limbs(A), na, dev_limbs(B), nb, NULL), "rns_local_core");
    if (nc > na + nb) {
        if (hipMemset(dev_limbs(C) + na + nb, 0, (nc - na - nb) * 8) != hipSuccess)
            ec_fatal(EC_RC_FATAL, "rns_local_core: clearing %zu limbs failed\n", nc - na - nb);
    }
    C->n = nc;
    db_norm(C);
}

/* the band product C = (A B mod B^w) with the limbs below lowcut approximate (E/rns_mul.h:66-73) */
static void band(dbig *C, const dbig *A, const dbig *B, size_t lowcut, size_t w, const char *what)
{
    const size_t na = A->n, nb = B->n, nout = na + nb < w ? na + nb : w;
    db_reserve(C, nout ? nout : 1);
    nttm_ck(nttm_mul_cut(dev_limbs(C), dev_limbs(A), na, dev_limbs(B), nb, lowcut < nout ? lowcut : nout, w, NULL), what);
    C->n = nout;
    db_norm(C);
}
void rns_mul_low_db(dbig *C, const dbig *A, const dbig *B, size_t w) { band(C, A, B, 0, w, "rns_mul_low_db"); }
void rns_mul_high_db(dbig *C, const dbig *A, const dbig *B, size_t cut) { band(C, A, B, cut, SIZE_MAX, "rns_mul_high_db"); }
void rns_mul_band_db(dbig *C, const dbig *A, const dbig *B, size_t lowcut, size_t w) { band(C, A, B, lowcut, w, "rns_mul_band_db"); }
void rns_mul_local(dbig *C, const dbig *A, const dbig *B, size_t lowcut, size_t w) { band(C, A, B, lowcut, w, "rns_mul_local"); }
void rns_mul_dist_db(dbig *C, const dbig *A, const dbig *B) { band(C, A, B, 0, SIZE_MAX, "rns_mul_dist_db"); }
