/* ntt.hip - transform kernels: the tiled pass (forward DIF / inverse DIT, optional fused pointwise product and final
 * scale), operand load (limbs -> residues of every prime), standalone pointwise product and format conversions.
 *
 * Reference counterparts (E/ = ecalc/): the passes replace k_b16r / k_b16 / k_b1r (E/ntt.c:71-359) with one tile
 * kernel of runtime geometry (plan.h); the fused pointwise product replaces ntt_inv_pw_y's MODE 2 (E/ntt.c:591-596);
 * k_load replaces k_scatter4 / k_lload / ntt_load (E/rns_mul.c:430-443,1051-1055; E/ntt.c:376-380,624-628).
 */
#include <string.h>
#include "internal.h"
#include "plan.h"
/* the round loops carry an unroll request that applies to the fixed-shape instantiations only */
#pragma clang diagnostic ignored "-Wpass-failed"


/* ---- the pass kernel -------------------------------------------------------------------------------------------------
 * A tile of T = W x 2^m x H points (T <= 2^13) is handled by T / 16 threads holding 16 points each in registers.  The m
 * stages run in rounds of up to 4 stages (radix 2..16 in registers); between rounds the tile goes through LDS once.
 * In a round over row bits [s_lo, s_lo + logR) (L = logw + s_lo), thread t owns G = 16 / R groups of R points; group g
 * has the base element
 *     eb = (tau mod 2^L) | (tau div 2^L) << (L + logR),   tau = g * nthr + t,
 * and point j of the group is element eb + j 2^L.  For every point the lanes of a wave touch consecutive LDS words
 * (and, in the first and last round, consecutive global addresses of a row).  Since the round's bits lie inside the
 * row field, point j sits at global offset ob + j 2^(s_lo + lo) and at bit-reversed row kb + brev(j) 2^(m - s_lo - logR):
 * addresses and boundary-twiddle exponents of a group are a base plus compile-time multiples.  Each thread reads and
 * writes exactly the elements of its groups within a round, so one barrier separates consecutive rounds.
 * Lazy reduction (bounds of arith.h): the DIF reduces its sum path at odd levels and at a round's last level (values
 * entering a butterfly stay within [-2p, 2p]); the DIT reduces both outputs at a round's last level only (after levels
 * 0..3 the outputs are within 1.75p, 2.5p, 3.5p, 4.5p + 2, and every mul_shoup input within 4p; with the trivial
 * first level of a lowest-bits round, 2p, 2.75p + 1, 3.75p + 2, 4.75p + 3, mul_shoup inputs within 3.75p + 2).
 * Where the round's low row bits are known to be zero at compile time, twiddles w^0 = 1 are not multiplied: the DIF
 * difference then follows the sum path's rule, the DIT skips them at its first level (where all of them are 1). */
#define PASS_PTS 16               /* points per thread of full tiles; tiles below 16 points use one thread */

template <int BITS>
__device__ __forceinline__ constexpr uint32_t brev_c(uint32_t j)
{
    uint32_t r = 0;
    for (int i = 0; i < BITS; i++) r |= ((j >> i) & 1u) << (BITS - 1 - i);
    return r;
}

/* rlow == 0 known at compile time (the round over the lowest bits of a fixed-shape tile): the butterflies with
 * jj = 0 then have the twiddle w^0 = 1 */
#define NTTM_TRIVIAL_W(rlow, jj) ((jj) == 0 && __builtin_constant_p(rlow) && (rlow) == 0)

/* a stage table in global memory, read through a buffer resource: the table's base in scalar registers, the lane's
 * part of an index as the 32-bit vector offset and its compile-time part as the scalar offset, so a twiddle's address
 * costs no vector instruction (a 64-bit pointer sum costs two, in kernels bound by vector issue) */
struct GTab {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ explicit GTab(const double2 *p)
        : r(__builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, 0x7fffffff, 0x00020000)) {}
};
/* a stage table in LDS (or any pointer): restrict, so that its loads move freely around the tile's LDS stores */
typedef const double2 *__restrict__ LTab;
/* entry lane + fixed of a stage table (lane: the part that varies across lanes; fixed: a compile-time part) */
__device__ __forceinline__ double2 tw_at(LTab tw, uint32_t lane, uint32_t fixed) { return tw[lane + fixed]; }
__device__ __forceinline__ double2 tw_at(const GTab &tw, uint32_t lane, uint32_t fixed)
{
    typedef int v4i __attribute__((ext_vector_type(4)));
    const v4i x = __builtin_amdgcn_raw_buffer_load_b128(tw.r, (int)(lane * 16u), (int)(fixed * 16u), 0);
    double2 d;
    __builtin_memcpy(&d, &x, 16);
    return d;
}

template <int LOGR, typename TW>
__device__ __forceinline__ void dif_group(double *v, int s_lo, uint32_t rlow, TW tw, double p, double pinv)
{
    constexpr int R = 1 << LOGR;
#pragma unroll
    for (int lev = 0; lev < LOGR; lev++) {
        const int k = LOGR - 1 - lev, h = 1 << k;
        const bool red = (lev & 1) || lev == LOGR - 1;
#pragma unroll
        for (int jj = 0; jj < h; jj++) {
            /* a trivial twiddle: the difference follows the sum path's bound and reduction rule */
            const bool one = NTTM_TRIVIAL_W(rlow, jj);
            const double2 w = one ? make_double2(1.0, 0.0) : tw_at(tw, rlow, (1u << (s_lo + k)) + ((uint32_t)jj << s_lo));
#pragma unroll
            for (int j = jj; j < R; j += 2 * h) {
                const double u = v[j], x = v[j + h];
                const double sm = u + x, df = u - x;
                v[j] = red ? reduce(sm, p, pinv) : sm;
                v[j + h] = one ? (red ? reduce(df, p, pinv) : df) : mul_shoup(df, w.x, w.y, p);
            }
        }
    }
}

template <int LOGR, typename TW>
__device__ __forceinline__ void dit_group(double *v, int s_lo, uint32_t rlow, TW tw, double p, double pinv)
{
    constexpr int R = 1 << LOGR;
#pragma unroll
    for (int lev = 0; lev < LOGR; lev++) {
        const int k = lev, h = 1 << k;
        const bool red = lev == LOGR - 1;
#pragma unroll
        for (int jj = 0; jj < h; jj++) {
            /* a trivial twiddle at the first level only (inputs within p: outputs within 2p, and the later bounds
             * below hold with 0.25p to spare) */
            const bool one = lev == 0 && NTTM_TRIVIAL_W(rlow, jj);
            const double2 w = one ? make_double2(1.0, 0.0) : tw_at(tw, rlow, (1u << (s_lo + k)) + ((uint32_t)jj << s_lo));
#pragma unroll
            for (int j = jj; j < R; j += 2 * h) {
                const double u = v[j], x = one ? v[j + h] : mul_shoup(v[j + h], w.x, w.y, p);
                const double s1 = u + x, s2 = u - x;
                v[j] = red ? reduce(s1, p, pinv) : s1;
                v[j + h] = red ? reduce(s2, p, pinv) : s2;
            }
        }
    }
}

/* geometry of the pass seen by one thread */
struct PassPos {
    int m, logw, lo, nthr, t;
};

/* f(slot, e, o, c, k) for every slot of a round: e the LDS index, o the global offset in the tile, c the column inside
 * the tile, k the bit-reversed row */
/* PAD: e is the padded LDS index (SwzLds below) of the element; eb and the slot offsets j 2^L have disjoint bits, so
 * pad(eb + j 2^L) = pad(eb) + pad(j 2^L), the second term a compile-time constant in the fixed-shape kernels */
template <int LOGR, int PTS, bool PAD = false, typename F>
__device__ __forceinline__ void for_slots(const PassPos &pp, int s_lo, F f)
{
    constexpr int R = 1 << LOGR, G = PTS / R;
    const int L = pp.logw + s_lo;
#pragma unroll
    for (int g = 0; g < G; g++) {
        const uint32_t tau = (uint32_t)(g * pp.nthr + pp.t);
        const uint32_t eb = (tau & ((1u << L) - 1)) | ((tau >> L) << (L + LOGR));
        const uint32_t pe = PAD ? eb + (eb >> 5) : eb;
        const uint32_t c = eb & ((1u << pp.logw) - 1), r = (eb >> pp.logw) & ((1u << pp.m) - 1);
        const uint32_t h = eb >> (pp.logw + pp.m);
        /* h << (lo + m) in 64 bits: a tile of the first pass of a 2^32-point transform has lo + m = 32 (and h = 0) */
        const uint32_t ob = (uint32_t)((uint64_t)h << (pp.lo + pp.m)) + (r << pp.lo) + c;
        const uint32_t kb = pp.m ? (__brev(r) >> (32 - pp.m)) : 0u;
        const int ksh = pp.m - s_lo - LOGR;
#pragma unroll
        for (int j = 0; j < R; j++) {
            const uint32_t dj = (uint32_t)j << L;
            f(g * R + j, pe + (PAD ? dj + (dj >> 5) : dj), ob + ((uint32_t)j << (s_lo + pp.lo)), c,
              kb + (brev_c<LOGR>(j) << ksh));
        }
    }
}

/* the radix step of one round on all PTS slots: PTS / R groups of R points */
/* TW: GTab (a global stage table) or LTab (a table in LDS, named explicitly so that it stays restrict) */
template <int LOGR, bool INV, int PTS, typename TW>
__device__ __forceinline__ void round_compute(double *v, const PassPos &pp, int s_lo, TW tw, double p, double pinv)
{
    constexpr int R = 1 << LOGR, G = PTS / R;
    const int L = pp.logw + s_lo;
#pragma unroll
    for (int g = 0; g < G; g++) {
        const uint32_t tau = (uint32_t)(g * pp.nthr + pp.t);
        const uint32_t rlow = (tau & ((1u << L) - 1)) >> pp.logw;      /* row bits below s_lo */
        if (INV) dit_group<LOGR, TW>(v + g * R, s_lo, rlow, tw, p, pinv);
        else dif_group<LOGR, TW>(v + g * R, s_lo, rlow, tw, p, pinv);
    }
}

/* dit_group with the inverse twiddles taken from the FORWARD stage table: w^-j of stage s is the mirrored forward entry
 * negated, w^-j = -w^(2^s - j) (w^(2^s) = -1), the negation folded into the butterfly (the product's sign flipped), and
 * w^0 = 1 taken from entry 1 (stage 0).  Lets a kernel keep one small table in LDS for both directions. */
template <int LOGR>
__device__ __forceinline__ void dit_group_mirror(double *v, int s_lo, uint32_t rlow, const double2 *tw, double p,
                                                 double pinv)
{
    constexpr int R = 1 << LOGR;
#pragma unroll
    for (int lev = 0; lev < LOGR; lev++) {
        const int k = lev, h = 1 << k, s = s_lo + k;
        const bool red = lev == LOGR - 1;
#pragma unroll
        for (int jj = 0; jj < h; jj++) {
            const bool one = lev == 0 && NTTM_TRIVIAL_W(rlow, jj);
            const uint32_t j = rlow + ((uint32_t)jj << s_lo);
            double2 w;                                   /* (w, wq) negated for a mirrored entry */
            if (jj == 0) {
                const bool z = rlow == 0;
                const double2 t = tw[z ? 1u : (2u << s) - j];
                w = z ? t : make_double2(-t.x, -t.y);
            } else {
                const double2 t = tw[(2u << s) - j];
                w = make_double2(-t.x, -t.y);
            }
#pragma unroll
            for (int i = jj; i < R; i += 2 * h) {
                const double u = v[i], x = one ? v[i + h] : mul_shoup(v[i + h], w.x, w.y, p);
                const double s1 = u + x, s2 = u - x;
                v[i] = red ? reduce(s1, p, pinv) : s1;
                v[i + h] = red ? reduce(s2, p, pinv) : s2;
            }
        }
    }
}
template <int LOGR, int PTS>
__device__ __forceinline__ void round_compute_mirror(double *v, const PassPos &pp, int s_lo, const double2 *tw, double p,
                                                     double pinv)
{
    constexpr int R = 1 << LOGR, G = PTS / R;
    const int L = pp.logw + s_lo;
#pragma unroll
    for (int g = 0; g < G; g++) {
        const uint32_t tau = (uint32_t)(g * pp.nthr + pp.t);
        const uint32_t rlow = (tau & ((1u << L) - 1)) >> pp.logw;
        dit_group_mirror<LOGR>(v + g * R, s_lo, rlow, tw, p, pinv);
    }
}

/* rounds of a pass: sizes of at most maxr bits (the log2 of the points per thread, at most 4), balanced; forward from
 * the top bits down, inverse from bit 0 up */
__device__ __forceinline__ int round_count(int m, int maxr) { return (m + maxr - 1) / maxr; }
__device__ __forceinline__ void round_split(int m, int i, int *s_lo, int *logR, bool inv, int maxr)
{
    const int nr = round_count(m, maxr), q = m / nr, rem = m % nr;
    int lo = 0;                                     /* rounds in ascending bit order: sizes q+1 first (rem of them) */
    int k = inv ? i : nr - 1 - i;
    for (int r = 0; r < k; r++) lo += q + (r < rem ? 1 : 0);
    *s_lo = lo; *logR = q + (k < rem ? 1 : 0);
}

/* call body.template operator()<LOGR>() with the runtime round size */
#define ROUND_SWITCH(logR, ...) \
    switch (logR) { \
    case 4: { constexpr int LR = 4; __VA_ARGS__ } break; \
    case 3: { constexpr int LR = 3; __VA_ARGS__ } break; \
    case 2: { constexpr int LR = 2; __VA_ARGS__ } break; \
    default: { constexpr int LR = 1; __VA_ARGS__ } break; }

/* the pass's input element: from the plane, or (the first forward pass) from the operand limbs, zero beyond them */
__device__ __forceinline__ double input_limb_raw(const PassArgs &a, uint64_t gi, double p, double pinv)
{
    const uint64_t b = gi >> a.logn, k = gi & (((uint64_t)1 << a.logn) - 1);
    /* the descriptor is copied as values: selecting between addresses (a field or the argument's pointer) would put
     * the kernel argument on the stack */
    const uint64_t *src; uint64_t nl;
    if (a.ldesc) { const LoadDesc d = a.ldesc[b]; src = a.which ? d.b : d.a; nl = a.which ? d.nb : d.na; }
    else { src = a.limbs; nl = a.nlimbs; }
    return k < nl ? limb_to_res(src[k], p, pinv) : 0.0;
}
__device__ __forceinline__ double input_limb(const PassArgs &a, uint64_t gi, double p, double pinv)
{
    double x = input_limb_raw(a, gi, p, pinv);
    if (a.do_in_scale) x = mul_shoup(x, a.in_scale.x, a.in_scale.y, p);
    return x;
}
/* input_limb with the operand of the last transform looked up kept: consecutive elements of one transform (the slots
 * of a group) read its descriptor once */
struct LimbCache {
    uint64_t b = ~(uint64_t)0, nl = 0;
    const uint64_t *src = nullptr;
    __device__ __forceinline__ double at(const PassArgs &a, uint64_t gi, double p, double pinv)
    {
        const uint64_t k = gi & (((uint64_t)1 << a.logn) - 1);
        if (a.ldesc) {
            const uint64_t bb = gi >> a.logn;
            if (bb != b) { const LoadDesc d = a.ldesc[bb]; src = a.which ? d.b : d.a; nl = a.which ? d.nb : d.na; b = bb; }
        } else { src = a.limbs; nl = a.nlimbs; }
        double x = k < nl ? limb_to_res(src[k], p, pinv) : 0.0;
        if (a.do_in_scale) x = mul_shoup(x, a.in_scale.x, a.in_scale.y, p);
        return x;
    }
};
__device__ __forceinline__ uint64_t y_index(const PassArgs &a, uint64_t gi)
{
    const uint64_t nmask = ((uint64_t)1 << a.logn) - 1;
    if (a.r3y) {                  /* transform t = 3 item + third: y holds whole items of 3 thirds */
        const uint64_t t = gi >> a.logn, item = t / 3, third = t - 3 * item;
        const uint64_t yt = a.ymode == NTTM_Y_FULL ? t : a.ymode == NTTM_Y_BCAST ? third : (item >> 1) * 3 + third;
        return (yt << a.logn) | (gi & nmask);
    }
    return a.ymode == NTTM_Y_FULL ? gi : a.ymode == NTTM_Y_BCAST ? (gi & nmask)
                                       : ((((gi >> a.logn) >> 1) << a.logn) | (gi & nmask));
}

/* The tile in LDS, padded: element e at word e + e / 32 (LDS_WORDS(T) words for T elements).  A round over the lowest
 * row bits of a contiguous tile has lanes 16 elements apart, which unpadded all hit the same bank pair; padded, every
 * round, the linear staging and the strided tiles' rounds spread over all banks.  For a group's slots
 * e = eb + j 2^L the padded index is pad(eb) + a compile-time multiple of j (L >= 5, or the slot bits below bit 5 with
 * eb's bits clear), so the offsets stay immediate. */
#define LDS_WORDS(T) ((T) + ((T) >> 5))
struct SwzLds {
    double *p;
    __device__ __forceinline__ double &operator[](uint32_t e) const { return p[e + (e >> 5)]; }
    __device__ __forceinline__ double &at(uint32_t pe) const { return p[pe]; }   /* an index padded already */
};

enum { PM_FWD = 0, PM_FWD_LIMB = 1, PM_INV = 2, PM_INV_PW = 3 };

/* GM, GW, GH >= 0 fix the tile shape (m, logw, logh) at compile time: round splits, LDS strides and shifts fold to
 * constants; -1 takes them from the arguments */
/* w^E for a 32-bit exponent fraction E from the four-level table (bq[k 2^8 + x] = w32^(x 2^(8k))), balanced
 * (|result| <= p/2 + 1) */
__device__ __forceinline__ double bq_twiddle(const double2 *bq, uint32_t E, double p, double pinv)
{
    constexpr uint32_t M = (1u << NTTM_BQ_LOG) - 1;
    double x = bq[(3 << NTTM_BQ_LOG) + (E >> 24)].x;
    const double2 t2 = bq[(2 << NTTM_BQ_LOG) + ((E >> 16) & M)];
    const double2 t1 = bq[(1 << NTTM_BQ_LOG) + ((E >> 8) & M)];
    const double2 t0 = bq[E & M];
    x = mul_shoup(x, t2.x, t2.y, p);
    x = mul_shoup(x, t1.x, t1.y, p);
    x = mul_shoup(x, t0.x, t0.y, p);
    return reduce(x, p, pinv);
}

/* the boundary twiddles w_N^(col k) (N = 2^lom) of the round over row bits [0, LOGR) (s_lo = 0), multiplied into v:
 * group g of a thread has column c and bit-reversed rows k_j = kb + brev(j) 2^(m - LOGR), so its twiddles are
 * base * step^brev(j) with base = w^(c kb), step = w^(c 2^(m - LOGR)) -- two table twiddles and a chain of Shoup
 * products per group instead of a table lookup per point */
template <int LOGR, int PTS>
__device__ __forceinline__ void bound_chain(double *v, const PassPos &pp, int lom, uint32_t colbase, const double2 *bq,
                                            double p, double pinv, bool do_scale = false,
                                            double2 scale = make_double2(1.0, 0.0))
{
    constexpr int R = 1 << LOGR, G = PTS / R;
    const int L = pp.logw, sh = 32 - lom, ksh = pp.m - LOGR;
#pragma unroll
    for (int g = 0; g < G; g++) {
        const uint32_t tau = (uint32_t)(g * pp.nthr + pp.t);
        const uint32_t eb = (tau & ((1u << L) - 1)) | ((tau >> L) << (L + LOGR));
        const uint32_t c = eb & ((1u << pp.logw) - 1), r = (eb >> pp.logw) & ((1u << pp.m) - 1);
        const uint32_t kb = pp.m ? (__brev(r) >> (32 - pp.m)) : 0u, col = colbase + c;
        double base = bq_twiddle(bq, (col * kb) << sh, p, pinv);
        if (do_scale) base = mul_shoup(base, scale.x, scale.y, p);      /* a factor common to all points, folded in */
        const double step = bq_twiddle(bq, (col << ksh) << sh, p, pinv), stepq = step / p;
        double w[R];
        w[0] = base;
#pragma unroll
        for (int i = 1; i < R; i++) w[i] = mul_shoup(w[i - 1], step, stepq, p);
#pragma unroll
        for (int j = 0; j < R; j++) v[g * R + j] = mul_barrett(v[g * R + j], w[brev_c<LOGR>(j)], p, pinv);
    }
}

template <int MODE, int PTS, int GM = -1, int GW = -1, int GH = -1>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_pass(PassArgs a)
{
    const int gm = GM >= 0 ? GM : a.m, gw = GW >= 0 ? GW : a.logw, gh = GH >= 0 ? GH : a.logh;
    constexpr bool INV = MODE >= PM_INV;
    extern __shared__ double lds_raw[];                         /* LDS_WORDS(T) doubles, sized at launch */
    const SwzLds lds{lds_raw};
    PassPos pp; pp.m = gm; pp.logw = gw; pp.lo = a.lo;
    const int T = 1 << (gm + gw + gh);
    pp.nthr = T / PTS; pp.t = threadIdx.x;
    const bool active = pp.t < pp.nthr;
    const double p = a.pk.p, pinv = a.pk.pinv;
    const GTab gtw(a.tw.stage);
    constexpr int MAXR = PTS >= 16 ? 4 : PTS == 8 ? 3 : PTS == 4 ? 2 : 1;
    const int lom = a.lo + gm, nr = round_count(gm, MAXR);
    /* the tile: column block cb (W columns below bit lo) of block group hb (H blocks of 2^lom points) */
    const uint64_t tile = blockIdx.x;
    const uint64_t cb = tile & ((((uint64_t)1 << a.lo) >> gw) - 1), hb = tile >> (a.lo - gw);
    const uint64_t gbase = (hb << (lom + gh)) + (cb << gw);
    const uint32_t colbase = (uint32_t)(cb << gw);
    double *xt = a.x + gbase;
    /* a limb-input tile inside one transform (lo + m + logh <= logn): its operand, read once per tile */
    const bool limb_one = MODE == PM_FWD_LIMB && a.lo + gm + gh <= a.logn;
    const uint64_t *lsrc = nullptr;
    uint64_t lnl = 0, lk0 = 0;
    if (limb_one) {
        const uint64_t tr = gbase >> a.logn;
        lk0 = gbase & (((uint64_t)1 << a.logn) - 1);
        if (a.ldesc) { const LoadDesc d = a.ldesc[tr]; lsrc = a.which ? d.b : d.a; lnl = a.which ? d.nb : d.na; }
        else { lsrc = a.limbs; lnl = a.nlimbs; }
    }
    /* a strided limb pass folds the input scale into its boundary twiddles */
    const bool scale_in_chain = MODE == PM_FWD_LIMB && a.lo > 0 && a.do_in_scale;
    auto limb_at = [&](uint32_t o) -> double {
        double x;
        if (limb_one) { const uint64_t k = lk0 + o; x = k < lnl ? limb_to_res(lsrc[k], p, pinv) : 0.0; }
        else x = input_limb_raw(a, gbase + o, p, pinv);
        if (a.do_in_scale && !scale_in_chain) x = mul_shoup(x, a.in_scale.x, a.in_scale.y, p);
        return x;
    };

    /* the boundary twiddles of a strided pass (lo > 0) go with the round over row bits [0, logR) -- the last forward
     * round, the first inverse one -- as base * step^i chains (bound_chain), applied to v between load and round */
    auto load_at = [&](uint32_t o, uint32_t, uint32_t) -> double {
        double x = MODE == PM_FWD_LIMB ? limb_at(o) : xt[o];
        if (MODE == PM_INV_PW) x = mul_barrett(x, a.y[y_index(a, gbase + o)], p, pinv);
        return x;
    };
    auto store_at = [&](uint32_t o, uint32_t, uint32_t, double x) {
        if (a.do_scale) x = mul_shoup(x, a.scale.x, a.scale.y, p);
        xt[o] = x;
    };

    double v[PTS];
    int s_lo, logR;
    /* first round: straight from global when its lanes cover runs of >= 16 consecutive elements; else (a contiguous
     * pass, lo = 0, where the offset of element e is e) a coalesced linear load through LDS */
    round_split(gm, 0, &s_lo, &logR, INV, MAXR);
    if (gw + s_lo >= 4) {
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS>(pp, s_lo, [&](int sl, uint32_t, uint32_t o, uint32_t c,
                                                                       uint32_t k) { v[sl] = load_at(o, c, k); });)
    } else {
        if (active) {
#pragma unroll
            for (int sl = 0; sl < PTS; sl++) { const uint32_t e = sl * pp.nthr + pp.t; lds[e] = load_at(e, 0, 0); }
        }
        __syncthreads();
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                                       uint32_t) { v[sl] = lds.at(e); });)
    }
    if (INV && a.lo > 0 && active) ROUND_SWITCH(logR, bound_chain<LR, PTS>(v, pp, lom, colbase, a.tw.bq, p, pinv);)
#pragma unroll
    for (int i = 0; i < (GM >= 0 ? (GM + MAXR - 1) / MAXR : nr); i++) {
        round_split(gm, i, &s_lo, &logR, INV, MAXR);
        if (i > 0) {
            __syncthreads();
            if (active) ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                                           uint32_t) { v[sl] = lds.at(e); });)
        }
        if (active) ROUND_SWITCH(logR, round_compute<LR, INV, PTS>(v, pp, s_lo, gtw, p, pinv);)
        if (i + 1 < nr && active)
            ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                               uint32_t) { lds.at(e) = v[sl]; });)
    }
    if (!INV && a.lo > 0 && active)
        ROUND_SWITCH(logR, bound_chain<LR, PTS>(v, pp, lom, colbase, a.tw.bq, p, pinv, scale_in_chain, a.in_scale);)
    /* last round: the same choice for the store (each thread rewrites only the LDS words it read in this round) */
    if (gw + s_lo >= 4) {
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS>(pp, s_lo, [&](int sl, uint32_t, uint32_t o, uint32_t c,
                                                                       uint32_t k) { store_at(o, c, k, v[sl]); });)
    } else {
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                                       uint32_t) { lds.at(e) = v[sl]; });)
        __syncthreads();
        if (active) {
#pragma unroll
            for (int sl = 0; sl < PTS; sl++) { const uint32_t e = sl * pp.nthr + pp.t; store_at(e, 0, 0, lds[e]); }
        }
    }
}

/* ---- the streaming strided pass ---------------------------------------------------------------------------------------
 * Strided passes (lo >= GW) of the fixed tile shapes 2^GM rows x 2^GW columns, GM + GW = 13 with GM = 5 .. 9 (H = 1,
 * 2^13 points, 16 per thread; 9-bit rows are 128-byte runs in three rounds of 3 bits, the others two rounds): one
 * block per CU walks the tiles blockIdx.x, + gridDim.x, ...  The next tile is fetched straight into the other of two
 * LDS buffers (global_load_lds, 16 bytes per lane, the tile in natural row-major order) while the current one is
 * transformed, so HBM streams during the butterflies.  Vector-memory operations of a wave complete in issue order, so
 * the loop issues no global load it must wait on besides the fetch:
 *   - the stage twiddles (entries below 2^GM) and the four-level boundary table (4 x 2^8 entries) are copied into LDS
 *     once per block;
 *   - the boundary twiddles w^(c k_j) of a thread's 16 points share the column c and have k_j = kb + i_j 2^(GM-4)
 *     (i_j = brev4(j)), so they are base * step^i with base = w^(c kb), step = w^(c 2^(GM-4)), each the product of four
 *     LDS table entries, and the powers a chain of Shoup products by step (exact integers, no error growth);
 *   - at the top of an iteration at most the previous tile's 16 stores may remain outstanding (vmcnt 16), which
 *     leaves exactly the fetch of the tile about to be transformed complete.
 * The boundary twiddle is applied as a Barrett product (both factors variable, |x|, |w| <= p), the forward on the
 * store (after the last round, bits [0, 4)), the inverse on the load (before the first round, the same bits).  The
 * other shapes have one or two groups per thread in that round and take their chains from bound_chain (the same
 * base * step^i construction per group, from the LDS copy of the boundary table).
 * A limb pass reads one operand (a batch of one) or, per tile, the operand of the tile's transform from the batch's
 * descriptors; descriptors and the operand's last limb are read through the scalar unit, whose loads the hand-counted
 * vector-memory waits do not see. */
#define STREAM_T 512
/* a workgroup barrier for LDS only: the wave's LDS operations complete, then s_barrier.  The compiler places a wait
 * for every outstanding vector-memory operation (an LDS fetch included) before a barrier it sees; written as one asm
 * statement (with a memory clobber, so no memory access moves across it) the fetch in flight stays in flight. */
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}


template <int MODE, int GM, int GW>
__global__ __launch_bounds__(STREAM_T) void k_pass_stream(PassArgs a, uint32_t ntiles)
{
    constexpr bool INV = MODE == PM_INV, LIMB = MODE == PM_FWD_LIMB;
    constexpr int PTS = 16, T = 1 << (GM + GW), W = 1 << GW, NBQ = 4 << NTTM_BQ_LOG;
    static_assert(T == PTS * STREAM_T, "2^13-point tiles, 16 points per thread");
    static_assert(GM >= 1 && GM <= 9, "rows of 1 to 9 bits");
    /* rounds of at most 4 bits (round_split's balanced split): NR of them; LR0 bits in the round over bits [0, LR0)
     * (the boundary twiddles' round), LRT in the top round (the first forward round) */
    constexpr int NR = (GM + 3) / 4, LR0 = GM / NR + (GM % NR > 0 ? 1 : 0), LRT = GM / NR + (GM % NR > NR - 1 ? 1 : 0);
    __shared__ double buf0[T], buf1[T];
    __shared__ double2 stw[1 << GM], bqs[NBQ];
    const int t = threadIdx.x;
    const uint32_t wv = __builtin_amdgcn_readfirstlane((uint32_t)t >> 6);
    const double p = a.pk.p, pinv = a.pk.pinv;
    const int lo = a.lo, lom = a.lo + GM;
    const uint32_t ncb_mask = (uint32_t)((((uint64_t)1 << lo) >> GW) - 1);
    PassPos pp; pp.m = GM; pp.logw = GW; pp.lo = lo; pp.nthr = STREAM_T; pp.t = t;
    /* descriptors and limbs read through the scalar unit (constant address space, uniform per tile): they leave the
     * vector-memory counter to the tile fetches and stores alone */
    typedef const __attribute__((address_space(4))) LoadDesc *CDesc;
    typedef const __attribute__((address_space(4))) uint64_t *CLimb;

    /* the thread's group in the round over bits [0, 4) of the 8-bit shape: column c, bit-reversed row kb of slot 0 */
    const uint32_t c_in = (uint32_t)t & (W - 1);
    const uint32_t r0 = (((uint32_t)t >> GW) << 4) & ((1u << GM) - 1);
    const uint32_t kb = __brev(r0) >> (32 - GM);

    auto tile_base = [&](uint32_t tile, uint32_t *colbase) -> uint64_t {
        const uint32_t cb = tile & ncb_mask;
        const uint64_t hb = (uint64_t)(tile >> (lo - GW));
        *colbase = cb << GW;
        return (hb << lom) + ((uint64_t)cb << GW);
    };
    /* a limb-input tile: its operand (one a.limbs of a batch of one, or the descriptor of the tile's transform -- a
     * strided tile lies inside one transform), the operand's limb at the tile's first element, and how many of the
     * tile's elements (counted from the first, saturated; offsets within a tile are below 2^lom <= 2^31) lie inside it */
    struct LimbTile { const uint64_t *base; uint64_t nl, gloc; uint32_t lim; };
    auto limb_tile = [&](uint64_t gb) -> LimbTile {
        LimbTile r;
        if (a.ldesc) {
            const CDesc d = (CDesc)a.ldesc + (gb >> a.logn);
            r.base = a.which ? d->b : d->a; r.nl = a.which ? d->nb : d->na;
            r.gloc = gb & (((uint64_t)1 << a.logn) - 1);
        } else { r.base = a.limbs; r.nl = a.nlimbs; r.gloc = gb; }
        r.lim = r.nl <= r.gloc ? 0u : r.nl - r.gloc >= 0xffffffffULL ? 0xffffffffu : (uint32_t)(r.nl - r.gloc);
        return r;
    };
    auto fetch = [&](uint32_t tile, double *dst) {
        uint32_t cbase;
        const uint64_t gb = tile_base(tile, &cbase);
        LimbTile lt;
        if (LIMB) lt = limb_tile(gb);
        const uint32_t lim = LIMB ? lt.lim : 0u;
        const double *src = LIMB ? (const double *)(lt.base + lt.gloc) : a.x + gb;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t q = (uint32_t)(i * STREAM_T + t);           /* 16-byte chunk: elements 2q, 2q + 1 */
            const uint32_t off = ((q >> (GW - 1)) << lo) + ((q & (W / 2 - 1)) << 1);
            /* a limb chunk is fetched when both limbs lie inside the operand; the one that straddles its end (odd
             * length) is read from the operand's last limb */
            if (!LIMB || off + 1 < lim)
                __builtin_amdgcn_global_load_lds((const void *)(src + off),
                                                 (__attribute__((address_space(3))) void *)(dst + (i * STREAM_T + wv * 64) * 2),
                                                 16, 0, 0);
        }
    };

    if (t < (1 << GM)) stw[t] = a.tw.stage[t];
    for (int i = t; i < NBQ; i += STREAM_T) bqs[i] = a.tw.bq[i];
    uint32_t tile = blockIdx.x;
    if (tile < ntiles) fetch(tile, buf0);
    __builtin_amdgcn_s_waitcnt(0);                                     /* vmcnt(0) expcnt(0) lgkmcnt(0) */
    for (uint32_t it = 0; tile < ntiles; it++, tile += gridDim.x) {
        double *cur = (it & 1) ? buf1 : buf0, *nxt = (it & 1) ? buf0 : buf1;
        /* this tile's fetch has landed (only the previous tile's 16 stores may be outstanding); every thread is past
         * the previous tile's LDS reads of nxt */
        __builtin_amdgcn_s_waitcnt(0x4F70);                             /* vmcnt(16) */
        __syncthreads();
        if (tile + gridDim.x < ntiles) fetch(tile + gridDim.x, nxt);
        uint32_t colbase;
        const uint64_t gbase = tile_base(tile, &colbase);
        double *xt = a.x + gbase;

        double v[PTS], w[PTS];
        const bool bound_out = !INV && !(LIMB && a.dfr_out);           /* the output's boundary twiddles here */
        if (GM == 8 && (INV || bound_out)) {   /* boundary twiddles of the bits-[0, 4) round: w[i] = base * step^i (a limb pass's n^-1 folded into base) */
            const uint32_t col = colbase + c_in, sh = 32 - lom;
            double base = bq_twiddle(bqs, (col * kb) << sh, p, pinv);
            if (LIMB && a.do_in_scale) base = mul_shoup(base, a.in_scale.x, a.in_scale.y, p);
            const double step = bq_twiddle(bqs, (col << (GM - 4)) << sh, p, pinv);
            const double stepq = step / p;
            w[0] = base;
#pragma unroll
            for (int i = 1; i < PTS; i++) w[i] = mul_shoup(w[i - 1], step, stepq, p);
        }
        int s_lo, logR;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            round_split(GM, i, &s_lo, &logR, INV, 4);
            if (i > 0) lds_barrier();
            if (LIMB && i == 0) {                   /* limbs to residues, zero beyond the operand */
                const LimbTile lt = limb_tile(gbase);
                const uint32_t lim = lt.lim;
                const uint64_t last = lt.nl ? ((CLimb)lt.base)[lt.nl - 1] : 0;
                const uint64_t *cl = (const uint64_t *)cur;
                /* slot j of a group of the first round holds the top LRT row bits = j, so rows >= 2^(GM-1) are the
                 * group's slots j >= 2^(LRT-1); tiles whose rows split into an all-inside and an all-outside half (the
                 * zero-padded upper half of a product's operand) convert only the inside slots */
                const uint32_t half = 1u << (GM - 1 + lo);
                if (lim <= half && lim > half - (1u << lo) + W) {
                    for_slots<LRT, PTS>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t, uint32_t) {
                        v[sl] = (sl & ((1 << LRT) - 1)) < (1 << (LRT - 1)) ? limb_to_res(cl[e], p, pinv) : 0.0; });
                } else {
                    for_slots<LRT, PTS>(pp, s_lo, [&](int sl, uint32_t e, uint32_t o, uint32_t, uint32_t) {
                        v[sl] = o < lim ? limb_to_res(o + 1 == lim ? last : cl[e], p, pinv) : 0.0; });
                }
            } else
                ROUND_SWITCH(logR, for_slots<LR, PTS>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t, uint32_t) {
                    v[sl] = cur[e]; });)
            if (INV && i == 0) {
                if (GM == 8) {
#pragma unroll
                    for (int j = 0; j < PTS; j++) v[j] = mul_barrett(v[j], w[brev_c<4>(j)], p, pinv);
                } else bound_chain<LR0, PTS>(v, pp, lom, colbase, bqs, p, pinv);
            }
            if (MODE == PM_FWD && GM >= 4 && i == 0 && a.dfr_in) {
                /* the first pass's boundary twiddles: slot (g, j) holds row r = (g << (9 - GW)) + (t >> GW) +
                 * (j << s_lo) of this tile (rows of at least 4 bits, so that 2^GW divides the 512 threads), i.e.
                 * column c1 = (r << lo) + col of the first pass, whose stored row h is
                 * this tile's block index (mod 2^dfr_gm): w^(K c1) = base_g * step^j with K = brev(h),
                 * base_g = w^(K ((rb_g << lo) + col)), step = w^(K 2^(s_lo + lo)) (one per tile) */
                const uint32_t sh1 = 32 - a.dfr_lom;
                const uint32_t hb = (uint32_t)(gbase >> lom) & ((1u << a.dfr_gm) - 1);
                const uint32_t K = __brev(hb) >> (32 - a.dfr_gm);
                const uint32_t col = colbase + c_in;
                const double step = bq_twiddle(bqs, (K << (s_lo + lo)) << sh1, p, pinv);
                const double stepq = step * pinv;
                constexpr int RT = 1 << LRT, GT = PTS / RT;
#pragma unroll
                for (int g = 0; g < GT; g++) {
                    const uint32_t rb = ((uint32_t)g << (GW < 9 ? 9 - GW : 0)) + ((uint32_t)t >> GW);
                    double wj = bq_twiddle(bqs, (K * ((rb << lo) + col)) << sh1, p, pinv);
                    if (a.dfr_do_scale) wj = mul_shoup(wj, a.dfr_scale.x, a.dfr_scale.y, p);
#pragma unroll
                    for (int j = 0; j < RT; j++) {
                        if (j > 0) wj = mul_shoup(wj, step, stepq, p);
                        v[g * RT + j] = mul_barrett(v[g * RT + j], wj, p, pinv);
                    }
                }
            }
            ROUND_SWITCH(logR, round_compute<LR, INV, PTS, LTab>(v, pp, s_lo, stw, p, pinv);)
            if (i + 1 < NR)
                ROUND_SWITCH(logR, for_slots<LR, PTS>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t, uint32_t) {
                    cur[e] = v[sl]; });)
        }
        if (bound_out) {
            if (GM == 8) {
#pragma unroll
                for (int j = 0; j < PTS; j++) v[j] = mul_barrett(v[j], w[brev_c<4>(j)], p, pinv);
            } else bound_chain<LR0, PTS>(v, pp, lom, colbase, bqs, p, pinv, LIMB && a.do_in_scale, a.in_scale);
        }
        if (INV && a.do_scale) {                                       /* the inverse's final n^-1 */
#pragma unroll
            for (int j = 0; j < PTS; j++) v[j] = mul_shoup(v[j], a.scale.x, a.scale.y, p);
        }
        ROUND_SWITCH(logR, for_slots<LR, PTS>(pp, s_lo, [&](int sl, uint32_t, uint32_t o, uint32_t, uint32_t) {
            xt[o] = v[sl]; });)
    }
    __builtin_amdgcn_s_waitcnt(0);                                     /* no LDS fetch in flight past the block's end */
}

/* ---- the fused last forward pass + pointwise product + first inverse pass ------------------------------------------------
 * For the contiguous low-bit pass (lo = 0, m + logh <= 12): the DIF rounds of the forward transform, the product with
 * y in a linear pass over LDS (y read coalesced), the DIT rounds of the inverse, one read and one write of x.
 * LIMB: x comes from the operand limbs (a transform that is a single pass). */
#define FUSED_LOG 12
template <bool LIMB, int PTS>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_fused(PassArgs a)
{
    __shared__ double xs_raw[LDS_WORDS(1 << FUSED_LOG)];
    const SwzLds xs{xs_raw};
    constexpr int MAXR = PTS >= 16 ? 4 : 3;
    PassPos pp; pp.m = a.m; pp.logw = 0; pp.lo = 0;
    const int T = 1 << (a.m + a.logh);
    pp.nthr = T / PTS; pp.t = threadIdx.x;
    const bool active = pp.t < pp.nthr;
    const double p = a.pk.p, pinv = a.pk.pinv;
    const uint64_t gbase = (uint64_t)blockIdx.x << (a.m + a.logh);  /* lo = 0, W = 1: the tile is contiguous */
    double *xt = a.x + gbase;
    const double *yt = a.y + gbase;                                   /* FULL layout (the product path) */
    const int nr = round_count(a.m, MAXR);
    double v[PTS];
    int s_lo, logR;
    auto in_at = [&](uint32_t o) -> double { return LIMB ? input_limb(a, gbase + o, p, pinv) : xt[o]; };

    round_split(a.m, 0, &s_lo, &logR, false, MAXR);
    if (s_lo >= 4) {
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS>(pp, s_lo, [&](int sl, uint32_t, uint32_t o, uint32_t,
                                                                            uint32_t) { v[sl] = in_at(o); });)
    } else {
        if (active) {
#pragma unroll
            for (int sl = 0; sl < PTS; sl++) { const uint32_t e = sl * pp.nthr + pp.t; xs[e] = in_at(e); }
        }
        __syncthreads();
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                                            uint32_t) { v[sl] = xs.at(e); });)
    }
    for (int i = 0; i < nr; i++) {                               /* forward, top bits first */
        round_split(a.m, i, &s_lo, &logR, false, MAXR);
        if (i > 0) {
            __syncthreads();
            if (active) ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t,
                                                                                uint32_t, uint32_t) { v[sl] = xs.at(e); });)
        }
        if (active) ROUND_SWITCH(logR, round_compute<LR, false, PTS>(v, pp, s_lo, GTab(a.tw.stage), p, pinv);)
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                                            uint32_t) { xs.at(e) = v[sl]; });)
    }
    __syncthreads();
    if (active) {                                                /* the pointwise product, linear: y read coalesced */
#pragma unroll
        for (int sl = 0; sl < PTS; sl++) {
            const uint32_t e = sl * pp.nthr + pp.t;
            xs[e] = mul_barrett(xs[e], yt[e], p, pinv);
        }
    }
    for (int i = 0; i < nr; i++) {                               /* inverse, bit 0 first */
        round_split(a.m, i, &s_lo, &logR, true, MAXR);
        __syncthreads();
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                                            uint32_t) { v[sl] = xs.at(e); });)
        if (active) ROUND_SWITCH(logR, round_compute<LR, true, PTS>(v, pp, s_lo, GTab(a.twi.stage), p, pinv);)
        if (i + 1 < nr && active)
            ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                                    uint32_t) { xs.at(e) = v[sl]; });)
    }
    if (s_lo >= 4) {
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS>(pp, s_lo, [&](int sl, uint32_t, uint32_t o, uint32_t,
                                                                            uint32_t) {
            double x = v[sl];
            if (a.do_scale) x = mul_shoup(x, a.scale.x, a.scale.y, p);
            xt[o] = x; });)
    } else {
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                                            uint32_t) { xs.at(e) = v[sl]; });)
        __syncthreads();
        if (active) {
#pragma unroll
            for (int sl = 0; sl < PTS; sl++) {
                const uint32_t e = sl * pp.nthr + pp.t;
                double x = xs[e];
                if (a.do_scale) x = mul_shoup(x, a.scale.x, a.scale.y, p);
                xt[e] = x;
            }
        }
    }
}

/* ---- the double-fused last pass: B's last forward pass + A's last forward pass + pointwise + A's first inverse pass ---
 * Same geometry as k_fused (lo = 0, m + logh <= 12).  The forward's last round and the inverse's first round cover the
 * same bits [0, logR) with the same slot layout, so B's transformed tile stays in registers (bv) and is never written;
 * the product v * bv and the first inverse round follow A's last forward round without an LDS round trip.
 * One read of A (or its limbs), one read of B (or its limbs), one write of A. */
/* stw: the forward stage table's first 2^STW_LOG entries in LDS (nullptr: none); rounds whose stages all lie below
 * STW_LOG read their twiddles there */
#define STW_LOG 8
/* LT: stw is the LDS copy of the stage table's low entries, for the rounds after the first */
template <int PTS, int MAXR, bool LIMBS, bool LT>
__device__ __forceinline__ void tile_forward(const PassArgs &q, const PassPos &pp, bool active, uint64_t gbase,
                                             const SwzLds &xs, double *v, double p, double pinv, const double2 *stw)
{
    const bool limb = LIMBS && (q.limbs || q.ldesc);
    /* a shared B: the tile (inside one transform) is read from the transform the map names */
    const double *xt = q.bmap ? q.x + (((uint64_t)q.bmap[gbase >> q.logn] << q.logn) | (gbase & (((uint64_t)1 << q.logn) - 1)))
                              : q.x + gbase;
    const int nr = round_count(q.m, MAXR);
    int s_lo, logR;
    round_split(q.m, 0, &s_lo, &logR, false, MAXR);
    LimbCache lc;
    if (s_lo >= 4) {
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS>(pp, s_lo, [&](int sl, uint32_t, uint32_t o, uint32_t,
                                                                       uint32_t) {
            v[sl] = limb ? lc.at(q, gbase + o, p, pinv) : xt[o]; });)
    } else {
        if (active) {
#pragma unroll
            for (int sl = 0; sl < PTS; sl++) {
                const uint32_t e = sl * pp.nthr + pp.t;
                xs[e] = limb ? lc.at(q, gbase + e, p, pinv) : xt[e];
            }
        }
        __syncthreads();
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                                       uint32_t) { v[sl] = xs.at(e); });)
    }
#pragma unroll
    for (int i = 0; i < nr; i++) {                               /* top bits first; the last round stays in v */
        round_split(q.m, i, &s_lo, &logR, false, MAXR);
        if (i > 0) {
            __syncthreads();
            if (active) ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                                           uint32_t) { v[sl] = xs.at(e); });)
        }
        /* with stw (the 4096-point tile in rounds of 4 bits), rounds 1 and 2 cover stages 7..0 */
        if (LT && i > 0) {
            if (active) ROUND_SWITCH(logR, round_compute<LR, false, PTS, LTab>(v, pp, s_lo, stw, p, pinv);)
        } else {
            if (active) ROUND_SWITCH(logR, round_compute<LR, false, PTS>(v, pp, s_lo, GTab(q.tw.stage), p, pinv);)
        }
        if (i + 1 < nr && active)
            ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                               uint32_t) { xs.at(e) = v[sl]; });)
    }
}

template <int PTS, int GM = -1, int GH = -1, bool LIMBS = true>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_fused2(PassArgs a, PassArgs bq)
{
    __shared__ double xs_raw[LDS_WORDS(1 << FUSED_LOG)];
    /* the fixed shapes (4096-point tiles of 2^GM-point blocks, GM = 10, 11 or 12) keep the stage table's low entries in
     * LDS (the inverse mirrors them): the forward rounds after the first and the first two inverse rounds cover stages
     * below 2^STW_LOG for each of them (rounds of 4, 3, 3 / 4, 4, 3 / 4, 4, 4 bits from bit 0) */
    constexpr bool LT = GM >= 0 && PTS >= 16;
    static_assert(!LT || (GM >= 10 && GM <= 12 && GM + GH == FUSED_LOG),
                  "the LDS stage table serves the 4096-point tiles of 2^10, 2^11 and 2^12-point blocks");
    __shared__ double2 stw_lds[LT ? 1 << STW_LOG : 1];
    const SwzLds xs{xs_raw};
    constexpr int MAXR = PTS >= 16 ? 4 : 3;
    if (LT) {
        for (int i = threadIdx.x; i < (1 << STW_LOG); i += blockDim.x) stw_lds[i] = a.tw.stage[i];
        /* visible to the rounds after the first round's barrier; the first round's stages lie above STW_LOG */
    }
    if (GM >= 0) { a.m = GM; bq.m = GM; }
    if (GH >= 0) { a.logh = GH; bq.logh = GH; }
    PassPos pp; pp.m = a.m; pp.logw = 0; pp.lo = 0;
    const int T = 1 << (a.m + a.logh);
    pp.nthr = T / PTS; pp.t = threadIdx.x;
    const bool active = pp.t < pp.nthr;
    const double p = a.pk.p, pinv = a.pk.pinv;
    const uint64_t gbase = (uint64_t)blockIdx.x << (a.m + a.logh);
    const int nr = round_count(a.m, MAXR);
    double *xt = a.x + gbase;
    double v[PTS], bv[PTS];
    int s_lo, logR;

    const double2 *stw = LT ? (const double2 *)stw_lds : nullptr;
    tile_forward<PTS, MAXR, LIMBS, LT>(bq, pp, active, gbase, xs, bv, p, pinv, stw);
    __syncthreads();                                             /* B's last LDS reads before A's first writes */
    tile_forward<PTS, MAXR, LIMBS, LT>(a, pp, active, gbase, xs, v, p, pinv, stw);
    if (active) {
#pragma unroll
        for (int sl = 0; sl < PTS; sl++) v[sl] = mul_barrett(v[sl], bv[sl], p, pinv);
    }
#pragma unroll
    for (int i = 0; i < nr; i++) {                               /* inverse, bit 0 first; round 0 is the layout in v */
        round_split(a.m, i, &s_lo, &logR, true, MAXR);
        if (i > 0) {
            __syncthreads();
            if (active) ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                                           uint32_t) { v[sl] = xs.at(e); });)
        }
        if (LT && i < 2) {                                       /* inverse rounds 0 and 1: stages 0..7 */
            if (active) ROUND_SWITCH(logR, round_compute_mirror<LR, PTS>(v, pp, s_lo, stw, p, pinv);)
        } else {
            if (active) ROUND_SWITCH(logR, round_compute<LR, true, PTS>(v, pp, s_lo, GTab(a.twi.stage), p, pinv);)
        }
        if (i + 1 < nr && active)
            ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                               uint32_t) { xs.at(e) = v[sl]; });)
    }
    if (s_lo >= 4) {
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS>(pp, s_lo, [&](int sl, uint32_t, uint32_t o, uint32_t,
                                                                       uint32_t) { xt[o] = v[sl]; });)
    } else {
        /* each thread rewrites only the LDS words it read in this round */
        if (active) ROUND_SWITCH(logR, for_slots<LR, PTS, true>(pp, s_lo, [&](int sl, uint32_t e, uint32_t, uint32_t,
                                                                       uint32_t) { xs.at(e) = v[sl]; });)
        __syncthreads();
        if (active) {
#pragma unroll
            for (int sl = 0; sl < PTS; sl++) { const uint32_t e = sl * pp.nthr + pp.t; xt[e] = xs[e]; }
        }
    }
}

/* limbs -> residues of np primes, zero-extended to npoints; optional per-prime scale (n^-1 folded into one operand).
 * Batched form: product j of a length group takes points [j n, (j+1) n) of every plane, operand a or b of desc[j]. */
struct LoadArgs {
    double *dst;                 /* plane i at dst + i * plane */
    uint64_t plane;
    const uint64_t *src;         /* single operand, or null with desc */
    uint64_t nlimbs, npoints;
    const LoadDesc *desc; int which, logn;
    int np, do_scale, out_u64;
    PrimeK pk[NTTM_MAX_PRIMES];
    double2 scale[NTTM_MAX_PRIMES];
};
__global__ __launch_bounds__(NTTM_THREADS) void k_load(LoadArgs a)
{
    const uint64_t stride = (uint64_t)gridDim.x * NTTM_THREADS;
    const uint64_t nmask = ((uint64_t)1 << a.logn) - 1;
    for (uint64_t i = (uint64_t)blockIdx.x * NTTM_THREADS + threadIdx.x; i < a.npoints; i += stride) {
        uint64_t x;
        if (a.desc) {
            const LoadDesc &d = a.desc[i >> a.logn];
            const uint64_t k = i & nmask;
            const uint64_t nl = a.which ? d.nb : d.na;
            x = k < nl ? (a.which ? d.b : d.a)[k] : 0;
        } else {
            x = i < a.nlimbs ? a.src[i] : 0;
        }
        for (int k = 0; k < a.np; k++) {
            double v = limb_to_res(x, a.pk[k].p, a.pk[k].pinv);
            if (a.do_scale) v = mul_shoup(v, a.scale[k].x, a.scale[k].y, a.pk[k].p);
            if (a.out_u64) reinterpret_cast<uint64_t *>(a.dst)[k * a.plane + i] = res_canon(v, a.pk[k].p, a.pk[k].pinv);
            else a.dst[k * a.plane + i] = v;
        }
    }
}

__global__ __launch_bounds__(NTTM_THREADS) void k_cvt_u64_f64(double *dst, const uint64_t *src, uint64_t count)
{
    const uint64_t stride = (uint64_t)gridDim.x * NTTM_THREADS;
    for (uint64_t i = (uint64_t)blockIdx.x * NTTM_THREADS + threadIdx.x; i < count; i += stride) {
        const uint64_t u = src[i];
        dst[i] = (double)u;
    }
}
__global__ __launch_bounds__(NTTM_THREADS) void k_f64_to_canon(double *x, uint64_t count, PrimeK pk)
{
    const uint64_t stride = (uint64_t)gridDim.x * NTTM_THREADS;
    for (uint64_t i = (uint64_t)blockIdx.x * NTTM_THREADS + threadIdx.x; i < count; i += stride) {
        const uint64_t u = res_canon(x[i], pk.p, pk.pinv);
        reinterpret_cast<uint64_t *>(x)[i] = u;
    }
}
/* canonical x, y in [0, p) -> canonical x y */
__global__ __launch_bounds__(NTTM_THREADS) void k_pw_u64(uint64_t *x, const uint64_t *y, uint64_t count, PrimeK pk)
{
    const uint64_t stride = (uint64_t)gridDim.x * NTTM_THREADS;
    for (uint64_t i = (uint64_t)blockIdx.x * NTTM_THREADS + threadIdx.x; i < count; i += stride)
        x[i] = res_canon(mul_barrett((double)x[i], (double)y[i], pk.p, pk.pinv), pk.p, pk.pinv);
}

/* ---- transforms of 3 2^k points: the radix-3 layer -----------------------------------------------------------------
 * A transform of n = 3m points (m = 2^logk) with the root w of order n (E/ntt.c:671-779, our own kernels):
 * forward, one DIF radix-3 layer over the thirds x[j], x[j+m], x[j+2m]:
 *     y_r[j] = (x[j] + w3^r x[j+m] + w3^(2r) x[j+2m]) w^(r j),     w3 = w^m (a cube root of unity),
 * stored in third r, then the m-point transforms of the three thirds (the 2^k engine, batch 3, root w^3): third r,
 * position brev_k(t), holds X[3t + r].  With 1 + w3 + w3^2 = 0: y_1 = (x0 - x2) + w3 (x1 - x2), y_2 = (x0 - x1) -
 * w3 (x1 - x2), one product by w3 for the three outputs.  The inverse runs the m-point inverses on the thirds and
 * then the mirrored layer: y'_r = y_r w^(-r j), x_0 = y'_0 + y'_1 + y'_2, x_1 = (y'_0 - y'_1) + w3 (y'_2 - y'_1),
 * x_2 = (y'_0 - y'_2) - w3 (y'_2 - y'_1) (the forward w3 in both directions).  w^j comes from the four-level table of
 * W3 (order 3 2^32): w^j = W3^E with E = j 2^(32 - logk); w^(2j) = (w^j)^2.  A thread handles one j of one transform. */
struct R3Args {
    double *x;                      /* batch transforms of 3m points, transform t at x + t 3m */
    uint64_t batch;
    int logk;
    PrimeK pk;
    const double2 *b3;              /* the direction's table (TwTables::b3) */
    const uint64_t *limbs; uint64_t nlimbs;     /* forward from limbs: a single operand (batch 1) ... */
    const LoadDesc *ldesc; int which;           /* ... or operand a / b of ldesc[t] */
    double2 scale; int do_scale;                /* multiply every output by scale */
};
__device__ __forceinline__ double r3_input(const R3Args &a, uint64_t t, uint64_t k)
{
    const uint64_t *src; uint64_t nl;
    if (a.ldesc) { const LoadDesc d = a.ldesc[t]; src = a.which ? d.b : d.a; nl = a.which ? d.nb : d.na; }
    else { src = a.limbs; nl = a.nlimbs; }
    return k < nl ? limb_to_res(src[k], a.pk.p, a.pk.pinv) : 0.0;
}
/* w^j and w^(2j) (times the scale when given), |.| <= p */
__device__ __forceinline__ void r3_twiddles(const R3Args &a, uint64_t j, double *t1, double *t2)
{
    const double p = a.pk.p, pinv = a.pk.pinv;
    const uint32_t E = (uint32_t)(j << (32 - a.logk));           /* logk >= 1 here, j < 2^logk */
    double w1 = bq_twiddle(a.b3, E, p, pinv);
    double w2 = mul_barrett(w1, w1, p, pinv);
    if (a.do_scale) { w1 = mul_shoup(w1, a.scale.x, a.scale.y, p); w2 = mul_shoup(w2, a.scale.x, a.scale.y, p); }
    *t1 = w1; *t2 = w2;
}
template <bool LIMB>
__global__ __launch_bounds__(NTTM_THREADS) void k_r3_fwd(R3Args a)
{
    const uint64_t m = (uint64_t)1 << a.logk, total = a.batch << a.logk;
    const uint64_t stride = (uint64_t)gridDim.x * NTTM_THREADS;
    const double p = a.pk.p, pinv = a.pk.pinv;
    const double2 w3 = a.b3[4 << NTTM_BQ_LOG];
    for (uint64_t i = (uint64_t)blockIdx.x * NTTM_THREADS + threadIdx.x; i < total; i += stride) {
        const uint64_t t = i >> a.logk, j = i & (m - 1);
        double *v = a.x + t * 3 * m;
        double x0, x1, x2;
        if (LIMB) { x0 = r3_input(a, t, j); x1 = r3_input(a, t, j + m); x2 = r3_input(a, t, j + 2 * m); }
        else { x0 = v[j]; x1 = v[j + m]; x2 = v[j + 2 * m]; }       /* |x| <= p */
        const double u = mul_shoup(x1 - x2, w3.x, w3.y, p);         /* |u| <= 5p/8 */
        double y0 = reduce(x0 + x1 + x2, p, pinv);
        const double y1 = reduce((x0 - x2) + u, p, pinv), y2 = reduce((x0 - x1) - u, p, pinv);
        double t1, t2;
        r3_twiddles(a, j, &t1, &t2);
        if (a.do_scale) y0 = mul_shoup(y0, a.scale.x, a.scale.y, p);
        v[j] = y0;
        v[j + m] = mul_barrett(y1, t1, p, pinv);
        v[j + 2 * m] = mul_barrett(y2, t2, p, pinv);
    }
}
__global__ __launch_bounds__(NTTM_THREADS) void k_r3_inv(R3Args a)
{
    const uint64_t m = (uint64_t)1 << a.logk, total = a.batch << a.logk;
    const uint64_t stride = (uint64_t)gridDim.x * NTTM_THREADS;
    const double p = a.pk.p, pinv = a.pk.pinv;
    const double2 w3 = a.b3[4 << NTTM_BQ_LOG];
    for (uint64_t i = (uint64_t)blockIdx.x * NTTM_THREADS + threadIdx.x; i < total; i += stride) {
        const uint64_t t = i >> a.logk, j = i & (m - 1);
        double *v = a.x + t * 3 * m;
        double y0 = v[j];
        double t1, t2;
        r3_twiddles(a, j, &t1, &t2);                                 /* w^-j, w^-2j (scaled) */
        if (a.do_scale) y0 = mul_shoup(y0, a.scale.x, a.scale.y, p);
        const double y1 = mul_barrett(v[j + m], t1, p, pinv), y2 = mul_barrett(v[j + 2 * m], t2, p, pinv);
        const double u = mul_shoup(y2 - y1, w3.x, w3.y, p);         /* |u| <= 5p/8 */
        v[j] = reduce(y0 + y1 + y2, p, pinv);
        v[j + m] = reduce((y0 - y1) + u, p, pinv);
        v[j + 2 * m] = reduce((y0 - y2) - u, p, pinv);
    }
}

/* ---- host launchers ------------------------------------------------------------------------------------------------ */
static unsigned grid_for(uint64_t count)
{
    uint64_t b = (count + NTTM_THREADS - 1) / NTTM_THREADS;
    return (unsigned)(b < 8192 ? (b ? b : 1) : 8192);
}

/* checks that every tile of the pass stays inside the batch (the kernel's own index map, on the host) */
static int pass_geom_ok(const PassGeom &g, int logn, uint64_t batch)
{
    if (g.m < 1 || g.m > NTTM_TILE_LOG || g.logw < 0 || g.logh < 0 || g.m + g.logw + g.logh > NTTM_TILE_LOG) return 0;
    if (g.logw > g.lo || g.lo + g.m > logn) return 0;
    if (g.lo + g.m + g.logh > 32) return 0;                       /* element offsets inside a tile fit 32 bits */
    const uint64_t total = batch << logn, T = (uint64_t)1 << (g.m + g.logw + g.logh);
    if (total % T) return 0;
    const uint64_t tiles = total / T;
    if (plan_gidx(g, tiles - 1, (int)T - 1) != total - 1) return 0;   /* the last element of the last tile */
    if (plan_gidx(g, 0, 0) != 0) return 0;
    return 1;
}

static int g_pts = PASS_PTS;                                   /* points per thread of full tiles: 16 or 8 */
void nttm_set_pts(int pts) { g_pts = pts == 8 ? 8 : PASS_PTS; }
int nttm_get_pts(void) { return g_pts; }
static int g_fixed_shapes = 1;                                  /* use the fixed-shape pass kernels where they apply */
void nttm_set_fixed_shapes(int on) { g_fixed_shapes = on != 0; }
static int g_stream = 1, g_ncu = 256;                           /* streaming strided passes; CUs of the device */
void nttm_set_stream(int on, int ncu) { g_stream = on != 0; g_ncu = ncu > 0 ? ncu : 256; }

/* whether nttm_run_pass runs this forward pass as the streamed kernel (the condition in nttm_run_pass below) */
bool nttm_fwd_pass_streams(const PassArgs &a, const PassGeom &g)
{
    if (!pass_geom_ok(g, a.logn, a.batch) || a.y) return false;
    const bool limb = a.limbs || a.ldesc;
    const int T = 1 << (g.m + g.logw + g.logh);
    const bool fixed = g_fixed_shapes && g.logh == 0 && T >= PASS_PTS;
    const bool shape = g.m >= 1 && g.m <= 9 && g.m + g.logw == 13;
    return fixed && g_stream && shape && g.lo >= g.logw && !a.do_scale &&
           (!limb || ((a.ldesc || a.batch == 1) && g.lo + g.m < 32));
}
int nttm_run_pass(bool inv, const PassArgs &a, const PassGeom &g, hipStream_t s)
{
    if (!pass_geom_ok(g, a.logn, a.batch)) return nttm_fail(NTTM_EINVAL, "internal: bad pass geometry");
    const uint64_t tiles = plan_tiles(g, a.logn, a.batch);
    if (tiles > 0x7fffffffULL) return nttm_fail(NTTM_EINVAL, "internal: too many tiles");
    PassArgs b = a; b.lo = g.lo; b.m = g.m; b.logw = g.logw; b.logh = g.logh;
    const bool limb = !inv && (b.limbs || b.ldesc);
    const int mode = !inv ? (limb ? PM_FWD_LIMB : PM_FWD) : (b.y ? PM_INV_PW : PM_INV);
    const double pts = (double)(a.batch << a.logn);
    /* algorithmic bytes: read and write the plane once (16 B per point), plus the y plane of a fused pointwise (8 B);
     * a limb-input pass reads the limbs (8 B per point of operand) instead of the plane */
    const double bytes = mode == PM_FWD_LIMB ? pts * 8.0 + 8.0 * (double)(b.ldesc ? 0 : b.nlimbs)
                       : pts * (mode == PM_INV_PW ? 24.0 : 16.0);
    static const char *names[4] = {"k_pass<fwd>", "k_pass<fwd,limb>", "k_pass<inv>", "k_pass<inv,pw>"};
    prof_begin(names[mode], bytes, s);
    const int T = 1 << (g.m + g.logw + g.logh);
    /* 8 points per thread only where the tile needs between 64 and 512 threads (the kernel's launch bound) */
    const int ppt = (T >= 64 * g_pts && T <= 512 * g_pts) ? g_pts : PASS_PTS;
    const unsigned nthr = (unsigned)(T / ppt < 64 ? 64 : T / ppt);
    if (nthr > 512) return nttm_fail(NTTM_EINVAL, "internal: %u threads per tile", nthr);
    const dim3 grid((unsigned)tiles), blk(nthr);
    const size_t lds = (size_t)LDS_WORDS(T) * sizeof(double);
#define NTTM_PASS_LAUNCH(P) \
    switch (mode) { \
    case PM_FWD: hipLaunchKernelGGL((k_pass<PM_FWD, P>), grid, blk, lds, s, b); break; \
    case PM_FWD_LIMB: hipLaunchKernelGGL((k_pass<PM_FWD_LIMB, P>), grid, blk, lds, s, b); break; \
    case PM_INV: hipLaunchKernelGGL((k_pass<PM_INV, P>), grid, blk, lds, s, b); break; \
    default: hipLaunchKernelGGL((k_pass<PM_INV_PW, P>), grid, blk, lds, s, b); break; }
#define NTTM_PASS_LAUNCH_G(P, M, W, H) \
    switch (mode) { \
    case PM_FWD: hipLaunchKernelGGL((k_pass<PM_FWD, P, M, W, H>), grid, blk, lds, s, b); break; \
    case PM_FWD_LIMB: hipLaunchKernelGGL((k_pass<PM_FWD_LIMB, P, M, W, H>), grid, blk, lds, s, b); break; \
    case PM_INV: hipLaunchKernelGGL((k_pass<PM_INV, P, M, W, H>), grid, blk, lds, s, b); break; \
    default: hipLaunchKernelGGL((k_pass<PM_INV_PW, P, M, W, H>), grid, blk, lds, s, b); break; }
    /* the strided tile shapes of large transforms, compiled with their geometry fixed */
    const bool fixed = g_fixed_shapes && g.logh == 0 && T >= PASS_PTS;
    /* the streamed shapes: strided 2^13-point tiles of 1 to 9-bit rows (2^(13 - m) columns), limbs from one operand
     * (a batch of one) or per transform from the descriptors */
    const bool shape = g.m >= 1 && g.m <= 9 && g.m + g.logw == 13;
    const bool stream = fixed && g_stream && shape && g.lo >= g.logw && mode != PM_INV_PW &&
                        (mode == PM_INV || !b.do_scale) &&
                        (mode != PM_FWD_LIMB || ((b.ldesc || b.batch == 1) && g.lo + g.m < 32));
    if ((b.dfr_out && (!stream || mode != PM_FWD_LIMB)) || (b.dfr_in && (!stream || mode != PM_FWD || g.m < 4)))
        return nttm_fail(NTTM_EINVAL, "internal: deferred boundary twiddles on a pass that does not stream");
    if (stream) {
        const unsigned nb = (unsigned)(tiles < (uint64_t)g_ncu ? tiles : (uint64_t)g_ncu);
        const uint32_t nt = (uint32_t)tiles;
#define NTTM_STREAM_LAUNCH(M, W) \
        if (mode == PM_FWD) hipLaunchKernelGGL((k_pass_stream<PM_FWD, M, W>), dim3(nb), dim3(STREAM_T), 0, s, b, nt); \
        else if (mode == PM_FWD_LIMB) hipLaunchKernelGGL((k_pass_stream<PM_FWD_LIMB, M, W>), dim3(nb), dim3(STREAM_T), 0, s, b, nt); \
        else hipLaunchKernelGGL((k_pass_stream<PM_INV, M, W>), dim3(nb), dim3(STREAM_T), 0, s, b, nt);
        switch (g.m) {
        case 1: NTTM_STREAM_LAUNCH(1, 12) break;
        case 2: NTTM_STREAM_LAUNCH(2, 11) break;
        case 3: NTTM_STREAM_LAUNCH(3, 10) break;
        case 4: NTTM_STREAM_LAUNCH(4, 9) break;
        case 5: NTTM_STREAM_LAUNCH(5, 8) break;
        case 6: NTTM_STREAM_LAUNCH(6, 7) break;
        case 7: NTTM_STREAM_LAUNCH(7, 6) break;
        case 8: NTTM_STREAM_LAUNCH(8, 5) break;
        default: NTTM_STREAM_LAUNCH(9, 4) break;
        }
#undef NTTM_STREAM_LAUNCH
    }
    else if (fixed && ppt == PASS_PTS && g.m == 8 && g.logw == 5) { NTTM_PASS_LAUNCH_G(PASS_PTS, 8, 5, 0) }
    else if (T >= PASS_PTS && ppt == 8) { NTTM_PASS_LAUNCH(8) }
    else if (T >= PASS_PTS) { NTTM_PASS_LAUNCH(PASS_PTS) }
    else if (T == 8) { NTTM_PASS_LAUNCH(8) }                   /* tiles of 2..8 points (transforms below 2^4, odd */
    else if (T == 4) { NTTM_PASS_LAUNCH(4) }                   /* batch): one thread */
    else { NTTM_PASS_LAUNCH(2) }
#undef NTTM_PASS_LAUNCH
#undef NTTM_PASS_LAUNCH_G
    prof_end(s);
    return nttm_hip_check(hipGetLastError(), "k_pass launch");
}

/* the fused last forward pass + pointwise + first inverse pass; g must be a contiguous pass (lo = 0) of at most 2^12
 * points per tile and at least 16 */
int nttm_run_fused(const PassArgs &a, const PassGeom &g, hipStream_t s)
{
    if (!pass_geom_ok(g, a.logn, a.batch) || g.lo != 0 || g.logw != 0 || g.m + g.logh > FUSED_LOG ||
        g.m + g.logh < 4 || !a.y || a.ymode != NTTM_Y_FULL)
        return nttm_fail(NTTM_EINVAL, "internal: bad fused pass geometry");
    const uint64_t tiles = plan_tiles(g, a.logn, a.batch);
    if (tiles > 0x7fffffffULL) return nttm_fail(NTTM_EINVAL, "internal: too many tiles");
    PassArgs b = a; b.lo = 0; b.m = g.m; b.logw = 0; b.logh = g.logh;
    const bool limb = b.limbs || b.ldesc;
    const double pts = (double)(a.batch << a.logn);
    /* read x (or its limbs), read y, write x */
    prof_begin(limb ? "k_fused<limb>" : "k_fused", pts * (limb ? 16.0 : 24.0), s);
    const int T = 1 << (g.m + g.logh);
    const int ppt = (T >= 64 * g_pts && T <= 512 * g_pts) ? g_pts : PASS_PTS;
    const unsigned nthr = (unsigned)(T / ppt < 64 ? 64 : T / ppt);
    if (nthr > 512) return nttm_fail(NTTM_EINVAL, "internal: %u threads per tile", nthr);
    if (ppt == 8) {
        if (limb) hipLaunchKernelGGL((k_fused<true, 8>), dim3((unsigned)tiles), dim3(nthr), 0, s, b);
        else hipLaunchKernelGGL((k_fused<false, 8>), dim3((unsigned)tiles), dim3(nthr), 0, s, b);
    } else {
        if (limb) hipLaunchKernelGGL((k_fused<true, PASS_PTS>), dim3((unsigned)tiles), dim3(nthr), 0, s, b);
        else hipLaunchKernelGGL((k_fused<false, PASS_PTS>), dim3((unsigned)tiles), dim3(nthr), 0, s, b);
    }
    prof_end(s);
    return nttm_hip_check(hipGetLastError(), "k_fused launch");
}

/* the double-fused last pass: a = A's pass arguments (x = A's plane), bq = B's (x = B's plane before its last pass) */
int nttm_run_fused2(const PassArgs &a, const PassArgs &bq, const PassGeom &g, hipStream_t s)
{
    if (!pass_geom_ok(g, a.logn, a.batch) || g.lo != 0 || g.logw != 0 || g.m + g.logh > FUSED_LOG ||
        g.m + g.logh < 4 || !bq.x || bq.logn != a.logn || bq.batch != a.batch || a.do_scale || a.bmap ||
        (bq.bmap && (g.m + g.logh > a.logn || bq.limbs || bq.ldesc)))
        return nttm_fail(NTTM_EINVAL, "internal: bad double-fused pass geometry");
    const uint64_t tiles = plan_tiles(g, a.logn, a.batch);
    if (tiles > 0x7fffffffULL) return nttm_fail(NTTM_EINVAL, "internal: too many tiles");
    PassArgs A = a; A.lo = 0; A.m = g.m; A.logw = 0; A.logh = g.logh;
    PassArgs Bq = bq; Bq.lo = 0; Bq.m = g.m; Bq.logw = 0; Bq.logh = g.logh;
    const double pts = (double)(a.batch << a.logn);
    const bool alimb = A.limbs || A.ldesc, blimb = Bq.limbs || Bq.ldesc;
    /* read A (or its limbs), read B (or its limbs), write A */
    prof_begin("k_fused2", pts * (24.0 - (alimb ? 8.0 : 0.0) - (blimb ? 8.0 : 0.0)), s);
    const int T = 1 << (g.m + g.logh);
    const int ppt = (T >= 64 * g_pts && T <= 512 * g_pts) ? g_pts : PASS_PTS;
    const unsigned nthr = (unsigned)(T / ppt < 64 ? 64 : T / ppt);
    if (nthr > 512) return nttm_fail(NTTM_EINVAL, "internal: %u threads per tile", nthr);
    /* the fixed shapes: 4096-point tiles of 2^12-point blocks, or of 4 x 2^10 and 2 x 2^11 (the last pass of a
     * transform of more than one pass, or a whole transform of 2^10 to 2^12 points, which reads the limbs); whole
     * transforms of 2^8 and 2^9 points from the limbs in tiles of 16 and 8 blocks (8 points per thread) */
    const bool fixed_shape = g_fixed_shapes && g.m + g.logh == FUSED_LOG && (g.m >= 10 || (g.m >= 8 && ppt == 8));
    const bool fixed = fixed_shape && !alimb && !blimb && g.m >= 10;
    const dim3 grid((unsigned)tiles), blk(nthr);
    if (fixed_shape && !fixed && (alimb || blimb)) {            /* a single-pass transform: both operands from limbs */
        if (ppt == 8) {
            if (g.m == 12) hipLaunchKernelGGL((k_fused2<8, 12, 0, true>), grid, blk, 0, s, A, Bq);
            else if (g.m == 11) hipLaunchKernelGGL((k_fused2<8, 11, 1, true>), grid, blk, 0, s, A, Bq);
            else if (g.m == 10) hipLaunchKernelGGL((k_fused2<8, 10, 2, true>), grid, blk, 0, s, A, Bq);
            else if (g.m == 9) hipLaunchKernelGGL((k_fused2<8, 9, 3, true>), grid, blk, 0, s, A, Bq);
            else hipLaunchKernelGGL((k_fused2<8, 8, 4, true>), grid, blk, 0, s, A, Bq);
        } else {
            if (g.m == 12) hipLaunchKernelGGL((k_fused2<PASS_PTS, 12, 0, true>), grid, blk, 0, s, A, Bq);
            else if (g.m == 11) hipLaunchKernelGGL((k_fused2<PASS_PTS, 11, 1, true>), grid, blk, 0, s, A, Bq);
            else hipLaunchKernelGGL((k_fused2<PASS_PTS, 10, 2, true>), grid, blk, 0, s, A, Bq);
        }
    } else if (fixed && ppt == 8) {
        if (g.m == 12) hipLaunchKernelGGL((k_fused2<8, 12, 0, false>), grid, blk, 0, s, A, Bq);
        else if (g.m == 11) hipLaunchKernelGGL((k_fused2<8, 11, 1, false>), grid, blk, 0, s, A, Bq);
        else hipLaunchKernelGGL((k_fused2<8, 10, 2, false>), grid, blk, 0, s, A, Bq);
    } else if (fixed) {
        if (g.m == 12) hipLaunchKernelGGL((k_fused2<PASS_PTS, 12, 0, false>), grid, blk, 0, s, A, Bq);
        else if (g.m == 11) hipLaunchKernelGGL((k_fused2<PASS_PTS, 11, 1, false>), grid, blk, 0, s, A, Bq);
        else hipLaunchKernelGGL((k_fused2<PASS_PTS, 10, 2, false>), grid, blk, 0, s, A, Bq);
    } else if (ppt == 8) hipLaunchKernelGGL((k_fused2<8>), grid, blk, 0, s, A, Bq);
    else hipLaunchKernelGGL((k_fused2<PASS_PTS>), grid, blk, 0, s, A, Bq);
    prof_end(s);
    return nttm_hip_check(hipGetLastError(), "k_fused2 launch");
}

static void load_common(LoadArgs &a, int np, const PrimeK *pk, const double2 *scale)
{
    a.np = np; a.do_scale = scale != nullptr;
    for (int k = 0; k < NTTM_MAX_PRIMES; k++) {
        a.pk[k] = k < np ? pk[k] : pk[0];
        a.scale[k] = scale && k < np ? scale[k] : make_double2(1.0, 0.0);
    }
}
int nttm_run_load_planes(double *dst, uint64_t plane, const uint64_t *src, uint64_t nlimbs, uint64_t npoints, int np,
                         const PrimeK *pk, const double2 *scale, int out_u64, hipStream_t s)
{
    LoadArgs a; memset(&a, 0, sizeof a);
    a.dst = dst; a.plane = plane; a.src = src; a.nlimbs = nlimbs; a.npoints = npoints; a.desc = nullptr; a.logn = 0;
    a.out_u64 = out_u64;
    load_common(a, np, pk, scale);
    prof_begin("k_load", 8.0 * (double)nlimbs + 8.0 * np * (double)npoints, s);
    hipLaunchKernelGGL(k_load, dim3(grid_for(npoints)), dim3(NTTM_THREADS), 0, s, a);
    prof_end(s);
    return nttm_hip_check(hipGetLastError(), "k_load launch");
}
int nttm_run_load_batch(double *dst, uint64_t plane, const void *desc, int which, uint64_t count, int logn, int np,
                        const PrimeK *pk, const double2 *scale, hipStream_t s)
{
    LoadArgs a; memset(&a, 0, sizeof a);
    a.dst = dst; a.plane = plane; a.src = nullptr; a.npoints = count << logn; a.desc = (const LoadDesc *)desc;
    a.which = which; a.logn = logn; a.out_u64 = 0;
    load_common(a, np, pk, scale);
    prof_begin("k_load", 16.0 * np * (double)a.npoints, s);
    hipLaunchKernelGGL(k_load, dim3(grid_for(a.npoints)), dim3(NTTM_THREADS), 0, s, a);
    prof_end(s);
    return nttm_hip_check(hipGetLastError(), "k_load launch");
}
/* the radix-3 layer of batch transforms of 3 2^logk points (logk >= 1): forward (from the plane or from limbs) or
 * inverse, in place */
int nttm_run_r3(bool inv, double *x, uint64_t batch, int logk, const PrimeK &pk, const double2 *b3,
                const uint64_t *limbs, uint64_t nlimbs, const LoadDesc *ldesc, int which, const double2 *scale,
                hipStream_t s)
{
    if (logk < 1 || logk > NTTM_LOGN_MAX || !batch) return nttm_fail(NTTM_EINVAL, "internal: radix-3 layer of 3 2^%d", logk);
    if (inv && (limbs || ldesc)) return nttm_fail(NTTM_EINVAL, "internal: inverse radix-3 layer from limbs");
    if (limbs && batch != 1) return nttm_fail(NTTM_EINVAL, "internal: one limb operand for %llu transforms",
                                              (unsigned long long)batch);
    R3Args a; memset(&a, 0, sizeof a);
    a.x = x; a.batch = batch; a.logk = logk; a.pk = pk; a.b3 = b3;
    a.limbs = limbs; a.nlimbs = nlimbs; a.ldesc = ldesc; a.which = which;
    a.do_scale = scale != nullptr; a.scale = scale ? *scale : make_double2(1.0, 0.0);
    const bool limb = limbs || ldesc;
    const uint64_t pts = 3 * (batch << logk);
    prof_begin(inv ? "k_r3_inv" : limb ? "k_r3_fwd<limb>" : "k_r3_fwd",
               limb ? 8.0 * (double)pts + 8.0 * (double)(ldesc ? 0 : nlimbs) : 16.0 * (double)pts, s);
    const unsigned grid = grid_for(batch << logk);
    if (inv) hipLaunchKernelGGL(k_r3_inv, dim3(grid), dim3(NTTM_THREADS), 0, s, a);
    else if (limb) hipLaunchKernelGGL(k_r3_fwd<true>, dim3(grid), dim3(NTTM_THREADS), 0, s, a);
    else hipLaunchKernelGGL(k_r3_fwd<false>, dim3(grid), dim3(NTTM_THREADS), 0, s, a);
    prof_end(s);
    return nttm_hip_check(hipGetLastError(), "k_r3 launch");
}
int nttm_run_u64_to_f64(double *x, uint64_t count, hipStream_t s)
{
    hipLaunchKernelGGL(k_cvt_u64_f64, dim3(grid_for(count)), dim3(NTTM_THREADS), 0, s, x, (const uint64_t *)x, count);
    return nttm_hip_check(hipGetLastError(), "k_cvt_u64_f64 launch");
}
int nttm_run_cvt_u64_f64(double *dst, const uint64_t *src, uint64_t count, hipStream_t s)
{
    hipLaunchKernelGGL(k_cvt_u64_f64, dim3(grid_for(count)), dim3(NTTM_THREADS), 0, s, dst, src, count);
    return nttm_hip_check(hipGetLastError(), "k_cvt_u64_f64 launch");
}
int nttm_run_f64_to_canon(double *x, uint64_t count, const PrimeK &pk, hipStream_t s)
{
    hipLaunchKernelGGL(k_f64_to_canon, dim3(grid_for(count)), dim3(NTTM_THREADS), 0, s, x, count, pk);
    return nttm_hip_check(hipGetLastError(), "k_f64_to_canon launch");
}
int nttm_run_pw_u64(uint64_t *x, const uint64_t *y, uint64_t count, const PrimeK &pk, hipStream_t s)
{
    hipLaunchKernelGGL(k_pw_u64, dim3(grid_for(count)), dim3(NTTM_THREADS), 0, s, x, y, count, pk);
    return nttm_hip_check(hipGetLastError(), "k_pw_u64 launch");
}
