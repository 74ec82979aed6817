/* crt.hip - Chinese remaindering of the residue planes into base-10^18 limbs, with the carry resolved on the device.
 *
 * Coefficient k of the convolution is rebuilt from its np residues by Garner's mixed radix (modular steps in the
 * FP64 arithmetic of arith.h), and its value X = sum_i t_i p_0 ... p_{i-1} is split into np base-10^18 digits
 * d_0..d_{np-1} from the products of the t_i with the base-10^18 digits of the prime prefixes (mulmod_b), and
 *     limb_k = d_0[k] + d_1[k-1] + ... + d_{np-1}[k-np+1] + x_k + carry_k            (E/rns_mul.c:332-405)
 * The top digit is small (X < 2^(50 np) gives d_{np-1} < 2^31), so v_k = sum of the digits + x_k < np B + 2^31.  With
 * q_k = floor(v_k / B) <= np, r_k = v_k mod B and w_k = r_k + q_{k-1} (<= B + np - 1), the carry into limb k+1 is
 * q_k + o_k where
 * the overflow bit o_k = [w_k + o_{k-1} >= B] = g_k | (p_k & o_{k-1}), g_k = [w_k >= B], p_k = [w_k = B - 1]: a binary
 * carry-lookahead chain.  For 64 consecutive limbs it is resolved by one 64-bit addition of the ballot masks
 * X = G, Y = G | P (the carries into the bits of X + Y + cin are the o's, the textbook adder identity).  Words chain
 * inside a wave and waves inside a block; each block then stores its limbs for a carry in of 0 together with its
 * (cin -> cout) map of one bit.  One workgroup scans the maps of all blocks (restarting at every product's first
 * block) into each block's carry in, and a last pass adds 1 to the limbs of the blocks that receive a carry: the
 * limbs for cin = 1 are those for cin = 0 plus one, a ripple through the leading limbs equal to B - 1.  No block
 * waits on another, and the scan reads one word per 2048 limbs.
 * The reference instead ripples each 256-coefficient chunk on one thread of a wave64 ballot scan, writes a 4-limb spill
 * per stripe and merges the spills serially on the host or on one thread per product (E/rns_mul.c:452-470;
 * E/dbig.c:1007-1052).  The final carry beyond limb nc-1 must be zero; a non-zero one sets an error flag.
 */
#include "internal.h"

#define CRT_T 256
#define CRT_WIDE 8                       /* limbs per thread of a block (2048-limb blocks); groups of small products
                                            use 1 (256-limb blocks, less of a block idle past a product's end) */
#define CRT_HALO 4                       /* digits of coefficients k0-4 .. k0-1: v_{k0-1} with four primes */
#define CRT_ID_MAP 0x2u                  /* one-bit carry maps: bit c = cout(cin = c); the identity is 0b10 */
#define CRT_CTL_WORDS 4                  /* control words before the block status words (16-byte aligned) */
/* block status word: map (2 bits) | error for cin 0 << 2 | error for cin 1 << 3 | carry in << 4 | first of product << 5 */
#define CRT_ST_E0 (1u << 2)
#define CRT_ST_E1 (1u << 3)
#define CRT_ST_CIN (1u << 4)
#define CRT_ST_FIRST (1u << 5)
#define CRT_SCAN_T 1024

struct CrtDesc {                         /* one product; must match api.hip */
    const double *pl; uint64_t plane;    /* residues of prime j, coefficient k at pl[j * plane + k] */
    uint64_t ncoef;
    const uint64_t *x; uint64_t nx;
    uint64_t *c; uint64_t nc;
    uint64_t first_block;                /* the product's first block in the launch */
    uint64_t trunc;                      /* nonzero: c is the value mod B^nc (nothing above it is checked) */
};
struct CrtArgs {
    const CrtDesc *desc; uint64_t nprod;
    uint32_t *status;                    /* one word per block (CRT_ST_*) */
    uint32_t *ctl;                       /* [1] error flags */
    CrtConst cc;
};

__device__ __forceinline__ uint32_t map_apply(uint32_t f, uint32_t c) { return (f >> c) & 1u; }
/* (g o f)(c) = g(f(c)) for maps {0,1} -> {0,1} */
__device__ __forceinline__ uint32_t map_compose(uint32_t g, uint32_t f)
{
    return map_apply(g, f & 1u) | (map_apply(g, (f >> 1) & 1u) << 1);
}
/* carry out of X + Y + cin over 64 bits */
__device__ __forceinline__ uint32_t add_cout(uint64_t X, uint64_t Y, uint32_t cin)
{
    const uint64_t s = X + Y;
    const uint32_t c1 = s < X;
    const uint64_t s2 = s + cin;
    return c1 | (uint32_t)(s2 < s);
}

template <int NP>
__device__ __forceinline__ void crt_digits(const CrtConst &cc, const double *r, uint64_t *d)
{
    /* Garner's digits t_i in [0, p_i), for residues |r_i| < p_i (balanced or canonical):
     *   t_0 = r_0 mod p_0;  t_i = (r_i - t_0 - sum_{0<j<i} t_j (p_0 ... p_{j-1})) (p_0 ... p_{i-1})^-1 mod p_i.
     * The primes differ by less than 2^40, so t_j < p_j <= 1.0007 p_i: each t_j m_ij (|m_ij| <= p_i / 2) stays within
     * the Barrett product's bound (|result| < 0.63 p_i), the sum r_i - acc within 4 p_i (one reduce), and no
     * intermediate reduction of t_j mod p_i is needed. */
    uint64_t t[NP];
    double td[NP];                                       /* the same digits as doubles (exact, below 2^50) */
    const double p0 = cc.pk[0].p;
    { const double v = r[0] < 0.0 ? r[0] + p0 : r[0]; td[0] = v; t[0] = f64_to_u52(v); }
#pragma unroll
    for (int i = 1; i < NP; i++) {
        const double p = cc.pk[i].p, pinv = cc.pk[i].pinv;
        double acc = td[0];
#pragma unroll
        for (int j = 1; j < i; j++) acc += mul_barrett(td[j], cc.mmod[i][j], p, pinv);
        const double dd = reduce(r[i] - acc, p, pinv);
        double ti = mul_barrett(dd, cc.cinv[i], p, pinv);
        ti = ti < 0.0 ? ti + p : ti;
        td[i] = ti; t[i] = f64_to_u52(ti);
    }
    /* X = sum_i t_i M_i with M_i = p_0 ... p_{i-1} = sum_j m_ij B^j: every t_i m_ij = a_ij B + b_ij (mulmod_b), then
     * base-B columns col_j = sum_i b_ij + sum_i a_i(j-1) (+ t_0 in column 0) with their small carries.  Column sums
     * stay below NP B + NP 2^50 < 2^62; X < 2^(50 NP) bounds the top digit (NP = 3: < 2^31, NP = 4: < 2^21). */
    uint64_t acc[NP + 1];
#pragma unroll
    for (int j = 0; j <= NP; j++) acc[j] = 0;
    acc[0] = t[0];
#pragma unroll
    for (int i = 1; i < NP; i++)
#pragma unroll
        for (int j = 0; j < i; j++) {
            uint64_t ah, bl;
            mulmod_b(t[i], cc.mb[i][j], cc.mbp[i][j], &ah, &bl);
            acc[j] += bl; acc[j + 1] += ah;
        }
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < NP - 1; j++) {
        /* column j holds NP - 1 - j terms below B, the rest (a's, t_0, the carry in) below 2^52 < B: its carry is at
         * most NP - 1 - j */
        uint64_t v = acc[j] + c;
        c = 0;
#pragma unroll
        for (int k = 0; k < NP - 1 - j; k++) { const int ge = v >= NTTM_B; v = ge ? v - NTTM_B : v; c += (uint64_t)ge; }
        d[j] = v;
    }
    d[NP - 1] = acc[NP - 1] + c;
}

template <int NP, int PER>
__global__ __launch_bounds__(CRT_T) void k_crt(CrtArgs a)
{
    constexpr int CH = CRT_T * PER;                                  /* limbs per block */
    __shared__ uint64_t D[NP][CH + CRT_HALO];
    __shared__ uint32_t s_prod, s_err;
    __shared__ uint32_t s_wave[CRT_T / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t blk = blockIdx.x;
    if (tid == 0) {
        uint64_t lo = 0, hi = a.nprod - 1;                  /* the product whose block range holds blk */
        while (lo < hi) { const uint64_t mid = (lo + hi + 1) >> 1; if (a.desc[mid].first_block <= blk) lo = mid; else hi = mid - 1; }
        s_prod = (uint32_t)lo; s_err = 0;
    }
    __syncthreads();
    const CrtDesc dsc = a.desc[s_prod];
    const uint32_t first = (uint32_t)dsc.first_block, local = blk - first;
    const uint64_t k0 = (uint64_t)local * CH;

    /* digits of coefficients k0 - HALO .. k0 + CH - 1 into LDS (coalesced plane reads, all of a thread's residues
     * requested before the first is used) */
    uint64_t xk[PER];
    {
        const bool halo = tid < CRT_HALO && k0 + tid >= CRT_HALO;   /* coefficient k0 - HALO + tid exists */
        /* the planes as global-space pointers: generic (flat) loads would also count against the LDS counter, whose
         * limit of 15 outstanding operations would serialize the prefetch */
        const __attribute__((address_space(1))) double *pl = (const __attribute__((address_space(1))) double *)dsc.pl;
        const __attribute__((address_space(1))) uint64_t *xg = (const __attribute__((address_space(1))) uint64_t *)dsc.x;
        double rr[PER + 1][NP];
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const uint64_t kk = k0 + i * CRT_T + tid;
#pragma unroll
            for (int j = 0; j < NP; j++) rr[i][j] = kk < dsc.ncoef ? pl[j * dsc.plane + kk] : 0.0;
        }
        {
            const uint64_t kk = k0 + tid - CRT_HALO;
#pragma unroll
            for (int j = 0; j < NP; j++) rr[PER][j] = halo && kk < dsc.ncoef ? pl[j * dsc.plane + kk] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < PER; i++) {                          /* x at the limbs of the limb phase below */
            const uint64_t k = k0 + wave * 64 * PER + i * 64 + lane;
            xk[i] = k < dsc.nx && k < dsc.nc ? xg[k] : 0;
        }
#pragma unroll
        for (int i = 0; i < PER; i++) {
            uint64_t d[NP];
            crt_digits<NP>(a.cc, rr[i], d);                          /* zero residues give zero digits */
#pragma unroll
            for (int j = 0; j < NP; j++) D[j][CRT_HALO + i * CRT_T + tid] = d[j];
        }
        if (tid < CRT_HALO) {
            uint64_t d[NP];
            crt_digits<NP>(a.cc, rr[PER], d);
#pragma unroll
            for (int j = 0; j < NP; j++) D[j][tid] = d[j];
        }
    }
    __syncthreads();

    /* limb k = k0 + wave 64 PER + i 64 + lane for i < PER: one ballot per i gives 64 consecutive limbs */
    /* v_k = sum_j d_j[k - j] + x_k < NP B + 2^31: its quotient by B is at most NP */
    auto vsum = [&](uint64_t k, int li, uint64_t xv) -> uint64_t {
        uint64_t v = 0;
        if (k < dsc.nc) {
#pragma unroll
            for (int j = 0; j < NP; j++) v += D[j][li - j];
            v += xv;
        }
        return v;
    };
    auto quot = [](uint64_t v) -> uint32_t {
        uint32_t q = 0;
#pragma unroll
        for (int c = 1; c <= NP; c++) q += v >= (uint64_t)c * NTTM_B;
        return q;
    };
    uint64_t wv[PER];
    uint64_t GX[PER], GY[PER];
    uint32_t qlast = 0, qprev = 0;
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int off = wave * 64 * PER + i * 64 + lane;          /* offset of limb k inside the block */
        const uint64_t k = k0 + off;
        const int li = CRT_HALO + off;
        const uint64_t v = vsum(k, li, xk[i]);
        const uint32_t q = quot(v);
        /* q_{k-1}: from the lane below; lane 0 from the previous word's lane 63, or (first word) computed */
        uint32_t qp = __shfl_up(q, 1, 64);
        if (i > 0) { if (lane == 0) qp = qprev; }
        else if (lane == 0) qp = k >= 1 ? quot(vsum(k - 1, li - 1, k - 1 < dsc.nx ? dsc.x[k - 1] : 0)) : 0u;
        qprev = __builtin_amdgcn_readlane(q, 63);
        const uint64_t w = v - (uint64_t)q * NTTM_B + qp;            /* <= B + NP */
        wv[i] = w;
        const uint64_t g = __ballot(w >= NTTM_B), pr = __ballot(w == NTTM_B - 1);
        GX[i] = g; GY[i] = g | pr;
        if (k + 1 == dsc.nc) qlast = q;
    }
    /* the wave's map cin -> cout over its PER words */
    uint32_t F = CRT_ID_MAP;
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const uint32_t f = add_cout(GX[i], GY[i], 0) | (add_cout(GX[i], GY[i], 1) << 1);
        F = map_compose(f, F);
    }
    if (lane == 0) s_wave[wave] = F;
    __syncthreads();
    uint32_t P = CRT_ID_MAP, A = CRT_ID_MAP;
    for (int w = 0; w < CRT_T / 64; w++) {
        if (w < wave) P = map_compose(s_wave[w], P);
        A = map_compose(s_wave[w], A);
    }

    /* final limbs for a carry of 0 into the block: out_k = w_k + o_{k-1} - B o_k, stored directly (lanes hold
     * consecutive limbs).  c1 follows the carries for a carry of 1 into the block, for the check of the last limb. */
    __attribute__((address_space(1))) uint64_t *cg = (__attribute__((address_space(1))) uint64_t *)dsc.c;
    uint32_t c = map_apply(P, 0), c1 = map_apply(P, 1);              /* the carries into this wave's first word */
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const uint64_t S = GX[i] + GY[i] + c;
        const uint32_t cout = add_cout(GX[i], GY[i], c);
        const uint64_t CI = S ^ GX[i] ^ GY[i];                      /* bit b: the carry into bit b (o_{k-1}) */
        const uint32_t oin = (uint32_t)(CI >> lane) & 1u;
        const uint32_t oout = lane == 63 ? cout : (uint32_t)(CI >> (lane + 1)) & 1u;
        const int off = wave * 64 * PER + i * 64 + lane;
        const uint64_t k = k0 + off;
        if (k < dsc.nc) cg[k] = wv[i] + oin - (oout ? NTTM_B : 0);
        const uint32_t cout1 = add_cout(GX[i], GY[i], c1);
        if (k + 1 == dsc.nc && !dsc.trunc) {                       /* the last limb: nothing may remain beyond it */
            const uint64_t CI1 = (GX[i] + GY[i] + c1) ^ GX[i] ^ GY[i];
            const uint32_t oout1 = lane == 63 ? cout1 : (uint32_t)(CI1 >> (lane + 1)) & 1u;
            uint64_t extra = qlast;
            const int li = CRT_HALO + off;
#pragma unroll
            for (int e = 1; e < NP; e++)
#pragma unroll
                for (int j = e; j < NP; j++) extra += D[j][li + e - j];
            s_err = (extra + oout ? CRT_ST_E0 : 0u) | (extra + oout1 ? CRT_ST_E1 : 0u);
        }
        c = cout; c1 = cout1;
    }
    __syncthreads();
    if (tid == 0) a.status[blk] = A | s_err | (local == 0 ? CRT_ST_FIRST : 0u);
}

/* carry in of every block: maps composed in block order, restarting at each product's first block.  One workgroup of
 * 16 waves, each over a contiguous run of the status words (four per lane and load); the error flag of a product's
 * last block is taken for its actual carry in. */
__device__ __forceinline__ uint32_t crt_step_map(uint32_t st)
{
    const uint32_t m = st & 3u;                                     /* a first block ignores the carry before it */
    return (st & CRT_ST_FIRST) ? (m & 1u) * 3u : m;
}
__global__ __launch_bounds__(CRT_SCAN_T) void k_crt_scan(uint32_t *status, uint64_t n, uint32_t *ctl)
{
    __shared__ uint32_t s_map[CRT_SCAN_T / 64], s_err;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int NW = CRT_SCAN_T / 64, U = 4;                      /* waves; loads of 4 words in flight per lane */
    if (tid == 0) s_err = 0;
    const uint64_t per = ((n + NW - 1) / NW + 256 * U - 1) / (256 * U) * (256 * U);
    const uint64_t lo = (uint64_t)wave * per, hi = lo + per < n ? lo + per : n;
    auto load4 = [&](uint64_t i, uint32_t *w) {                     /* words i..i+3 (i a multiple of 4) */
        if (i + 4 <= hi) {
            const uint4 v = *(const uint4 *)(status + i);
            w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) w[j] = i + j < hi ? status[i + j] : CRT_ID_MAP;
        }
    };
    /* pass 1: the map of the wave's run */
    uint32_t F = CRT_ID_MAP;
    for (uint64_t b = lo; b < hi; b += 256 * U) {
        uint32_t w[U][4];
#pragma unroll
        for (int u = 0; u < U; u++) load4(b + u * 256 + lane * 4, w[u]);
#pragma unroll
        for (int u = 0; u < U; u++) {
            uint32_t m = CRT_ID_MAP;
#pragma unroll
            for (int j = 0; j < 4; j++) m = map_compose(crt_step_map(w[u][j]), m);
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {                      /* lane 0: m_63 o ... o m_0 */
                const uint32_t o = __shfl_down(m, d, 64);
                if ((lane & (2 * d - 1)) == 0) m = map_compose(o, m);
            }
            F = map_compose(__builtin_amdgcn_readfirstlane(m), F);
        }
    }
    if (lane == 0) s_map[wave] = F;
    __syncthreads();
    uint32_t cw = 0;                                                /* the carry into the wave's run */
    for (int v = 0; v < wave; v++) cw = map_apply(s_map[v], cw);
    /* pass 2: carries into the blocks */
    uint32_t err = 0;
    for (uint64_t b = lo; b < hi; b += 256 * U) {
        uint32_t w[U][4];
#pragma unroll
        for (int u = 0; u < U; u++) load4(b + u * 256 + lane * 4, w[u]);
#pragma unroll
        for (int u = 0; u < U; u++) {
            uint32_t m = CRT_ID_MAP;
#pragma unroll
            for (int j = 0; j < 4; j++) m = map_compose(crt_step_map(w[u][j]), m);
            uint32_t inc = m;                                       /* inclusive scan over the lanes */
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t o = __shfl_up(inc, d, 64);
                if (lane >= d) inc = map_compose(inc, o);
            }
            uint32_t ex = __shfl_up(inc, 1, 64);
            if (lane == 0) ex = CRT_ID_MAP;
            uint32_t ci = map_apply(ex, cw);
            const uint64_t i0 = b + u * 256 + lane * 4;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t st = w[u][j];
                if (st & CRT_ST_FIRST) ci = 0;
                err |= st & (ci ? CRT_ST_E1 : CRT_ST_E0);
                w[u][j] = (st & ~CRT_ST_CIN) | (ci ? CRT_ST_CIN : 0u);
                ci = map_apply(st & 3u, ci);
            }
            if (i0 + 4 <= hi) *(uint4 *)(status + i0) = make_uint4(w[u][0], w[u][1], w[u][2], w[u][3]);
            else {
#pragma unroll
                for (int j = 0; j < 4; j++) if (i0 + j < hi) status[i0 + j] = w[u][j];
            }
            cw = map_apply(__builtin_amdgcn_readlane(inc, 63), cw);
        }
    }
    if (__ballot(err != 0) && lane == 0) atomicOr(&s_err, 1u);
    __syncthreads();
    if (tid == 0) { ctl[0] = 0; ctl[1] = s_err; }
}

/* the limbs of a block whose carry in is 1: its stored limbs (for carry 0) plus one, i.e. the leading limbs equal to
 * B - 1 become 0 and the first other one grows by 1.  A workgroup scouts 64 status words; each of its 4 waves takes
 * every 4th block found and rewrites it (CH / 64 limbs per lane, all loads in flight at once). */
#define CRT_FIX_W 64
template <int PER>
__global__ __launch_bounds__(256) void k_crt_fix(const CrtDesc *desc, uint64_t nprod, const uint32_t *status, uint64_t n)
{
    constexpr int CH = CRT_T * PER;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t b0 = (uint64_t)blockIdx.x * CRT_FIX_W;
    const uint32_t st = b0 + lane < n ? status[b0 + lane] : 0u;
    uint64_t todo = __ballot((st & CRT_ST_CIN) != 0);
    for (int r = 0; todo; r++) {
        const int f = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        if ((r & 3) != wave) continue;
        const uint64_t blk = b0 + f;
        uint64_t lo = 0, hi = nprod - 1;
        while (lo < hi) { const uint64_t mid = (lo + hi + 1) >> 1; if (desc[mid].first_block <= blk) lo = mid; else hi = mid - 1; }
        const CrtDesc d = desc[lo];
        const uint64_t k0 = (blk - d.first_block) * CH, end = k0 + CH < d.nc ? k0 + CH : d.nc;
        uint64_t v[CH / 64];
#pragma unroll
        for (int j = 0; j < CH / 64; j++) { const uint64_t k = k0 + j * 64 + lane; v[j] = k < end ? d.c[k] : 0; }
        int first_ne = CH;                                      /* offset of the first limb below B - 1 */
#pragma unroll
        for (int j = CH / 64 - 1; j >= 0; j--) {
            const uint64_t ne = __ballot(k0 + j * 64 + lane < end && v[j] != NTTM_B - 1);
            if (ne) first_ne = j * 64 + __ffsll((long long)ne) - 1;
        }
#pragma unroll
        for (int j = 0; j < CH / 64; j++) {
            const int o = j * 64 + lane;
            const uint64_t k = k0 + o;
            if (k < end && o <= first_ne) d.c[k] = o < first_ne ? 0 : v[j] + 1;
        }
    }
}

/* c[0 .. n) += *carry (0 or 1), the carry rippling through the leading limbs equal to B - 1 (one workgroup, 2048
 * limbs per step; a carry beyond limb n - 1 is dropped: the caller's bound makes it zero) */
__global__ __launch_bounds__(256) void k_add_carry(uint64_t *c, uint64_t n, const uint64_t *carry)
{
    __shared__ uint32_t s_first;
    const int tid = threadIdx.x;
    const uint64_t cin = *carry;
    if (!cin) return;
    for (uint64_t base = 0; base < n; base += 256 * 8) {
        if (tid == 0) s_first = 0xffffffffu;
        __syncthreads();
        uint64_t v[8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t off = (uint32_t)(i * 256 + tid);
            const uint64_t k = base + off;
            v[i] = k < n ? c[k] : 0;
            if (k < n && v[i] != NTTM_B - 1) atomicMin(&s_first, off);
        }
        __syncthreads();
        const uint32_t f = s_first;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t off = (uint32_t)(i * 256 + tid);
            const uint64_t k = base + off;
            if (k < n && off <= f) c[k] = off < f ? 0 : v[i] + cin;
        }
        if (f != 0xffffffffu) return;
        __syncthreads();
    }
}
int nttm_run_add_carry(uint64_t *c, uint64_t n, const uint64_t *carry, hipStream_t s)
{
    if (!n) return 0;
    hipLaunchKernelGGL(k_add_carry, dim3(1), dim3(256), 0, s, c, n, carry);
    return nttm_hip_check(hipGetLastError(), "k_add_carry launch");
}

__global__ void k_norm(const CrtDesc *d, uint64_t nprod, uint64_t *ncn)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nprod) return;
    uint64_t n = d[i].nc;
    while (n && d[i].c[n - 1] == 0) n--;
    ncn[i] = n;
}

/* blocks of a product of nc limbs at per limbs per thread (CRT_WIDE, or 1) */
uint64_t nttm_crt_blocks(uint64_t nc, int per) { const uint64_t ch = (uint64_t)CRT_T * (per == 1 ? 1 : CRT_WIDE); return (nc + ch - 1) / ch; }

uint64_t nttm_crt_scratch_words(uint64_t nblocks) { return nblocks + CRT_CTL_WORDS; }

/* desc_dev: nprod descriptors in device memory whose block ranges tile [0, nblocks); scratch: 16-byte aligned,
 * nttm_crt_scratch_words(nblocks) words, of which word 1 receives the error flags */
int nttm_crt_launch_batch(const void *desc_dev, uint64_t nprod, uint64_t nblocks, int np, const CrtConst &cc,
                          uint32_t *scratch, hipStream_t s, double alg_bytes, int per)
{
    if (per != 1 && per != CRT_WIDE) return nttm_fail(NTTM_EINVAL, "crt: %d limbs per thread", per);
    if (np < 3 || np > NTTM_MAX_PRIMES) return nttm_fail(NTTM_EINVAL, "crt: np = %d", np);
    if (!nprod || !nblocks) return 0;
    if (nblocks > 0x7fffffffULL) return nttm_fail(NTTM_EINVAL, "crt: too many blocks");
    if ((uintptr_t)scratch & 15) return nttm_fail(NTTM_EINVAL, "crt: scratch not 16-byte aligned");
    CrtArgs a;
    a.desc = (const CrtDesc *)desc_dev; a.nprod = nprod;
    a.ctl = scratch; a.status = scratch + CRT_CTL_WORDS; a.cc = cc;
    prof_begin("k_crt", alg_bytes, s);
    if (per == 1) {
        if (np == 3) hipLaunchKernelGGL((k_crt<3, 1>), dim3((unsigned)nblocks), dim3(CRT_T), 0, s, a);
        else hipLaunchKernelGGL((k_crt<4, 1>), dim3((unsigned)nblocks), dim3(CRT_T), 0, s, a);
    } else {
        if (np == 3) hipLaunchKernelGGL((k_crt<3, CRT_WIDE>), dim3((unsigned)nblocks), dim3(CRT_T), 0, s, a);
        else hipLaunchKernelGGL((k_crt<4, CRT_WIDE>), dim3((unsigned)nblocks), dim3(CRT_T), 0, s, a);
    }
    prof_end(s);
    int rc = nttm_hip_check(hipGetLastError(), "k_crt launch");
    if (rc) return rc;
    prof_begin("k_crt_scan", 8.0 * (double)nblocks, s);
    hipLaunchKernelGGL(k_crt_scan, dim3(1), dim3(CRT_SCAN_T), 0, s, a.status, nblocks, a.ctl);
    prof_end(s);
    rc = nttm_hip_check(hipGetLastError(), "k_crt_scan launch");
    if (rc) return rc;
    prof_begin("k_crt_fix", 4.0 * (double)nblocks, s);
    const dim3 gfix((unsigned)((nblocks + CRT_FIX_W - 1) / CRT_FIX_W));
    if (per == 1) hipLaunchKernelGGL((k_crt_fix<1>), gfix, dim3(256), 0, s, a.desc, nprod, (const uint32_t *)a.status, nblocks);
    else hipLaunchKernelGGL((k_crt_fix<CRT_WIDE>), gfix, dim3(256), 0, s, a.desc, nprod, (const uint32_t *)a.status, nblocks);
    prof_end(s);
    return nttm_hip_check(hipGetLastError(), "k_crt_fix launch");
}

/* ---- addition and subtraction of limb arrays, with the carry resolved like the CRT's -------------------------------
 * r[0 .. nout) = a B^sa + b B^sb, or a B^sa - b B^sb computed as a B^sa + (B^nout - 1 - b B^sb) + 1 (b's limbs
 * complemented, 1 added at limb 0): its carry out of limb nout - 1 is 1 exactly when the difference is >= 0.
 * b may instead be B^bpow (bpow >= 0).  Each 2048-limb block writes its limbs for a carry in of 0 and its one-bit
 * carry map, with the same status words as k_crt; k_crt_scan then gives every block its carry in (and, from the last
 * block's error bits, the final carry out) and k_addsub_fix adds the carries.  Limbs past nout count as B - 1 so that
 * the last block's map is the carry out of limb nout - 1. */
struct AddArgs {
    const uint64_t *a; uint64_t na, sa;
    const uint64_t *b; uint64_t nb, sb;
    int64_t apow, bpow;
    uint64_t *r; uint64_t nout, nblocks;
    int sub;
    uint32_t *status;
};
__device__ __forceinline__ uint64_t add_limb(const uint64_t *p, uint64_t n, uint64_t s, uint64_t k)
{
    return (k >= s && k - s < n) ? p[k - s] : 0;
}
__global__ __launch_bounds__(CRT_T) void k_addsub(AddArgs a)
{
    constexpr int PER = CRT_WIDE, CH = CRT_T * PER;
    __shared__ uint32_t s_wave[CRT_T / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t blk = blockIdx.x, k0 = blk * CH;
    uint64_t wv[PER], GX[PER], GY[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const uint64_t k = k0 + wave * 64 * PER + i * 64 + lane;
        uint64_t w = NTTM_B - 1;
        if (k < a.nout) {
            const uint64_t x = a.apow >= 0 ? (k == (uint64_t)a.apow ? 1u : 0u) : add_limb(a.a, a.na, a.sa, k);
            const uint64_t y = a.bpow >= 0 ? (k == (uint64_t)a.bpow ? 1u : 0u) : add_limb(a.b, a.nb, a.sb, k);
            w = x + (a.sub ? NTTM_B - 1 - y : y) + (a.sub && k == 0 ? 1u : 0u);          /* <= 2B - 1 */
        }
        wv[i] = w;
        const uint64_t g = __ballot(w >= NTTM_B), pr = __ballot(w == NTTM_B - 1);
        GX[i] = g; GY[i] = g | pr;
    }
    uint32_t F = CRT_ID_MAP;
#pragma unroll
    for (int i = 0; i < PER; i++) F = map_compose(add_cout(GX[i], GY[i], 0) | (add_cout(GX[i], GY[i], 1) << 1), F);
    if (lane == 0) s_wave[wave] = F;
    __syncthreads();
    uint32_t P = CRT_ID_MAP, A = CRT_ID_MAP;
    for (int w = 0; w < CRT_T / 64; w++) {
        if (w < wave) P = map_compose(s_wave[w], P);
        A = map_compose(s_wave[w], A);
    }
    uint32_t c = map_apply(P, 0);
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const uint64_t S = GX[i] + GY[i] + c;
        const uint32_t cout = add_cout(GX[i], GY[i], c);
        const uint64_t CI = S ^ GX[i] ^ GY[i];
        const uint32_t oin = (uint32_t)(CI >> lane) & 1u;
        const uint32_t oout = lane == 63 ? cout : (uint32_t)(CI >> (lane + 1)) & 1u;
        const uint64_t k = k0 + wave * 64 * PER + i * 64 + lane;
        if (k < a.nout) a.r[k] = wv[i] + oin - (oout ? NTTM_B : 0);
        c = cout;
    }
    if (tid == 0) {
        uint32_t st = A | (blk == 0 ? CRT_ST_FIRST : 0u);
        if (blk + 1 == a.nblocks) st |= ((A & 1u) ? CRT_ST_E0 : 0u) | ((A & 2u) ? CRT_ST_E1 : 0u);
        a.status[blk] = st;
    }
}
/* the blocks whose carry in is 1: their limbs plus one (a ripple through the leading limbs equal to B - 1) */
__global__ __launch_bounds__(64) void k_addsub_fix(uint64_t *r, uint64_t nout, const uint32_t *status, uint64_t nblocks)
{
    constexpr int CH = CRT_T * CRT_WIDE;
    const int lane = threadIdx.x;
    const uint64_t blk = blockIdx.x;
    if (blk >= nblocks || !(status[blk] & CRT_ST_CIN)) return;
    const uint64_t k0 = blk * CH, end = k0 + CH < nout ? k0 + CH : nout;
    for (uint64_t base = k0; base < end; base += 64) {
        const uint64_t k = base + lane;
        const uint64_t v = k < end ? r[k] : 0;
        const uint64_t ne = __ballot(k < end && v != NTTM_B - 1);
        const int f = ne ? __ffsll((long long)ne) - 1 : 64;                    /* first limb below B - 1 */
        if (k < end && lane <= f) r[k] = lane < f ? 0 : v + 1;
        if (ne) return;
    }
}
int nttm_run_addsub(uint64_t *r, uint64_t nout, const uint64_t *a, uint64_t na, uint64_t sa, int64_t apow, const uint64_t *b,
                    uint64_t nb, uint64_t sb, int64_t bpow, int sub, uint32_t *scratch, hipStream_t s)
{
    if (!nout) return 0;
    const uint64_t nblocks = nttm_crt_blocks(nout, CRT_WIDE);
    if (nblocks > 0x7fffffffULL) return nttm_fail(NTTM_EINVAL, "addsub: too many blocks");
    if ((uintptr_t)scratch & 15) return nttm_fail(NTTM_EINVAL, "addsub: scratch not 16-byte aligned");
    AddArgs A;
    A.a = a; A.na = a ? na : 0; A.sa = sa; A.apow = apow; A.b = b; A.nb = b ? nb : 0; A.sb = sb; A.bpow = bpow;
    A.r = r; A.nout = nout; A.nblocks = nblocks; A.sub = sub; A.status = scratch + CRT_CTL_WORDS;
    hipLaunchKernelGGL(k_addsub, dim3((unsigned)nblocks), dim3(CRT_T), 0, s, A);
    int rc = nttm_hip_check(hipGetLastError(), "k_addsub launch");
    if (rc) return rc;
    hipLaunchKernelGGL(k_crt_scan, dim3(1), dim3(CRT_SCAN_T), 0, s, A.status, nblocks, scratch);
    if ((rc = nttm_hip_check(hipGetLastError(), "k_crt_scan launch"))) return rc;
    hipLaunchKernelGGL(k_addsub_fix, dim3((unsigned)nblocks), dim3(64), 0, s, r, nout, (const uint32_t *)A.status, nblocks);
    return nttm_hip_check(hipGetLastError(), "k_addsub_fix launch");
}

/* the normalized length of x[0 .. n): one more than the index of the highest nonzero limb */
/* one more than the index of the top non-zero limb of x[0 .. n) (0 if none), by atomicMax into *out (cleared first).
 * Chunk c covers [n - (c + 1) C, n - c C): the chunks are taken from the top down, block b taking b, b + G, b + 2G, ...
 * (G blocks, at most one per CU), and a block stops as soon as *out reaches the top of its next chunk (nothing below
 * can raise the maximum).  A number with a few leading zero limbs costs one chunk per block; one with many streams
 * its zero top at full bandwidth until the first non-zero chunk. */
#define TOPNZ_PER 16
#define TOPNZ_CHUNK (256 * TOPNZ_PER)
#define TOPNZ_GRID 256
__global__ __launch_bounds__(256) void k_top_nonzero(const uint64_t *x, uint64_t n, unsigned long long *out)
{
    const uint64_t nchunk = (n + TOPNZ_CHUNK - 1) / TOPNZ_CHUNK;
    for (uint64_t c = blockIdx.x; c < nchunk; c += gridDim.x) {
        const uint64_t hi = n - c * TOPNZ_CHUNK, lo = hi > TOPNZ_CHUNK ? hi - TOPNZ_CHUNK : 0;
        if (__hip_atomic_load(out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= hi) return;
        unsigned long long best = 0;
#pragma unroll
        for (int j = 0; j < TOPNZ_PER; j++) {
            const uint64_t k = lo + (uint64_t)j * 256 + threadIdx.x;
            if (k < hi && x[k]) best = k + 1;
        }
        for (int d = 32; d; d >>= 1) {
            const unsigned long long o = __shfl_down(best, d, 64);
            best = o > best ? o : best;
        }
        if ((threadIdx.x & 63) == 0 && best) atomicMax(out, best);
    }
}
int nttm_run_top_nonzero(const uint64_t *x, uint64_t n, uint64_t *out_dev, hipStream_t s)
{
    int rc = nttm_hip_check(hipMemsetAsync(out_dev, 0, 8, s), "clear length");
    if (rc || !n) return rc;
    const uint64_t chunks = (n + TOPNZ_CHUNK - 1) / TOPNZ_CHUNK;
    const unsigned blocks = (unsigned)(chunks < TOPNZ_GRID ? chunks : TOPNZ_GRID);
    hipLaunchKernelGGL(k_top_nonzero, dim3(blocks), dim3(256), 0, s, x, n, (unsigned long long *)out_dev);
    return nttm_hip_check(hipGetLastError(), "k_top_nonzero launch");
}
/* the normalized length of x[0 .. n) into *len_dev (device pointers), for tests */
extern "C" int nttm_debug_top_nonzero(const uint64_t *x, size_t n, uint64_t *len_dev, void *stream)
{
    if (!len_dev || (n && !x)) return nttm_fail(NTTM_EINVAL, "nttm_debug_top_nonzero arguments");
    return nttm_run_top_nonzero(x, n, len_dev, (hipStream_t)stream);
}

/* products with an empty operand: c[0 .. nc) = x zero-extended (a kernel, not a memset, so a captured graph holds it) */
__global__ __launch_bounds__(256) void k_fill_x(const CrtDesc *d, uint64_t nprod)
{
    const CrtDesc q = d[blockIdx.x];
    for (uint64_t k = threadIdx.x; k < q.nc; k += 256) q.c[k] = k < q.nx ? q.x[k] : 0;
}
int nttm_run_fill_x(const void *desc_dev, uint64_t nprod, hipStream_t s)
{
    if (!nprod) return 0;
    if (nprod > 0x7fffffffULL) return nttm_fail(NTTM_EINVAL, "fill: too many products");
    hipLaunchKernelGGL(k_fill_x, dim3((unsigned)nprod), dim3(256), 0, s, (const CrtDesc *)desc_dev, nprod);
    return nttm_hip_check(hipGetLastError(), "k_fill_x launch");
}

int nttm_run_norm(const void *desc_dev, uint64_t nprod, uint64_t *ncn_dev, hipStream_t s)
{
    if (!nprod) return 0;
    hipLaunchKernelGGL(k_norm, dim3((unsigned)((nprod + 255) / 256)), dim3(256), 0, s, (const CrtDesc *)desc_dev, nprod, ncn_dev);
    return nttm_hip_check(hipGetLastError(), "k_norm launch");
}
