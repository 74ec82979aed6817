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
 * X = G, Y = G | P (the carries into the bits of X + Y + cin are the o's, the textbook adder identity); words chain
 * inside a wave, waves inside a block, and blocks by a decoupled look-back over (cin -> cout) maps of one bit
 * (blocks take consecutive tickets and publish their map before looking back, so every wait is on a running block).
 * The reference instead ripples each 256-coefficient chunk on one thread of a wave64 ballot scan, writes a 4-limb spill
 * per stripe and merges the spills serially on the host or on one thread per product (E/rns_mul.c:452-470;
 * E/dbig.c:1007-1052).  The final carry beyond limb nc-1 must be zero; a non-zero one sets an error flag.
 */
#include "internal.h"

#define CRT_T 256
#define CRT_PER 8
#define CRT_CH (CRT_T * CRT_PER)
#define CRT_HALO 4                       /* digits of coefficients k0-4 .. k0-1: v_{k0-1} with four primes */
#define CRT_ID_MAP 0x2u                  /* one-bit carry maps: bit c = cout(cin = c); the identity is 0b10 */
#define CRT_ST_AGG 1u
#define CRT_ST_INC 2u
#define CRT_SPIN_MAX (1u << 22)

struct CrtDesc {                         /* one product; must match api.hip */
    const double *pl; uint64_t plane;    /* residues of prime j, coefficient k at pl[j * plane + k] */
    uint64_t ncoef;
    const uint64_t *x; uint64_t nx;
    uint64_t *c; uint64_t nc;
    uint64_t first_block;                /* the product's first block in the launch */
};
struct CrtArgs {
    const CrtDesc *desc; uint64_t nprod;
    uint32_t *status;                    /* one word per block: state << 30 | carry out << 2 | aggregate map (2 bits) */
    uint32_t *ctl;                       /* [0] ticket counter, [1] error flags */
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

template <int NP>
__global__ __launch_bounds__(CRT_T) void k_crt(CrtArgs a)
{
    __shared__ uint64_t D[NP][CRT_CH + CRT_HALO];
    __shared__ uint32_t s_blk, s_cin, s_prod;
    __shared__ uint32_t s_wave[CRT_T / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        const uint32_t b = atomicAdd(&a.ctl[0], 1u);
        uint64_t lo = 0, hi = a.nprod - 1;                  /* the product whose block range holds ticket b */
        while (lo < hi) { const uint64_t mid = (lo + hi + 1) >> 1; if (a.desc[mid].first_block <= b) lo = mid; else hi = mid - 1; }
        s_blk = b; s_prod = (uint32_t)lo;
    }
    __syncthreads();
    const uint32_t blk = s_blk;
    const CrtDesc dsc = a.desc[s_prod];
    const uint32_t first = (uint32_t)dsc.first_block, local = blk - first;
    const uint64_t k0 = (uint64_t)local * CRT_CH;

    /* digits of coefficients k0 - HALO .. k0 + CH - 1 into LDS (coalesced plane reads, all of a thread's residues
     * requested before the first is used) */
    uint64_t xk[CRT_PER];
    {
        const bool halo = tid < CRT_HALO && k0 + tid >= CRT_HALO;   /* coefficient k0 - HALO + tid exists */
        /* the planes as global-space pointers: generic (flat) loads would also count against the LDS counter, whose
         * limit of 15 outstanding operations would serialize the prefetch */
        const __attribute__((address_space(1))) double *pl = (const __attribute__((address_space(1))) double *)dsc.pl;
        const __attribute__((address_space(1))) uint64_t *xg = (const __attribute__((address_space(1))) uint64_t *)dsc.x;
        double rr[CRT_PER + 1][NP];
#pragma unroll
        for (int i = 0; i < CRT_PER; i++) {
            const uint64_t kk = k0 + i * CRT_T + tid;
#pragma unroll
            for (int j = 0; j < NP; j++) rr[i][j] = kk < dsc.ncoef ? pl[j * dsc.plane + kk] : 0.0;
        }
        {
            const uint64_t kk = k0 + tid - CRT_HALO;
#pragma unroll
            for (int j = 0; j < NP; j++) rr[CRT_PER][j] = halo && kk < dsc.ncoef ? pl[j * dsc.plane + kk] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < CRT_PER; i++) {                          /* x at the limbs of the limb phase below */
            const uint64_t k = k0 + wave * 64 * CRT_PER + i * 64 + lane;
            xk[i] = k < dsc.nx && k < dsc.nc ? xg[k] : 0;
        }
#pragma unroll
        for (int i = 0; i < CRT_PER; i++) {
            uint64_t d[NP];
            crt_digits<NP>(a.cc, rr[i], d);                          /* zero residues give zero digits */
#pragma unroll
            for (int j = 0; j < NP; j++) D[j][CRT_HALO + i * CRT_T + tid] = d[j];
        }
        if (tid < CRT_HALO) {
            uint64_t d[NP];
            crt_digits<NP>(a.cc, rr[CRT_PER], d);
#pragma unroll
            for (int j = 0; j < NP; j++) D[j][tid] = d[j];
        }
    }
    __syncthreads();

    /* limb k = k0 + wave 512 + i 64 + lane for i < 8: one ballot per i gives 64 consecutive limbs */
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
    uint64_t wv[CRT_PER];
    uint64_t GX[CRT_PER], GY[CRT_PER];
    uint32_t qlast = 0, qprev = 0;
#pragma unroll
    for (int i = 0; i < CRT_PER; i++) {
        const int off = wave * 64 * CRT_PER + i * 64 + lane;          /* offset of limb k inside the block */
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
    /* the wave's map cin -> cout over its 8 words */
    uint32_t F = CRT_ID_MAP;
#pragma unroll
    for (int i = 0; i < CRT_PER; i++) {
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

    /* carry into the block: decoupled look-back over the earlier blocks of the product, 64 at a time by wave 0.  The
     * status word carries its whole payload (map and carry), so the atomics are relaxed at agent scope: no cache
     * write-back or invalidate is needed around them.  Lane i
     * reads block blk-1-i; the nearest block with an inclusive carry ends the walk, the aggregate maps of the blocks
     * in between are composed (nearest outermost).  Blocks before the product's first act as an inclusive carry 0. */
    if (wave == 0) {
        if (lane == 0)
            __hip_atomic_store(&a.status[blk], local == 0 ? ((CRT_ST_INC << 30) | (map_apply(A, 0) << 2) | A)
                                                          : ((CRT_ST_AGG << 30) | A),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t cin = 0;
        if (local != 0) {
            uint32_t G = CRT_ID_MAP, spins = 0;
            int64_t base = (int64_t)blk - 1;
            for (;;) {
                const int64_t idx = base - lane;
                const uint32_t st = idx >= (int64_t)first
                    ? __hip_atomic_load(&a.status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                    : ((CRT_ST_INC << 30) | CRT_ID_MAP);
                const uint32_t state = st >> 30;
                const uint64_t inc = __ballot(state == CRT_ST_INC), wait = __ballot(state == 0);
                const int fi = inc ? __ffsll((long long)inc) - 1 : 64;
                const int fw = wait ? __ffsll((long long)wait) - 1 : 64;
                if (fw < fi) {                               /* a block between us and the nearest inclusive one */
                    if (++spins > CRT_SPIN_MAX) { if (lane == 0) atomicOr(&a.ctl[1], 2u); cin = 0; break; }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                uint32_t m = lane < fi ? (st & 3u) : CRT_ID_MAP;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {           /* m_0 o m_1 o ... o m_63 into lane 0 */
                    const uint32_t o = __shfl_down(m, d, 64);
                    if (lane + d < 64) m = map_compose(m, o);
                }
                m = __shfl(m, 0, 64);
                G = map_compose(G, m);
                if (fi < 64) {
                    const uint32_t c_inc = (__shfl(st, fi, 64) >> 2) & 1u;
                    cin = map_apply(G, c_inc);
                    break;
                }
                base -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&a.status[blk], (CRT_ST_INC << 30) | (map_apply(A, cin) << 2) | A, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) s_cin = cin;
    }
    __syncthreads();

    /* final limbs: out_k = w_k + o_{k-1} - B o_k, stored directly (lanes hold consecutive limbs) */
    __attribute__((address_space(1))) uint64_t *cg = (__attribute__((address_space(1))) uint64_t *)dsc.c;
    uint32_t c = map_apply(P, s_cin);                               /* the carry into this wave's first word */
#pragma unroll
    for (int i = 0; i < CRT_PER; i++) {
        const uint64_t S = GX[i] + GY[i] + c;
        const uint32_t cout = add_cout(GX[i], GY[i], c);
        const uint64_t CI = S ^ GX[i] ^ GY[i];                      /* bit b: the carry into bit b (o_{k-1}) */
        const uint32_t oin = (uint32_t)(CI >> lane) & 1u;
        const uint32_t oout = lane == 63 ? cout : (uint32_t)(CI >> (lane + 1)) & 1u;
        const int off = wave * 64 * CRT_PER + i * 64 + lane;
        const uint64_t k = k0 + off;
        if (k < dsc.nc) cg[k] = wv[i] + oin - (oout ? NTTM_B : 0);
        if (k + 1 == dsc.nc) {                                     /* the last limb: nothing may remain beyond it */
            uint64_t extra = qlast + oout;
            const int li = CRT_HALO + off;
#pragma unroll
            for (int e = 1; e < NP; e++)
#pragma unroll
                for (int j = e; j < NP; j++) extra += D[j][li + e - j];
            if (extra) atomicOr(&a.ctl[1], 1u);
        }
        c = cout;
    }
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_norm(const CrtDesc *d, uint64_t nprod, uint64_t *ncn)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nprod) return;
    uint64_t n = d[i].nc;
    while (n && d[i].c[n - 1] == 0) n--;
    ncn[i] = n;
}

uint64_t nttm_crt_blocks(uint64_t nc) { return (nc + CRT_CH - 1) / CRT_CH; }

/* desc_dev: nprod descriptors in device memory whose block ranges tile [0, nblocks); scratch: nblocks + 2 words */
int nttm_crt_launch_batch(const void *desc_dev, uint64_t nprod, uint64_t nblocks, int np, const CrtConst &cc,
                          uint32_t *scratch, hipStream_t s, double alg_bytes)
{
    if (np < 3 || np > NTTM_MAX_PRIMES) return nttm_fail(NTTM_EINVAL, "crt: np = %d", np);
    if (!nprod || !nblocks) return 0;
    if (nblocks > 0x7fffffffULL) return nttm_fail(NTTM_EINVAL, "crt: too many blocks");
    CrtArgs a;
    a.desc = (const CrtDesc *)desc_dev; a.nprod = nprod;
    a.ctl = scratch; a.status = scratch + 2; a.cc = cc;
    const uint64_t fill = nblocks + 2;
    const unsigned fg = (unsigned)((fill + 255) / 256 < 4096 ? (fill + 255) / 256 : 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(fg), dim3(256), 0, s, scratch, fill, 0u);
    prof_begin("k_crt", alg_bytes, s);
    if (np == 3) hipLaunchKernelGGL(k_crt<3>, dim3((unsigned)nblocks), dim3(CRT_T), 0, s, a);
    else hipLaunchKernelGGL(k_crt<4>, dim3((unsigned)nblocks), dim3(CRT_T), 0, s, a);
    prof_end(s);
    return nttm_hip_check(hipGetLastError(), "k_crt launch");
}

int nttm_run_norm(const void *desc_dev, uint64_t nprod, uint64_t *ncn_dev, hipStream_t s)
{
    if (!nprod) return 0;
    hipLaunchKernelGGL(k_norm, dim3((unsigned)((nprod + 255) / 256)), dim3(256), 0, s, (const CrtDesc *)desc_dev, nprod, ncn_dev);
    return nttm_hip_check(hipGetLastError(), "k_norm launch");
}
