/* crt.hip - Chinese remaindering of the residue planes into base-10^18 limbs, with the carry resolved on the device.
 *
 * Coefficient k of the convolution is rebuilt from its np residues by Garner's mixed radix (modular steps in the
 * FP64 arithmetic of arith.h, the integer value in np words), split into np base-10^18 digits d_0..d_{np-1}, and
 *     limb_k = d_0[k] + d_1[k-1] + ... + d_{np-1}[k-np+1] + x_k + carry_k            (E/rns_mul.c:332-405)
 * The top digit is small (X < 2^(50 np) gives d_{np-1} < 2^31), so v_k < np B + 2^31 <= 4 B + 2^31, and the carry into
 * the next limb, f_k(carry_k) = floor((v_k + carry_k) / B), maps {0..7} into {0..4}: a map fits 8 x 3 bits.  Maps
 * compose, so the carries are a parallel scan:
 * inside a block by wave shuffles and LDS, across blocks by a decoupled look-back (blocks take consecutive tickets, a
 * block publishes its aggregate map before looking back, so every wait is on a block that is already running).
 * The reference instead ripples each 256-coefficient chunk on one thread of a wave64 ballot scan, writes a 4-limb spill
 * per stripe and merges the spills serially on the host or on one thread per product (E/rns_mul.c:452-470;
 * E/dbig.c:1007-1052).  The final carry beyond limb nc-1 must be zero; a non-zero one sets an error flag.
 */
#include "internal.h"

#define CRT_T 256
#define CRT_PER 8
#define CRT_CH (CRT_T * CRT_PER)
#define CRT_HALO 3
#define CRT_ID_MAP 0xFAC688u             /* the identity map c -> c, 3 bits per c (c = 0..7) */
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
    uint32_t *status;                    /* one word per block: state << 30 | carry << 24 | aggregate map */
    uint32_t *ctl;                       /* [0] ticket counter, [1] error flags */
    CrtConst cc;
};

__device__ __forceinline__ uint32_t map_apply(uint32_t f, uint32_t c) { return (f >> (3 * c)) & 7u; }
/* (g o f)(c) = g(f(c)) */
__device__ __forceinline__ uint32_t map_compose(uint32_t g, uint32_t f)
{
    uint32_t h = 0;
#pragma unroll
    for (uint32_t c = 0; c < 8; c++) h |= map_apply(g, (f >> (3 * c)) & 7u) << (3 * c);
    return h;
}

template <int NP>
__device__ __forceinline__ void crt_digits(const CrtConst &cc, const double *r, uint64_t *d)
{
    uint64_t t[NP];
    const double p0 = cc.pk[0].p, pinv0 = cc.pk[0].pinv;
    { double v = reduce(r[0], p0, pinv0); v = v < 0.0 ? v + p0 : v; t[0] = (uint64_t)v; }
#pragma unroll
    for (int i = 1; i < NP; i++) {
        const double p = cc.pk[i].p, pinv = cc.pk[i].pinv;
        double acc = reduce((double)t[0], p, pinv);
#pragma unroll
        for (int j = 1; j < i; j++)
            acc = reduce(acc + mul_barrett(reduce((double)t[j], p, pinv), cc.mmod[i][j], p, pinv), p, pinv);
        double dd = reduce(r[i] - acc, p, pinv);
        double ti = mul_barrett(dd, cc.cinv[i], p, pinv);
        ti = ti < 0.0 ? ti + p : ti;
        t[i] = (uint64_t)ti;
    }
    uint64_t X[NP];
#pragma unroll
    for (int w = 0; w < NP; w++) X[w] = 0;
    X[0] = t[0];
#pragma unroll
    for (int i = 1; i < NP; i++) {                       /* X += t_i * (p_0 ... p_{i-1}) */
        uint64_t carry = 0;
#pragma unroll
        for (int w = 0; w < NP; w++) {
            const uint64_t mw = w < i ? cc.mprod[i][w] : 0;
            const uint64_t lo = t[i] * mw, hi = umulhi64(t[i], mw);
            uint64_t s = X[w] + lo; uint64_t c1 = s < lo;
            uint64_t s2 = s + carry; c1 += s2 < s;
            X[w] = s2; carry = hi + c1;
        }
    }
#pragma unroll
    for (int k = 0; k < NP; k++) {                       /* long division by B, least significant digit first */
        uint64_t rem = 0;
#pragma unroll
        for (int w = NP - 1; w >= 0; w--) { uint64_t q; div_b(rem, X[w], &q, &rem); X[w] = q; }
        d[k] = rem;
    }
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

    /* digits of coefficients k0 - HALO .. k0 + CH - 1 into LDS (coalesced plane reads) */
    for (int i = -1; i < CRT_PER; i++) {
        int li; uint64_t kk;
        if (i < 0) { if (tid >= CRT_HALO) continue; li = tid; kk = k0 + tid - CRT_HALO; if (k0 < CRT_HALO - tid) { for (int j = 0; j < NP; j++) D[j][li] = 0; continue; } }
        else { li = CRT_HALO + i * CRT_T + tid; kk = k0 + i * CRT_T + tid; }
        uint64_t d[NP];
        if (kk < dsc.ncoef) {
            double r[NP];
#pragma unroll
            for (int j = 0; j < NP; j++) r[j] = dsc.pl[j * dsc.plane + kk];
            crt_digits<NP>(a.cc, r, d);
        } else {
#pragma unroll
            for (int j = 0; j < NP; j++) d[j] = 0;
        }
#pragma unroll
        for (int j = 0; j < NP; j++) D[j][li] = d[j];
    }
    __syncthreads();

    /* per thread: 8 consecutive limb positions, their sums and carry maps */
    uint64_t rr[CRT_PER]; uint32_t qq[CRT_PER];
    uint32_t F = CRT_ID_MAP;
    const uint64_t kb = k0 + (uint64_t)tid * CRT_PER;
#pragma unroll
    for (int i = 0; i < CRT_PER; i++) {
        const uint64_t k = kb + i;
        const int li = CRT_HALO + tid * CRT_PER + i;
        uint64_t v = 0;
        if (k < dsc.nc) {
#pragma unroll
            for (int j = 0; j < NP; j++) v += D[j][li - j];
            if (k < dsc.nx) v += dsc.x[k];
        }
        const uint32_t q = (v >= NTTM_B) + (v >= 2 * NTTM_B) + (v >= 3 * NTTM_B) + (v >= 4 * NTTM_B);
        const uint64_t r = v - (uint64_t)q * NTTM_B;
        rr[i] = r; qq[i] = q;
        uint32_t f = 0;
#pragma unroll
        for (uint32_t c = 0; c < 8; c++) f |= (q + (r + c >= NTTM_B ? 1u : 0u)) << (3 * c);
        F = map_compose(f, F);
    }

    /* block scan of the maps: inclusive within the wave, then across the waves */
    uint32_t I = F;
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
        uint32_t o = __shfl_up(I, dd, 64);
        if (lane >= dd) I = map_compose(I, o);
    }
    uint32_t E = __shfl_up(I, 1, 64);
    if (lane == 0) E = CRT_ID_MAP;
    if (lane == 63) s_wave[wave] = I;
    __syncthreads();
    uint32_t P = CRT_ID_MAP, A = CRT_ID_MAP;
    for (int w = 0; w < CRT_T / 64; w++) {
        if (w < wave) P = map_compose(s_wave[w], P);
        A = map_compose(s_wave[w], A);
    }
    E = map_compose(E, P);                                  /* maps of every earlier position of the block */

    /* carry into the block: decoupled look-back over the earlier blocks */
    if (tid == 0) {
        uint32_t cin = 0;
        if (local == 0) {
            __hip_atomic_store(&a.status[blk], (CRT_ST_INC << 30) | (map_apply(A, 0) << 24) | A, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&a.status[blk], (CRT_ST_AGG << 30) | A, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t G = CRT_ID_MAP, spins = 0;
            int64_t i = (int64_t)blk - 1;
            while (i >= (int64_t)first) {
                const uint32_t st = __hip_atomic_load(&a.status[i], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t state = st >> 30;
                if (state == CRT_ST_INC) { cin = map_apply(G, (st >> 24) & 7u); break; }
                if (state == CRT_ST_AGG) { G = map_compose(G, st & 0xffffffu); i--; continue; }
                if (++spins > CRT_SPIN_MAX) { atomicOr(&a.ctl[1], 2u); cin = 0; break; }
                __builtin_amdgcn_s_sleep(1);
            }
            __hip_atomic_store(&a.status[blk], (CRT_ST_INC << 30) | (map_apply(A, cin) << 24) | A, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        s_cin = cin;
    }
    __syncthreads();

    /* final limbs into LDS (plane 0's digit array is no longer read), then a coalesced store */
    uint32_t c = map_apply(E, s_cin);
    uint64_t *out = &D[0][0];
#pragma unroll
    for (int i = 0; i < CRT_PER; i++) {
        const uint64_t s = rr[i] + c;
        const uint32_t ge = s >= NTTM_B;
        out[tid * CRT_PER + i] = ge ? s - NTTM_B : s;
        c = qq[i] + ge;
        if (kb + i + 1 == dsc.nc) {                           /* the last limb: nothing may remain beyond it */
            uint64_t extra = c;
            const int li = CRT_HALO + tid * CRT_PER + i;
#pragma unroll
            for (int e = 1; e < NP; e++)
#pragma unroll
                for (int j = e; j < NP; j++) extra += D[j][li + e - j];
            if (extra) atomicOr(&a.ctl[1], 1u);
        }
    }
    __syncthreads();
    for (int i = 0; i < CRT_PER; i++) {
        const uint64_t kk = k0 + (uint64_t)i * CRT_T + tid;
        if (kk < dsc.nc) dsc.c[kk] = out[i * CRT_T + tid];
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
