/* mfma.hip - the split-precision radix-16 stage on INT8 matrix cores (v_mfma_i32_16x16x64_i8), exact.
 *
 * The first four levels of a 2^logn DIF transform are one radix-16 stage: for each j0 < n/16 the 16 points
 * x[j0 + t n/16] give Y_k = (sum_t W^(t k) x[j0 + t n/16]) w^(j0 k), written to position bitrev4(k) n/16 + j0 (W the
 * 16th root of unity, w the n-th; the rest of the transform is 16 independent DIF transforms of n/16 points).
 * The 16-point DFT runs on the matrix cores.  A residue (< 2^50) is 8 digits of 7 bits, d_b = (x >> 7 b) & 127, and so
 * is each W^(tk) mod p (digits e_c).  With kappa = 8 t + b (128 values) and the 16 groups j0 of one wave as the columns:
 *     T_s[k][g] = sum_{t, b} E_s[k][(t, b)] D[(t, b)][g],   E_s[k][(t, b)] = e_{s - b}(W^(t k)) (0 outside 0..7),
 *     Y_k = sum_{s=0..14} T_s 2^(7 s)  (mod p),
 * one 16 x 128 by 128 x 16 product per shift s: two MFMA of K = 64 each, 30 per 256 points.  T_s < 128 127^2 < 2^21,
 * exact in int32.  The recombination takes U_i = sum of five T_s 2^(7 (s - 5 i)) (< 2^50, exact in FP64) and
 * Y = U_0 + U_1 2^35 + U_2 2^70 with two Shoup products; the twiddle w^(j0 k) follows by Barrett products from a
 * two-level table of w^j0.
 * Operand layout (checked on the device): lane l holds A row l % 16 and B column l % 16 at K positions 16 (l / 16) + 0..15
 * (byte s of the lane's 16), and C rows 4 (l / 16) + 0..3 of column l % 16; A and B share the K order, so any
 * bijection of kappa onto (l / 16, byte) works: here kappa = 16 (l / 16) + byte, i.e. t = 2 (l / 16) + byte / 8 (+ 8 in
 * the second K half), b = byte % 8. */
#include "internal.h"

typedef int v4i __attribute__((ext_vector_type(4)));

struct Dft16Args {
    const uint64_t *x;           /* canonical residues, batch transforms of n points */
    double *y;                   /* output (balanced residues), same layout */
    const v4i *etab;             /* [15 shifts][2 halves][64 lanes] A operands */
    const double *tlo, *thi;     /* w^i, w^(4096 i), balanced */
    uint64_t n, units;           /* points per transform; units of 256 points (16 groups) in all */
    int logn;
    double p, pinv;
    double c35, c35q, c70, c70q; /* 2^35, 2^70 mod p (balanced) and their Shoup factors */
};

__device__ __forceinline__ v4i digits2(uint64_t a, uint64_t b)
{
    v4i r;
    r.x = (int)((a & 127) | ((a >> 7) & 127) << 8 | ((a >> 14) & 127) << 16 | ((a >> 21) & 127) << 24);
    r.y = (int)(((a >> 28) & 127) | ((a >> 35) & 127) << 8 | ((a >> 42) & 127) << 16 | ((a >> 49) & 127) << 24);
    r.z = (int)((b & 127) | ((b >> 7) & 127) << 8 | ((b >> 14) & 127) << 16 | ((b >> 21) & 127) << 24);
    r.w = (int)(((b >> 28) & 127) | ((b >> 35) & 127) << 8 | ((b >> 42) & 127) << 16 | ((b >> 49) & 127) << 24);
    return r;
}

__global__ __launch_bounds__(256) void k_dft16_mfma(Dft16Args a)
{
    __shared__ v4i s_e[15 * 2 * 64];                               /* 30 KiB: every shift's A operand */
    for (int i = threadIdx.x; i < 15 * 2 * 64; i += 256) s_e[i] = a.etab[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, col = lane & 15, q = lane >> 4;
    const uint64_t n16 = a.n >> 4, units_per = n16 >> 4;
    const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t u = wave0; u < a.units; u += nwaves) {
        const uint64_t tb = u / units_per, j0 = ((u % units_per) << 4) + col;
        const uint64_t *xb = a.x + tb * a.n + j0;
        const uint64_t t0 = 2 * q;
        const v4i b0 = digits2(xb[t0 * n16], xb[(t0 + 1) * n16]);
        const v4i b1 = digits2(xb[(t0 + 8) * n16], xb[(t0 + 9) * n16]);
        double U[3][4];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) U[i][r] = 0.0;
#pragma unroll
        for (int s = 0; s < 15; s++) {
            v4i c = {0, 0, 0, 0};
            c = __builtin_amdgcn_mfma_i32_16x16x64_i8(s_e[(s * 2 + 0) * 64 + lane], b0, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_i32_16x16x64_i8(s_e[(s * 2 + 1) * 64 + lane], b1, c, 0, 0, 0);
            const double sc = (double)(1u << (7 * (s % 5)));          /* exact: U_i < 2^50 */
            U[s / 5][0] = nttm_fma((double)c.x, sc, U[s / 5][0]);
            U[s / 5][1] = nttm_fma((double)c.y, sc, U[s / 5][1]);
            U[s / 5][2] = nttm_fma((double)c.z, sc, U[s / 5][2]);
            U[s / 5][3] = nttm_fma((double)c.w, sc, U[s / 5][3]);
        }
        /* w^(j0 k) for this lane's k = 4 q + r: e = w^j0, then e^(4 q) and successive products by e */
        const uint64_t jl = j0 & 4095, jh = j0 >> 12;
        const double e = jh ? mul_barrett(a.thi[jh], a.tlo[jl], a.p, a.pinv) : a.tlo[jl];
        double e4 = mul_barrett(e, e, a.p, a.pinv);
        e4 = mul_barrett(e4, e4, a.p, a.pinv);
        double tw = 1.0;
        for (int i = 0; i < q; i++) tw = mul_barrett(tw, e4, a.p, a.pinv);
        double *yb = a.y + tb * a.n + j0;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const double r1 = mul_shoup(U[1][r], a.c35, a.c35q, a.p), r2 = mul_shoup(U[2][r], a.c70, a.c70q, a.p);
            const double v = reduce(U[0][r] + r1 + r2, a.p, a.pinv);
            const int k = 4 * q + r;
            const int kb = ((k & 1) << 3) | ((k & 2) << 1) | ((k & 4) >> 1) | ((k & 8) >> 3);   /* bitrev4(k) */
            yb[(uint64_t)kb * n16] = mul_barrett(v, tw, a.p, a.pinv);
            tw = mul_barrett(tw, e, a.p, a.pinv);
        }
    }
}

/* etab, tlo, thi: device memory; consts: HOST memory (2^35 and 2^70 mod p with their Shoup factors) */
int nttm_run_dft16_mfma(const uint64_t *x, double *y, int logn, uint64_t batch, const void *etab, const double *tlo,
                        const double *thi, const double *consts, double p, hipStream_t s)
{
    if (logn < 8) return nttm_fail(NTTM_EINVAL, "mfma stage: 2^%d points (at least 2^8)", logn);
    Dft16Args a;
    a.x = x; a.y = y; a.etab = (const v4i *)etab; a.tlo = tlo; a.thi = thi;
    a.n = (uint64_t)1 << logn; a.logn = logn; a.units = batch * (a.n >> 8);
    a.p = p; a.pinv = 1.0 / p;
    a.c35 = consts[0]; a.c35q = consts[1]; a.c70 = consts[2]; a.c70q = consts[3];
    const uint64_t waves = a.units, blocks = (waves + 3) / 4 < 8192 ? (waves + 3) / 4 : 8192;
    prof_begin("k_dft16_mfma", 16.0 * (double)(a.n * batch), s);
    hipLaunchKernelGGL(k_dft16_mfma, dim3((unsigned)blocks), dim3(256), 0, s, a);
    prof_end(s);
    return nttm_hip_check(hipGetLastError(), "k_dft16_mfma launch");
}
