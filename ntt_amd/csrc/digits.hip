/* digits.hip - the packed digit output of a base-10^18 number (the reference's ECPACK18 version-2 file,
 * E/packed_fmt.h:1-20,50-55) and the residues that certify it, on the device.
 *
 * The 60-bit form: the limbs in file order (most significant first, index j = 0 .. n-1) make one little-endian bit
 * stream, limb j in bits [60 j, 60 j + 60); a pair takes 15 bytes, a last odd limb 8 (zero-extended), so n limbs take
 * 15 (n / 2) + 8 (n & 1) bytes.  k_pack60 gives each thread 8 limbs = 480 bits = 60 bytes, built as 15 32-bit words from
 * compile-time bit offsets and stored as words (consecutive threads write consecutive 60-byte runs); the last
 * thread writes the short tail byte by byte.
 *
 * The residues (the header's dres: the number mod eight primes just above 2^62, E/verify.c:18): k_limbs_mod sums
 * a_k B^k mod q with thread g taking the limbs k = g, g + T, g + 2T, ... (coalesced), B^k advanced by B^T, in
 * Montgomery arithmetic (R = 2^64; q odd, below 2^63), then a block sum; the host adds the block sums.
 *
 * nttm_e_write_packed writes a whole single-part file for e's digits: X = 2 B^m + the fraction (so nl = m + 1 limbs,
 * pad = 17 leading zero characters, d = 18 m digits computed, d_out of them in the ASCII form). */
#include <stdio.h>
#include <string.h>
#include <stddef.h>
#include <vector>
#include <algorithm>
#include "internal.h"

#define P60_MASK ((1ULL << 60) - 1)
#define LMOD_T 256
#define LMOD_MAXQ 8

/* ---- the 60-bit stream ---------------------------------------------------------------------------------------------- */
__device__ __forceinline__ uint32_t p60_word(const uint64_t *L, int w)          /* bits [32 w, 32 w + 32) of 8 limbs */
{
    const int P = 32 * w, i = P / 60, o = P % 60;
    uint64_t v = L[i] >> o;
    if (o > 28 && i + 1 < 8) v |= L[i + 1] << (60 - o);
    return (uint32_t)v;
}
/* file limb j = x[hi - 1 - j], j < n = hi - lo; out: 4-byte aligned, 15 (n / 2) + 8 (n & 1) bytes */
__global__ __launch_bounds__(256) void k_pack60(uint8_t *out, const uint64_t *x, uint64_t hi, uint64_t n)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nfull = n / 8;
    if (t > nfull) return;
    uint64_t L[8];
    const uint64_t j0 = 8 * t;
#pragma unroll
    for (int i = 0; i < 8; i++) L[i] = j0 + i < n ? x[hi - 1 - (j0 + i)] & P60_MASK : 0;
    if (t < nfull) {
        uint32_t *o = (uint32_t *)(out + 60 * t);
#pragma unroll
        for (int w = 0; w < 15; w++) o[w] = p60_word(L, w);
        return;
    }
    const uint64_t r = n - j0, bytes = 15 * (r / 2) + 8 * (r & 1);             /* the tail: r < 8 limbs */
    uint8_t *o = out + 60 * t;
#pragma unroll
    for (int w = 0; w < 15; w++) {
        const uint32_t v = p60_word(L, w);
#pragma unroll
        for (int b = 0; b < 4; b++) if ((uint64_t)(4 * w + b) < bytes) o[4 * w + b] = (uint8_t)(v >> (8 * b));
    }
}

/* ---- residues mod primes near 2^62 ------------------------------------------------------------------------------- */
struct LmodQ {
    uint64_t q, qneg;          /* q; -q^-1 mod 2^64 */
    uint64_t one;              /* R mod q (1 in Montgomery form) */
    uint64_t bm;               /* B R mod q */
    uint64_t bTm;              /* B^T R mod q, T = the grid's thread count */
};
struct LmodArgs {
    LmodQ q[LMOD_MAXQ];
    int nq;
};
__device__ __forceinline__ uint64_t redc(uint64_t lo, uint64_t hi, uint64_t q, uint64_t qneg)
{
    const uint64_t m = lo * qneg;
    const uint64_t mlo = m * q, mhi = __umul64hi(m, q);
    const uint64_t s = lo + mlo;                                   /* low word is 0 by construction; its carry */
    const uint64_t c = s < lo;
    uint64_t t = hi + mhi + c;                                     /* < 2q < 2^64 */
    return t >= q ? t - q : t;
}
__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b, uint64_t q, uint64_t qneg)
{
    return redc(a * b, __umul64hi(a, b), q, qneg);
}
__global__ __launch_bounds__(LMOD_T) void k_limbs_mod(const uint64_t *x, uint64_t n, LmodArgs A, uint64_t *partial)
{
    __shared__ uint64_t s_acc[LMOD_T / 64][LMOD_MAXQ];
    const uint64_t g = (uint64_t)blockIdx.x * LMOD_T + threadIdx.x, T = (uint64_t)gridDim.x * LMOD_T;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t pw[LMOD_MAXQ], acc[LMOD_MAXQ];
#pragma unroll
    for (int j = 0; j < LMOD_MAXQ; j++) {
        acc[j] = 0; pw[j] = 0;
        if (j < A.nq) {
            const uint64_t q = A.q[j].q, qn = A.q[j].qneg;
            uint64_t p = A.q[j].one, base = A.q[j].bm;             /* B^g R by square and multiply */
            for (uint64_t e = g; e; e >>= 1) {
                if (e & 1) p = mont_mul(p, base, q, qn);
                base = mont_mul(base, base, q, qn);
            }
            pw[j] = p;
        }
    }
    for (uint64_t k = g; k < n; k += T) {                          /* every limb read once */
        const uint64_t a = x[k];
#pragma unroll
        for (int j = 0; j < LMOD_MAXQ; j++) {
            if (j < A.nq) {
                const uint64_t q = A.q[j].q, qn = A.q[j].qneg;
                const uint64_t v = mont_mul(a, pw[j], q, qn);        /* a_k B^k (plain) */
                acc[j] += v; if (acc[j] >= q) acc[j] -= q;
                pw[j] = mont_mul(pw[j], A.q[j].bTm, q, qn);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < LMOD_MAXQ; j++) {
        if (j < A.nq) {
            const uint64_t q = A.q[j].q;
            uint64_t v = acc[j];
            for (int d = 32; d; d >>= 1) {
                const uint64_t o = __shfl_down(v, d, 64);
                v += o; if (v >= q) v -= q;
            }
            if (lane == 0) s_acc[wave][j] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)A.nq) {
        const int j = threadIdx.x;
        const uint64_t q = A.q[j].q;
        uint64_t s = 0;
        for (int w = 0; w < LMOD_T / 64; w++) { s += s_acc[w][j]; if (s >= q) s -= q; }
        partial[(uint64_t)blockIdx.x * LMOD_MAXQ + j] = s;
    }
}

/* host side of the Montgomery constants */
static uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((unsigned __int128)a * b % q); }
static int lmod_prepare(LmodArgs *A, const uint64_t *q, int nq, uint64_t T)
{
    if (nq < 1 || nq > LMOD_MAXQ) return nttm_fail(NTTM_EINVAL, "limbs_mod: %d moduli (1 .. %d)", nq, LMOD_MAXQ);
    A->nq = nq;
    for (int j = 0; j < nq; j++) {
        const uint64_t Q = q[j];
        if (!(Q & 1) || Q >> 63 || Q <= NTTM_B) return nttm_fail(NTTM_EINVAL, "limbs_mod: modulus %llu must be odd, in (10^18, 2^63)", (unsigned long long)Q);
        uint64_t inv = Q;                                          /* Q^-1 mod 2^64 by Newton: 5 doublings from 3 bits */
        for (int i = 0; i < 5; i++) inv *= 2 - Q * inv;
        LmodQ &L = A->q[j];
        L.q = Q; L.qneg = (uint64_t)0 - inv;
        const uint64_t r = (uint64_t)((((unsigned __int128)1) << 64) % Q);
        L.one = r;
        L.bm = h_mulmod(NTTM_B % Q, r, Q);
        uint64_t bt = 1, b = NTTM_B % Q;
        for (uint64_t e = T; e; e >>= 1) { if (e & 1) bt = h_mulmod(bt, b, Q); b = h_mulmod(b, b, Q); }
        L.bTm = h_mulmod(bt, r, Q);
    }
    return 0;
}

/* ---- residues of many limb arrays, for the products' post-condition ------------------------------------------------
 * Each chunk (at most RES_CHUNK limbs of one array) is one block: thread t sums x_k B^k for k = t, t + 256, ... of the
 * chunk (B^k relative to the chunk's start) mod q in Montgomery arithmetic, a block sum gives the chunk's residue.  The
 * caller weights chunk j of an array by B^(j RES_CHUNK) and adds. */
#define RES_CHUNK 65536
struct ResChunk { const uint64_t *x; uint64_t n; };
__global__ __launch_bounds__(LMOD_T) void k_res_chunks(const ResChunk *ch, LmodQ Q, uint64_t *out)
{
    __shared__ uint64_t s_acc[LMOD_T / 64];
    const ResChunk c = ch[blockIdx.x];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint64_t q = Q.q, qn = Q.qneg;
    uint64_t pw = Q.one, base = Q.bm;                               /* B^t R by square and multiply */
    for (uint32_t e = (uint32_t)t; e; e >>= 1) {
        if (e & 1) pw = mont_mul(pw, base, q, qn);
        base = mont_mul(base, base, q, qn);
    }
    uint64_t acc = 0;
    for (uint64_t k = (uint64_t)t; k < c.n; k += LMOD_T) {
        acc += mont_mul(c.x[k], pw, q, qn); if (acc >= q) acc -= q;
        pw = mont_mul(pw, Q.bTm, q, qn);
    }
    for (int d = 32; d; d >>= 1) { const uint64_t o = __shfl_down(acc, d, 64); acc += o; if (acc >= q) acc -= q; }
    if (lane == 0) s_acc[wave] = acc;
    __syncthreads();
    if (t == 0) {
        uint64_t v = 0;
        for (int w = 0; w < LMOD_T / 64; w++) { v += s_acc[w]; if (v >= q) v -= q; }
        out[blockIdx.x] = v;
    }
}
uint64_t nttm_res_chunk_limbs(void) { return RES_CHUNK; }
/* out_dev[j] = the residue mod q of chunk j (chunks_dev: nchunks ResChunk records, each n <= RES_CHUNK) */
int nttm_run_res_chunks(const void *chunks_dev, uint64_t nchunks, uint64_t q, uint64_t *out_dev, hipStream_t s)
{
    if (!nchunks) return 0;
    if (nchunks > 0x7fffffffULL) return nttm_fail(NTTM_EINVAL, "residues: %llu chunks", (unsigned long long)nchunks);
    LmodArgs A;
    RCK(lmod_prepare(&A, &q, 1, LMOD_T));
    hipLaunchKernelGGL(k_res_chunks, dim3((unsigned)nchunks), dim3(LMOD_T), 0, s, (const ResChunk *)chunks_dev, A.q[0], out_dev);
    return nttm_hip_check(hipGetLastError(), "k_res_chunks launch");
}

extern "C" int nttm_limbs_mod(uint64_t *res, const uint64_t *x, size_t n, const uint64_t *q, int nq, void *stream)
{
    if (!res || !q || (n && !x)) return nttm_fail(NTTM_EINVAL, "nttm_limbs_mod arguments");
    RCK(nttm_require_device(x, n * 8, "x"));
    hipStream_t s = (hipStream_t)stream;
    const unsigned blocks = n ? (unsigned)std::min<uint64_t>(1024, (n + LMOD_T - 1) / LMOD_T) : 1;
    LmodArgs A;
    RCK(lmod_prepare(&A, q, nq, (uint64_t)blocks * LMOD_T));
    for (int j = 0; j < nq; j++) res[j] = 0;
    if (!n) return 0;
    uint64_t *dp = nullptr;
    HIPCK(hipMalloc(&dp, (size_t)blocks * LMOD_MAXQ * 8));
    hipLaunchKernelGGL(k_limbs_mod, dim3(blocks), dim3(LMOD_T), 0, s, x, (uint64_t)n, A, dp);
    int rc = nttm_hip_check(hipGetLastError(), "k_limbs_mod launch");
    std::vector<uint64_t> hp((size_t)blocks * LMOD_MAXQ);
    if (!rc) rc = nttm_hip_check(hipMemcpyAsync(hp.data(), dp, hp.size() * 8, hipMemcpyDeviceToHost, s), "copy residues");
    if (!rc) rc = nttm_hip_check(hipStreamSynchronize(s), "k_limbs_mod");
    (void)hipFree(dp);
    if (rc) return rc;
    for (unsigned b = 0; b < blocks; b++)
        for (int j = 0; j < nq; j++) { res[j] += hp[(size_t)b * LMOD_MAXQ + j]; if (res[j] >= q[j]) res[j] -= q[j]; }
    return 0;
}

/* ---- synthetic operands on the device --------------------------------------------------------------------------------
 * limb j of an operand is value j + 1 of the splitmix64 stream seeded `seed`, reduced mod B, and a zero top limb is
 * set to 1 -- the reference's operand generator (E/tests/harness.h:55-81: splitmix64, limb = next() % B, top limb
 * nonzero).  Lets a product of up to 2^31 x 2^31 limbs be made and checked (by residues) without a host copy. */
__global__ __launch_bounds__(256) void k_gen_limbs(uint64_t *x, uint64_t n, uint64_t seed)
{
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += stride) {
        uint64_t z = seed + (j + 1) * 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        uint64_t q, r;
        div_b(0, z, &q, &r);
        x[j] = (j + 1 == n && r == 0) ? 1 : r;
    }
}
extern "C" int nttm_debug_gen_limbs(uint64_t *x, size_t n, uint64_t seed, void *stream)
{
    if (!n) return 0;
    if (!x) return nttm_fail(NTTM_EINVAL, "nttm_debug_gen_limbs: null x");
    RCK(nttm_require_device(x, n * 8, "x"));
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_gen_limbs, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, x, (uint64_t)n, seed);
    return nttm_hip_check(hipGetLastError(), "k_gen_limbs launch");
}

extern "C" uint64_t nttm_pack60_bytes(size_t n) { return 15 * (uint64_t)(n / 2) + 8 * (uint64_t)(n & 1); }

extern "C" int nttm_pack60(uint8_t *out, const uint64_t *x, size_t lo, size_t hi, void *stream)
{
    if (!out || lo > hi || (hi > lo && !x)) return nttm_fail(NTTM_EINVAL, "nttm_pack60 arguments");
    if ((uintptr_t)out & 3) return nttm_fail(NTTM_EINVAL, "nttm_pack60: the output must be 4-byte aligned");
    const uint64_t n = hi - lo;
    if (!n) return 0;
    RCK(nttm_require_device(x + lo, n * 8, "x"));
    RCK(nttm_require_device(out, nttm_pack60_bytes(n), "out"));
    const uint64_t threads = n / 8 + 1;
    hipLaunchKernelGGL(k_pack60, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, (hipStream_t)stream, out, x,
                       (uint64_t)hi, n);
    return nttm_hip_check(hipGetLastError(), "k_pack60 launch");
}

/* ---- the file (E/packed_fmt.h:32-48: the header's fields, in this order and these widths) ---------------------------- */
struct PackedHeader {
    char magic[8];
    uint32_t version, hdr_bytes, endian, limb_bytes, limb_digits, order, part, nparts, reserved0;
    uint64_t d, d_out, nl, pad, lo, hi, k0, k1, complete, nq;
    uint64_t q[8], dres[8];
};
static_assert(offsetof(PackedHeader, d) == 48 && offsetof(PackedHeader, q) == 128 && sizeof(PackedHeader) == 256,
              "the packed header's layout");
#define PK_HDR_BYTES 4096
#define PK_TEXT_OFF 1024
static const uint64_t PK_T1[8] = {                               /* the first eight primes above 2^62 (E/verify.c:18) */
    4611686018427388039ULL, 4611686018427388073ULL, 4611686018427388081ULL, 4611686018427388091ULL,
    4611686018427388093ULL, 4611686018427388097ULL, 4611686018427388157ULL, 4611686018427388181ULL};

extern "C" int nttm_e_write_packed(const char *path, const uint64_t *frac, size_t m, uint64_t d_out, void *stream)
{
    if (!path || !frac || !m || d_out > 18 * (uint64_t)m) return nttm_fail(NTTM_EINVAL, "nttm_e_write_packed arguments");
    hipStream_t s = (hipStream_t)stream;
    const uint64_t nl = m + 1, nbytes = nttm_pack60_bytes(nl);
    uint64_t *dx = nullptr;
    uint8_t *dpk = nullptr;
    if (hipMalloc(&dx, nl * 8) != hipSuccess || hipMalloc(&dpk, (nbytes + 63) & ~63ULL) != hipSuccess) {
        (void)hipGetLastError();
        if (dx) (void)hipFree(dx);
        return nttm_fail(NTTM_ENOMEM, "nttm_e_write_packed: %llu limbs", (unsigned long long)nl);
    }
    PackedHeader h;
    memset(&h, 0, sizeof h);
    const uint64_t two = 2;
    int rc = nttm_hip_check(hipMemcpyAsync(dx, frac, m * 8, hipMemcpyHostToDevice, s), "copy fraction");
    if (!rc) rc = nttm_hip_check(hipMemcpyAsync(dx + m, &two, 8, hipMemcpyHostToDevice, s), "copy top limb");
    if (!rc) rc = nttm_limbs_mod(h.dres, dx, nl, PK_T1, 8, stream);   /* synchronizes */
    if (!rc) rc = nttm_pack60(dpk, dx, 0, nl, stream);
    std::vector<uint8_t> body;
    if (!rc) {
        body.resize(nbytes);
        rc = nttm_hip_check(hipMemcpyAsync(body.data(), dpk, nbytes, hipMemcpyDeviceToHost, s), "copy packed limbs");
        if (!rc) rc = nttm_hip_check(hipStreamSynchronize(s), "k_pack60");
    }
    (void)hipFree(dx);
    (void)hipFree(dpk);
    if (rc) return rc;
    memcpy(h.magic, "ECPACK18", 8);
    h.version = 2; h.hdr_bytes = PK_HDR_BYTES; h.endian = 0x01020304u; h.limb_bytes = 8; h.limb_digits = 18;
    h.order = 1; h.part = 0; h.nparts = 1; h.reserved0 = 60;
    h.d = 18 * (uint64_t)m; h.d_out = d_out; h.nl = nl; h.pad = 17; h.lo = 0; h.hi = nl;
    h.k0 = 0; h.k1 = h.d + 1; h.complete = 1; h.nq = 8;
    memcpy(h.q, PK_T1, sizeof h.q);
    std::vector<char> hdr(PK_HDR_BYTES, 0);
    memcpy(hdr.data(), &h, sizeof h);
    snprintf(hdr.data() + PK_TEXT_OFF, PK_HDR_BYTES - PK_TEXT_OFF,
             "digits of e, packed: X = floor(e 10^%llu) as %llu base-10^18 limbs, top limb first, 60 bits per limb in a "
             "little-endian bit stream after this %d-byte header; the text form is \"2.\" and the first %llu digits "
             "after the point, then a newline. Written by libnttm.\n",
             (unsigned long long)h.d, (unsigned long long)nl, PK_HDR_BYTES, (unsigned long long)d_out);
    FILE *f = fopen(path, "wb");
    if (!f) return nttm_fail(NTTM_EIO, "nttm_e_write_packed: cannot open %s", path);
    const bool ok = fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size() && fwrite(body.data(), 1, nbytes, f) == nbytes;
    const bool closed = fclose(f) == 0;
    if (!ok || !closed) return nttm_fail(NTTM_EIO, "nttm_e_write_packed: writing %s failed", path);
    return 0;
}
