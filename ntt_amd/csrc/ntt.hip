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

__device__ __forceinline__ uint32_t brev_m(uint32_t r, int m) { return m ? (__brev(r) >> (32 - m)) : 0u; }

/* w_N^(e) for N = 2^(lo+m): exponent scaled to the order-2^32 tables and split 16/16 */
__device__ __forceinline__ double twiddle_apply(double v, uint64_t e, int lom, const TwTables &tw, double p)
{
    uint32_t E = (uint32_t)(e << (32 - lom));
    double2 h = tw.bh[E >> NTTM_BT_LOG];
    double2 l = tw.bl[E & ((1u << NTTM_BT_LOG) - 1)];
    v = mul_shoup(v, h.x, h.y, p);
    return mul_shoup(v, l.x, l.y, p);
}

template <bool INV, bool PW>
__global__ __launch_bounds__(NTTM_THREADS) void k_pass(PassArgs a)
{
    __shared__ double t[1 << NTTM_TILE_LOG];
    PassGeom g; g.lo = a.lo; g.m = a.m; g.logw = a.logw; g.logh = a.logh;
    const int T = 1 << (g.m + g.logw + g.logh);
    const int M = 1 << g.m;
    const uint64_t tile = blockIdx.x;
    const double p = a.pk.p, pinv = a.pk.pinv;
    const uint64_t nmask = ((uint64_t)1 << a.logn) - 1;
    const int lom = g.lo + g.m;

    for (int e = threadIdx.x; e < T; e += NTTM_THREADS) {
        const uint64_t gi = plan_gidx(g, tile, e);
        double v = a.x[gi];
        if (PW) {
            uint64_t yi = a.ymode == NTTM_Y_FULL ? gi
                        : a.ymode == NTTM_Y_BCAST ? (gi & nmask)
                        : ((((gi >> a.logn) >> 1) << a.logn) | (gi & nmask));
            v = mul_barrett(v, a.y[yi], p, pinv);
        }
        if (INV && g.lo > 0) {
            const uint32_t r = (e >> g.logw) & (M - 1);
            v = twiddle_apply(v, plan_col(g, tile, e) * brev_m(r, g.m), lom, a.tw, p);
        }
        t[e] = v;
    }
    __syncthreads();

    const int half_count = T >> 1;
    if (!INV) {
        for (int s = g.m - 1; s >= 0; --s) {
            const int sh = g.logw + s, hs = 1 << s;
            for (int j = threadIdx.x; j < half_count; j += NTTM_THREADS) {
                const int e1 = ((j >> sh) << (sh + 1)) | (j & ((1 << sh) - 1)), e2 = e1 | (1 << sh);
                const int r = (e1 >> g.logw) & (M - 1);
                const double2 w = a.tw.stage[hs + (r & (hs - 1))];
                const double u = t[e1], v = t[e2];
                t[e1] = reduce(u + v, p, pinv);
                t[e2] = mul_shoup(u - v, w.x, w.y, p);
            }
            __syncthreads();
        }
    } else {
        for (int s = 0; s < g.m; ++s) {
            const int sh = g.logw + s, hs = 1 << s;
            for (int j = threadIdx.x; j < half_count; j += NTTM_THREADS) {
                const int e1 = ((j >> sh) << (sh + 1)) | (j & ((1 << sh) - 1)), e2 = e1 | (1 << sh);
                const int r = (e1 >> g.logw) & (M - 1);
                const double2 w = a.tw.stage[hs + (r & (hs - 1))];
                const double u = t[e1], v = mul_shoup(t[e2], w.x, w.y, p);
                t[e1] = reduce(u + v, p, pinv);
                t[e2] = reduce(u - v, p, pinv);
            }
            __syncthreads();
        }
    }

    for (int e = threadIdx.x; e < T; e += NTTM_THREADS) {
        double v = t[e];
        if (!INV && g.lo > 0) {
            const uint32_t r = (e >> g.logw) & (M - 1);
            v = twiddle_apply(v, plan_col(g, tile, e) * brev_m(r, g.m), lom, a.tw, p);
        }
        if (a.do_scale) v = mul_shoup(v, a.scale.x, a.scale.y, p);
        a.x[plan_gidx(g, tile, e)] = v;
    }
}

/* limbs -> residues of np primes, zero-extended to npoints; optional per-prime scale (n^-1 folded into one operand).
 * Batched form: product j of a length group takes points [j n, (j+1) n) of every plane, operand a or b of desc[j]. */
struct LoadDesc {                /* must match api.hip */
    const uint64_t *a; uint64_t na;
    const uint64_t *b; uint64_t nb;
};
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
    const uint64_t total = batch << logn, T = (uint64_t)1 << (g.m + g.logw + g.logh);
    if (total % T) return 0;
    const uint64_t tiles = total / T;
    if (plan_gidx(g, tiles - 1, (int)T - 1) != total - 1) return 0;   /* the last element of the last tile */
    if (plan_gidx(g, 0, 0) != 0) return 0;
    return 1;
}

int nttm_run_pass(bool inv, const PassArgs &a, const PassGeom &g, hipStream_t s)
{
    if (!pass_geom_ok(g, a.logn, a.batch)) return nttm_fail(NTTM_EINVAL, "internal: bad pass geometry");
    const uint64_t tiles = plan_tiles(g, a.logn, a.batch);
    if (tiles > 0x7fffffffULL) return nttm_fail(NTTM_EINVAL, "internal: too many tiles");
    PassArgs b = a; b.lo = g.lo; b.m = g.m; b.logw = g.logw; b.logh = g.logh;
    const double pts = (double)(a.batch << a.logn);
    /* algorithmic bytes: read and write the plane once (16 B per point), plus the y plane of a fused pointwise (8 B) */
    prof_begin(!inv ? "k_pass<fwd>" : b.y ? "k_pass<inv,pw>" : "k_pass<inv>", pts * (b.y ? 24.0 : 16.0), s);
    if (!inv) hipLaunchKernelGGL((k_pass<false, false>), dim3((unsigned)tiles), dim3(NTTM_THREADS), 0, s, b);
    else if (b.y) hipLaunchKernelGGL((k_pass<true, true>), dim3((unsigned)tiles), dim3(NTTM_THREADS), 0, s, b);
    else hipLaunchKernelGGL((k_pass<true, false>), dim3((unsigned)tiles), dim3(NTTM_THREADS), 0, s, b);
    prof_end(s);
    return nttm_hip_check(hipGetLastError(), "k_pass launch");
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
