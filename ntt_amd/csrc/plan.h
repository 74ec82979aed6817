/* plan.h - the pass plan of a transform and the tile -> element map its kernels use (host + device).
 *
 * A transform of n = 2^logn points is done as passes over contiguous bit ranges of the index, high bits first in the
 * forward (DIF) direction: a pass over bits [lo, lo + m) runs, inside every block of N = 2^(lo+m) points, the 2^m-point
 * DFT along the stride-2^lo rows, then multiplies by w_N^(col * bitrev_m(row)) (the four-step twiddle; col < 2^lo).
 * The inverse runs the passes in the other order: twiddle by w_N^-(col * bitrev_m(row)), then the DIT 2^m-point DFT.
 * A tile is W low columns x 2^m rows x H consecutive blocks (W <= 2^lo), so strided passes read W-element runs
 * (W >= 16 FP64 = 128 bytes in every strided pass the planner makes) and the last pass (lo = 0) reads whole rows.
 */
#ifndef NTTM_PLAN_H
#define NTTM_PLAN_H
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define NTTM_PHD __host__ __device__ __forceinline__
#else
#define NTTM_PHD static inline
#endif

#define NTTM_MAX_PASSES 6
/* bits of a strided pass: at most tile_log - 4, so its tiles read runs of >= 16 columns (128 bytes) */

struct PassGeom { int lo, m, logw, logh; };
struct Plan { int logn, npass; PassGeom p[NTTM_MAX_PASSES]; };

NTTM_PHD int plan_ctz64(uint64_t x) { int c = 0; if (!x) return 63; while (!(x & 1)) { x >>= 1; c++; } return c; }

/* passes in forward order (highest bits first); tile_log = log2 of points per tile of the strided passes, last_log
 * that of the last (contiguous) pass, last_bits the bits that pass takes (at most last_log; its tiles then hold
 * 2^(last_log - last_bits) consecutive blocks) */
NTTM_PHD void plan_make3(Plan *pl, int logn, uint64_t batch, int tile_log, int last_bits, int last_log)
{
    int bz = plan_ctz64(batch);
    if (bz > 40) bz = 40;
    int m0 = logn < last_bits ? logn : last_bits;
    int R = logn - m0;
    const int smax = tile_log - 4;
    int k = (R + smax - 1) / smax;
    pl->logn = logn; pl->npass = k + 1;
    int top = logn;
    for (int i = 0; i < k; i++) {
        int m = (R + (k - i) - 1) / (k - i);              /* balanced split of the remaining bits */
        R -= m;
        PassGeom g; g.m = m; g.lo = top - m;
        g.logw = tile_log - m < g.lo ? tile_log - m : g.lo;
        int room = tile_log - m - g.logw, blocks = (logn - g.lo - m) + bz;
        g.logh = room < blocks ? room : blocks;
        pl->p[i] = g;
        top -= m;
    }
    PassGeom g; g.lo = 0; g.m = m0; g.logw = 0;
    int room = last_log - m0, blocks = (logn - m0) + bz;
    g.logh = room < blocks ? room : blocks;
    pl->p[k] = g;
}
NTTM_PHD void plan_make2(Plan *pl, int logn, uint64_t batch, int tile_log, int last_log)
{
    plan_make3(pl, logn, batch, tile_log, last_log, last_log);
}
NTTM_PHD void plan_make(Plan *pl, int logn, uint64_t batch, int tile_log) { plan_make2(pl, logn, batch, tile_log, tile_log); }

NTTM_PHD uint64_t plan_tiles(const PassGeom &g, int logn, uint64_t batch)
{
    return (batch << logn) >> (g.m + g.logw + g.logh);
}

/* global index (within the batch) of element e of tile `tile` */
NTTM_PHD uint64_t plan_gidx(const PassGeom &g, uint64_t tile, int e)
{
    const uint64_t ncb = ((uint64_t)1 << g.lo) >> g.logw;
    const uint64_t cb = tile % ncb, hb = tile / ncb;
    const uint64_t N = (uint64_t)1 << (g.lo + g.m);
    const int c = e & ((1 << g.logw) - 1);
    const int r = (e >> g.logw) & ((1 << g.m) - 1);
    const int h = e >> (g.logw + g.m);
    return hb * (N << g.logh) + cb * ((uint64_t)1 << g.logw) + (uint64_t)h * N + ((uint64_t)r << g.lo) + (uint64_t)c;
}

/* the column (index within the block below bit lo) of element e of tile `tile` */
NTTM_PHD uint64_t plan_col(const PassGeom &g, uint64_t tile, int e)
{
    const uint64_t ncb = ((uint64_t)1 << g.lo) >> g.logw;
    return (tile % ncb) * ((uint64_t)1 << g.logw) + (uint64_t)(e & ((1 << g.logw) - 1));
}
#endif
