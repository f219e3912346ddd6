// Tile arithmetic of the level-3 GEMM launches, shared by the kernels (tile
// numbering), the launcher (grid size) and the host's implementation scoring.
#pragma once
#include <hip/hip_runtime.h>

namespace icl {

// Tiles of tile-row tm that meet the triangle a level-3 update writes
// (write_tri 1 = ICL_TRI_LOWER: m >= n, 2 = ICL_TRI_UPPER: m <= n): columns
// [lo, lo + cnt).
__host__ __device__ __attribute__((always_inline)) inline void tri_tile_row(int write_tri, int tm, int tiles_n, int bm, int bn, int& lo, int& cnt)
{
    if (write_tri == 1) {
        lo = 0;
        const int hi = (tm * bm + bm - 1) / bn;
        cnt = hi + 1 < tiles_n ? hi + 1 : tiles_n;
    } else {
        const int x = tm * bm - bn + 1;
        lo = x <= 0 ? 0 : (x + bn - 1) / bn;
        cnt = tiles_n - lo > 0 ? tiles_n - lo : 0;
    }
}

// Number of tiles of an M x N output that meet the written triangle.
__host__ __device__ inline long tri_tile_count(int M, int N, int write_tri, int bm, int bn)
{
    const int tiles_m = (M + bm - 1) / bm, tiles_n = (N + bn - 1) / bn;
    long total = 0;
    for (int tm = 0; tm < tiles_m; ++tm) {
        int lo, cnt;
        tri_tile_row(write_tri, tm, tiles_n, bm, bn, lo, cnt);
        total += cnt;
    }
    return total;
}

} // namespace icl
