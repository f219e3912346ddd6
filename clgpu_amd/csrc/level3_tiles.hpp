// Tile arithmetic of the level-3 GEMM launches, shared by the kernels (tile
// numbering), the launcher (grid size) and the host's implementation scoring.
#pragma once
#include <hip/hip_runtime.h>

namespace icl {

// Tiles of tile-row tm that meet the triangle a level-3 update writes
// (write_tri & 3: 1 = ICL_TRI_LOWER, m >= n; 2 = ICL_TRI_UPPER, m <= n): columns
// [lo, lo + cnt). With ICL_TRI_OFF_BAND (4) or ICL_TRI_BAND (8) only those
// outside / inside the diagonal 256 x 256 block of the tile row (bm and bn
// divide 256, so a tile row lies in one block and a block spans whole tiles).
__host__ __device__ __attribute__((always_inline)) inline void tri_tile_row(int write_tri, int tm, int tiles_n, int bm, int bn, int& lo, int& cnt)
{
    int a, b; // column tiles [a, b) meeting the triangle
    if ((write_tri & 3) == 1) {
        a = 0;
        const int hi = (tm * bm + bm - 1) / bn;
        b = hi + 1 < tiles_n ? hi + 1 : tiles_n;
    } else {
        const int x = tm * bm - bn + 1;
        a = x <= 0 ? 0 : (x + bn - 1) / bn;
        b = tiles_n;
    }
    if (write_tri & 12) {
        const int blk = tm * bm / 256, c0 = blk * 256 / bn, c1 = (blk + 1) * 256 / bn; // the block's column tiles
        if (write_tri & 8) {
            a = a > c0 ? a : c0;
            b = b < c1 ? b : c1;
        } else if ((write_tri & 3) == 1) {
            b = b < c0 ? b : c0;
        } else {
            a = a > c1 ? a : c1;
        }
    }
    lo = a;
    cnt = b > a ? b - a : 0;
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
