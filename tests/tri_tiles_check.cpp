// Host check of the level-3 triangle tile selection (clgpu_amd/csrc/level3_tiles.hpp):
// for square outputs that are and are not multiples of 256, both triangles and
// every pair of tile sizes, the ICL_TRI_OFF_BAND launch and the ICL_TRI_BAND
// launch together write every element of the triangle exactly once (the kernels'
// element mask applied to the tiles each selects), and tri_tile_count agrees with
// the rows. Exits non-zero on a mismatch. Built and run by tests/test_tri_tiles_cpu.py.
#include "level3_tiles.hpp"
#include "kernel_abi.h"
#include <cstdio>
#include <vector>
int main() {
    long bad = 0, cases = 0;
    const int sizes[] = {256, 320, 512, 1000, 1024, 2048};
    const int tiles[][2] = {{64, 64}, {128, 128}, {256, 256}, {256, 128}, {128, 256}, {128, 64}};
    for (int n : sizes)
        for (int tri = 1; tri <= 2; ++tri)
            for (auto& to : tiles)
                for (auto& tb : tiles) {
                    std::vector<int> cov((size_t)n * n, 0);
                    for (int part = 0; part < 2; ++part) {
                        const int wt = tri | (part == 0 ? ICL_TRI_OFF_BAND : ICL_TRI_BAND);
                        const int bm = part == 0 ? to[0] : tb[0], bn = part == 0 ? to[1] : tb[1];
                        const int tm_n = (n + bm - 1) / bm, tn_n = (n + bn - 1) / bn;
                        long count = 0;
                        for (int tm = 0; tm < tm_n; ++tm) {
                            int lo, cnt;
                            icl::tri_tile_row(wt, tm, tn_n, bm, bn, lo, cnt);
                            count += cnt;
                            for (int j = lo; j < lo + cnt; ++j)
                                for (int r = tm * bm; r < tm * bm + bm && r < n; ++r)
                                    for (int c = j * bn; c < j * bn + bn && c < n; ++c)
                                        if (tri == 1 ? r >= c : r <= c) ++cov[(size_t)r * n + c];
                        }
                        if (count != icl::tri_tile_count(n, n, wt, bm, bn)) ++bad;
                    }
                    for (int r = 0; r < n; ++r)
                        for (int c = 0; c < n; ++c)
                            if (cov[(size_t)r * n + c] != ((tri == 1 ? r >= c : r <= c) ? 1 : 0)) ++bad;
                    ++cases;
                }
    std::printf("%ld cases, %ld mismatches\n", cases, bad);
    return bad != 0;
}
