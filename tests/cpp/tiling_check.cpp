// Host check of the work decomposition of reef3d_amd/csrc/tiling.h: every interior cell
// of a tank is produced by exactly one (item, tile, thread), no thread maps outside the
// tank, and the eight block groups partition the item range.
#include <cstdio>
#include <vector>

#include "../../reef3d_amd/csrc/tiling.h"

using namespace reef;

static int check(int nx, int ny, int nz, int ty, int tz, int cpt)
{
    const TileShape s = make_shape(nx, ny, nz, ty, tz, cpt);
    std::vector<int> hits((size_t)nx * ny * nz, 0);
    std::vector<int> item_hits(s.nitems, 0);
    for(int x = 0; x < 8; ++x)
    {
        int lo, hi;
        group_range(s.nitems, x, lo, hi);
        for(int it = lo; it < hi; ++it) item_hits[it]++;
    }
    for(int it = 0; it < s.nitems; ++it)
        if(item_hits[it] != 1) return printf("item %d covered %d times\n", it, item_hits[it]), 1;
    for(int it = 0; it < s.nitems; ++it)
        for(int t = it * cpt; t < (it + 1) * cpt && t < s.ntiles; ++t)
            for(int th = 0; th < 256; ++th)
            {
                int i, j, k;
                if(!tile_cell_of(s, nx, ny, nz, t, th, i, j, k)) continue;
                if(i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz)
                    return printf("out of range %d %d %d\n", i, j, k), 1;
                hits[(size_t)i + (size_t)nx * (j + (size_t)ny * k)]++;
            }
    for(size_t q = 0; q < hits.size(); ++q)
        if(hits[q] != 1) return printf("cell %zu hit %d times (%d %d %d, %d %d, cpt %d)\n", q, hits[q], nx, ny, nz, ty, tz, cpt), 1;
    return 0;
}

int main()
{
    const int shapes[][3] = {{1, 1, 1}, {64, 1, 4}, {65, 1, 7}, {150, 21, 13}, {512, 8, 12}, {2048, 1, 256},
                             {130, 5, 3}, {63, 4, 1}, {200, 1, 97}};
    const int tiles[][2] = {{4, 1}, {2, 2}, {1, 4}};
    const int cpts[] = {1, 3, 8, 16};
    int n = 0;
    for(auto& sh : shapes)
        for(auto& tl : tiles)
            for(int c : cpts)
            {
                if(check(sh[0], sh[1], sh[2], tl[0], tl[1], c)) return 1;
                ++n;
            }
    printf("ok %d\n", n);
    return 0;
}
