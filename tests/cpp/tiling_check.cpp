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
    // the position -> item map is a permutation of the items; in z-slab-major order one
    // group's share of the positions is a z-slab of every column
    std::vector<int> perm_hits(s.nitems, 0);
    for(int l = 0; l < s.nitems; ++l)
    {
        const int it = item_at(s, l);
        if(it < 0 || it >= s.nitems) return printf("position %d -> item %d out of range\n", l, it), 1;
        perm_hits[it]++;
        if(s.zslab)
        {
            const int ncol = s.nbx * s.nby;
            const int col = it / s.zslab, zc = it % s.zslab;
            if(col != l % ncol || zc != l / ncol) return printf("position %d -> item %d: not z-slab-major\n", l, it), 1;
        }
    }
    for(int it = 0; it < s.nitems; ++it)
        if(perm_hits[it] != 1) return printf("item %d at %d positions\n", it, perm_hits[it]), 1;
    if(s.zslab != ((ty == 4 && tz == 1 && s.nbz % cpt == 0 && s.nbz / cpt > 1) ? s.nbz / cpt : 0))
        return printf("zslab %d for %d %d %d, %d %d, cpt %d\n", s.zslab, nx, ny, nz, ty, tz, cpt), 1;
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
                             {130, 5, 3}, {63, 4, 1}, {200, 1, 97}, {512, 64, 128}, {256, 64, 48}};
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
    // the 512 x 64 x 128 wave tank's own plan: 16-plane items, 8 z-slabs, XCD group x
    // walks z-slab x of all 128 columns
    {
        const TileShape s = make_shape(512, 64, 128, 4, 1, 16);
        if(s.zslab != 8 || s.nitems != 1024) return printf("wave tank plan: zslab %d items %d\n", s.zslab, s.nitems), 1;
        for(int x = 0; x < 8; ++x)
        {
            int lo, hi;
            group_range(s.nitems, x, lo, hi);
            for(int l = lo; l < hi; ++l)
                if(item_at(s, l) % 8 != x) return printf("group %d walks item %d\n", x, item_at(s, l)), 1;
        }
    }
    printf("ok %d\n", n);
    return 0;
}
