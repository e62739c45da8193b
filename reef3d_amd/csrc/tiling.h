// Work decomposition shared by every grid kernel, host and device.
//
// A tank (one box domain) is cut into tiles of 64 x ty x tz cells (ty * tz = 4, one
// 64-wide wavefront per row of the tile, 256 threads per block).  Tiles are numbered
// z fastest: t -> column col = t / nbz (bx, by), bz = t % nbz, so consecutive tiles of a
// column share their k-1 / k+1 planes.  Consecutive runs of cpt tiles form a work item
// (a chunk); reductions produce one partial per chunk, so the summation order of a
// tank's dot products depends only on the tank and its chunk length, never on which
// other tanks share a launch.
//
// A launch covers the items of every tank of a batch: item ids are global, tank t owns
// items [item0, item0 + nitems).  The fixed grid of a launch is split into 8 groups of
// blocks (blockIdx % 8, the XCD a block is dispatched to); group x walks the items of
// the x-th eighth of the item range, so neighbouring chunks share an L2.
#pragma once

#ifdef __HIPCC__
#define REEF_HD __host__ __device__ __forceinline__
#else
#define REEF_HD inline
#endif

namespace reef {

constexpr int TILE_X = 64;

// blocks per CU of the plane-walk solver kernels (256 CUs)
#define REEF_OCC_PW 4
constexpr int PW_BLOCKS = 256 * REEF_OCC_PW;

struct TileShape
{
    int ty, tz;          // tile rows in y and z, ty * tz == 4
    int tzl;             // log2(tz)
    int nbx, nby, nbz;   // tiles per axis
    int ntiles;
    int cpt;             // tiles per work item
    int nitems;
    // 0, or the z-chunks per column when the items are walked z-slab-major: 64 x 4 x 1
    // tiles whose columns split into whole items (nbz % cpt == 0, more than one chunk).
    // Item position l then stands for item (l % columns) * zslab + l / columns, so one XCD
    // group's contiguous share of the positions is a z-slab of every column: the rows of
    // the y-neighbour tiles a plane walk stages in its halo belong to blocks of the same
    // XCD walking the same planes at the same time, and are served by that XCD's L2.
    // Items and their reduction slots are unchanged, only the block that walks them is.
    int zslab;
};

REEF_HD TileShape make_shape(int nx, int ny, int nz, int ty, int tz, int cpt)
{
    TileShape s;
    s.ty = ty;
    s.tz = tz;
    s.tzl = tz == 4 ? 2 : tz == 2 ? 1 : 0;
    s.nbx = (nx + TILE_X - 1) / TILE_X;
    s.nby = (ny + ty - 1) / ty;
    s.nbz = (nz + tz - 1) / tz;
    s.ntiles = s.nbx * s.nby * s.nbz;
    s.cpt = cpt;
    s.nitems = (s.ntiles + cpt - 1) / cpt;
    s.zslab = (ty == 4 && tz == 1 && s.nbz % cpt == 0 && s.nbz / cpt > 1) ? s.nbz / cpt : 0;
    return s;
}

// thread `lane_row` (threadIdx.x) of tile t -> cell (i, j, k); false outside the tank
REEF_HD bool tile_cell_of(const TileShape& s, int nx, int ny, int nz, int t, int thread, int& i, int& j, int& k)
{
    const int col = t / s.nbz;
    const int bz = t - col * s.nbz;
    const int by = col / s.nbx;
    const int bx = col - by * s.nbx;
    const int row = thread >> 6;               // 0..3: one wavefront per row
    const int ry = row >> s.tzl, rz = row & (s.tz - 1);
    i = bx * TILE_X + (thread & (TILE_X - 1));
    j = by * s.ty + ry;
    k = bz * s.tz + rz;
    return i < nx && j < ny && k < nz;
}

// the item at position l of a tank's item range (see TileShape::zslab)
REEF_HD int item_at(const TileShape& s, int l)
{
    if(s.zslab == 0) return l;
    const int ncol = s.nbx * s.nby;
    const int zc = l / ncol;
    return (l - zc * ncol) * s.zslab + zc;
}

// items [lo, hi) walked by block group `xcd` (0..7) of a launch over nitems items
REEF_HD void group_range(int nitems, int xcd, int& lo, int& hi)
{
    lo = (int)(((long)nitems * xcd) / 8);
    hi = (int)(((long)nitems * (xcd + 1)) / 8);
}

}  // namespace reef
