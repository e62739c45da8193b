// Kernels of the batched engine.  A launch covers a set of tanks (TankSet, by value).
// Stencil and elementwise kernels run one 256-cell tile per block; the tiles of the set
// are numbered tank by tank, z fastest, and block b takes tile (b % 8) * per + b / 8, so
// the blocks the hardware sends to XCD x (b % 8 == x) work on one contiguous eighth of
// the tiles and share that XCD's L2 for their halos.  Reductions run a persistent grid
// (a multiple of 8 blocks) over the set's work items (runs of cpt tiles of one tank):
// block group x walks the x-th eighth of the item range.  Every reduction writes one
// partial per work item and a per-tank finalize block adds a tank's partials in item
// order, so a tank's arithmetic does not depend on the other tanks of a launch.
//
// The per-cell arithmetic is cell_ops.h's (the reference's operation order), so the
// stencil results are bit-identical to the reference whatever the set.
#include "batch.h"
#include "cell_ops.h"

#include <algorithm>
#include <mutex>

namespace reef {
namespace bk {

// ------------------------------------------------------------------ work walk
// The tank table is never written by a kernel: reading it through the constant address
// space lets the compiler use scalar loads (the tank's pointers and sizes stay in SGPRs).
// Kernels take the tank by reference there and copy only its DevGrid: indexing a local
// copy of the pointer table with a run-time array id would keep the whole table in
// registers.
// (the host compilation pass of this file never runs device code and has no such
// address space)
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(4))) const TankDev CTank;
#else
typedef const TankDev CTank;
#endif

// index p with a[p] <= g < a[p + 1] in a set's prefix array a[0..n]
__device__ __forceinline__ int set_find(const int* a, int n, int g)
{
    int lo = 0, hi = n;
    while(hi - lo > 1)
    {
        const int mid = (lo + hi) >> 1;
        if(a[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ CTank& tank_at(const TankSet& b, int p) { return ((CTank*)b.tanks)[b.slot[p]]; }

// persistent walk over the set's work items: q the item, p its tank in the set
struct ItemIter
{
    int q, hi, step, p;
};

// items [first, first + count) of the set, split into 8 XCD groups
__device__ __forceinline__ ItemIter set_items(const TankSet& b, int first, int count)
{
    int lo, hi;
    group_range(count, (int)(blockIdx.x & 7), lo, hi);
    lo += first;
    hi += first;
    ItemIter it;
    it.q = lo + (int)(blockIdx.x >> 3);
    it.hi = hi;
    it.step = (int)(gridDim.x >> 3);
    it.p = it.q < hi ? set_find(b.item0, b.n, it.q) : 0;
    return it;
}

__device__ __forceinline__ ItemIter set_items(const TankSet& b) { return set_items(b, 0, b.nitems); }

// all blocks stride over all items, so item q runs on XCD q % 8 (blocks go to the XCDs
// in turn) and the items of any one tank spread over every XCD: for kernels without
// neighbour reuse between items, where only some tanks of the set may have work
__device__ __forceinline__ ItemIter set_items_spread(const TankSet& b)
{
    ItemIter it;
    it.q = (int)blockIdx.x;
    it.hi = b.nitems;
    it.step = (int)gridDim.x;
    it.p = it.q < it.hi ? set_find(b.item0, b.n, it.q) : 0;
    return it;
}

// the items of the plane-walked tanks (a prefix of the set) or of the others
template<bool PW>
__device__ __forceinline__ ItemIter set_items_of(const TankSet& b)
{
    return PW ? set_items(b, 0, b.pw_items) : set_items(b, b.pw_items, b.nitems - b.pw_items);
}

__device__ __forceinline__ void next_item(ItemIter& it, const TankSet& b)
{
    it.q += it.step;
    if(it.q < it.hi)
        while(b.item0[it.p + 1] <= it.q) ++it.p;
}

// a tank the walk skips: go on to this block's last item of it (the tank's items are
// contiguous in the set), so the next step leaves the tank after a single state check
__device__ __forceinline__ void skip_tank(ItemIter& it, const TankSet& b)
{
    const int end = b.item0[it.p + 1];
    if(it.q < end) it.q += ((end - 1 - it.q) / it.step) * it.step;
}

// the one tile of a block of a per-tile launch
struct TileIter
{
    bool ok;
    int p, t;
};

__device__ __forceinline__ TileIter set_tile(const TankSet& b)
{
    TileIter it;
    const int per = (b.ntiles + 7) >> 3;
    const int g = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    it.ok = g < b.ntiles;
    it.p = it.ok ? set_find(b.tile0, b.n, g) : 0;
    it.t = it.ok ? g - b.tile0[it.p] : 0;
    return it;
}

// tiles [t, t1) of the walk; gi = reduction slot of an item walk
struct TileWalk
{
    int t, t1, gi;
    TileShape s;
};

__device__ __forceinline__ TileWalk item_walk(CTank& T, const TankSet& b, const ItemIter& q)
{
    const int li = item_at(T.shape, q.q - b.item0[q.p]);
    TileWalk w;
    w.s = T.shape;
    w.t = li * T.shape.cpt;
    w.t1 = min(w.t + T.shape.cpt, T.shape.ntiles);
    w.gi = T.item0 + li;
    return w;
}

// reduction slot of the current item of a persistent walk
__device__ __forceinline__ int item_slot(CTank& T, const TankSet& b, const ItemIter& q)
{
    return T.item0 + item_at(T.shape, q.q - b.item0[q.p]);
}

__device__ __forceinline__ TileWalk tile_walk(CTank& T, int t)
{
    TileWalk w;
    w.s = T.shape;
    w.t = t;
    w.t1 = t + 1;
    w.gi = -1;
    return w;
}

__device__ __forceinline__ bool tile_cell(const DevGrid& g, const TileWalk& w, int t, int& i, int& j, int& k, long& c)
{
    if(!tile_cell_of(w.s, g.nx, g.ny, g.nz, t, (int)threadIdx.x, i, j, k)) return false;
    c = g.base + i + g.sx * j + g.sxy * k;
    return true;
}

// tile_cell for the per-cell stencils: a wavefront covers one row of the tile, so j and k
// are the same in all its lanes and are made scalar; the y / z grid tables they index then
// load through the scalar unit
__device__ __forceinline__ bool tile_cell_row(const DevGrid& g, const TileWalk& w, int t, int& i, int& j, int& k,
                                              long& c)
{
    tile_cell_of(w.s, g.nx, g.ny, g.nz, t, (int)threadIdx.x, i, j, k);
    j = __builtin_amdgcn_readfirstlane(j);
    k = __builtin_amdgcn_readfirstlane(k);
    if(i >= g.nx || j >= g.ny || k >= g.nz) return false;
    c = g.base + i + g.sx * j + g.sxy * k;
    return true;
}

// block sum of M values (wavefront butterfly, lane 0's sum, 4-wave combine) into dst[0..M)
template<int M>
__device__ __forceinline__ void block_sum_to(double (&v)[M], double* __restrict__ dst)
{
    __shared__ double ws[M][NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for(int m = 0; m < M; ++m)
        for(int off = 32; off > 0; off >>= 1) v[m] += __shfl_xor(v[m], off);
    if(lane == 0)
        for(int m = 0; m < M; ++m) ws[m][wave] = v[m];
    __syncthreads();
    if(threadIdx.x == 0)
        for(int m = 0; m < M; ++m) dst[m] = ((ws[m][0] + ws[m][1]) + ws[m][2]) + ws[m][3];
    __syncthreads();
}

// a tank's partials: items [item0, item0 + nitems), M values each
template<int M>
__device__ __forceinline__ void tank_totals(CTank& T, const TankSet& b, double (&tot)[M])
{
    sum_partials<M>(b.partials + (size_t)T.item0 * M, T.shape.nitems, tot);
}

// ------------------------------------------------------------------ reliable-update outcome
// The state that a reliable update (mode 3) or restart (mode 1) of the mixed solve leaves,
// from the FP64 sums of its residual phase tot = {|r|^2, (r0, r)} (kb_mix_ru).  The phase
// has no finalize launch of its own: the blocks of the next v phase evaluate the outcome
// from the same sums (they need beta and the mode), kb_fin_v stores it, and
// kb_fin_mix_ru stores it after the last iteration of a batch, so the states the host
// reads are settled.  An idle phase then costs one launch, not two.
struct RuNext
{
    double rn, residual, r_j, norm_r0, beta;
    int iter, done, mode;
};

__device__ __forceinline__ void tank_totals_ru(CTank& T, const TankSet& b, double (&tot)[2])
{
    sum_partials<2>(b.partials_ru + (size_t)T.item0 * 2, T.shape.nitems, tot);
}

__device__ __forceinline__ RuNext ru_next(const BiState* st, const double (&tot)[2])
{
    RuNext n;
    n.rn = sqrt(tot[0]);
    n.residual = n.rn / double(st->ncell);
    n.iter = st->iter + (st->mode == 1 ? 1 : 0);   // a restart counts as an iteration (bicgstab_ijk.cpp:183-191)
    n.done = !((n.residual >= st->stop) && (n.iter < st->maxiter));
    n.mode = st->mode == 1 ? 2 : 0;
    n.r_j = tot[1];
    n.norm_r0 = st->norm_r0;
    n.beta = st->beta;
    if(st->mode == 1)
        n.norm_r0 = sqrt(tot[1]);
    else
    {
        const double wr = st->w * st->r_j_prev;
        n.beta = st->alpha * tot[1] / (wr == 0 ? 1.0e-15 : wr);
    }
    return n;
}

// the outcome as the v phase needs it: mode (0 or 2; 4 for done) in the high word, the FP32
// beta's bits in the low word.  Out of line, so that the walk of the calling kernel keeps
// its registers.
__device__ __noinline__ unsigned long long ru_outcome(const double* __restrict__ parts, int nitems, const BiState* st)
{
    double tot[2];
    sum_partials<2>(parts, nitems, tot);
    const RuNext n = ru_next(st, tot);
    return ((unsigned long long)(n.done ? 4u : (unsigned)n.mode) << 32) | __builtin_bit_cast(unsigned, (float)n.beta);
}

// The FP64 solve's restart (kb_bicg_p in mode 1) has no finalize launch of its own either:
// the v phase that follows it in the same iteration needs only its stop decision, taken
// from its sum |r|^2 (restart_done), and kb_fin_v stores its outcome (restart_store), as
// kb_fin_p does for the explicit-operator solve (bicgstab_ijk.cpp:183-191).
__device__ __noinline__ int restart_done(const double* __restrict__ parts, int nitems, const BiState* st)
{
    double tot[1];
    sum_partials<1>(parts, nitems, tot);
    const double residual = sqrt(tot[0]) / double(st->ncell);
    return !((residual >= st->stop) && (st->iter + 1 < st->maxiter));
}

__device__ __forceinline__ void restart_store(BiState* st, double rr)
{
    st->r_j = rr;
    st->rr = rr;
    st->norm_r0 = sqrt(rr);
    st->residual = sqrt(rr) / double(st->ncell);
    st->iter += 1;
    if(!((st->residual >= st->stop) && (st->iter < st->maxiter))) st->done = 1;
    st->mode = 2;
}

// stores the outcome (one thread)
__device__ __forceinline__ void ru_store(BiState* st, const RuNext& n)
{
    st->residual = n.residual;
    st->nru += 1;
    st->pending = 1;
    st->rmax = n.rn;
    st->iter = n.iter;
    if(n.done)
    {
        st->done = 1;
        return;
    }
    st->r_j = n.r_j;
    st->norm_r0 = n.norm_r0;
    st->beta = n.beta;
    st->mode = n.mode;
}

// ------------------------------------------------------------------ per-precision views
template<class V>
struct Prec;

template<>
struct Prec<double>
{
    using Op = SolverOp64;
    __device__ static SolverOp64 op(CTank& T)
    {
        SolverOp64 o;
        o.rfx = T.f[A_RFX];
        o.rfy = T.f[A_RFY];
        o.rfz = T.f[A_RFZ];
        return o;
    }
    __device__ static BiVec<double> vec(CTank& T, int x, int rhs)
    {
        BiVec<double> v;
        v.x = T.f[x];
        v.rhs = T.f[rhs];
        v.r = T.f[A_R];
        v.r0 = T.f[A_R0];
        v.p = T.f[A_P];
        v.ph = T.f[A_PH];
        v.v = T.f[A_V];
        v.t = T.f[A_T];
        v.aii = T.f[A_AII];
        v.partials = nullptr;
        v.st = T.st;
        return v;
    }
};

template<>
struct Prec<float>
{
    using Op = SolverOp32;
    __device__ static SolverOp32 op(CTank& T)
    {
        SolverOp32 o;
        o.kx = T.s[S_KX];
        o.ky = T.s[S_KY];
        o.kz = T.s[S_KZ];
        o.iDXN = T.iDXN32;
        o.iDYN = T.iDYN32;
        o.iDZN = T.iDZN32;
        return o;
    }
    __device__ static BiVec<float> vec(CTank& T, int, int)
    {
        const int h = T.st->buf;
        BiVec<float> v;
        v.x = T.s[S_X];
        v.rhs = nullptr;
        v.r = T.s[S_R];
        v.r0 = T.s[S_R0];
        v.p = T.s[h ? S_P2 : S_P];
        v.ph = T.s[S_PH];
        v.v = T.s[h ? S_V2 : S_V];
        v.t = T.s[S_T];
        v.aii = T.s[S_AII];
        v.partials = nullptr;
        v.st = T.st;
        return v;
    }
};

// ------------------------------------------------------------------ simple kernels
__global__ void __launch_bounds__(NT) kb_copy(TankSet b, int src, int dst)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const double* __restrict__ s = T.f[src];
        double* __restrict__ d = T.f[dst];
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(tile_cell(g, tw, t, i, j, k, c)) d[c] = s[c];
        }
    }
}

// every element of array id (ghosts and padding included) of every active tank
__global__ void __launch_bounds__(NT) kb_fill(TankSet b, int id, double v)
{
    for(int tk = 0; tk < b.n; ++tk)
    {
        CTank& T = tank_at(b, tk);
        double* __restrict__ f = T.f[id];
        const long n = T.g.nalloc;
        for(long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) f[q] = v;
    }
}

// compiled ghost program prog of every active tank on array id
__global__ void __launch_bounds__(NT) kb_ghost(TankSet b, int prog, int id)
{
    for(int tk = 0; tk < b.n; ++tk)
    {
        CTank& T = tank_at(b, tk);
        const GhostOp* __restrict__ ops = T.gops[prog];
        const int n = T.gnops[prog];
        double* __restrict__ f = T.f[id];
        for(int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
        {
            const GhostOp op = ops[q];
            double val;
            if(op.src < 0)
                val = 0.0;
            else if(op.src & (1 << 30))
                val = 0.0 + op.w * f[op.src & ~(1 << 30)];
            else
                val = f[op.src];
            f[op.dst] = val;
        }
    }
}

// ------------------------------------------------------------------ level set, reinit
// HJ-WENO transport stage fused with the TVD-RK3 combination (levelset_RK3.cpp:118-157,
// momentum_FC3.cpp:132-358); cdt = cfac * dt of the tank; stage 0 accumulates out += L
__global__ void __launch_bounds__(NT, 4) kb_hj_stage(TankSet b, int stage, int phi_s_, int phi_n_, int u_, int v_,
                                                   int w_, int out_, double cfac)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const double* __restrict__ phi_s = T.f[phi_s_];
        const double* __restrict__ phi_n = phi_n_ >= 0 ? T.f[phi_n_] : nullptr;
        const double* __restrict__ u = T.f[u_];
        const double* __restrict__ v = T.f[v_];
        const double* __restrict__ w = T.f[w_];
        double* __restrict__ out = T.f[out_];
        const double cdt = cfac * T.tstep[0];
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell_row(g, tw, t, i, j, k, c)) continue;
            const double L = hj_L(g, phi_s, u, v, w, c, i, j, k);
            const double La = 0.0 + L;
            if(stage == 0)
                out[c] += L;
            else if(stage == 1)
                out[c] = phi_s[c] + cdt * La;
            else if(stage == 2)
                out[c] = 0.75 * phi_n[c] + 0.25 * phi_s[c] + cdt * La;
            else
                out[c] = (1.0 / 3.0) * phi_n[c] + (2.0 / 3.0) * phi_s[c] + cdt * La;
        }
    }
}

// reinit sub-stage (reini_RK3.cpp:114-139) with the per-cell pseudo time step
__global__ void __launch_bounds__(NT, 4) kb_reini_stage(TankSet b, int stage, int fs_, int f0_, int out_, double F43,
                                                      int fsf)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const double* __restrict__ f = T.f[fs_];
        const double* __restrict__ f0 = T.f[f0_];
        double* __restrict__ out = T.f[out_];
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell_row(g, tw, t, i, j, k, c)) continue;
            const double lsv = f[c];
            const double L = reini_L(g, f, c, i, j, k, fsf);
            const double dtc = reini_dtc(g, i, j, k, F43);
            if(stage == 1)
                out[c] = f0[c] + dtc * L;
            else if(stage == 2)
                out[c] = 0.75 * f0[c] + 0.25 * lsv + 0.25 * dtc * L;
            else
                out[c] = (1.0 / 3.0) * f0[c] + (2.0 / 3.0) * lsv + (2.0 / 3.0) * dtc * L;
        }
    }
}

// ------------------------------------------------------------------ momentum
// WENO flux divergence accumulated into acc (convection::start, ipol 1..3)
__global__ void __launch_bounds__(NT, 4) kb_flux_acc(TankSet b, int ipol, int q_, int u_, int v_, int w_, int acc_)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        if(ipol == 2 && g.jdir == 0) continue;
        const double* __restrict__ qf = T.f[q_];
        const double* __restrict__ u = T.f[u_];
        const double* __restrict__ v = T.f[v_];
        const double* __restrict__ w = T.f[w_];
        double* __restrict__ acc = T.f[acc_];
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell_row(g, tw, t, i, j, k, c) || !comp_cell(g, ipol, i, j, k)) continue;
            acc[c] += flux_div(g, ipol, qf, u, v, w, c, i, j, k);
        }
    }
}

// velocity predictor of one component (momentum_FC3.cpp:160-208 and the later stages);
// cdt = cfac * dt of the tank.  MODE 0: diff_void, 1: ediff2's explicit viscous terms added
// to F, 2 (idiff2_FS, D20 = 2): only F (irhs + convection) into out, the implicit solve and
// the stage combination follow (km_idiff, km_combine)
template<int comp, int MODE>
__global__ void __launch_bounds__(NT, 2) kb_momentum(TankSet b, int stage, int qin_, int qold_, int u_, int v_, int w_,
                                                   double grav, double cfac, int out_, double rw, double ra)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        if(comp == 2 && g.jdir == 0) continue;
        const double* __restrict__ qin = T.f[qin_];
        const double* __restrict__ qold = T.f[qold_];
        const double* __restrict__ u = T.f[u_];
        const double* __restrict__ v = T.f[v_];
        const double* __restrict__ w = T.f[w_];
        const double* __restrict__ press = T.f[REEF_PRESS];
        const double* __restrict__ phi = T.f[REEF_PHI];
        double* __restrict__ out = T.f[out_];
        const double cdt = cfac * T.tstep[0];
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell_row(g, tw, t, i, j, k, c) || !comp_cell(g, comp, i, j, k)) continue;
            double F = momentum_F(g, comp, qin, u, v, w, press, phi, grav, rw, ra, c, i, j, k);
            if constexpr(MODE == 1) F += ediff_term(g, comp, u, v, w, T.f[REEF_VISC], c, i, j, k);
            if constexpr(MODE == 2)
            {
                out[c] = F;
                continue;
            }
            const double df = qin[c];
            if(stage == 1)
                out[c] = df + cdt * 1.0 * F;
            else if(stage == 2)
                out[c] = 0.75 * qold[c] + 0.25 * df + cdt * 1.0 * F;
            else
                out[c] = (1.0 / 3.0) * qold[c] + (2.0 / 3.0) * df + cdt * 1.0 * F;
        }
    }
}

// ------------------------------------------------------------------ fluid update
// fluid_update_fsf::start (fluid_update_fsf.cpp:39-74): rho, nu and the volume partials
__global__ void __launch_bounds__(NT) kb_heaviside(TankSet b, double rw, double vw, double ra, double va)
{
    #pragma unroll 1
    for(ItemIter q = set_items(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const double* __restrict__ phi = T.f[REEF_PHI];
        double* __restrict__ ro = T.f[REEF_RO];
        double* __restrict__ visc = T.f[REEF_VISC];
        double vals[2] = {0.0, 0.0};
        const TileWalk tw = item_walk(T, b, q);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            const double H = heaviside_H(phi[c], g.psi);
            ro[c] = rw * H + ra * (1.0 - H);
            visc[c] = vw * H + va * (1.0 - H);
            const double vol = g.DXN[i + MARGE] * g.DYN[j + MARGE] * g.DZN[k + MARGE];
            vals[0] += vol * (H - (1.0 - 1.0));
            vals[1] += vol * (1.0 - H - (1.0 - 1.0));
        }
        block_sum_to<2>(vals, b.partials + (size_t)item_slot(T, b, q) * 2);
    }
}

// volume1, volume2 of each tank of the set: into the tank's volume slot, or with out set
// into out[2 * (tank's index in the set) + 0, 1]
__global__ void __launch_bounds__(NT) kb_heaviside_fin(TankSet b, double* __restrict__ out)
{
    CTank& T = tank_at(b, blockIdx.x);
    double tot[2];
    tank_totals<2>(T, b, tot);
    double* dst = out ? out + 2 * blockIdx.x : T.vol;
    if(threadIdx.x == 0)
    {
        dst[0] = tot[0];
        dst[1] = tot[1];
    }
}

// ------------------------------------------------------------------ projection
__global__ void __launch_bounds__(NT) kb_poisson_rhs(TankSet b, int u_, int v_, int w_, double afac)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const double* __restrict__ u = T.f[u_];
        const double* __restrict__ v = T.f[v_];
        const double* __restrict__ w = T.f[w_];
        double* __restrict__ rhs = T.f[A_RHS];
        const double adt = afac * T.tstep[0];
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(tile_cell(g, tw, t, i, j, k, c)) rhs[c] = poisson_rhs_cell(g, u, v, w, adt, c, i, j, k);
        }
    }
}

// face densities of the "+" faces of the interior and of the -1 layer, optionally the
// Jacobi diagonal aii = -1/(M.p + 1e-19) (needs all faces: second pass)
__global__ void __launch_bounds__(NT) kb_face_rho(TankSet b, double rw, double ra)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const double* __restrict__ phi = T.f[REEF_PHI];
        double* __restrict__ rfx = T.f[A_RFX];
        double* __restrict__ rfy = T.f[A_RFY];
        double* __restrict__ rfz = T.f[A_RFZ];
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            rfx[c] = roface(phi[c], phi[c + 1], g.psi, rw, ra);
            rfy[c] = roface(phi[c], phi[c + g.sx], g.psi, rw, ra);
            rfz[c] = roface(phi[c], phi[c + g.sxy], g.psi, rw, ra);
            if(i == 0) rfx[c - 1] = roface(phi[c], phi[c - 1], g.psi, rw, ra);
            if(j == 0) rfy[c - g.sx] = roface(phi[c], phi[c - g.sx], g.psi, rw, ra);
            if(k == 0) rfz[c - g.sxy] = roface(phi[c], phi[c - g.sxy], g.psi, rw, ra);
        }
    }
}

__global__ void __launch_bounds__(NT) kb_aii(TankSet b)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const double* __restrict__ rfx = T.f[A_RFX];
        const double* __restrict__ rfy = T.f[A_RFY];
        const double* __restrict__ rfz = T.f[A_RFZ];
        double* __restrict__ aii = T.f[A_AII];
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            const Coupling cq = couplings(g, rfx, rfy, rfz, c, i, j, k);
            aii[c] = -1.0 / (diag_of(g, cq) + 1e-19);
        }
    }
}

__global__ void __launch_bounds__(NT) kb_correct(TankSet b, double rw, double ra, double afac, int u_, int v_,
                                                  int w_)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const double adt = afac * T.tstep[0];
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(tile_cell(g, tw, t, i, j, k, c))
                correct_cell(g, T.f[REEF_PCORR], T.f[REEF_PHI], rw, ra, adt, T.f[u_], T.f[v_], T.f[w_],
                             T.f[REEF_PRESS], c, i, j, k);
        }
    }
}

// ------------------------------------------------------------------ BiCGStab
// Occupancy of the persistent solver kernels (waves per SIMD = blocks per CU): 8 for the
// plane walk and the FP32 / elementwise kernels; 5 for the kernels that apply the FP64
// operator cell by cell (about 100 registers), whose grids are sized to match.
constexpr int OCC_FULL = 8, OCC_A64 = 5;
// plane walk: 4 blocks per CU for 4-byte operands (two planes of loads in flight per
// block); the default plan of a 3-D tank gives it one work item per block.  8-byte operands
// keep PW64_DEPTH planes of loads in flight at OCC_PW64 blocks per CU: the FP64 walk is
// bound by the latency of its loads, and a deeper prefetch at a lower occupancy keeps more
// bytes in flight per CU than one plane at four blocks per CU
#ifndef REEF_PW64_DEPTH
#define REEF_PW64_DEPTH 1
#endif
#ifndef REEF_OCC_PW64
#define REEF_OCC_PW64 4
#endif
constexpr int OCC_PW = REEF_OCC_PW;
constexpr int PW64_DEPTH = REEF_PW64_DEPTH;
constexpr int OCC_PW64 = REEF_OCC_PW64;
template<class V>
constexpr int occ_pw() { return sizeof(V) == 4 ? OCC_PW : OCC_PW64; }

// The state machine of bicgstab_ijk::solve (bicgstab_ijk.cpp:125-278) per tank: mode 1 = a
// restart is pending, 2 = p and ph were just set from the residual, 3 = reliable update
// pending (mixed precision).  Dot products are accumulated in FP64 for both precisions.

// plane walk of the 3-D stencil kernels (tanks with 64 x 4 x 1 tiles): a block walks up
// the k planes of its item's column; the stencil input of plane k (rows j-1 .. j+4,
// columns i-1 .. i+64) and the +x / +y faces sit in LDS, planes k-1, k+1 of the own cell
// and the -z face stay in registers, so every array crosses HBM about once per cell.
// The LDS planes are double buffered: one barrier per plane.
template<class V>
struct PlaneLDS
{
    V x[TY + 2][TX + 2];
    V fx[TY][TX + 1];
    V fy[TY + 1][TX];
};

// element e of an array through a 32-bit byte offset from its (uniform) base, so the load
// takes the scalar-base + 32-bit vector-offset form (a tank's arrays are below 4 GiB)
template<class T>
__device__ __forceinline__ T ld32(const T* __restrict__ p, unsigned e)
{
    return *(const T*)((const char*)p + e * (unsigned)sizeof(T));
}
template<class T>
__device__ __forceinline__ void st32(T* __restrict__ p, unsigned e, T v)
{
    *(T*)((char*)p + e * (unsigned)sizeof(T)) = v;
}

// Loads of one plane of the walk, issued one plane ahead: the operand's raw inputs at one
// halo cell of the block and at the cell above (k + 1), the faces, and the epilogue's
// inputs.  Halo cells are spread over the threads: thread h < 64 loads the south row
// (j0-1, i0+h), 64 <= h < 128 the north row (j0+TY, i0+h-64), 128 <= h < 132 the west
// column (j0+h-128, i0-1), 132 <= h < 136 the east column (j0+h-132, i0+TX); the halo
// faces: thread h < 64 the -y face of column i0+h, 128 <= h < 132 the -x face of row
// j0+h-128.
template<class V, class R, class E, class AF>
struct PlaneIn
{
    R h, xu;
    V f0, f2, f4, fh;
    AF az;
    E e;
};

struct NoEpi
{
};

// raw(cc) -> R: the loads of the stencil operand at cc; val(R, V& aux) -> V: the operand
// (aux: what the epilogue needs of the centre); pre(c) -> E: the epilogue's loads;
// post(c, Ax, aux, E): the epilogue.  The loads of planes t+1 .. t+D are in flight while
// plane t is staged, synchronised and computed: D = 2 for 4-byte operands, 1 for 8-byte
// ones.  Every load of a plane is issued unconditionally at a clamped (valid) address
// and masked afterwards, and the loop is unrolled over D+1 plane buffers, so no plane's
// loads wait for the next plane's to be issued.
template<class V, class Op, class Raw, class Val, class Pre, class Post>
__device__ __forceinline__ void plane_walk(const DevGrid& g, const Op& op, const TileWalk& tw, Raw raw, Val val, Pre pre,
                                           Post post)
{
    using R = decltype(raw(0u));
    using E = decltype(pre(0u));
    using AF = AxisFactor<Op>;
    using In = PlaneIn<V, R, E, AF>;
    constexpr int D = sizeof(V) == 4 ? 2 : PW64_DEPTH;
    static_assert(D == 1 || D == 2, "plane walk prefetch depth");
    __shared__ PlaneLDS<V> L[2];
    const V* __restrict__ fxa = face_x(op);
    const V* __restrict__ fya = face_y(op);
    const V* __restrict__ fza = face_z(op);
    const int h = threadIdx.x;
    // one tile row per wavefront: the row (and with it j and the column's y grid factors)
    // is wave-uniform and lives in scalar registers
    const int lane = h & (TX - 1), row = __builtin_amdgcn_readfirstlane(h >> 6);
    // the halo cell of thread h relative to the tile origin (h < 136), and its LDS slot:
    // h < 64 the south row (j0-1, i0+h), 64 <= h < 128 the north row (j0+TY, i0+h-64),
    // 128 <= h < 132 the west column (j0+h-128, i0-1), 132 <= h < 136 the east column
    // (j0+h-132, i0+TX); the halo faces: h < 64 the -y face of column i0+h,
    // 128 <= h < 132 the -x face of row j0+h-128
    int hi = 0, hj = 0, hrow = 0, hcol = 0;
    if(h < TX)
    {
        hi = h;
        hj = -1;
        hrow = 0;
        hcol = h + 1;
    }
    else if(h < 2 * TX)
    {
        hi = h - TX;
        hj = TY;
        hrow = TY + 1;
        hcol = h - TX + 1;
    }
    else if(h < 2 * TX + TY)
    {
        hi = -1;
        hj = h - 2 * TX;
        hrow = h - 2 * TX + 1;
        hcol = 0;
    }
    else if(h < 2 * TX + 2 * TY)
    {
        hi = TX;
        hj = h - 2 * TX - TY;
        hrow = h - 2 * TX - TY + 1;
        hcol = TX + 1;
    }
    const bool has_halo = h < 2 * TX + 2 * TY;
    const bool fh_y = h < TX, fh_x = h >= 2 * TX && h < 2 * TX + TY;
    // position of a tile of the walk: plane k of column col = (bx, by).  The walk moves one
    // tile at a time, so positions advance by increments (one division per item, at its
    // start, instead of one per plane)
    struct Pos
    {
        int k, col, bx, by;
    };
    auto pos_of = [&](int tt) {
        Pos q;
        q.col = tt / tw.s.nbz;
        q.k = tt - q.col * tw.s.nbz;
        q.by = q.col / tw.s.nbx;
        q.bx = q.col - q.by * tw.s.nbx;
        return q;
    };
    auto adv = [&](Pos q) {
        if(++q.k == tw.s.nbz)
        {
            q.k = 0;
            ++q.col;
            if(++q.bx == tw.s.nbx)
            {
                q.bx = 0;
                ++q.by;
            }
        }
        return q;
    };
    // per-column state: the thread's cell, which halo/face loads are real
    int i0, j0, k, col, i, j;
    bool in = false, hval = false, fhval = false;
    unsigned coff = 0, hoff = 0, fhoff = 0;   // element offsets within plane k = 0
    auto column_setup = [&]() {
        i = i0 + lane;
        j = j0 + row;
        in = i < g.nx && j < g.ny;
        const int ic = min(i, g.nx - 1), jc = min(j, g.ny - 1);
        coff = (unsigned)(g.base + ic + g.sx * jc);
        const int ih = i0 + hi, jh = j0 + hj;
        hval = has_halo && ih >= 0 && ih < g.nx && jh >= 0 && jh < g.ny;
        hoff = (unsigned)(g.base + max(0, min(ih, g.nx - 1)) + g.sx * max(0, min(jh, g.ny - 1)));
        const int jr = j0 + (h - 2 * TX);
        fhval = (fh_y && i < g.nx && j0 < g.ny) || (fh_x && jr < g.ny);
        if(fh_y)
            fhoff = (unsigned)(g.base + ic + g.sx * (min(j0, g.ny - 1) - 1));
        else if(fh_x)
            fhoff = (unsigned)(g.base + i0 - 1 + g.sx * min(jr, g.ny - 1));
        else
            fhoff = coff;
    };
    auto fetch = [&](int kk, In& P) {
        const unsigned pk = (unsigned)(g.sxy * kk);
        const unsigned c = coff + pk;
        P.h = raw(hoff + pk);
        P.fh = ld32(fh_y ? fya : fxa, fhoff + pk);
        P.f0 = ld32(fxa, c);
        P.f2 = ld32(fya, c);
        P.f4 = ld32(fza, c);
        P.xu = raw(c + (unsigned)g.sxy);   // k + 1 <= nz: a ghost plane at worst
        P.az = axis_factor(g, op, 2, kk);
        P.e = pre(c);
    };
    // k-window of the operand: xd (k-1), xc (k) and the centre's aux; fzd: the -z face;
    // ax, ay: the column's x and y grid factors
    V xd, xc, auxc, fzd;
    AF ax, ay;
    int t = tw.t;
    if(t >= tw.t1) return;
    Pos pc = pos_of(t);   // plane t
    i0 = pc.bx * TX;
    j0 = pc.by * TY;
    k = pc.k;
    col = pc.col;
    bool v1 = false;   // D = 2: the buffer after the current one holds plane t+1
    auto start = [&](In& P0, In& P1) {
        column_setup();
        const unsigned c = coff + (unsigned)(g.sxy * k);
        V d;
        xd = (in && k > 0) ? val(raw(c - (unsigned)g.sxy), d) : V(0);
        auxc = V(0);
        xc = in ? val(raw(c), auxc) : V(0);
        fzd = in ? ld32(fza, c - (unsigned)g.sxy) : V(0);
        ax = axis_factor(g, op, 0, min(i, g.nx - 1));
        ay = axis_factor(g, op, 1, min(j, g.ny - 1));
        fetch(k, P0);
        v1 = false;
        if constexpr(D == 2)
            if(t + 1 < tw.t1)
            {
                const Pos q = D == 1 ? adv(pc) : pos_of(t + 1);
                if(q.col == col)
                {
                    fetch(q.k, P1);
                    v1 = true;
                }
            }
    };
    int buf = 0;
    // one plane: cur is plane t, n1 plane t+1 (D = 2), nf receives plane t+D; false at the end
    auto step = [&](In& cur, In& n1, In& nf) -> bool {
        PlaneLDS<V>& P = L[buf];
        {
            V d;
            const V hv = val(cur.h, d);
            P.x[row + 1][lane + 1] = xc;
            if(has_halo) P.x[hrow][hcol] = hval ? hv : V(0);
            P.fx[row][lane + 1] = in ? cur.f0 : V(0);
            P.fy[row + 1][lane] = in ? cur.f2 : V(0);
            const V fh = fhval ? cur.fh : V(0);
            if(fh_y)
                P.fy[0][h] = fh;
            else if(fh_x)
                P.fx[h - 2 * TX][0] = fh;
        }
        bool vf = false;
        if(t + D < tw.t1 && (D == 1 || v1))
        {
            const Pos q = D == 1 ? adv(pc) : pos_of(t + 2);
            if(q.col == col)
            {
                fetch(q.k, nf);
                vf = true;
            }
        }
        __syncthreads();
        V auxu = V(0);
        const V xu = (in && k + 1 < g.nz) ? val(cur.xu, auxu) : V(0);
        if(!(in && k + 1 < g.nz)) auxu = V(0);
        if(in)
        {
            const unsigned c = coff + (unsigned)(g.sxy * k);
            const CouplingV<V> cq =
                coupling_axes(op, cur.f0, P.fx[row][lane], cur.f2, P.fy[row][lane], cur.f4, fzd, ax, ay, cur.az);
            const V Ax = stencil7(g, cq, xc, P.x[row + 1][lane + 2], P.x[row + 1][lane], P.x[row + 2][lane + 1],
                                  P.x[row][lane + 1], xu, xd, i, j, k);
            post(c, Ax, auxc, cur.e);
        }
        buf ^= 1;
        if(t + 1 >= tw.t1) return false;
        ++t;
        pc = D == 1 ? adv(pc) : pos_of(t);
        if(pc.col == col)
        {
            xd = xc;
            xc = xu;
            auxc = auxu;
            fzd = in ? cur.f4 : V(0);
            k = pc.k;
            v1 = vf;
        }
        else
        {
            i0 = pc.bx * TX;
            j0 = pc.by * TY;
            k = pc.k;
            col = pc.col;
            // the next step's current buffer (and its successor) restart the column
            if constexpr(D == 1)
                start(nf, n1);
            else
                start(n1, nf);
        }
        return true;
    };
    if constexpr(D == 1)
    {
        // one buffer ahead: the next plane lands in nx and is moved into cur (registers
        // allow no second buffer for 8-byte operands at this occupancy)
        In cur, nx;
        start(cur, nx);
        #pragma unroll 1
        for(;;)
        {
            if(!step(cur, nx, nx)) break;
            cur = nx;
        }
    }
    else
    {
        In X, Y, Z;
        start(X, Y);
        #pragma unroll 1
        for(;;)
        {
            if(!step(X, Y, Z)) break;
            if(!step(Y, Z, X)) break;
            if(!step(Z, X, Y)) break;
        }
    }
}

// Row walk of the 2-D (x-z) solver kernels: the item is a run of 64 x 1 x 4 tiles of one
// x-column (z rows k0 .. k0 + 4 cpt); wave w of the block walks rows
// [k0 + w cpt, k0 + (w+1) cpt) up z with the operand of rows k-1, k, k+1 in registers;
// the i+-1 neighbours come from the neighbouring lanes (wavefront shuffles), lanes 0 and
// 63 load the one outside the wavefront.  No block barrier until the item's reduction.
// The y couplings of a 2-D tank are multiplied by 0 in the operator, so they are 0 here.
__device__ __forceinline__ bool row_walk_ok(const DevGrid& g, const TileWalk& tw)
{
    return g.jdir == 0 && tw.s.ty == 1 && tw.s.tz == 4 && tw.t / tw.s.nbz == (tw.t1 - 1) / tw.s.nbz;
}

template<class V>
__device__ __forceinline__ V shfl_dn1(V x)
{
    return __shfl_down(x, 1, 64);
}
template<class V>
__device__ __forceinline__ V shfl_up1(V x)
{
    return __shfl_up(x, 1, 64);
}

template<class V, class Op, class Raw, class Val, class Pre, class Post>
__device__ __forceinline__ void row_walk(const DevGrid& g, const Op& op, const TileWalk& tw, Raw raw, Val val, Pre pre,
                                         Post post)
{
    using AF = AxisFactor<Op>;
    const V* __restrict__ fxa = face_x(op);
    const V* __restrict__ fza = face_z(op);
    const int lane = threadIdx.x & (TX - 1), wave = threadIdx.x >> 6;
    const int col = tw.t / tw.s.nbz;
    const int bz0 = tw.t - col * tw.s.nbz;
    const int per = tw.t1 - tw.t;                 // rows per wave (4 waves, 4 rows per tile)
    const int k0 = bz0 * 4 + wave * per;
    const int k1 = min(k0 + per, g.nz);
    const int i = col * TX + lane;
    const bool in = i < g.nx;
    const int ic = min(i, g.nx - 1);
    const unsigned c0 = (unsigned)(g.base + ic);   // (ic, 0, 0)
    const unsigned sxy = (unsigned)g.sxy;
    const AF ax = axis_factor(g, op, 0, ic);
    if(k0 >= k1) return;
    V d, auxc = V(0);
    unsigned c = c0 + (unsigned)k0 * sxy;
    V xd = k0 > 0 ? val(raw(c - sxy), d) : V(0);
    V xc = val(raw(c), auxc);
    V fzd = ld32(fza, c - sxy);
    #pragma unroll 1
    for(int k = k0; k < k1; ++k, c += sxy)
    {
        V auxu = V(0);
        const V xu = k + 1 < g.nz ? val(raw(c + sxy), auxu) : V(0);
        const auto e = pre(c);
        const V f0 = ld32(fxa, c), f4 = ld32(fza, c);
        // the -x face and the x neighbours: from the neighbouring lanes, the lanes at the
        // wavefront's edges load theirs
        V f1 = shfl_up1(f0), xs = shfl_up1(xc), xn = shfl_dn1(xc);
        if(lane == 0)
        {
            f1 = ld32(fxa, c - 1u);
            xs = i > 0 ? val(raw(c - 1u), d) : V(0);
        }
        if(lane == TX - 1) xn = i + 1 < g.nx ? val(raw(c + 1u), d) : V(0);
        if(in)
        {
            const AF az = axis_factor(g, op, 2, k);
            post(c, stencil5(g, coupling_xz(op, f0, f1, f4, fzd, ax, az), xc, xn, xs, xu, xd, i, k), auxc, e);
        }
        xd = xc;
        xc = xu;
        auxc = auxu;
        fzd = f4;
    }
}

__device__ __forceinline__ bool plane_walk_ok(CTank& T)
{
    return T.g.jdir == 1 && T.shape.ty == TY && T.shape.tz == 1;
}

// restart / initial residual of the iterated precision: r = b - A x, r0 = p = r,
// ph = aii p (restart label, bicgstab_ijk.cpp:136-148)
template<class V, class Op>
__device__ __forceinline__ void restart_body(const DevGrid& g, const Op& op, const BiVec<V>& v, const TileWalk& tw,
                                             double& rr)
{
    #pragma unroll 1
    for(int t = tw.t; t < tw.t1; ++t)
    {
        int i, j, k;
        long c;
        if(!tile_cell(g, tw, t, i, j, k, c)) continue;
        const CouplingV<V> cq = coupling_of(g, op, c, i, j, k);
        const V rv = v.rhs[c] - (apply_A(g, cq, (const V*)v.x, c, i, j, k));
        v.r[c] = rv;
        v.r0[c] = rv;
        v.p[c] = rv;
        v.ph[c] = rv * v.aii[c];
        rr += double(rv) * double(rv);
    }
}

__global__ void kb_setup(TankSet b, double stop, int maxiter, int mixed, double delta)
{
    const int tk = blockIdx.x * blockDim.x + threadIdx.x;
    if(tk >= b.n) return;
    CTank& T = tank_at(b, tk);
    BiState* st = T.st;
    st->r_j = st->norm_r0 = st->alpha = st->w = st->beta = 0.0;
    st->sigma = st->norm_vj = st->norm_rj = st->norm_sj = st->w1 = st->w2 = st->r_j1 = 0.0;
    st->residual = 1.0e9;
    st->stop = stop;
    st->iter = 0;
    st->maxiter = maxiter;
    st->done = 0;
    st->mode = 0;
    st->branch = 0;
    st->mixed = mixed;
    st->r_j_prev = st->rmax = st->res0 = st->rr = 0.0;
    st->delta = delta;
    st->pending = 0;
    st->nru = 0;
    st->buf = 0;
    st->ncell = (long)T.g.nx * T.g.ny * T.g.nz;
}

template<class V>
__global__ void __launch_bounds__(NT, OCC_A64) kb_bicg_init(TankSet b, int x, int rhs)
{
    #pragma unroll 1
    for(ItemIter q = set_items(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        double vals[1] = {0.0};
        restart_body(g, Prec<V>::op(T), Prec<V>::vec(T, x, rhs), item_walk(T, b, q), vals[0]);
        block_sum_to<1>(vals, b.partials + (size_t)item_slot(T, b, q));
    }
}

__global__ void kb_fin_init(TankSet b)
{
    CTank& T = tank_at(b, blockIdx.x);
    BiState* st = T.st;
    double tot[1];
    tank_totals<1>(T, b, tot);
    if(threadIdx.x == 0)
    {
        st->r_j = tot[0];
        st->rr = tot[0];
        st->norm_r0 = sqrt(tot[0]);
        // a negative stop asks for ||r||_2 / N < |stop| * ||b - A x0||_2 / N
        if(st->stop < 0.0) st->stop = -st->stop * sqrt(tot[0]) / double(st->ncell);
        if(!((st->residual >= st->stop) && (st->iter < st->maxiter))) st->done = 1;
        st->mode = 2;
    }
}

// start of an iteration: the pending restart, or p = r + beta (p - w v), ph = aii p
// (bicgstab_ijk.cpp:234-235 and precon_solve)
template<class V>
__global__ void __launch_bounds__(NT, OCC_A64) kb_bicg_p(TankSet b, int x, int rhs)
{
    #pragma unroll 1
    for(ItemIter q = set_items_spread(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const BiState* st = T.st;
        if(st->done || st->mode == 2)
        {
            skip_tank(q, b);
            continue;
        }
        const BiVec<V> v = Prec<V>::vec(T, x, rhs);
        const TileWalk tw = item_walk(T, b, q);
        if(st->mode == 1)
        {
            double vals[1] = {0.0};
            restart_body(g, Prec<V>::op(T), v, tw, vals[0]);
            block_sum_to<1>(vals, b.partials_ru + (size_t)item_slot(T, b, q));
            continue;
        }
        const V beta = (V)st->beta, w = (V)st->w;
        const V* __restrict__ r = v.r;
        const V* __restrict__ vv = v.v;
        const V* __restrict__ aii = v.aii;
        V* __restrict__ p = v.p;
        V* __restrict__ ph = v.ph;
        // independent tiles: unrolled so that several tiles' loads are in flight
        #pragma unroll 4
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            const V pn = r[c] + beta * (p[c] - w * vv[c]);
            p[c] = pn;
            ph[c] = pn * aii[c];
        }
    }
}

// totals of a restart of the explicit-operator solve (km_p): ++solveriter, residual = |r|/N
// (bicgstab_ijk.cpp:183-191); the Poisson solve's restarts are finalized by its v phase
__global__ void kb_fin_p(TankSet b)
{
    CTank& T = tank_at(b, blockIdx.x);
    BiState* st = T.st;
    if(st->done || st->mode != 1) return;
    double tot[1];
    tank_totals<1>(T, b, tot);
    if(threadIdx.x == 0)
    {
        st->r_j = tot[0];
        st->rr = tot[0];
        st->norm_r0 = sqrt(tot[0]);
        st->residual = sqrt(tot[0]) / double(st->ncell);
        st->iter += 1;
        if(!((st->residual >= st->stop) && (st->iter < st->maxiter))) st->done = 1;
        st->mode = 2;
    }
}

// v = A ph; sigma = (v, r0), |v|^2 (bicgstab_ijk.cpp:151-160).  |r|^2 of the same loop is
// the square norm of the residual the x phase (or the restart) has just formed and summed:
// kb_fin_v takes it from the state instead of reading r again here.
// PW: items of the tanks walked by the plane walk, else the others (direct stencil)
template<class V, bool PW>
__global__ void __launch_bounds__(NT, PW ? occ_pw<V>() : (sizeof(V) == 4 ? OCC_FULL : OCC_A64)) kb_bicg_v(TankSet b, int x,
                                                                                               int rhs)
{
    #pragma unroll 1
    for(ItemIter q = set_items_of<PW>(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const BiState* st = T.st;
        if(st->done || st->mode == 3 || plane_walk_ok(T) != PW)
        {
            skip_tank(q, b);
            continue;
        }
        // mode 1: this iteration's p phase has restarted the tank (ph = aii r, fresh)
        if(st->mode == 1
           && (sizeof(V) != 8
               || __builtin_amdgcn_readfirstlane(restart_done(b.partials_ru + (size_t)T.item0, T.shape.nitems, st))))
        {
            skip_tank(q, b);
            continue;
        }
        const BiVec<V> bv = Prec<V>::vec(T, x, rhs);
        const typename Prec<V>::Op op = Prec<V>::op(T);
        const TileWalk tw = item_walk(T, b, q);
        double vals[3] = {0.0, 0.0, 0.0};
        const V* __restrict__ ph = bv.ph;
        const V* __restrict__ r0 = bv.r0;
        V* __restrict__ vo = bv.v;
        auto epilogue = [&](long c, V v) {
            vo[c] = v;
            const double vd = v;
            vals[0] += vd * double(r0[c]);
            vals[1] += vd * vd;
        };
        if constexpr(PW)
        {
            struct EpiIn
            {
                V r0;
            };
            plane_walk<V>(
                g, op, tw, [&](unsigned cc) { return ld32(ph, cc); }, [](V x, V&) { return x; },
                [&](unsigned c) { return EpiIn{ld32(r0, c)}; },
                [&](unsigned c, V Ax, V, const EpiIn& e) {
                    st32(vo, c, Ax);
                    const double vd = Ax;
                    vals[0] += vd * double(e.r0);
                    vals[1] += vd * vd;
                });
        }
        else if(row_walk_ok(g, tw))
        {
            struct EpiIn
            {
                V r0;
            };
            row_walk<V>(
                g, op, tw, [&](unsigned cc) { return ld32(ph, cc); }, [](V x, V&) { return x; },
                [&](unsigned c) { return EpiIn{ld32(r0, c)}; },
                [&](unsigned c, V Ax, V, const EpiIn& e) {
                    st32(vo, c, Ax);
                    const double vd = Ax;
                    vals[0] += vd * double(e.r0);
                    vals[1] += vd * vd;
                });
        }
        else
            #pragma unroll 1
            for(int t = tw.t; t < tw.t1; ++t)
            {
                int i, j, k;
                long c;
                if(!tile_cell(g, tw, t, i, j, k, c)) continue;
                const CouplingV<V> cq = coupling_of(g, op, c, i, j, k);
                epilogue(c, apply_A(g, cq, ph, c, i, j, k));
            }
        block_sum_to<3>(vals, b.partials + (size_t)item_slot(T, b, q) * 3);
    }
}

// Start of an FP32 iteration fused with v = A ph (mixed precision): the pending
// x64 += x32 after a reliable update, and, in mode 0, p = r + beta (p - w v), ph = aii p
// evaluated at every stencil point (the FP32 p update, bicgstab_ijk.cpp:234-235 with
// precon_solve) so that no separate p pass runs; in mode 2 (p, ph fresh from the start
// or a restart) ph is read.  The new p and v go to the other half of their double
// buffers (kb_fin_v swaps the halves), ph is written at the cell for the x phase;
// (v, r0), |v|^2, |r|^2 as kb_bicg_v.  A tank still in mode 1 / 3 has just had its
// reliable-update phase: the walk runs on that phase's outcome (RuNext).
template<bool PW>
__global__ void __launch_bounds__(NT, PW ? OCC_PW : OCC_FULL) kb_mix_v(TankSet b, int x)
{
    #pragma unroll 1
    for(ItemIter q = set_items_of<PW>(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const BiState* st = T.st;
        if(st->done || plane_walk_ok(T) != PW)
        {
            skip_tank(q, b);
            continue;
        }
        int mode = st->mode, pend = st->pending;
        float beta = (float)st->beta;
        if(mode == 1 || mode == 3)
        {
            // the reliable-update phase ran last: its outcome from its sums (every thread
            // holds the same totals; the decisions are made scalar)
            const unsigned long long o = ru_outcome(b.partials_ru + (size_t)T.item0 * 2, T.shape.nitems, st);
            const int om = __builtin_amdgcn_readfirstlane((int)(o >> 32));
            if(om == 4)
            {
                skip_tank(q, b);
                continue;
            }
            mode = om;
            beta = __builtin_bit_cast(float, (unsigned)o);
            pend = 1;
        }
        const int upd = mode == 0, hsel = st->buf;
        const float w = (float)st->w;
        const SolverOp32 op = Prec<float>::op(T);
        const float* __restrict__ r = T.s[S_R];
        const float* __restrict__ r0 = T.s[S_R0];
        const float* __restrict__ aii = T.s[S_AII];
        const float* __restrict__ pc = T.s[hsel ? S_P2 : S_P];
        const float* __restrict__ vc = T.s[hsel ? S_V2 : S_V];
        float* __restrict__ pn = T.s[hsel ? S_P : S_P2];
        float* __restrict__ vn = T.s[hsel ? S_V : S_V2];
        float* __restrict__ ph = T.s[S_PH];
        float* __restrict__ x32 = T.s[S_X];
        double* __restrict__ x64 = T.f[x];
        const TileWalk tw = item_walk(T, b, q);
        double vals[3] = {0.0, 0.0, 0.0};
        struct Epi
        {
            float r0, r, a, p;
        };
        auto pre = [&](unsigned c) { return Epi{ld32(r0, c), ld32(r, c), ld32(aii, c), upd ? 0.0f : ld32(pc, c)}; };
        // aux: the new p at the centre (mode 0)
        auto post = [&](unsigned c, float Ax, float pnew, const Epi& e) {
            st32(vn, c, Ax);
            if(upd)
            {
                st32(pn, c, pnew);
                st32(ph, c, pnew * e.a);
            }
            else
                st32(pn, c, e.p);
            if(pend)
            {
                st32(x64, c, ld32(x64, c) + (double)ld32(x32, c));
                st32(x32, c, 0.0f);
            }
            const double vd = Ax, rc = e.r;
            vals[0] += vd * double(e.r0);
            vals[1] += vd * vd;
            vals[2] += rc * rc;
        };
        struct Rp
        {
            float r, p, v, a;
        };
        auto raw_u = [&](unsigned cc) { return Rp{ld32(r, cc), ld32(pc, cc), ld32(vc, cc), ld32(aii, cc)}; };
        auto val_u = [&](const Rp& z, float& pnew) {
            pnew = z.r + beta * (z.p - w * z.v);
            return pnew * z.a;
        };
        auto raw_f = [&](unsigned cc) { return ld32(ph, cc); };
        auto val_f = [](float z, float&) { return z; };
        if constexpr(PW)
        {
            if(upd)
                plane_walk<float>(g, op, tw, raw_u, val_u, pre, post);
            else
                plane_walk<float>(g, op, tw, raw_f, val_f, pre, post);
        }
        else if(row_walk_ok(g, tw))
        {
            if(upd)
                row_walk<float>(g, op, tw, raw_u, val_u, pre, post);
            else
                row_walk<float>(g, op, tw, raw_f, val_f, pre, post);
        }
        else
            #pragma unroll 1
            for(int t = tw.t; t < tw.t1; ++t)
            {
                int i, j, k;
                long c;
                if(!tile_cell(g, tw, t, i, j, k, c)) continue;
                const CouplingV<float> cq = coupling_of(g, op, c, i, j, k);
                float pnew = 0.0f, d;
                float Ax;
                if(upd)
                {
                    const float phc = val_u(raw_u((unsigned)c), pnew);
                    Ax = apply_A_fn(g, cq, [&](long cc) { return cc == c ? phc : val_u(raw_u((unsigned)cc), d); }, c,
                                    i, j, k);
                }
                else
                    Ax = apply_A(g, cq, (const float*)ph, c, i, j, k);
                post((unsigned)c, Ax, pnew, pre((unsigned)c));
            }
        block_sum_to<3>(vals, b.partials + (size_t)item_slot(T, b, q) * 3);
    }
}

// alpha and the restart tests (bicgstab_ijk.cpp:162-191); rr_state: |r|^2 from the state
// (kb_bicg_v), else the phase's third sum (kb_mix_v, km_v).  A mixed-precision tank in
// mode 1 / 3 first takes the outcome of its reliable-update phase, as kb_mix_v did.
__global__ void kb_fin_v(TankSet b, int rr_state)
{
    CTank& T = tank_at(b, blockIdx.x);
    BiState* st = T.st;
    // the partials' loads go out together with the state's (one memory latency, not two)
    const int done = st->done, mode = st->mode;
    const bool ru = st->mixed && (mode == 1 || mode == 3);
    // the FP64 Poisson solve (rr_state): mode 1 is the restart of this iteration's p phase
    const bool rs = rr_state && mode == 1;
    double tot[3];
    tank_totals<3>(T, b, tot);
    double rtot[2];
    if(ru && !done) tank_totals_ru(T, b, rtot);
    double stot[1];
    if(rs && !done) sum_partials<1>(b.partials_ru + (size_t)T.item0, T.shape.nitems, stot);
    if(done || (!ru && !rs && (mode == 1 || mode == 3))) return;
    if(threadIdx.x == 0)
    {
        if(ru)
        {
            const RuNext n = ru_next(st, rtot);
            ru_store(st, n);
            if(n.done) return;
        }
        if(rs)
        {
            restart_store(st, stot[0]);
            if(st->done) return;
        }
        const double sigma = tot[0];
        const double norm_vj = sqrt(tot[1]);
        const double norm_rj = sqrt(rr_state ? st->rr : tot[2]);
        const double alpha = st->r_j / sigma;
        st->sigma = sigma;
        st->norm_vj = norm_vj;
        st->norm_rj = norm_rj;
        st->alpha = alpha;
        if(st->mixed)
        {
            // the fused v phase wrote p and v into the other half and added x32
            st->buf ^= 1;
            st->pending = 0;
        }
        if(fabs(sigma) <= (1.0e-12 * (norm_vj * st->norm_r0)))
            st->mode = 1;
        else if((fabs(alpha) * norm_vj / (norm_rj == 0 ? 1.0e-15 : norm_rj)) <= 0.08)
            st->mode = 1;
        else
            st->mode = 0;
    }
}

// s = r - alpha v, sh = aii s, t = A sh; (t,s), (t,t), (s,s) (bicgstab_ijk.cpp:193-221)
template<class V, bool PW>
__global__ void __launch_bounds__(NT, PW ? occ_pw<V>() : (sizeof(V) == 4 ? OCC_FULL : OCC_A64)) kb_bicg_t(TankSet b, int x,
                                                                                               int rhs)
{
    #pragma unroll 1
    for(ItemIter q = set_items_of<PW>(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const BiState* st = T.st;
        if(st->done || st->mode == 1 || st->mode == 3 || plane_walk_ok(T) != PW)
        {
            skip_tank(q, b);
            continue;
        }
        const BiVec<V> bv = Prec<V>::vec(T, x, rhs);
        const typename Prec<V>::Op op = Prec<V>::op(T);
        const TileWalk tw = item_walk(T, b, q);
        double vals[3] = {0.0, 0.0, 0.0};
        const V alpha = (V)st->alpha;
        const V* __restrict__ r = bv.r;
        const V* __restrict__ v = bv.v;
        const V* __restrict__ a = bv.aii;
        V* __restrict__ to = bv.t;
        auto epilogue = [&](long c, V t, V s) {
            to[c] = t;
            const double td = t, sd = s;
            vals[0] += td * sd;
            vals[1] += td * td;
            vals[2] += sd * sd;
        };
        auto shv = [&](long cc, V& sv) {
            sv = r[cc] - alpha * v[cc];
            return sv * a[cc];
        };
        if constexpr(PW)
        {
            struct Rin
            {
                V r, v, a;
            };
            plane_walk<V>(
                g, op, tw, [&](unsigned cc) { return Rin{ld32(r, cc), ld32(v, cc), ld32(a, cc)}; },
                [&](const Rin& q, V& sv) {
                    sv = q.r - alpha * q.v;
                    return sv * q.a;
                },
                [](unsigned) { return NoEpi{}; },
                [&](unsigned c, V Ax, V s, const NoEpi&) {
                    st32(to, c, Ax);
                    const double td = Ax, sd = s;
                    vals[0] += td * sd;
                    vals[1] += td * td;
                    vals[2] += sd * sd;
                });
        }
        else if(row_walk_ok(g, tw))
        {
            struct Rin
            {
                V r, v, a;
            };
            row_walk<V>(
                g, op, tw, [&](unsigned cc) { return Rin{ld32(r, cc), ld32(v, cc), ld32(a, cc)}; },
                [&](const Rin& q, V& sv) {
                    sv = q.r - alpha * q.v;
                    return sv * q.a;
                },
                [](unsigned) { return NoEpi{}; },
                [&](unsigned c, V Ax, V sc, const NoEpi&) {
                    st32(to, c, Ax);
                    const double td = Ax, sd = sc;
                    vals[0] += td * sd;
                    vals[1] += td * td;
                    vals[2] += sd * sd;
                });
        }
        else
            #pragma unroll 1
            for(int tt = tw.t; tt < tw.t1; ++tt)
            {
                int i, j, k;
                long c;
                if(!tile_cell(g, tw, tt, i, j, k, c)) continue;
                const CouplingV<V> cq = coupling_of(g, op, c, i, j, k);
                V s;
                const V shc = shv(c, s);
                V dummy;
                const V t = apply_A_fn(g, cq, [&](long cc) { return cc == c ? shc : shv(cc, dummy); }, c, i, j, k);
                epilogue(c, t, s);
            }
        block_sum_to<3>(vals, b.partials + (size_t)item_slot(T, b, q) * 3);
    }
}

__global__ void kb_fin_t(TankSet b)
{
    CTank& T = tank_at(b, blockIdx.x);
    BiState* st = T.st;
    const int done = st->done, mode = st->mode;
    double tot[3];
    tank_totals<3>(T, b, tot);
    if(done || mode == 1 || mode == 3) return;
    if(threadIdx.x == 0)
    {
        st->w1 = tot[0];
        st->w2 = tot[1];
        st->norm_sj = sqrt(tot[2]);
        st->branch = st->norm_sj > st->stop ? 1 : 0;
        st->w = st->branch ? tot[0] / (tot[1] == 0 ? 1.0e-15 : tot[1]) : st->w;
    }
}

// x and r updates, (r, r0), |r|^2 (bicgstab_ijk.cpp:223-268)
template<class V>
__global__ void __launch_bounds__(NT, OCC_FULL) kb_bicg_x(TankSet b, int x, int rhs)
{
    #pragma unroll 1
    for(ItemIter q = set_items_spread(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const BiState* st = T.st;
        if(st->done || st->mode == 1 || st->mode == 3)
        {
            skip_tank(q, b);
            continue;
        }
        const BiVec<V> bv = Prec<V>::vec(T, x, rhs);
        const TileWalk tw = item_walk(T, b, q);
        const V alpha = (V)st->alpha, w = (V)st->w;
        const int branch = st->branch;
        double vals[2] = {0.0, 0.0};
        V* __restrict__ xv = bv.x;
        V* __restrict__ r = bv.r;
        const V* __restrict__ vv = bv.v;
        const V* __restrict__ aii = bv.aii;
        const V* __restrict__ ph = bv.ph;
        const V* __restrict__ tv = bv.t;
        const V* __restrict__ r0 = bv.r0;
        // independent tiles (the dot sums keep their order): unrolled so that several
        // tiles' loads are in flight
        #pragma unroll 4
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            const V s = r[c] - alpha * vv[c];
            V rn;
            if(branch)
            {
                const V sh = s * aii[c];
                xv[c] += alpha * ph[c] + w * sh;
                rn = s - w * tv[c];
            }
            else
            {
                xv[c] += alpha * ph[c];
                rn = s;
            }
            r[c] = rn;
            const double rd = rn;
            vals[0] += rd * double(r0[c]);
            vals[1] += rd * rd;
        }
        block_sum_to<2>(vals, b.partials + (size_t)item_slot(T, b, q) * 2);
    }
}

// beta, residual, iteration count; the mixed form asks for reliable updates instead of
// ending on the iterated residual
__global__ void kb_fin_x(TankSet b)
{
    CTank& T = tank_at(b, blockIdx.x);
    BiState* st = T.st;
    const int done = st->done, mode = st->mode;
    const double alpha = st->alpha, w = st->w;
    const int branch = st->branch;
    double tot[2];
    tank_totals<2>(T, b, tot);
    if(done || mode == 1 || mode == 3) return;
    if(threadIdx.x == 0)
    {
        const double r_j1 = tot[0];
        st->r_j_prev = st->r_j;
        if(branch) st->beta = alpha * r_j1 / (w * st->r_j == 0 ? 1.0e-15 : (w * st->r_j));
        st->r_j1 = r_j1;
        st->r_j = r_j1;
        st->rr = tot[1];
        st->residual = sqrt(tot[1]) / double(st->ncell);
        st->iter += 1;
        st->mode = 0;
        if(!st->mixed)
        {
            if(!((st->residual >= st->stop) && (st->iter < st->maxiter))) st->done = 1;
            return;
        }
        const double rn = sqrt(tot[1]);
        if(rn > st->rmax) st->rmax = rn;
        if(!branch)
            st->mode = 1;
        else if(!(st->residual >= st->stop) || st->iter >= st->maxiter || rn < st->delta * st->rmax)
            st->mode = 3;
    }
}

// ------------------------------------------------------------------ mixed precision
// FP32 face conductances kx = 1/(rho_f DXP) (+x faces of the interior and the -1 layer,
// likewise y, z) and aii = -1/(M.p + 1e-19) from the FP64 diagonal, rounded once
__global__ void __launch_bounds__(NT) kb_coef32(TankSet b)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const double* __restrict__ rfx = T.f[A_RFX];
        const double* __restrict__ rfy = T.f[A_RFY];
        const double* __restrict__ rfz = T.f[A_RFZ];
        float* __restrict__ kx = T.s[S_KX];
        float* __restrict__ ky = T.s[S_KY];
        float* __restrict__ kz = T.s[S_KZ];
        float* __restrict__ aii = T.s[S_AII];
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            const int ip = i + MARGE, jp = j + MARGE, kp = k + MARGE;
            kx[c] = (float)(1.0 / (rfx[c] * g.DXP[ip]));
            ky[c] = (float)(1.0 / (rfy[c] * g.DYP[jp]));
            kz[c] = (float)(1.0 / (rfz[c] * g.DZP[kp]));
            if(i == 0) kx[c - 1] = (float)(1.0 / (rfx[c - 1] * g.DXP[ip - 1]));
            if(j == 0) ky[c - g.sx] = (float)(1.0 / (rfy[c - g.sx] * g.DYP[jp - 1]));
            if(k == 0) kz[c - g.sxy] = (float)(1.0 / (rfz[c - g.sxy] * g.DZP[kp - 1]));
            const Coupling cq = couplings(g, rfx, rfy, rfz, c, i, j, k);
            aii[c] = (float)(-1.0 / (diag_of(g, cq) + 1e-19));
        }
    }
}

// initial residual in FP64: r = b - A64 x64; r0 = p = r, ph = aii r, x32 = 0
__global__ void __launch_bounds__(NT, OCC_A64) kb_mix_start(TankSet b, int x, int rhs)
{
    #pragma unroll 1
    for(ItemIter q = set_items(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const SolverOp64 op = Prec<double>::op(T);
        const BiVec<float> v = Prec<float>::vec(T, x, rhs);
        const double* __restrict__ x64 = T.f[x];
        const double* __restrict__ b64 = T.f[rhs];
        const TileWalk tw = item_walk(T, b, q);
        double vals[2] = {0.0, 0.0};
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            const Coupling cq = coupling_of(g, op, c, i, j, k);
            const double rv = b64[c] - apply_A(g, cq, x64, c, i, j, k);
            const float r = (float)rv;
            v.r[c] = r;
            v.r0[c] = r;
            v.p[c] = r;
            v.ph[c] = r * v.aii[c];
            v.x[c] = 0.0f;
            vals[0] += rv * rv;
            vals[1] += double(r) * double(r);
        }
        block_sum_to<2>(vals, b.partials + (size_t)item_slot(T, b, q) * 2);
    }
}

__global__ void kb_fin_mix_start(TankSet b)
{
    CTank& T = tank_at(b, blockIdx.x);
    BiState* st = T.st;
    double tot[2];
    tank_totals<2>(T, b, tot);
    if(threadIdx.x == 0)
    {
        const double rn = sqrt(tot[0]);
        st->r_j = tot[1];
        st->norm_r0 = sqrt(tot[1]);
        if(st->stop < 0.0) st->stop = -st->stop * rn / double(st->ncell);
        st->residual = rn / double(st->ncell);
        st->rmax = rn;
        st->res0 = rn;
        if(!((st->residual >= st->stop) && (st->iter < st->maxiter))) st->done = 1;
        st->mode = 2;
    }
}

// reliable update (mode 3) or restart (mode 1): r = b - A64 (x64 + x32) in FP64, rounded to
// the FP32 residual; a restart also sets r0 = p = r, ph = aii r
__global__ void __launch_bounds__(NT, OCC_A64) kb_mix_ru(TankSet b, int x, int rhs)
{
    #pragma unroll 1
    for(ItemIter q = set_items_spread(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const BiState* st = T.st;
        if(st->done || (st->mode != 1 && st->mode != 3))
        {
            skip_tank(q, b);
            continue;
        }
        const int restart = st->mode == 1;
        const DevGrid g = T.g;
        const SolverOp64 op = Prec<double>::op(T);
        const BiVec<float> v = Prec<float>::vec(T, x, rhs);
        const double* __restrict__ x64 = T.f[x];
        const double* __restrict__ b64 = T.f[rhs];
        const float* __restrict__ x32 = v.x;
        const TileWalk tw = item_walk(T, b, q);
        double vals[2] = {0.0, 0.0};
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            const Coupling cq = coupling_of(g, op, c, i, j, k);
            const double rv =
                b64[c] - apply_A_fn(g, cq, [&](long cc) { return x64[cc] + (double)x32[cc]; }, c, i, j, k);
            const float r = (float)rv;
            v.r[c] = r;
            float r0v;
            if(restart)
            {
                v.r0[c] = r;
                v.p[c] = r;
                v.ph[c] = r * v.aii[c];
                r0v = r;
            }
            else
                r0v = v.r0[c];
            vals[0] += rv * rv;
            vals[1] += double(r0v) * double(r);
        }
        block_sum_to<2>(vals, b.partials_ru + (size_t)item_slot(T, b, q) * 2);
    }
}

// the outcome of a reliable-update phase that no v phase has taken yet (launched after the
// last iteration of a batch)
__global__ void kb_fin_mix_ru(TankSet b)
{
    CTank& T = tank_at(b, blockIdx.x);
    BiState* st = T.st;
    if(st->done || (st->mode != 1 && st->mode != 3)) return;
    double tot[2];
    tank_totals_ru(T, b, tot);
    if(threadIdx.x == 0) ru_store(st, ru_next(st, tot));
}

// ------------------------------------------------------------------ explicit 7-diagonal operator
// bicgstab_ijk::start / bicgstab_ijk_2D::start (bicgstab_ijk.cpp:51-123, bicgstab_ijk_2D.cpp:
// 49-94) on a caller's matrix_diag: M.p, M.n, M.s, M.w, M.e, M.t, M.b and the right-hand
// side of every FLEXLOOP cell sit in field layout in TankDev::m[0..7], the solver's private
// x in m[8].  The FLEXLOOP cells of var (looping.h:55-57,159) on a box are i < nx - ulast,
// j < ny - vlast, k < nz - wlast (ulast / vlast / wlast = 1 for var 1 / 2 / 3, whose
// flag1..3 drop the last row of their grid; var 4, 5: every interior cell).  As in the
// reference, fillxvec copies the field into the private x on those cells only: its other
// interior cells keep what the calls of other vars left there (they are read by the
// restart's A x), its ghosts stay 0.  ph and sh are 0 outside the FLEXLOOP cells (the
// reference zeroes them after every solve, bicgstab_ijk.cpp:271-275), so their neighbours
// there contribute M * 0.0 and are not read.  The iteration is the state machine of the
// Poisson solve (the same finalize kernels); 3-D tanks use the 7-point product of
// matvec_std (bicgstab_ijk.cpp:296-310), 2-D tanks the 5-point one of bicgstab_ijk_2D.cpp:
// 308-320.
struct MatMask
{
    int ie, je, ke;    // exclusive ends of the FLEXLOOP box
};

__device__ __forceinline__ MatMask mat_mask(const DevGrid& g, int var)
{
    MatMask m;
    m.ie = g.nx - (var == 1 ? 1 : 0);
    m.je = g.ny - (var == 2 ? 1 : 0);
    m.ke = g.nz - (var == 3 ? 1 : 0);
    return m;
}

struct MatOp
{
    const double *p, *n, *s, *w, *e, *t, *b;
};

// the private x of the solve: m[8] for var 1..3 (psolv, the diffusion solves), m[9] for
// var 4, 5 (ppoissonsolv, a separate bicgstab_ijk object in the driver)
__device__ __forceinline__ int mat_xslot(int var) { return var <= 3 ? 8 : 9; }

__device__ __forceinline__ MatOp mat_op(CTank& T)
{
    MatOp o;
    o.p = T.m[0];
    o.n = T.m[1];
    o.s = T.m[2];
    o.w = T.m[3];
    o.e = T.m[4];
    o.t = T.m[5];
    o.b = T.m[6];
    return o;
}

// (A y)(c); MASK: y(cc) = yv(cc) inside the FLEXLOOP box and 0 outside it (ph, sh), else
// yv(cc) everywhere (the private x)
template<bool MASK, class YV>
__device__ __forceinline__ double mat_apply(const DevGrid& g, const MatOp& M, const MatMask& mk, YV yv, long c, int i,
                                            int j, int k)
{
    const double yn = !MASK || i + 1 < mk.ie ? yv(c + 1) : 0.0;
    const double ys = !MASK || i > 0 ? yv(c - 1) : 0.0;
    const double yt = !MASK || k + 1 < mk.ke ? yv(c + g.sxy) : 0.0;
    const double yb = !MASK || k > 0 ? yv(c - g.sxy) : 0.0;
    if(g.jdir == 1)
    {
        const double yw = !MASK || j + 1 < mk.je ? yv(c + g.sx) : 0.0;
        const double ye = !MASK || j > 0 ? yv(c - g.sx) : 0.0;
        return M.p[c] * yv(c) + M.n[c] * yn + M.s[c] * ys + M.w[c] * yw + M.e[c] * ye + M.t[c] * yt + M.b[c] * yb;
    }
    return M.p[c] * yv(c) + M.n[c] * yn + M.s[c] * ys + M.t[c] * yt + M.b[c] * yb;
}

__device__ __forceinline__ bool mat_cell(const DevGrid& g, const MatMask& mk, const TileWalk& tw, int t, int& i, int& j,
                                         int& k, long& c)
{
    return tile_cell(g, tw, t, i, j, k, c) && i < mk.ie && j < mk.je && k < mk.ke;
}

// restart label (bicgstab_ijk.cpp:134-148): r = rhs - A x, r0 = p = r, ph = aii p; on the
// first pass also precon_setup (:356-364), aii = -1/(M.p + 1e-19)
__device__ __forceinline__ void mat_restart(CTank& T, const DevGrid& g, const MatMask& mk, int var,
                                            const TileWalk& tw, bool setup, double& rr)
{
    const double* __restrict__ x = T.m[mat_xslot(var)];
    const MatOp M = mat_op(T);
    const double* __restrict__ rhs = T.m[7];
    double* __restrict__ r = T.f[A_R];
    double* __restrict__ r0 = T.f[A_R0];
    double* __restrict__ p = T.f[A_P];
    double* __restrict__ ph = T.f[A_PH];
    double* __restrict__ aii = T.f[A_AII];
    #pragma unroll 1
    for(int t = tw.t; t < tw.t1; ++t)
    {
        int i, j, k;
        long c;
        if(!mat_cell(g, mk, tw, t, i, j, k, c)) continue;
        if(setup) aii[c] = -1.0 / (M.p[c] + 1.0e-19);
        const double rv = rhs[c] - (mat_apply<false>(g, M, mk, [&](long cc) { return x[cc]; }, c, i, j, k));
        r[c] = rv;
        r0[c] = rv;
        p[c] = rv;
        ph[c] = rv * aii[c];
        rr += rv * rv;
    }
}

// fillxvec (bicgstab_ijk.cpp:366-378): private x = field f on the FLEXLOOP cells (dir 0);
// finalize (:380-384): f = private x there (dir 1)
__global__ void __launch_bounds__(NT) km_copy_x(TankSet b, int f_, int var, int dir)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const MatMask mk = mat_mask(g, var);
        double* __restrict__ f = T.f[f_];
        double* __restrict__ x = T.m[mat_xslot(var)];
        const TileWalk tw = tile_walk(T, q.t);
        int i, j, k;
        long c;
        if(!mat_cell(g, mk, tw, q.t, i, j, k, c)) continue;
        if(dir == 0)
            x[c] = f[c];
        else
            f[c] = x[c];
    }
}

__global__ void __launch_bounds__(NT, OCC_A64) km_init(TankSet b, int var)
{
    #pragma unroll 1
    for(ItemIter q = set_items(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        double vals[1] = {0.0};
        mat_restart(T, g, mat_mask(g, var), var, item_walk(T, b, q), true, vals[0]);
        block_sum_to<1>(vals, b.partials + (size_t)item_slot(T, b, q));
    }
}

// the pending restart, or p = r + beta (p - w v), ph = aii p (bicgstab_ijk.cpp:234-235, 151)
__global__ void __launch_bounds__(NT, OCC_A64) km_p(TankSet b, int var)
{
    #pragma unroll 1
    for(ItemIter q = set_items(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const BiState* st = T.st;
        if(st->done || st->mode == 2)
        {
            skip_tank(q, b);
            continue;
        }
        const MatMask mk = mat_mask(g, var);
        const TileWalk tw = item_walk(T, b, q);
        if(st->mode == 1)
        {
            double vals[1] = {0.0};
            mat_restart(T, g, mk, var, tw, false, vals[0]);
            block_sum_to<1>(vals, b.partials + (size_t)item_slot(T, b, q));
            continue;
        }
        const double beta = st->beta, w = st->w;
        const double* __restrict__ r = T.f[A_R];
        const double* __restrict__ vv = T.f[A_V];
        const double* __restrict__ aii = T.f[A_AII];
        double* __restrict__ p = T.f[A_P];
        double* __restrict__ ph = T.f[A_PH];
        #pragma unroll 4
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!mat_cell(g, mk, tw, t, i, j, k, c)) continue;
            const double pn = r[c] + beta * (p[c] - w * vv[c]);
            p[c] = pn;
            ph[c] = pn * aii[c];
        }
    }
}

// v = A ph; sigma = (v, r0), |v|^2, |r|^2 (bicgstab_ijk.cpp:151-160)
__global__ void __launch_bounds__(NT, OCC_A64) km_v(TankSet b, int var)
{
    #pragma unroll 1
    for(ItemIter q = set_items(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const BiState* st = T.st;
        if(st->done || st->mode == 1 || st->mode == 3)
        {
            skip_tank(q, b);
            continue;
        }
        const MatMask mk = mat_mask(g, var);
        const MatOp M = mat_op(T);
        const TileWalk tw = item_walk(T, b, q);
        const double* __restrict__ ph = T.f[A_PH];
        const double* __restrict__ r = T.f[A_R];
        const double* __restrict__ r0 = T.f[A_R0];
        double* __restrict__ vo = T.f[A_V];
        double vals[3] = {0.0, 0.0, 0.0};
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!mat_cell(g, mk, tw, t, i, j, k, c)) continue;
            const double v = mat_apply<true>(g, M, mk, [&](long cc) { return ph[cc]; }, c, i, j, k);
            vo[c] = v;
            const double rc = r[c];
            vals[0] += v * r0[c];
            vals[1] += v * v;
            vals[2] += rc * rc;
        }
        block_sum_to<3>(vals, b.partials + (size_t)item_slot(T, b, q) * 3);
    }
}

// s = r - alpha v, sh = aii s (at the stencil points), t = A sh; (t, s), (t, t), (s, s)
// (bicgstab_ijk.cpp:193-221)
__global__ void __launch_bounds__(NT, OCC_A64) km_t(TankSet b, int var)
{
    #pragma unroll 1
    for(ItemIter q = set_items(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const BiState* st = T.st;
        if(st->done || st->mode == 1 || st->mode == 3)
        {
            skip_tank(q, b);
            continue;
        }
        const MatMask mk = mat_mask(g, var);
        const MatOp M = mat_op(T);
        const TileWalk tw = item_walk(T, b, q);
        const double alpha = st->alpha;
        const double* __restrict__ r = T.f[A_R];
        const double* __restrict__ v = T.f[A_V];
        const double* __restrict__ a = T.f[A_AII];
        double* __restrict__ to = T.f[A_T];
        double vals[3] = {0.0, 0.0, 0.0};
        #pragma unroll 1
        for(int tt = tw.t; tt < tw.t1; ++tt)
        {
            int i, j, k;
            long c;
            if(!mat_cell(g, mk, tw, tt, i, j, k, c)) continue;
            const double sc = r[c] - alpha * v[c];
            const double t =
                mat_apply<true>(g, M, mk, [&](long cc) { return (r[cc] - alpha * v[cc]) * a[cc]; }, c, i, j, k);
            to[c] = t;
            vals[0] += t * sc;
            vals[1] += t * t;
            vals[2] += sc * sc;
        }
        block_sum_to<3>(vals, b.partials + (size_t)item_slot(T, b, q) * 3);
    }
}

// x and r updates, (r, r0), |r|^2 (bicgstab_ijk.cpp:223-268)
__global__ void __launch_bounds__(NT, OCC_FULL) km_x(TankSet b, int var)
{
    #pragma unroll 1
    for(ItemIter q = set_items(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const BiState* st = T.st;
        if(st->done || st->mode == 1 || st->mode == 3)
        {
            skip_tank(q, b);
            continue;
        }
        const MatMask mk = mat_mask(g, var);
        const TileWalk tw = item_walk(T, b, q);
        const double alpha = st->alpha, w = st->w;
        const int branch = st->branch;
        double* __restrict__ xv = T.m[mat_xslot(var)];
        double* __restrict__ r = T.f[A_R];
        const double* __restrict__ vv = T.f[A_V];
        const double* __restrict__ aii = T.f[A_AII];
        const double* __restrict__ ph = T.f[A_PH];
        const double* __restrict__ tv = T.f[A_T];
        const double* __restrict__ r0 = T.f[A_R0];
        double vals[2] = {0.0, 0.0};
        #pragma unroll 4
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!mat_cell(g, mk, tw, t, i, j, k, c)) continue;
            const double sv = r[c] - alpha * vv[c];
            double rn;
            if(branch)
            {
                const double sh = sv * aii[c];
                xv[c] += alpha * ph[c] + w * sh;
                rn = sv - w * tv[c];
            }
            else
            {
                xv[c] += alpha * ph[c];
                rn = sv;
            }
            r[c] = rn;
            vals[0] += rn * r0[c];
            vals[1] += rn * rn;
        }
        block_sum_to<2>(vals, b.partials + (size_t)item_slot(T, b, q) * 2);
    }
}

// ------------------------------------------------------------------ implicit diffusion (D20 = 2)
// pdiff->diff_u / diff_v / diff_w of idiff2_FS (idiff2_FS_2D on 2-D tanks) inside a
// momentum_FC3 stage (momentum_FC3.cpp:170-204, 272-306, 370-404): the operator rows and
// right-hand side of component comp into TankDev::m[0..7], and fillxvec of the diff field
// (= u_in on the loop, idiff2_FSu.cpp:47-48) into the private x m[8].  The direct-forcing
// flags are set in TankDev::tstep[5] after the first stage-1 predictor of the run.
template<int comp>
__global__ void __launch_bounds__(NT, 2) km_idiff(TankSet b, double alpha, int qin_, int u_, int v_, int w_)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        if(comp == 2 && g.jdir == 0) continue;
        const MatMask mk = mat_mask(g, comp);
        const double* __restrict__ vin = T.f[qin_];
        const double* __restrict__ u = T.f[u_];
        const double* __restrict__ v = T.f[v_];
        const double* __restrict__ w = T.f[w_];
        const double dt = T.tstep[0];
        const bool df = T.tstep[5] != 0.0;
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell_row(g, tw, t, i, j, k, c) || !comp_cell(g, comp, i, j, k)) continue;
            const MatRow r = idiff_row(g, comp, alpha, dt, vin, u, v, w, T.f[REEF_U], T.f[REEF_V], T.f[REEF_VISC], df,
                                       mk.ie, mk.je, mk.ke, c, i, j, k);
            T.m[0][c] = r.p;
            T.m[1][c] = r.n;
            T.m[2][c] = r.s;
            T.m[3][c] = r.w;
            T.m[4][c] = r.e;
            T.m[5][c] = r.t;
            T.m[6][c] = r.b;
            T.m[7][c] = r.rhs;
            T.m[8][c] = vin[c];
        }
    }
}

// the stage combination of component comp with the solved diff (the private x, which
// finalize would copy into diff): stage 1 urk1 = udiff + dt F, stage 2 urk2 = 0.75 u +
// 0.25 udiff + 0.25 dt F, stage 3 u = u / 3 + 2/3 udiff + 2/3 dt F (momentum_FC3.cpp:
// 172-174, 274-276, 372-374); F is the component's F / G / H
template<int comp>
__global__ void __launch_bounds__(NT) km_combine(TankSet b, int stage, int qold_, int F_, double cfac, int out_)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        if(comp == 2 && g.jdir == 0) continue;
        const double* __restrict__ qold = T.f[qold_];
        const double* __restrict__ F = T.f[F_];
        const double* __restrict__ df = T.m[8];
        double* __restrict__ out = T.f[out_];
        const double cdt = cfac * T.tstep[0];
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell_row(g, tw, t, i, j, k, c) || !comp_cell(g, comp, i, j, k)) continue;
            if(stage == 1)
                out[c] = df[c] + cdt * 1.0 * F[c];
            else if(stage == 2)
                out[c] = 0.75 * qold[c] + 0.25 * df[c] + cdt * 1.0 * F[c];
            else
                out[c] = (1.0 / 3.0) * qold[c] + (2.0 / 3.0) * df[c] + cdt * 1.0 * F[c];
        }
    }
}

// ghostcell::solid_forcing_flag_update at the end of momentum_forcing_start
// (momentum_forcing.cpp:102): from here on the diffusion rows fold their wall couplings
__global__ void km_df_set(TankSet b)
{
    const int tk = blockIdx.x * blockDim.x + threadIdx.x;
    if(tk >= b.n) return;
    tank_at(b, tk).tstep[5] = 1.0;
}

// ------------------------------------------------------------------ adaptive time step
// ietimestep::start (ietimestep.cpp:38-207) and ::ini (:209-304) with N50 = 1, first pass:
// per work item the largest |u| over the u loop (i < nx - 1), |v| over the v loop
// (j < ny - 1) and |w| over the w loop (k < nz - 1): ULOOP / VLOOP / WLOOP
// (looping.h:149-151) without the last rows of the staggered grids
__global__ void __launch_bounds__(NT) kb_velmax(TankSet b, int u_, int v_, int w_)
{
    __shared__ double ws[3][NT / 64];
    #pragma unroll 1
    for(ItemIter q = set_items_spread(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const double* __restrict__ u = T.f[u_];
        const double* __restrict__ v = T.f[v_];
        const double* __restrict__ w = T.f[w_];
        const TileWalk tw = item_walk(T, b, q);
        double m[3] = {0.0, 0.0, 0.0};
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            if(i < g.nx - 1) m[0] = fmax(m[0], fabs(u[c]));
            if(j < g.ny - 1) m[1] = fmax(m[1], fabs(v[c]));
            if(k < g.nz - 1) m[2] = fmax(m[2], fabs(w[c]));
        }
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for(int a = 0; a < 3; ++a)
            for(int off = 32; off > 0; off >>= 1) m[a] = fmax(m[a], __shfl_xor(m[a], off));
        if(lane == 0)
            for(int a = 0; a < 3; ++a) ws[a][wave] = m[a];
        __syncthreads();
        if(threadIdx.x == 0)
            for(int a = 0; a < 3; ++a)
                b.partials[(size_t)item_slot(T, b, q) * 3 + a] =
                    fmax(fmax(ws[a][0], ws[a][1]), fmax(ws[a][2], ws[a][3]));
        __syncthreads();
    }
}

// second pass, one block per tank: the maxima over the tank's items, then
//   dt = N47 * MIN over cells of 2 / (|u|max / dx + sqrt(4 |maxF,G,H| / dx)), dx = MIN3(DXN,
//   DYN, DZN) of the cell, with maxF = maxG = maxH = 0 (only the EARSM closure sets them)
// (::ini: umax at least 2, dt * 0.25, at least 1e-6).  The term is monotone in dx, so the
// minimum over the cells is the term at the smallest cell width (tstep[4], from the grid).
__global__ void kb_fin_dt(TankSet b, double N47, int ini)
{
    __shared__ double sh[3][NT];
    CTank& T = tank_at(b, blockIdx.x);
    const double* __restrict__ part = b.partials + (size_t)T.item0 * 3;
    double m[3] = {0.0, 0.0, 0.0};
    for(int it = threadIdx.x; it < T.shape.nitems; it += NT)
        for(int a = 0; a < 3; ++a) m[a] = fmax(m[a], part[(size_t)it * 3 + a]);
    for(int a = 0; a < 3; ++a) sh[a][threadIdx.x] = m[a];
    __syncthreads();
    for(int s = NT / 2; s > 0; s >>= 1)
    {
        if(threadIdx.x < s)
            for(int a = 0; a < 3; ++a) sh[a][threadIdx.x] = fmax(sh[a][threadIdx.x], sh[a][threadIdx.x + s]);
        __syncthreads();
    }
    if(threadIdx.x == 0)
    {
        double umax = sh[0][0];
        const double vmax = sh[1][0], wmax = sh[2][0];
        if(ini) umax = umax > 2.0 ? umax : 2.0;
        const double dx = T.tstep[4], maxFGH = 0.0;
        const double c = 2.0 / ((sqrt(umax * umax + vmax * vmax + wmax * wmax)) / dx + sqrt((4.0 * fabs(maxFGH)) / dx));
        double cu = 1.0e10 < c ? 1.0e10 : c;
        double dt = N47 * cu;
        if(ini)
        {
            dt = N47 * cu * 0.25;
            dt = dt > 1.0e-6 ? dt : 1.0e-6;
        }
        T.tstep[0] = dt;
        T.tstep[1] = umax;
        T.tstep[2] = vmax;
        T.tstep[3] = wmax;
    }
}

// x64 += x32, x32 = 0
__global__ void __launch_bounds__(NT) kb_add32(TankSet b, int x)
{
    #pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        double* __restrict__ x64 = T.f[x];
        float* __restrict__ x32 = T.s[S_X];
        const TileWalk tw = tile_walk(T, q.t);
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            x64[c] += (double)x32[c];
            x32[c] = 0.0f;
        }
    }
}

// ||rhs - A x||_2 (res_calc, bicgstab_ijk.cpp:314-338, without 1/N)
__global__ void __launch_bounds__(NT, OCC_A64) kb_defect(TankSet b, int x, int rhs)
{
    #pragma unroll 1
    for(ItemIter q = set_items(b); q.q < q.hi; next_item(q, b))
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        const SolverOp64 op = Prec<double>::op(T);
        const double* __restrict__ xv = T.f[x];
        const double* __restrict__ bv = T.f[rhs];
        const TileWalk tw = item_walk(T, b, q);
        double vals[1] = {0.0};
        #pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            const Coupling cq = coupling_of(g, op, c, i, j, k);
            const double rv = bv[c] - apply_A(g, cq, xv, c, i, j, k);
            vals[0] += rv * rv;
        }
        block_sum_to<1>(vals, b.partials + (size_t)item_slot(T, b, q));
    }
}

__global__ void kb_fin_norm(TankSet b, double* __restrict__ norms)
{
    CTank& T = tank_at(b, blockIdx.x);
    double tot[1];
    tank_totals<1>(T, b, tot);
    if(threadIdx.x == 0) norms[b.slot[blockIdx.x]] = sqrt(tot[0]);
}

// solver states of the set's tanks into dst[slot] (one copy-out per iteration batch)
__global__ void kb_gather_states(TankSet b, BiState* __restrict__ dst)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= b.n) return;
    CTank& T = tank_at(b, p);
    dst[b.slot[p]] = *T.st;
}

// ------------------------------------------------------------------ level-set benchmark
// benchmark_vortex3D::start (benchmark_vortex3D.cpp:67-96): the prescribed deformation
// velocity at the face positions, at time simtime
__global__ void __launch_bounds__(NT) kb_vortex_velocity(TankSet b, int comp, int id, double simtime, double DXM)
{
#pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        double* __restrict__ qf = T.f[id];
        const TileWalk tw = tile_walk(T, q.t);
#pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            if((comp == 1 && i >= g.nx - 1) || (comp == 2 && j >= g.ny - 1) || (comp == 3 && k >= g.nz - 1)) continue;
            const double x = g.XP[i + MARGE], y = g.YP[j + MARGE], z = g.ZP[k + MARGE];
            const double ct = cos((PI_REF * simtime) / 3.0);
            if(comp == 1)
            {
                const double xc = x + 0.5 * DXM;
                qf[c] = 2.0 * pow(sin(PI_REF * xc), 2.0) * sin(2.0 * PI_REF * y) * sin(2.0 * PI_REF * z) * ct;
            }
            else if(comp == 2)
            {
                const double yc = y + 0.5 * DXM;
                qf[c] = -sin(2.0 * PI_REF * x) * pow(sin(PI_REF * yc), 2.0) * sin(2.0 * PI_REF * z) * ct;
            }
            else
            {
                const double zc = z + 0.5 * DXM;
                qf[c] = -sin(2.0 * PI_REF * x) * sin(2.0 * PI_REF * y) * pow(sin(PI_REF * zc), 2.0) * ct;
            }
        }
    }
}

// momentum_RK3 with void convection, diffusion and pressure (momentum_RK3.cpp:70-326):
// the three stages of one component with the constant forcing Fv
__global__ void __launch_bounds__(NT) kb_void_rk3(TankSet b, int comp, int id, double Fv)
{
#pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const DevGrid g = T.g;
        double* __restrict__ qf = T.f[id];
        const double dt = T.tstep[0];
        const TileWalk tw = tile_walk(T, q.t);
#pragma unroll 1
        for(int t = tw.t; t < tw.t1; ++t)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, t, i, j, k, c)) continue;
            if((comp == 1 && i >= g.nx - 1) || (comp == 2 && j >= g.ny - 1) || (comp == 3 && k >= g.nz - 1)) continue;
            const double q0 = qf[c];
            double a1 = q0 + dt * 1.0 * Fv;
            a1 += 1.0 * dt * 1.0 * 0.0;
            double a2 = 0.75 * q0 + 0.25 * a1 + 0.25 * dt * 1.0 * Fv;
            a2 += 0.25 * dt * 1.0 * 0.0;
            double qn = (1.0 / 3.0) * q0 + (2.0 / 3.0) * a2 + (2.0 / 3.0) * dt * 1.0 * Fv;
            qn += (2.0 / 3.0) * dt * 1.0 * 0.0;
            qf[c] = qn;
        }
    }
}

// ------------------------------------------------------------------ wave generation (iowave)
// Relaxation-zone wave generation (B98 = 2) and numerical beach (B99) of linear waves.  The
// static part (zones, relaxation factors, distances along the wave) comes per column from
// the host; the kernels evaluate the wave at the tank's simtime (tstep[6]), the value
// wavegen_precalc gives at the start of the step, where a relaxation reads it.

// iowave::ramp (iowave_relax_functions.cpp:239-254); ramp_T = B102 wT (B101 = 1) or B102
__device__ __forceinline__ double wave_ramp(const WaveDev& W, double t)
{
    double f = 1.0;
    if(W.ramp_mode == 1 && t < W.ramp_T) f = t / (W.ramp_T) - (1.0 / PI_REF) * sin(PI_REF * (t / (W.ramp_T)));
    if(W.ramp_mode == 2 && t < W.ramp_T) f = t / (W.ramp_T) - (1.0 / PI_REF) * sin(PI_REF * (t / (W.ramp_T)));
    return f;
}

// wave_interface (wave_interface.cpp:137-219) over wave_lib_linear (wave_lib_linear.cpp:
// 47-85): the start and end times are 0 and 1e20, the depth clamp is wD
__device__ __forceinline__ double wave_horz(const WaveDev& W, double x, double z, double t)
{
    const double teta = W.wk * x - W.ww * (t) + W.pshift;
    return W.ww * W.wa * (cosh(W.wk * (W.wdt + z)) / sinh(W.wk * W.wdt)) * cos(teta);
}
__device__ __forceinline__ double wave_target(const WaveDev& W, int comp, double x, double z, double t)
{
    z = z > -W.wD ? z : -W.wD;
    if(!(t >= 0.0 && t <= 1.0e20)) return 0.0;
    if(comp == 1) return W.cosg * wave_horz(W, x, z, t);
    if(comp == 2) return W.sing * wave_horz(W, x, z, t);
    const double teta = W.wk * x - W.ww * (t) + W.pshift;
    return W.ww * W.wa * (sinh(W.wk * (W.wdt + z)) / sinh(W.wk * W.wdt)) * sin(teta);
}

// eta(i, j) of wavegen_precalc_relax (iowave_precalc_relax.cpp:37-48) in the generation
// zone's columns; the other columns keep 0, as the reference's slice does.  One block per
// tank of the set.
__global__ void __launch_bounds__(NT) kb_wave_eta(TankSet b)
{
    CTank& T = tank_at(b, blockIdx.x);
    const WaveDev W = T.wave;
    if(!W.col || !W.gen) return;
    const long ncol = (long)T.g.nx * T.g.ny;
    const double t = T.tstep[6];
    const double* __restrict__ gen = W.col + WC_GEN * ncol;
    const double* __restrict__ xg = W.col + WC_XG4 * ncol;
    for(long q = threadIdx.x; q < ncol; q += NT)
        if(gen[q] != 0.0) W.eta[q] = (t >= 0.0 && t <= 1.0e20) ? W.wa * cos(W.wk * xg[q] - W.ww * (t) + W.pshift) : 0.0;
}

// iowave::u_relax, v_relax, w_relax, p_relax, phi_relax (iowave_relax.cpp:28-395) with the
// target values of wavegen_precalc_relax (iowave_precalc_relax.cpp:50-160), per cell of the
// zones' columns: velocity (u_ >= 0: the fields u_, v_, w_), pressure (p_ >= 0) and level set
// (f_ >= 0)
__global__ void __launch_bounds__(NT) kb_wave_relax(TankSet b, int u_, int v_, int w_, int p_, int f_)
{
#pragma unroll 1
    for(TileIter q = set_tile(b); q.ok; q.ok = false)
    {
        CTank& T = tank_at(b, q.p);
        const WaveDev W = T.wave;
        if(!W.col) continue;
        const DevGrid g = T.g;
        const long ncol = (long)g.nx * g.ny;
        const double t = T.tstep[6];
        const double ramp = wave_ramp(W, t);
        const double* __restrict__ phi = T.f[REEF_PHI];
        const TileWalk tw = tile_walk(T, q.t);
#pragma unroll 1
        for(int tt = tw.t; tt < tw.t1; ++tt)
        {
            int i, j, k;
            long c;
            if(!tile_cell(g, tw, tt, i, j, k, c)) continue;
            const long col = i + (long)g.nx * j;
            const bool gen = W.gen && W.col[WC_GEN * ncol + col] != 0.0;
            const bool beach = W.beach && W.col[WC_BEACH * ncol + col] != 0.0;
            if(!gen && !beach) continue;
            const double zp = g.ZP[k + MARGE];
            if(u_ >= 0)
                for(int comp = 1; comp <= 3; ++comp)
                {
                    if(!comp_cell(g, comp, i, j, k)) continue;
                    double* __restrict__ vel = T.f[comp == 1 ? u_ : comp == 2 ? v_ : w_];
                    const long off = comp == 1 ? 1 : comp == 2 ? g.sx : g.sxy;
                    const double phival = 0.5 * (phi[c] + phi[c + off]);
                    double H = 0.0, G = 0.0;
                    if(phival >= -W.psi) { H = 1.0; G = 1.0; }
                    if(phival < -W.epsi) { H = 0.0; G = 0.0; }
                    if(phival >= -W.epsi && phival < -W.psi)
                    {
                        H = 0.5 * (1.0 + phival / W.epsi + (1.0 / PI_REF) * sin((PI_REF * phival) / W.epsi));
                        G = H;
                    }
                    double v = vel[c];
                    if(gen)
                    {
                        // the target of wavegen_precalc_relax at this point
                        const double e0 = W.eta[col];
                        const double e1 = comp == 1 ? W.eta[col + 1] : comp == 2 ? W.eta[col + g.nx] : 0.0;
                        const double zloc = comp == 3 ? W.zn[k + 1 + MARGE] : zp;
                        const double emid = comp == 3 ? e0 : 0.5 * (e0 + e1);
                        const double fsfloc = emid + W.phimean;
                        const double z = zloc <= fsfloc ? zloc - W.phimean : emid;
                        const double xg = W.col[(comp == 1 ? WC_XG1 : comp == 2 ? WC_XG2 : WC_XG4) * ncol + col];
                        double val = 0.0;
                        if(zloc <= fsfloc + W.epsi) val = wave_target(W, comp, xg, z, t) + (comp == 1 ? W.Ui : 0.0);
                        const double r = W.col[(comp == 1 ? WC_R1WG : comp == 2 ? WC_R2WG : WC_R4WG) * ncol + col];
                        v = (1.0 - r) * ramp * val * H + r * H * v + (1.0 - G) * v;
                        if(comp == 1) v += (1.0 - r) * ramp * W.W50 * (1.0 - H) + r * (1.0 - H) * W.W50;
                    }
                    if(beach) v = W.col[(comp == 1 ? WC_R1NB : comp == 2 ? WC_R2NB : WC_R4NB) * ncol + col] * v;
                    vel[c] = v;
                }
            if(p_ >= 0 && beach)
            {
                double* __restrict__ press = T.f[p_];
                const double r = W.col[WC_R4NB * ncol + col];
                press[c] = (1.0 - r) * ((W.phimean - zp) * T.f[REEF_RO][c] * W.gabs) + r * press[c];
            }
            if(f_ >= 0)
            {
                double* __restrict__ f = T.f[f_];
                if(gen)
                {
                    const double r = W.col[WC_R4WG * ncol + col];
                    f[c] = (1.0 - r) * ramp * (W.eta[col] + W.phimean - zp) + r * f[c];
                }
                if(beach && W.beach_phi)
                {
                    const double r = W.col[WC_R4NB * ncol + col];
                    f[c] = (1.0 - r) * (W.phimean - zp) + r * f[c];
                }
            }
        }
    }
}

// the driver's p->simtime += p->dt after the step (driver_loop_cfd.cpp:100)
__global__ void kb_wave_time(TankSet b)
{
    CTank& T = tank_at(b, blockIdx.x);
    if(T.wave.col && threadIdx.x == 0) T.tstep[6] = T.tstep[6] + T.tstep[0];
}

}  // namespace bk

// ------------------------------------------------------------------ launchers
using namespace bk;

static inline dim3 tankgrid(const TankSet& b) { return dim3((unsigned)b.n); }
static inline dim3 setup_grid(const TankSet& b) { return dim3((unsigned)((b.n + 63) / 64)); }
// per-tile launch: 8 * ceil(tiles / 8) blocks
static inline dim3 tgrid(const TankSet& b) { return dim3((unsigned)(8 * std::max(1, (b.ntiles + 7) / 8))); }
// persistent launch over the work items: one block per work item, at most occ blocks per
// CU (the kernel's occupancy) on the 256 CUs, a multiple of 8
static inline dim3 pgrid_n(int nitems, int occ)
{
    return dim3((unsigned)std::max(8, std::min(256 * occ, (nitems + 7) / 8 * 8)));
}
static inline dim3 pgrid(const TankSet& b, int occ = OCC_FULL) { return pgrid_n(b.nitems, occ); }
// the launches of one kind of tank (set_items_of): plane-walked or the others
static inline dim3 pgrid_pw(const TankSet& b, int occ) { return pgrid_n(b.pw_items, occ); }
static inline dim3 pgrid_direct(const TankSet& b, int occ) { return pgrid_n(b.nitems - b.pw_items, occ); }

void bl_copy(const TankSet& b, int src, int dst, hipStream_t s)
{
    hipLaunchKernelGGL(kb_copy, tgrid(b), dim3(NT), 0, s, b, src, dst);
}
void bl_fill(const TankSet& b, int id, double v, hipStream_t s)
{
    hipLaunchKernelGGL(kb_fill, dim3(1024), dim3(NT), 0, s, b, id, v);
}
void bl_ghost(const TankSet& b, int prog, int id, hipStream_t s)
{
    hipLaunchKernelGGL(kb_ghost, dim3(256), dim3(NT), 0, s, b, prog, id);
}
void bl_hj_stage(const TankSet& b, int stage, int phi_s, int phi_n, int u, int v, int w, int out, double cfac,
                 hipStream_t s)
{
    hipLaunchKernelGGL(kb_hj_stage, tgrid(b), dim3(NT), 0, s, b, stage, phi_s, phi_n, u, v, w, out, cfac);
}
void bl_reini_stage(const TankSet& b, int stage, int fs, int f0, int out, double F43, int fsf, hipStream_t s)
{
    hipLaunchKernelGGL(kb_reini_stage, tgrid(b), dim3(NT), 0, s, b, stage, fs, f0, out, F43, fsf);
}
void bl_flux_acc(const TankSet& b, int ipol, int q, int u, int v, int w, int acc, hipStream_t s)
{
    hipLaunchKernelGGL(kb_flux_acc, tgrid(b), dim3(NT), 0, s, b, ipol, q, u, v, w, acc);
}
void bl_momentum(const TankSet& b, int comp, int stage, int qin, int qold, int u, int v, int w, double grav,
                 double cfac, int out, double rw, double ra, int diff, hipStream_t s)
{
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, tgrid(b), dim3(NT), 0, s, b, stage, qin, qold, u, v, w, grav, cfac, out, rw, ra);
    };
    if(comp == 1)
        diff == 2 ? go(kb_momentum<1, 2>) : diff ? go(kb_momentum<1, 1>) : go(kb_momentum<1, 0>);
    else if(comp == 2)
        diff == 2 ? go(kb_momentum<2, 2>) : diff ? go(kb_momentum<2, 1>) : go(kb_momentum<2, 0>);
    else
        diff == 2 ? go(kb_momentum<3, 2>) : diff ? go(kb_momentum<3, 1>) : go(kb_momentum<3, 0>);
}
void bl_heaviside(const TankSet& b, const double* prm4, hipStream_t s)
{
    hipLaunchKernelGGL(kb_heaviside, pgrid(b), dim3(NT), 0, s, b, prm4[0], prm4[1], prm4[2], prm4[3]);
    hipLaunchKernelGGL(kb_heaviside_fin, tankgrid(b), dim3(NT), 0, s, b, (double*)nullptr);
}

// ------------------------------------------------------------------ allocation-free scratch
// The reduction memory of the allocation-free fluid update (reef_fluid_update_noalloc): a
// file-scope device array instead of library-allocated memory.  Layout: the block partials
// (2 doubles per work item of the set, the slots kb_heaviside writes) from element 0, the
// set's volumes (2 doubles per tank) from element NOALLOC_VOL0.  Every GPU the library
// runs on holds its own copy until the process ends.  Host code never passes the array's
// name to a kernel (in host code the name is a stand-in, not a device address): it takes
// the address hipGetSymbolAddress gives for the current device, once per device.
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}

int noalloc_scratch(double** out)
{
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return REEF_E_HIP;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if(!g_scratch_ptr[dev])
    {
        void* ptr = nullptr;
        if(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return REEF_E_HIP;
        g_scratch_ptr[dev] = (double*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

void bl_heaviside_noalloc(const TankSet& b, const double* prm4, double* scratch_dev, hipStream_t s)
{
    TankSet t = b;
    t.partials = scratch_dev;
    hipLaunchKernelGGL(kb_heaviside, pgrid(t), dim3(NT), 0, s, t, prm4[0], prm4[1], prm4[2], prm4[3]);
    hipLaunchKernelGGL(kb_heaviside_fin, tankgrid(t), dim3(NT), 0, s, t, scratch_dev + NOALLOC_VOL0);
}
void bl_poisson_rhs(const TankSet& b, int u, int v, int w, double afac, hipStream_t s)
{
    hipLaunchKernelGGL(kb_poisson_rhs, tgrid(b), dim3(NT), 0, s, b, u, v, w, afac);
}
void bl_face_rho(const TankSet& b, double rw, double ra, int with_aii, hipStream_t s)
{
    hipLaunchKernelGGL(kb_face_rho, tgrid(b), dim3(NT), 0, s, b, rw, ra);
    if(with_aii) hipLaunchKernelGGL(kb_aii, tgrid(b), dim3(NT), 0, s, b);
}
void bl_correct(const TankSet& b, double rw, double ra, double afac, int u, int v, int w, hipStream_t s)
{
    hipLaunchKernelGGL(kb_correct, tgrid(b), dim3(NT), 0, s, b, rw, ra, afac, u, v, w);
}

void bl_bicg_init(const TankSet& b, int x, int rhs, double stop, int maxiter, hipStream_t s)
{
    hipLaunchKernelGGL(kb_setup, setup_grid(b), dim3(64), 0, s, b, stop, maxiter, 0, 0.0);
    hipLaunchKernelGGL(kb_bicg_init<double>, pgrid(b, OCC_A64), dim3(NT), 0, s, b, x, rhs);
    hipLaunchKernelGGL(kb_fin_init, tankgrid(b), dim3(NT), 0, s, b);
}

void bl_bicg_phase(const TankSet& b, int x, int rhs, int phase, int part, hipStream_t s)
{
    const dim3 grid = pgrid(b), grid64 = pgrid(b, OCC_A64);
    if(part == 0)
    {
        switch(phase)
        {
            case 0: hipLaunchKernelGGL(kb_bicg_p<double>, grid64, dim3(NT), 0, s, b, x, rhs); break;
            case 1:
                if(b.has_pw) hipLaunchKernelGGL((kb_bicg_v<double, true>), pgrid_pw(b, OCC_PW64), dim3(NT), 0, s, b, x, rhs);
                if(b.has_direct)
                    hipLaunchKernelGGL((kb_bicg_v<double, false>), pgrid_direct(b, OCC_A64), dim3(NT), 0, s, b, x, rhs);
                break;
            case 2:
                if(b.has_pw) hipLaunchKernelGGL((kb_bicg_t<double, true>), pgrid_pw(b, OCC_PW64), dim3(NT), 0, s, b, x, rhs);
                if(b.has_direct)
                    hipLaunchKernelGGL((kb_bicg_t<double, false>), pgrid_direct(b, OCC_A64), dim3(NT), 0, s, b, x, rhs);
                break;
            default: hipLaunchKernelGGL(kb_bicg_x<double>, grid, dim3(NT), 0, s, b, x, rhs); break;
        }
        return;
    }
    switch(phase)
    {
        case 0: break;   // a restart of the p phase is finalized by the v phase
        case 1: hipLaunchKernelGGL(kb_fin_v, tankgrid(b), dim3(NT), 0, s, b, 1); break;
        case 2: hipLaunchKernelGGL(kb_fin_t, tankgrid(b), dim3(NT), 0, s, b); break;
        default: hipLaunchKernelGGL(kb_fin_x, tankgrid(b), dim3(NT), 0, s, b); break;
    }
}

void bl_idiff(const TankSet& b, int comp, double alpha, int qin, int u, int v, int w, hipStream_t s)
{
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, tgrid(b), dim3(NT), 0, s, b, alpha, qin, u, v, w); };
    comp == 1 ? go(km_idiff<1>) : comp == 2 ? go(km_idiff<2>) : go(km_idiff<3>);
}

void bl_idiff_combine(const TankSet& b, int comp, int stage, int qold, int F, double cfac, int out, hipStream_t s)
{
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, tgrid(b), dim3(NT), 0, s, b, stage, qold, F, cfac, out); };
    comp == 1 ? go(km_combine<1>) : comp == 2 ? go(km_combine<2>) : go(km_combine<3>);
}

void bl_df_set(const TankSet& b, hipStream_t s)
{
    hipLaunchKernelGGL(km_df_set, setup_grid(b), dim3(64), 0, s, b);
}

void bl_mat_copy_x(const TankSet& b, int f, int var, int dir, hipStream_t s)
{
    hipLaunchKernelGGL(km_copy_x, tgrid(b), dim3(NT), 0, s, b, f, var, dir);
}

void bl_mat_init(const TankSet& b, int var, double stop, int maxiter, hipStream_t s)
{
    hipLaunchKernelGGL(kb_setup, setup_grid(b), dim3(64), 0, s, b, stop, maxiter, 0, 0.0);
    hipLaunchKernelGGL(km_init, pgrid(b, OCC_A64), dim3(NT), 0, s, b, var);
    hipLaunchKernelGGL(kb_fin_init, tankgrid(b), dim3(NT), 0, s, b);
}

void bl_mat_phase(const TankSet& b, int var, int phase, int part, hipStream_t s)
{
    const dim3 grid64 = pgrid(b, OCC_A64);
    if(part == 0)
    {
        switch(phase)
        {
            case 0: hipLaunchKernelGGL(km_p, grid64, dim3(NT), 0, s, b, var); break;
            case 1: hipLaunchKernelGGL(km_v, grid64, dim3(NT), 0, s, b, var); break;
            case 2: hipLaunchKernelGGL(km_t, grid64, dim3(NT), 0, s, b, var); break;
            default: hipLaunchKernelGGL(km_x, pgrid(b), dim3(NT), 0, s, b, var); break;
        }
        return;
    }
    switch(phase)
    {
        case 0: hipLaunchKernelGGL(kb_fin_p, tankgrid(b), dim3(NT), 0, s, b); break;
        case 1: hipLaunchKernelGGL(kb_fin_v, tankgrid(b), dim3(NT), 0, s, b, 0); break;
        case 2: hipLaunchKernelGGL(kb_fin_t, tankgrid(b), dim3(NT), 0, s, b); break;
        default: hipLaunchKernelGGL(kb_fin_x, tankgrid(b), dim3(NT), 0, s, b); break;
    }
}

void bl_coef32(const TankSet& b, hipStream_t s)
{
    hipLaunchKernelGGL(kb_coef32, tgrid(b), dim3(NT), 0, s, b);
}

void bl_mixed_init(const TankSet& b, int x, int rhs, double stop, int maxiter, double delta, hipStream_t s)
{
    hipLaunchKernelGGL(kb_setup, setup_grid(b), dim3(64), 0, s, b, stop, maxiter, 1, delta);
    hipLaunchKernelGGL(kb_mix_start, pgrid(b, OCC_A64), dim3(NT), 0, s, b, x, rhs);
    hipLaunchKernelGGL(kb_fin_mix_start, tankgrid(b), dim3(NT), 0, s, b);
}

void bl_mixed_phase(const TankSet& b, int x, int rhs, int phase, int part, hipStream_t s)
{
    const dim3 grid = pgrid(b);
    if(part == 0)
    {
        switch(phase)
        {
            case 0:
                if(b.has_pw) hipLaunchKernelGGL((kb_mix_v<true>), pgrid_pw(b, OCC_PW), dim3(NT), 0, s, b, x);
                if(b.has_direct) hipLaunchKernelGGL((kb_mix_v<false>), pgrid_direct(b, OCC_FULL), dim3(NT), 0, s, b, x);
                break;
            case 1:
                if(b.has_pw) hipLaunchKernelGGL((kb_bicg_t<float, true>), pgrid_pw(b, OCC_PW), dim3(NT), 0, s, b, x, rhs);
                if(b.has_direct)
                    hipLaunchKernelGGL((kb_bicg_t<float, false>), pgrid_direct(b, OCC_FULL), dim3(NT), 0, s, b, x, rhs);
                break;
            case 2: hipLaunchKernelGGL(kb_bicg_x<float>, grid, dim3(NT), 0, s, b, x, rhs); break;
            default: hipLaunchKernelGGL(kb_mix_ru, pgrid(b, OCC_A64), dim3(NT), 0, s, b, x, rhs); break;
        }
        return;
    }
    switch(phase)
    {
        case 0: hipLaunchKernelGGL(kb_fin_v, tankgrid(b), dim3(NT), 0, s, b, 0); break;
        case 1: hipLaunchKernelGGL(kb_fin_t, tankgrid(b), dim3(NT), 0, s, b); break;
        case 2: hipLaunchKernelGGL(kb_fin_x, tankgrid(b), dim3(NT), 0, s, b); break;
        default: break;   // the reliable-update phase is finalized by the next v phase
    }
}

void bl_mixed_settle(const TankSet& b, hipStream_t s)
{
    hipLaunchKernelGGL(kb_fin_mix_ru, tankgrid(b), dim3(NT), 0, s, b);
}

void bl_add32(const TankSet& b, int x, hipStream_t s)
{
    hipLaunchKernelGGL(kb_add32, tgrid(b), dim3(NT), 0, s, b, x);
}

void bl_timestep(const TankSet& b, int u, int v, int w, double N47, int ini, hipStream_t s)
{
    hipLaunchKernelGGL(kb_velmax, pgrid(b), dim3(NT), 0, s, b, u, v, w);
    hipLaunchKernelGGL(kb_fin_dt, tankgrid(b), dim3(NT), 0, s, b, N47, ini);
}

void bl_vortex_velocity(const TankSet& b, int comp, int id, double simtime, double DXM, hipStream_t s)
{
    hipLaunchKernelGGL(kb_vortex_velocity, tgrid(b), dim3(NT), 0, s, b, comp, id, simtime, DXM);
}
void bl_void_rk3(const TankSet& b, int comp, int id, double Fv, hipStream_t s)
{
    hipLaunchKernelGGL(kb_void_rk3, tgrid(b), dim3(NT), 0, s, b, comp, id, Fv);
}

void bl_wave_eta(const TankSet& b, hipStream_t s)
{
    hipLaunchKernelGGL(kb_wave_eta, tankgrid(b), dim3(NT), 0, s, b);
}
void bl_wave_relax(const TankSet& b, int u, int v, int w, int p, int f, hipStream_t s)
{
    hipLaunchKernelGGL(kb_wave_relax, tgrid(b), dim3(NT), 0, s, b, u, v, w, p, f);
}
void bl_wave_time(const TankSet& b, hipStream_t s)
{
    hipLaunchKernelGGL(kb_wave_time, tankgrid(b), dim3(64), 0, s, b);
}

void bl_gather_states(const TankSet& b, BiState* dst, hipStream_t s)
{
    hipLaunchKernelGGL(kb_gather_states, setup_grid(b), dim3(64), 0, s, b, dst);
}

void bl_residual(const TankSet& b, int x, int rhs, double* norms, hipStream_t s)
{
    hipLaunchKernelGGL(kb_defect, pgrid(b, OCC_A64), dim3(NT), 0, s, b, x, rhs);
    hipLaunchKernelGGL(kb_fin_norm, tankgrid(b), dim3(NT), 0, s, b, norms);
}

}  // namespace reef
