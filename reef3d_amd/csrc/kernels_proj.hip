// Ghost fill, fluid properties, pressure projection and the Jacobi-preconditioned
// BiCGStab of bicgstab_ijk on the device.
//
// Reductions: every block writes its partial sums, a one-block finalize kernel adds them
// in block order and updates the solver state in device memory, so the iteration never
// waits on the host.  Partial sums are formed in a fixed order, so results are
// run-to-run deterministic (they differ from the reference's serial i-j-k sum only by
// rounding).
#include "cell_ops.h"

namespace reef {


__device__ __forceinline__ bool tile_of_p(int nbx, int nby, int nbtot, int& bx, int& by, int& bz)
{
    const unsigned per = gridDim.x >> 3;
    const unsigned b = blockIdx.x;
    const unsigned lin = (b & 7u) * per + (b >> 3);
    if(lin >= (unsigned)nbtot) return false;
    bz = lin / (nbx * nby);
    const unsigned rem = lin - (unsigned)bz * nbx * nby;
    by = rem / nbx;
    bx = rem - (unsigned)by * nbx;
    return true;
}

static inline dim3 tiles(const DevGrid& g, int& nbx, int& nby, int& nbtot)
{
    nbx = (g.nx + TX - 1) / TX;
    nby = (g.ny + TY - 1) / TY;
    nbtot = nbx * nby * g.nz;
    return dim3((unsigned)((nbtot + 7) / 8 * 8));
}

// cell of this thread; false outside the interior or in padding blocks
__device__ __forceinline__ bool cell(const DevGrid& g, int nbx, int nby, int nbtot, int& i, int& j, int& k, long& c)
{
    int bx, by, bz;
    if(!tile_of_p(nbx, nby, nbtot, bx, by, bz)) return false;
    i = bx * TX + (threadIdx.x & (TX - 1));
    j = by * TY + (threadIdx.x >> 6);
    k = bz;
    if(i >= g.nx || j >= g.ny) return false;
    c = g.base + i + g.sx * j + g.sxy * k;
    return true;
}

// ------------------------------------------------------------------ ghost program
__global__ void k_ghost(const GhostOp* __restrict__ ops, int nops, double* __restrict__ f)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if(q >= nops) return;
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

void launch_ghost(const GhostOp* ops, int nops, double* f, hipStream_t s)
{
    if(nops <= 0) return;
    hipLaunchKernelGGL(k_ghost, dim3((nops + 255) / 256), dim3(256), 0, s, ops, nops, f);
}

// ------------------------------------------------------------------ simple kernels
__global__ void k_copy_interior(DevGrid g, int nbx, int nby, int nbtot, const double* __restrict__ src,
                                double* __restrict__ dst)
{
    int i, j, k;
    long c;
    if(!cell(g, nbx, nby, nbtot, i, j, k, c)) return;
    dst[c] = src[c];
}

void launch_copy_interior(const DevGrid& g, const double* src, double* dst, hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tiles(g, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_copy_interior, grid, dim3(NT), 0, s, g, nbx, nby, nbtot, src, dst);
}

__global__ void k_fill(double* f, long n, double v)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(q < n) f[q] = v;
}

void launch_fill(double* f, long n, double v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f, n, v);
}

// ------------------------------------------------------------------ fluid properties
// fluid_update_fsf::start (fluid_update_fsf.cpp:39-74): smoothed Heaviside, rho, nu and
// the volume sums (porosity 1)
__global__ void __launch_bounds__(NT) k_heaviside(DevGrid g, int nbx, int nby, int nbtot, const double* __restrict__ phi,
                                                  double* __restrict__ ro, double* __restrict__ visc, double rw,
                                                  double vw, double ra, double va, double* partials, double* vol_out,
                                                  unsigned int* ticket)
{
    int i, j, k;
    long c;
    double vals[2] = {0.0, 0.0};
    if(cell(g, nbx, nby, nbtot, i, j, k, c))
    {
        const double H = heaviside_H(phi[c], g.psi);
        ro[c] = rw * H + ra * (1.0 - H);
        visc[c] = vw * H + va * (1.0 - H);
        const double vol = g.DXN[i + MARGE] * g.DYN[j + MARGE] * g.DZN[k + MARGE];
        vals[0] = vol * (H - (1.0 - 1.0));
        vals[1] = vol * (1.0 - H - (1.0 - 1.0));
    }
    block_partials<2>(vals, partials);
}

__global__ void k_heaviside_fin(const double* __restrict__ partials, int nb, double* vol_out)
{
    double tot[2];
    sum_partials<2>(partials, nb, tot);
    if(threadIdx.x == 0)
    {
        vol_out[0] = tot[0];
        vol_out[1] = tot[1];
    }
}

void launch_heaviside(const DevGrid& g, const double* phi, double* ro, double* visc, const double* prm4,
                      double* partials, double* vol_out, unsigned int* ticket, hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tiles(g, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_heaviside, grid, dim3(NT), 0, s, g, nbx, nby, nbtot, phi, ro, visc, prm4[0], prm4[1], prm4[2],
                       prm4[3], partials, vol_out, ticket);
    hipLaunchKernelGGL(k_heaviside_fin, dim3(1), dim3(NT), 0, s, partials, (int)grid.x, vol_out);
}

// ------------------------------------------------------------------ projection
// pjm_corr::rhs (pjm_corr.cpp:151-176): -div(u*)/(alpha dt), in field layout
__global__ void k_poisson_rhs(DevGrid g, int nbx, int nby, int nbtot, const double* __restrict__ u,
                              const double* __restrict__ v, const double* __restrict__ w, double adt,
                              double* __restrict__ rhs)
{
    int i, j, k;
    long c;
    if(!cell(g, nbx, nby, nbtot, i, j, k, c)) return;
    rhs[c] = poisson_rhs_cell(g, u, v, w, adt, c, i, j, k);
}

void launch_poisson_rhs(const DevGrid& g, const double* u, const double* v, const double* w, double adt, double* rhs,
                        hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tiles(g, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_poisson_rhs, grid, dim3(NT), 0, s, g, nbx, nby, nbtot, u, v, w, adt, rhs);
}

// face densities of every "+" face of the interior cells and of the -1 layer faces, plus
// the Jacobi preconditioner aii = -1/(M.p + 1e-19) (bicgstab_ijk.cpp:356-364)
__global__ void k_face_rho(DevGrid g, int nbx, int nby, int nbtot, const double* __restrict__ phi, double rw,
                           double ra, double* __restrict__ rfx, double* __restrict__ rfy, double* __restrict__ rfz)
{
    int i, j, k;
    long c;
    if(!cell(g, nbx, nby, nbtot, i, j, k, c)) return;
    rfx[c] = roface(phi[c], phi[c + 1], g.psi, rw, ra);
    rfy[c] = roface(phi[c], phi[c + g.sx], g.psi, rw, ra);
    rfz[c] = roface(phi[c], phi[c + g.sxy], g.psi, rw, ra);
    if(i == 0) rfx[c - 1] = roface(phi[c], phi[c - 1], g.psi, rw, ra);
    if(j == 0) rfy[c - g.sx] = roface(phi[c], phi[c - g.sx], g.psi, rw, ra);
    if(k == 0) rfz[c - g.sxy] = roface(phi[c], phi[c - g.sxy], g.psi, rw, ra);
}

__global__ void k_aii(DevGrid g, int nbx, int nby, int nbtot, const double* __restrict__ rfx,
                      const double* __restrict__ rfy, const double* __restrict__ rfz, double* __restrict__ aii)
{
    int i, j, k;
    long c;
    if(!cell(g, nbx, nby, nbtot, i, j, k, c)) return;
    const Coupling q = couplings(g, rfx, rfy, rfz, c, i, j, k);
    aii[c] = -1.0 / (diag_of(g, q) + 1e-19);
}

void launch_face_rho(const DevGrid& g, const double* phi, double rw, double ra, double* rfx, double* rfy, double* rfz,
                     double* aii, hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tiles(g, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_face_rho, grid, dim3(NT), 0, s, g, nbx, nby, nbtot, phi, rw, ra, rfx, rfy, rfz);
    if(aii) hipLaunchKernelGGL(k_aii, grid, dim3(NT), 0, s, g, nbx, nby, nbtot, rfx, rfy, rfz, aii);
}

// pjm_corr::presscorr + ucorr/vcorr/wcorr (pjm_corr.cpp:123-149)
__global__ void k_correct(DevGrid g, int nbx, int nby, int nbtot, const double* __restrict__ pcorr,
                          const double* __restrict__ phi, double rw, double ra, double adt, double* __restrict__ u,
                          double* __restrict__ v, double* __restrict__ w, double* __restrict__ press)
{
    int i, j, k;
    long c;
    if(!cell(g, nbx, nby, nbtot, i, j, k, c)) return;
    correct_cell(g, pcorr, phi, rw, ra, adt, u, v, w, press, c, i, j, k);
}

void launch_correct(const DevGrid& g, const double* pcorr, const double* phi, double rw, double ra, double adt,
                    double* u, double* v, double* w, double* press, hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tiles(g, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_correct, grid, dim3(NT), 0, s, g, nbx, nby, nbtot, pcorr, phi, rw, ra, adt, u, v, w, press);
}

// ------------------------------------------------------------------ BiCGStab
// State machine of bicgstab_ijk::solve (bicgstab_ijk.cpp:125-278): mode 1 = a restart
// is pending (res_calc + restart label), 2 = p and ph were just set from the initial
// residual.
//
// The solver kernels run a fixed grid of RB blocks (8 per CU); block b works through a
// contiguous range of 64x4 tiles (z fastest) inside the slab of XCD b % 8, so stencil
// neighbours mostly hit L1 or that XCD's L2, and every reduction has RB partials.
//
// Every kernel is a template on the vector type V and the operator: V = double with the
// FP64 operator is the reference solver; V = float with the FP32 operator is the inner
// solve of the mixed-precision projection.  Dot products are accumulated in FP64 and the
// scalars (alpha, beta, omega) are formed in FP64 in both cases.
constexpr int RB = 2048;

struct TileWalk
{
    int t, t1, nbx, nby;
};

__device__ __forceinline__ TileWalk tile_walk(const DevGrid& g)
{
    TileWalk w;
    w.nbx = (g.nx + TX - 1) / TX;
    w.nby = (g.ny + TY - 1) / TY;
    const int ntiles = w.nbx * w.nby * g.nz;
    const int per_xcd_blocks = gridDim.x >> 3;
    const int xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
    const int tiles_xcd = (ntiles + 7) >> 3;
    const int tpb = (tiles_xcd + per_xcd_blocks - 1) / per_xcd_blocks;
    const int lo = xcd * tiles_xcd;
    const int hi = min(lo + tiles_xcd, ntiles);
    w.t = min(lo + loc * tpb, hi);
    w.t1 = min(w.t + tpb, hi);
    return w;
}

// tiles are numbered z fastest: a block walks up one 64x4 column, so the k-1 plane of a
// 7-point stencil is still in L1 from the previous tile
__device__ __forceinline__ bool tile_cell(const DevGrid& g, const TileWalk& w, int t, int& i, int& j, int& k, long& c)
{
    const int col = t / g.nz;
    k = t - col * g.nz;
    const int by = col / w.nbx, bx = col - by * w.nbx;
    i = bx * TX + (threadIdx.x & (TX - 1));
    j = by * TY + (threadIdx.x >> 6);
    if(i >= g.nx || j >= g.ny) return false;
    c = g.base + i + g.sx * j + g.sxy * k;
    return true;
}

__device__ __forceinline__ double sq(double a) { return a * a; }

// ---- plane walk of the 3-D stencil kernels.  A block's tiles are consecutive k planes
// of one 64 x 4 column (z fastest), so the block walks up the column: the stencil input
// of plane k (rows j-1 .. j+4, columns i-1 .. i+64) and the +x / +y face values sit in
// LDS, planes k-1 and k+1 of the own cell and the -z face stay in registers.  Every
// array is then read from HBM about once per cell; the LDS planes are double buffered,
// so one barrier per plane suffices.
template<class V>
struct PlaneLDS
{
    V x[TY + 2][TX + 2];   // stencil input, rows j-1..j+4, columns i-1..i+64
    V fx[TY][TX + 1];      // +x face of columns i-1..i+63
    V fy[TY + 1][TX];      // +y face of rows j-1..j+3
};

// load(cc, aux) returns the stencil input at element cc (aux: a by-product the cell
// epilogue wants, e.g. s = r - alpha v); cell(c, i, j, k, Ax, aux) consumes the result
template<class V, class Op, class Load, class Cell>
__device__ __forceinline__ void plane_walk(const DevGrid& g, const Op& op, Load load, Cell cell)
{
    __shared__ PlaneLDS<V> L[2];
    const V* __restrict__ fxa = face_x(op);
    const V* __restrict__ fya = face_y(op);
    const V* __restrict__ fza = face_z(op);
    const TileWalk tw = tile_walk(g);
    const int lane = threadIdx.x & (TX - 1), row = threadIdx.x >> 6;
    int curcol = -1, buf = 0;
    V xd = V(0), xc = V(0), fzd = V(0), auxc = V(0);
    for(int t = tw.t; t < tw.t1; ++t)
    {
        const int col = t / g.nz;
        const int k = t - col * g.nz;
        const int by = col / tw.nbx, bx = col - by * tw.nbx;
        const int i = bx * TX + lane, j = by * TY + row;
        const bool in = i < g.nx && j < g.ny;
        const long c = g.base + i + g.sx * j + g.sxy * k;
        if(col != curcol)
        {
            // first plane of this column in the block: prime the register window
            curcol = col;
            V dummy;
            xd = (in && k > 0) ? load(c - g.sxy, dummy) : V(0);
            xc = in ? load(c, auxc) : V(0);
            fzd = in ? fza[c - g.sxy] : V(0);
        }
        V auxu = V(0);
        const V xu = (in && k + 1 < g.nz) ? load(c + g.sxy, auxu) : V(0);
        PlaneLDS<V>& P = L[buf];
        P.x[row + 1][lane + 1] = xc;
        {
            V dummy;
            if(row == 0) P.x[0][lane + 1] = (i < g.nx && j > 0 && j - 1 < g.ny) ? load(c - g.sx, dummy) : V(0);
            if(row == TY - 1) P.x[TY + 1][lane + 1] = (i < g.nx && j + 1 < g.ny) ? load(c + g.sx, dummy) : V(0);
            if(lane == 0) P.x[row + 1][0] = (j < g.ny && i > 0) ? load(c - 1, dummy) : V(0);
            if(lane == TX - 1) P.x[row + 1][TX + 1] = (j < g.ny && i + 1 < g.nx) ? load(c + 1, dummy) : V(0);
        }
        const V f0 = in ? fxa[c] : V(0);
        const V f2 = in ? fya[c] : V(0);
        const V f4 = in ? fza[c] : V(0);
        P.fx[row][lane + 1] = f0;
        if(lane == 0) P.fx[row][0] = (j < g.ny) ? fxa[c - 1] : V(0);
        P.fy[row + 1][lane] = f2;
        if(row == 0) P.fy[0][lane] = (i < g.nx && j < g.ny) ? fya[c - g.sx] : V(0);
        __syncthreads();
        if(in)
        {
            const CouplingV<V> q = coupling_faces(g, op, f0, P.fx[row][lane], f2, P.fy[row][lane], f4, fzd, i, j, k);
            const V Ax = stencil7(g, q, xc, P.x[row + 1][lane + 2], P.x[row + 1][lane], P.x[row + 2][lane + 1],
                                  P.x[row][lane + 1], xu, xd, i, j, k);
            cell(c, i, j, k, Ax, auxc);
        }
        buf ^= 1;
        xd = xc;
        xc = xu;
        auxc = auxu;
        fzd = f4;
    }
}

// restart / initial residual: r = b - A x, r0 = p = r, ph = aii p, r_j = (r, r0)
// (restart label of bicgstab_ijk.cpp:136-148; res_calc :314-338 gives the same sums)
template<class V, class Op>
__device__ __forceinline__ void restart_tiles(const DevGrid& g, const Op& op, const BiVec<V>& b, double& rr)
{
    const V* __restrict__ rhs = b.rhs;
    const V* __restrict__ x = b.x;
    const V* __restrict__ aii = b.aii;
    V* __restrict__ r = b.r;
    V* __restrict__ r0 = b.r0;
    V* __restrict__ p = b.p;
    V* __restrict__ ph = b.ph;
    const TileWalk tw = tile_walk(g);
    for(int t = tw.t; t < tw.t1; ++t)
    {
        int i, j, k;
        long c;
        if(!tile_cell(g, tw, t, i, j, k, c)) continue;
        const CouplingV<V> q = coupling_of(g, op, c, i, j, k);
        const V rv = rhs[c] - (apply_A(g, q, x, c, i, j, k));
        r[c] = rv;
        r0[c] = rv;
        p[c] = rv;
        ph[c] = rv * aii[c];
        rr += sq(rv);
    }
}

template<class V, class Op>
__global__ void __launch_bounds__(NT, 8) k_bicg_init(DevGrid g, Op op, BiVec<V> b)
{
    double vals[1] = {0.0};
    restart_tiles(g, op, b, vals[0]);
    block_partials<1>(vals, b.partials);
}

__global__ void k_fin_init(const double* __restrict__ partials, BiState* st, int nb)
{
    double tot[1];
    sum_partials<1>(partials, nb, tot);
    if(threadIdx.x == 0)
    {
        st->r_j = tot[0];
        st->norm_r0 = sqrt(tot[0]);
        // a negative stop value asks for a criterion relative to the initial residual:
        // ||r||_2 / N < |stop| * ||b - A x0||_2 / N (= ||b||_2 / N for x0 = 0)
        if(st->stop < 0.0) st->stop = -st->stop * sqrt(tot[0]) / double(st->ncell);
        if(!((st->residual >= st->stop) && (st->iter < st->maxiter))) st->done = 1;
        st->mode = 2;
    }
}

// start of an iteration: either the pending restart, or p = r + beta (p - w v)
// (bicgstab_ijk.cpp:234-235); then ph = aii p (precon_solve)
template<class V, class Op>
__global__ void __launch_bounds__(NT, 8) k_bicg_p(DevGrid g, Op op, BiVec<V> b)
{
    BiState* st = b.st;
    if(st->done) return;
    if(st->mode == 1)
    {
        double vals[1] = {0.0};
        restart_tiles(g, op, b, vals[0]);
        block_partials<1>(vals, b.partials);
        return;
    }
    if(st->mode == 2) return;     // p = r and ph were set by the initial residual
    const V beta = (V)st->beta, w = (V)st->w;
    const V* __restrict__ r = b.r;
    const V* __restrict__ v = b.v;
    const V* __restrict__ aii = b.aii;
    V* __restrict__ p = b.p;
    V* __restrict__ ph = b.ph;
    const TileWalk tw = tile_walk(g);
    for(int t = tw.t; t < tw.t1; ++t)
    {
        int i, j, k;
        long c;
        if(!tile_cell(g, tw, t, i, j, k, c)) continue;
        const V pn = r[c] + beta * (p[c] - w * v[c]);
        p[c] = pn;
        ph[c] = pn * aii[c];
    }
}

// totals of a restart done by k_bicg_p: ++solveriter, residual = |r|/N (bicgstab_ijk.cpp:183-191)
__global__ void k_fin_p(const double* __restrict__ partials, BiState* st, int nb)
{
    if(st->done || st->mode != 1) return;
    double tot[1];
    sum_partials<1>(partials, nb, tot);
    if(threadIdx.x == 0)
    {
        st->r_j = tot[0];
        st->norm_r0 = sqrt(tot[0]);
        st->residual = sqrt(tot[0]) / double(st->ncell);
        st->iter += 1;
        if(!((st->residual >= st->stop) && (st->iter < st->maxiter))) st->done = 1;
        st->mode = 2;
    }
}

// v = A ph; sigma = (v, r0), |v|, |r|; alpha and the restart tests (bicgstab_ijk.cpp:151-191)
template<class V, class Op>
__global__ void __launch_bounds__(NT, 8) k_bicg_v(DevGrid g, Op op, BiVec<V> b)
{
    BiState* st = b.st;
    if(st->done || st->mode == 1 || st->mode == 3) return;
    double vals[3] = {0.0, 0.0, 0.0};
    const V* __restrict__ ph = b.ph;
    const V* __restrict__ r = b.r;
    const V* __restrict__ r0 = b.r0;
    V* __restrict__ vo = b.v;
    auto epilogue = [&](long c, V v) {
        vo[c] = v;
        const double vd = v, rc = r[c];
        vals[0] += vd * double(r0[c]);
        vals[1] += vd * vd;
        vals[2] += rc * rc;
    };
    if(g.jdir == 1)
    {
        plane_walk<V>(g, op, [&](long cc, V&) { return ph[cc]; },
                      [&](long c, int, int, int, V Ax, V) { epilogue(c, Ax); });
        block_partials_w<3>(vals, b.partials);
        return;
    }
    const TileWalk tw = tile_walk(g);
    for(int t = tw.t; t < tw.t1; ++t)
    {
        int i, j, k;
        long c;
        if(!tile_cell(g, tw, t, i, j, k, c)) continue;
        const CouplingV<V> q = coupling_of(g, op, c, i, j, k);
        epilogue(c, apply_A(g, q, ph, c, i, j, k));
    }
    block_partials_w<3>(vals, b.partials);
}

__global__ void k_fin_v(const double* __restrict__ partials, BiState* st, int nb)
{
    if(st->done || st->mode == 1 || st->mode == 3) return;
    double tot[3];
    sum_partials<3>(partials, nb, tot);
    if(threadIdx.x == 0)
    {
        const double sigma = tot[0];
        const double norm_vj = sqrt(tot[1]);
        const double norm_rj = sqrt(tot[2]);
        const double alpha = st->r_j / sigma;
        st->sigma = sigma;
        st->norm_vj = norm_vj;
        st->norm_rj = norm_rj;
        st->alpha = alpha;
        if(fabs(sigma) <= (1.0e-12 * (norm_vj * st->norm_r0)))
            st->mode = 1;
        else if((fabs(alpha) * norm_vj / (norm_rj == 0 ? 1.0e-15 : norm_rj)) <= 0.08)
            st->mode = 1;
        else
            st->mode = 0;
    }
}

// s = r - alpha v, sh = aii s, t = A sh; (t,s), (t,t), (s,s) (bicgstab_ijk.cpp:193-221)
template<class V, class Op>
__global__ void __launch_bounds__(NT, 8) k_bicg_t(DevGrid g, Op op, BiVec<V> b)
{
    BiState* st = b.st;
    if(st->done || st->mode == 1 || st->mode == 3) return;
    double vals[3] = {0.0, 0.0, 0.0};
    const V alpha = (V)st->alpha;
    const V* __restrict__ r = b.r;
    const V* __restrict__ v = b.v;
    const V* __restrict__ a = b.aii;
    V* __restrict__ to = b.t;
    auto epilogue = [&](long c, V t, V s) {
        to[c] = t;
        const double td = t, sd = s;
        vals[0] += td * sd;
        vals[1] += td * td;
        vals[2] += sd * sd;
    };
    // sh at an element (0 outside the domain is applied by the stencil)
    auto shv = [&](long cc, V& sv) {
        sv = r[cc] - alpha * v[cc];
        return sv * a[cc];
    };
    if(g.jdir == 1)
    {
        plane_walk<V>(g, op, shv, [&](long c, int, int, int, V Ax, V s) { epilogue(c, Ax, s); });
        block_partials_w<3>(vals, b.partials);
        return;
    }
    const TileWalk tw = tile_walk(g);
    for(int tt = tw.t; tt < tw.t1; ++tt)
    {
        int i, j, k;
        long c;
        if(!tile_cell(g, tw, tt, i, j, k, c)) continue;
        const CouplingV<V> q = coupling_of(g, op, c, i, j, k);
        V s;
        const V shc = shv(c, s);
        V dummy;
        const V t = apply_A_fn(g, q, [&](long cc) { return cc == c ? shc : shv(cc, dummy); }, c, i, j, k);
        epilogue(c, t, s);
    }
    block_partials_w<3>(vals, b.partials);
}

__global__ void k_fin_t(const double* __restrict__ partials, BiState* st, int nb)
{
    if(st->done || st->mode == 1 || st->mode == 3) return;
    double tot[3];
    sum_partials<3>(partials, nb, tot);
    if(threadIdx.x == 0)
    {
        st->w1 = tot[0];
        st->w2 = tot[1];
        st->norm_sj = sqrt(tot[2]);
        st->branch = st->norm_sj > st->stop ? 1 : 0;
        st->w = st->branch ? tot[0] / (tot[1] == 0 ? 1.0e-15 : tot[1]) : st->w;
    }
}

// x and r updates, r_j1 = (r, r0), beta, residual = |r| / N (bicgstab_ijk.cpp:223-268)
template<class V, bool BRANCH>
__device__ __forceinline__ void x_tiles(const DevGrid& g, const BiVec<V>& b, V alpha, V w, double (&vals)[2])
{
    const V* __restrict__ v = b.v;
    const V* __restrict__ ph = b.ph;
    const V* __restrict__ a = b.aii;
    const V* __restrict__ tt = b.t;
    const V* __restrict__ r0 = b.r0;
    V* __restrict__ r = b.r;
    V* __restrict__ x = b.x;
    const TileWalk tw = tile_walk(g);
    for(int t = tw.t; t < tw.t1; ++t)
    {
        int i, j, k;
        long c;
        if(!tile_cell(g, tw, t, i, j, k, c)) continue;
        const V s = r[c] - alpha * v[c];
        V rn;
        if(BRANCH)
        {
            const V sh = s * a[c];
            x[c] += alpha * ph[c] + w * sh;
            rn = s - w * tt[c];
        }
        else
        {
            x[c] += alpha * ph[c];
            rn = s;
        }
        r[c] = rn;
        const double rd = rn;
        vals[0] += rd * double(r0[c]);
        vals[1] += rd * rd;
    }
}

template<class V>
__global__ void __launch_bounds__(NT, 8) k_bicg_x(DevGrid g, BiVec<V> b)
{
    BiState* st = b.st;
    if(st->done || st->mode == 1 || st->mode == 3) return;
    double vals[2] = {0.0, 0.0};
    const V alpha = (V)st->alpha, w = (V)st->w;
    const int branch = st->branch;
    if(branch)
        x_tiles<V, true>(g, b, alpha, w, vals);
    else
        x_tiles<V, false>(g, b, alpha, w, vals);
    block_partials<2>(vals, b.partials);
}

__global__ void k_fin_x(const double* __restrict__ partials, BiState* st, int nb)
{
    if(st->done || st->mode == 1 || st->mode == 3) return;
    const double alpha = st->alpha, w = st->w;
    const int branch = st->branch;
    double tot[2];
    sum_partials<2>(partials, nb, tot);
    if(threadIdx.x == 0)
    {
        const double r_j1 = tot[0];
        st->r_j_prev = st->r_j;
        if(branch) st->beta = alpha * r_j1 / (w * st->r_j == 0 ? 1.0e-15 : (w * st->r_j));
        st->r_j1 = r_j1;
        st->r_j = r_j1;
        st->residual = sqrt(tot[1]) / double(st->ncell);
        st->iter += 1;
        st->mode = 0;
        if(!st->mixed)
        {
            if(!((st->residual >= st->stop) && (st->iter < st->maxiter))) st->done = 1;
            return;
        }
        // FP32 iteration: the solve only ends on a true FP64 residual.  A step that took
        // the norm_sj <= stop exit, or any restart, is continued by a restart from the true
        // residual; convergence of the iterated residual, the iteration limit, or a drop
        // below delta * (largest ||r|| since the last update) asks for a reliable update.
        const double rn = sqrt(tot[1]);
        if(rn > st->rmax) st->rmax = rn;
        if(!branch)
            st->mode = 1;
        else if(!(st->residual >= st->stop) || st->iter >= st->maxiter || rn < st->delta * st->rmax)
            st->mode = 3;
    }
}

__global__ void k_bicg_setup(BiState* st, double stop, int maxiter, long ncell)
{
    st->r_j = st->norm_r0 = st->alpha = st->w = st->beta = 0.0;
    st->sigma = st->norm_vj = st->norm_rj = st->norm_sj = st->w1 = st->w2 = st->r_j1 = 0.0;
    st->residual = 1.0e9;
    st->stop = stop;
    st->iter = 0;
    st->maxiter = maxiter;
    st->done = 0;
    st->mode = 0;
    st->branch = 0;
    st->mixed = 0;
    st->r_j_prev = st->rmax = st->delta = st->res0 = 0.0;
    st->pending = 0;
    st->nru = 0;
    st->ncell = ncell;
}

static inline unsigned rblocks(const DevGrid& g)
{
    const long ntiles = (long)((g.nx + TX - 1) / TX) * ((g.ny + TY - 1) / TY) * g.nz;
    long nb = ntiles < RB ? ntiles : RB;
    return (unsigned)((nb + 7) / 8 * 8);
}

template<class V, class Op>
static void bicg_init_t(const DevGrid& g, const Op& op, const BiVec<V>& b, double stop, int maxiter, hipStream_t s)
{
    const unsigned nb = rblocks(g);
    hipLaunchKernelGGL(k_bicg_setup, dim3(1), dim3(1), 0, s, b.st, stop, maxiter, (long)g.nx * g.ny * g.nz);
    hipLaunchKernelGGL((k_bicg_init<V, Op>), dim3(nb), dim3(NT), 0, s, g, op, b);
    hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(NT), 0, s, (const double*)b.partials, b.st, (int)nb);
}

// part 0 launches the phase's grid kernel, part 1 its one-block finalize kernel
template<class V, class Op>
static void bicg_phase_t(const DevGrid& g, const Op& op, const BiVec<V>& b, int phase, int part, hipStream_t s)
{
    const dim3 grid(rblocks(g));
    const int nb = (int)grid.x;
    const double* part_sums = b.partials;
    if(part == 0)
    {
        switch(phase)
        {
            case 0: hipLaunchKernelGGL((k_bicg_p<V, Op>), grid, dim3(NT), 0, s, g, op, b); break;
            case 1: hipLaunchKernelGGL((k_bicg_v<V, Op>), grid, dim3(NT), 0, s, g, op, b); break;
            case 2: hipLaunchKernelGGL((k_bicg_t<V, Op>), grid, dim3(NT), 0, s, g, op, b); break;
            default: hipLaunchKernelGGL((k_bicg_x<V>), grid, dim3(NT), 0, s, g, b); break;
        }
        return;
    }
    switch(phase)
    {
        case 0: hipLaunchKernelGGL(k_fin_p, dim3(1), dim3(NT), 0, s, part_sums, b.st, nb); break;
        case 1: hipLaunchKernelGGL(k_fin_v, dim3(1), dim3(NT), 0, s, part_sums, b.st, nb); break;
        case 2: hipLaunchKernelGGL(k_fin_t, dim3(1), dim3(NT), 0, s, part_sums, b.st, nb); break;
        default: hipLaunchKernelGGL(k_fin_x, dim3(1), dim3(NT), 0, s, part_sums, b.st, nb); break;
    }
}

void launch_bicg_init(const DevGrid& g, const SolverOp64& op, const BiVec<double>& b, double stop, int maxiter,
                      hipStream_t s)
{
    bicg_init_t(g, op, b, stop, maxiter, s);
}
void launch_bicg_init(const DevGrid& g, const SolverOp32& op, const BiVec<float>& b, double stop, int maxiter,
                      hipStream_t s)
{
    bicg_init_t(g, op, b, stop, maxiter, s);
}
void launch_bicg_phase(const DevGrid& g, const SolverOp64& op, const BiVec<double>& b, int phase, int part,
                       hipStream_t s)
{
    bicg_phase_t(g, op, b, phase, part, s);
}
void launch_bicg_phase(const DevGrid& g, const SolverOp32& op, const BiVec<float>& b, int phase, int part,
                       hipStream_t s)
{
    bicg_phase_t(g, op, b, phase, part, s);
}

// ------------------------------------------------------------------ mixed precision
// FP32 operator and preconditioner of the inner solve from the FP64 face densities:
// kx = 1/(rho_f DXP) on every "+x" face of the interior and of the -1 layer (ky, kz
// likewise), aii = -1/(M.p + 1e-19) from the FP64 diagonal, rounded once.
__global__ void k_coef32(DevGrid g, int nbx, int nby, int nbtot, const double* __restrict__ rfx,
                         const double* __restrict__ rfy, const double* __restrict__ rfz, float* __restrict__ kx,
                         float* __restrict__ ky, float* __restrict__ kz, float* __restrict__ aii)
{
    int i, j, k;
    long c;
    if(!cell(g, nbx, nby, nbtot, i, j, k, c)) return;
    const int ip = i + MARGE, jp = j + MARGE, kp = k + MARGE;
    kx[c] = (float)(1.0 / (rfx[c] * g.DXP[ip]));
    ky[c] = (float)(1.0 / (rfy[c] * g.DYP[jp]));
    kz[c] = (float)(1.0 / (rfz[c] * g.DZP[kp]));
    if(i == 0) kx[c - 1] = (float)(1.0 / (rfx[c - 1] * g.DXP[ip - 1]));
    if(j == 0) ky[c - g.sx] = (float)(1.0 / (rfy[c - g.sx] * g.DYP[jp - 1]));
    if(k == 0) kz[c - g.sxy] = (float)(1.0 / (rfz[c - g.sxy] * g.DZP[kp - 1]));
    const Coupling q = couplings(g, rfx, rfy, rfz, c, i, j, k);
    aii[c] = (float)(-1.0 / (diag_of(g, q) + 1e-19));
}

void launch_coef32(const DevGrid& g, const double* rfx, const double* rfy, const double* rfz, float* kx, float* ky,
                   float* kz, float* aii, hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tiles(g, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_coef32, grid, dim3(NT), 0, s, g, nbx, nby, nbtot, rfx, rfy, rfz, kx, ky, kz, aii);
}

// ||b - A x||_2 with the FP64 operator (res_calc, bicgstab_ijk.cpp:314-338, without 1/N)
__global__ void __launch_bounds__(NT, 8) k_defect(DevGrid g, SolverOp64 op, const double* __restrict__ x,
                                                   const double* __restrict__ rhs, double* __restrict__ partials)
{
    double vals[1] = {0.0};
    const TileWalk tw = tile_walk(g);
    for(int t = tw.t; t < tw.t1; ++t)
    {
        int i, j, k;
        long c;
        if(!tile_cell(g, tw, t, i, j, k, c)) continue;
        const Coupling q = coupling_of(g, op, c, i, j, k);
        const double rv = rhs[c] - apply_A(g, q, x, c, i, j, k);
        vals[0] += rv * rv;
    }
    block_partials<1>(vals, partials);
}

__global__ void k_fin_norm(const double* __restrict__ partials, int nb, double* __restrict__ out)
{
    double tot[1];
    sum_partials<1>(partials, nb, tot);
    if(threadIdx.x == 0) out[0] = sqrt(tot[0]);
}

void launch_defect(const DevGrid& g, const SolverOp64& op, const double* x, const double* rhs, double* partials,
                   double* norm_out, hipStream_t s)
{
    const unsigned nb = rblocks(g);
    hipLaunchKernelGGL(k_defect, dim3(nb), dim3(NT), 0, s, g, op, x, rhs, partials);
    hipLaunchKernelGGL(k_fin_norm, dim3(1), dim3(NT), 0, s, (const double*)partials, (int)nb, norm_out);
}

// ---- FP32 BiCGStab with FP64 reliable updates.  The iteration is bicgstab_ijk's on the
// FP32 operator; x32 holds the solution increment since the last reliable update, x64 the
// rest.  A reliable update replaces the iterated residual by the true one,
// r = b - A64 (x64 + x32), keeps the search direction and recomputes beta from the new
// (r0, r); a restart (the reference's tests, or the norm_sj <= stop exit) also resets
// r0 = p = r.  The solve ends only on a true FP64 residual.

// initial residual in FP64: r = b - A64 x64; r0 = p = r, ph = aii r, x32 = 0
__global__ void __launch_bounds__(NT, 8) k_mix_start(DevGrid g, SolverOp64 op, MixedVec m, BiVec<float> b)
{
    double vals[2] = {0.0, 0.0};
    const TileWalk tw = tile_walk(g);
    for(int t = tw.t; t < tw.t1; ++t)
    {
        int i, j, k;
        long c;
        if(!tile_cell(g, tw, t, i, j, k, c)) continue;
        const Coupling q = coupling_of(g, op, c, i, j, k);
        const double rv = m.b64[c] - apply_A(g, q, (const double*)m.x64, c, i, j, k);
        const float r = (float)rv;
        b.r[c] = r;
        b.r0[c] = r;
        b.p[c] = r;
        b.ph[c] = r * b.aii[c];
        b.x[c] = 0.0f;
        vals[0] += rv * rv;
        vals[1] += double(r) * double(r);
    }
    block_partials<2>(vals, b.partials);
}

__global__ void k_mix_setup(BiState* st, double stop, int maxiter, long ncell, double delta)
{
    st->r_j = st->norm_r0 = st->alpha = st->w = st->beta = 0.0;
    st->sigma = st->norm_vj = st->norm_rj = st->norm_sj = st->w1 = st->w2 = st->r_j1 = 0.0;
    st->residual = 1.0e9;
    st->stop = stop;
    st->iter = 0;
    st->maxiter = maxiter;
    st->done = 0;
    st->mode = 0;
    st->branch = 0;
    st->mixed = 1;
    st->r_j_prev = st->rmax = st->res0 = 0.0;
    st->delta = delta;
    st->pending = 0;
    st->nru = 0;
    st->ncell = ncell;
}

__global__ void k_fin_mix_start(const double* __restrict__ partials, BiState* st, int nb)
{
    double tot[2];
    sum_partials<2>(partials, nb, tot);
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

// start of an FP32 iteration: x64 += x32 after a reliable update, then the p update
__global__ void __launch_bounds__(NT, 8) k_mix_p(DevGrid g, MixedVec m, BiVec<float> b)
{
    BiState* st = b.st;
    if(st->done || st->mode == 1 || st->mode == 3) return;
    const int upd = st->mode == 0, pend = st->pending;
    if(!upd && !pend) return;
    const float beta = (float)st->beta, w = (float)st->w;
    const float* __restrict__ r = b.r;
    const float* __restrict__ v = b.v;
    const float* __restrict__ aii = b.aii;
    float* __restrict__ p = b.p;
    float* __restrict__ ph = b.ph;
    float* __restrict__ x32 = b.x;
    double* __restrict__ x64 = m.x64;
    const TileWalk tw = tile_walk(g);
    for(int t = tw.t; t < tw.t1; ++t)
    {
        int i, j, k;
        long c;
        if(!tile_cell(g, tw, t, i, j, k, c)) continue;
        if(pend)
        {
            x64[c] += (double)x32[c];
            x32[c] = 0.0f;
        }
        if(upd)
        {
            const float pn = r[c] + beta * (p[c] - w * v[c]);
            p[c] = pn;
            ph[c] = pn * aii[c];
        }
    }
}

__global__ void k_fin_mix_p(BiState* st)
{
    if(st->done || st->mode == 1 || st->mode == 3) return;
    st->pending = 0;
}

// reliable update (mode 3) or restart (mode 1): r = b - A64 (x64 + x32) in FP64, rounded
// to the FP32 residual; a restart also sets r0 = p = r, ph = aii r
__global__ void __launch_bounds__(NT, 8) k_mix_ru(DevGrid g, SolverOp64 op, MixedVec m, BiVec<float> b)
{
    BiState* st = b.st;
    if(st->done || (st->mode != 1 && st->mode != 3)) return;
    const int restart = st->mode == 1;
    double vals[2] = {0.0, 0.0};
    const double* __restrict__ x64 = m.x64;
    const float* __restrict__ x32 = b.x;
    const TileWalk tw = tile_walk(g);
    for(int t = tw.t; t < tw.t1; ++t)
    {
        int i, j, k;
        long c;
        if(!tile_cell(g, tw, t, i, j, k, c)) continue;
        const Coupling q = coupling_of(g, op, c, i, j, k);
        const double rv =
            m.b64[c] - apply_A_fn(g, q, [&](long cc) { return x64[cc] + (double)x32[cc]; }, c, i, j, k);
        const float r = (float)rv;
        b.r[c] = r;
        float r0v;
        if(restart)
        {
            b.r0[c] = r;
            b.p[c] = r;
            b.ph[c] = r * b.aii[c];
            r0v = r;
        }
        else
            r0v = b.r0[c];
        vals[0] += rv * rv;
        vals[1] += double(r0v) * double(r);
    }
    block_partials<2>(vals, b.partials);
}

__global__ void k_fin_mix_ru(const double* __restrict__ partials, BiState* st, int nb)
{
    if(st->done || (st->mode != 1 && st->mode != 3)) return;
    double tot[2];
    sum_partials<2>(partials, nb, tot);
    if(threadIdx.x == 0)
    {
        const double rn = sqrt(tot[0]);
        st->residual = rn / double(st->ncell);
        st->nru += 1;
        st->pending = 1;
        st->rmax = rn;
        if(st->mode == 1) st->iter += 1;     // a restart counts as an iteration (bicgstab_ijk.cpp:183-191)
        if(!((st->residual >= st->stop) && (st->iter < st->maxiter)))
        {
            st->done = 1;
            return;
        }
        if(st->mode == 1)
        {
            st->r_j = tot[1];
            st->norm_r0 = sqrt(tot[1]);
            st->mode = 2;
        }
        else
        {
            const double wr = st->w * st->r_j_prev;
            st->beta = st->alpha * tot[1] / (wr == 0 ? 1.0e-15 : wr);
            st->r_j = tot[1];
            st->mode = 0;
        }
    }
}

void launch_mixed_init(const DevGrid& g, const SolverOp64& op64, const MixedVec& m, const BiVec<float>& b32,
                       double stop, int maxiter, double delta, hipStream_t s)
{
    const unsigned nb = rblocks(g);
    hipLaunchKernelGGL(k_mix_setup, dim3(1), dim3(1), 0, s, b32.st, stop, maxiter, (long)g.nx * g.ny * g.nz, delta);
    hipLaunchKernelGGL(k_mix_start, dim3(nb), dim3(NT), 0, s, g, op64, m, b32);
    hipLaunchKernelGGL(k_fin_mix_start, dim3(1), dim3(NT), 0, s, (const double*)b32.partials, b32.st, (int)nb);
}

void launch_mixed_phase(const DevGrid& g, const SolverOp64& op64, const SolverOp32& op32, const MixedVec& m,
                        const BiVec<float>& b32, int phase, int part, hipStream_t s)
{
    const dim3 grid(rblocks(g));
    const int nb = (int)grid.x;
    if(phase == 0)
    {
        if(part == 0)
            hipLaunchKernelGGL(k_mix_p, grid, dim3(NT), 0, s, g, m, b32);
        else
            hipLaunchKernelGGL(k_fin_mix_p, dim3(1), dim3(1), 0, s, b32.st);
        return;
    }
    if(phase == 4)
    {
        if(part == 0)
            hipLaunchKernelGGL(k_mix_ru, grid, dim3(NT), 0, s, g, op64, m, b32);
        else
            hipLaunchKernelGGL(k_fin_mix_ru, dim3(1), dim3(NT), 0, s, (const double*)b32.partials, b32.st, nb);
        return;
    }
    bicg_phase_t(g, op32, b32, phase, part, s);
}

// x += d, d = 0 (FP64 accumulation of the FP32 increment)
__global__ void k_add32(DevGrid g, int nbx, int nby, int nbtot, double* __restrict__ x, float* __restrict__ d)
{
    int i, j, k;
    long c;
    if(!cell(g, nbx, nby, nbtot, i, j, k, c)) return;
    x[c] += (double)d[c];
    d[c] = 0.0f;
}

void launch_add32(const DevGrid& g, double* x, float* d, hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tiles(g, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_add32, grid, dim3(NT), 0, s, g, nbx, nby, nbtot, x, d);
}

// ------------------------------------------------------------------ level-set benchmark
// benchmark_vortex3D::start (benchmark_vortex3D.cpp:67-96)
__global__ void k_vortex_velocity(DevGrid g, int nbx, int nby, int nbtot, int comp, double* __restrict__ q,
                                  double simtime, double DXM)
{
    int i, j, k;
    long c;
    if(!cell(g, nbx, nby, nbtot, i, j, k, c)) return;
    if((comp == 1 && i >= g.nx - 1) || (comp == 2 && j >= g.ny - 1) || (comp == 3 && k >= g.nz - 1)) return;
    const double x = g.XP[i + MARGE], y = g.YP[j + MARGE], z = g.ZP[k + MARGE];
    const double ct = cos((PI_REF * simtime) / 3.0);
    if(comp == 1)
    {
        const double xc = x + 0.5 * DXM;
        q[c] = 2.0 * pow(sin(PI_REF * xc), 2.0) * sin(2.0 * PI_REF * y) * sin(2.0 * PI_REF * z) * ct;
    }
    else if(comp == 2)
    {
        const double yc = y + 0.5 * DXM;
        q[c] = -sin(2.0 * PI_REF * x) * pow(sin(PI_REF * yc), 2.0) * sin(2.0 * PI_REF * z) * ct;
    }
    else
    {
        const double zc = z + 0.5 * DXM;
        q[c] = -sin(2.0 * PI_REF * x) * sin(2.0 * PI_REF * y) * pow(sin(PI_REF * zc), 2.0) * ct;
    }
}

void launch_vortex_velocity(const DevGrid& g, int comp, double* q, double simtime, double DXM, hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tiles(g, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_vortex_velocity, grid, dim3(NT), 0, s, g, nbx, nby, nbtot, comp, q, simtime, DXM);
}

// momentum_RK3 with void convection, diffusion and pressure (momentum_RK3.cpp:70-326):
// the three stages and the zero forcing of one component, F = 0 + g
__global__ void k_void_rk3(DevGrid g, int nbx, int nby, int nbtot, int comp, double* __restrict__ q, double dt,
                           double Fv)
{
    int i, j, k;
    long c;
    if(!cell(g, nbx, nby, nbtot, i, j, k, c)) return;
    if((comp == 1 && i >= g.nx - 1) || (comp == 2 && j >= g.ny - 1) || (comp == 3 && k >= g.nz - 1)) return;
    const double q0 = q[c];
    double a1 = q0 + dt * 1.0 * Fv;
    a1 += 1.0 * dt * 1.0 * 0.0;
    double a2 = 0.75 * q0 + 0.25 * a1 + 0.25 * dt * 1.0 * Fv;
    a2 += 0.25 * dt * 1.0 * 0.0;
    double qn = (1.0 / 3.0) * q0 + (2.0 / 3.0) * a2 + (2.0 / 3.0) * dt * 1.0 * Fv;
    qn += (2.0 / 3.0) * dt * 1.0 * 0.0;
    q[c] = qn;
}

void launch_void_rk3(const DevGrid& g, int comp, double* q, double dt, double Fv, hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tiles(g, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_void_rk3, grid, dim3(NT), 0, s, g, nbx, nby, nbtot, comp, q, dt, Fv);
}

}  // namespace reef
