// FP64 stencil kernels of the level-set and momentum convection: HJ-WENO5 transport
// stages of phi, the reinitialisation sub-stages, and the conservative WENO5 flux of the
// staggered velocities fused with the momentum_FC3 stage update.
//
// Every kernel reproduces the reference's per-cell arithmetic in the reference's
// operation order (the library is built with -ffp-contract=off), so results are
// bit-identical to the CPU path on the same inputs.
#include "cell_ops.h"

namespace reef {

// Linear block id -> tile, remapped so that each of the 8 XCDs works on one contiguous
// slab of the domain (workgroups are dispatched to XCDs round robin).  Grids are padded
// to a multiple of 8 blocks; padding blocks return false.
__device__ __forceinline__ bool tile_of(int nbx, int nby, int nbtot, int& bx, int& by, int& bz)
{
    const unsigned nb8 = gridDim.x;
    const unsigned per = nb8 >> 3;
    const unsigned b = blockIdx.x;
    const unsigned lin = (b & 7u) * per + (b >> 3);
    if(lin >= (unsigned)nbtot) return false;
    bz = lin / (nbx * nby);
    const unsigned rem = lin - (unsigned)bz * nbx * nby;
    by = rem / nbx;
    bx = rem - (unsigned)by * nbx;
    return true;
}


static inline dim3 tile_grid(const DevGrid& g, int nx, int ny, int nz, int& nbx, int& nby, int& nbtot)
{
    nbx = (nx + TX - 1) / TX;
    nby = (ny + TY - 1) / TY;
    nbtot = nbx * nby * nz;
    return dim3((unsigned)((nbtot + 7) / 8 * 8));
}

// ------------------------------------------------------------------ HJ-WENO stage
// weno_hj_df_nug::aij (weno_hj_df_nug.cpp:85-108) with flux_HJ_CDS2 ipol 4 velocities,
// fused with the TVD-RK3 combination of levelset_RK3.cpp:118-157 / momentum_FC3.cpp.
__global__ void __launch_bounds__(256) k_hj_stage(DevGrid g, int nbx, int nby, int nbtot, int stage,
                                                  const double* __restrict__ phi_s, const double* __restrict__ phi_n,
                                                  const double* __restrict__ u, const double* __restrict__ v,
                                                  const double* __restrict__ w, double* __restrict__ out, double cdt)
{
    int bx, by, bz;
    if(!tile_of(nbx, nby, nbtot, bx, by, bz)) return;
    const int i = bx * TX + (threadIdx.x & (TX - 1));
    const int j = by * TY + (threadIdx.x >> 6);
    const int k = bz;
    if(i >= g.nx || j >= g.ny) return;
    const long c = g.base + i + g.sx * j + g.sxy * k;
    const double L = hj_L(g, phi_s, u, v, w, c, i, j, k);
    const double La = 0.0 + L;     // a->L zeroed, then += aij
    if(stage == 0)
        out[c] += L;
    else if(stage == 1)
        out[c] = phi_s[c] + cdt * La;
    else if(stage == 2)
        out[c] = 0.75 * phi_n[c] + 0.25 * phi_s[c] + cdt * La;
    else
        out[c] = (1.0 / 3.0) * phi_n[c] + (2.0 / 3.0) * phi_s[c] + cdt * La;
}

void launch_hj_stage(const DevGrid& g, int stage, const double* phi_s, const double* phi_n, const double* u,
                     const double* v, const double* w, double* out, double cdt, hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tile_grid(g, g.nx, g.ny, g.nz, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_hj_stage, grid, dim3(256), 0, s, g, nbx, nby, nbtot, stage, phi_s, phi_n, u, v, w, out, cdt);
}

// ------------------------------------------------------------------ reinitialisation
// reinidisc_f::disc (reinidisc_f.cpp:57-127) fused with the reini_RK3 sub-stage
// (reini_RK3.cpp:114-139) and the per-cell pseudo time step (reini_RK3.cpp:150-160).
// fsf = 1 is reinidisc_fsf::disc (reinidisc_fsf.cpp:66-141, F49 = 0): the same update
// on cells whose 7-point stencil has one sign, L = 0 elsewhere (ddweno_nug and
// ddweno_nug_sf in modus 0 compute the same gradients).
__global__ void __launch_bounds__(256) k_reini_stage(DevGrid g, int nbx, int nby, int nbtot, int stage,
                                                     const double* __restrict__ f, const double* __restrict__ f0,
                                                     double* __restrict__ out, double F43, int fsf)
{
    int bx, by, bz;
    if(!tile_of(nbx, nby, nbtot, bx, by, bz)) return;
    const int i = bx * TX + (threadIdx.x & (TX - 1));
    const int j = by * TY + (threadIdx.x >> 6);
    const int k = bz;
    if(i >= g.nx || j >= g.ny) return;
    const long c = g.base + i + g.sx * j + g.sxy * k;
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

void launch_reini_stage(const DevGrid& g, int stage, const double* fs, const double* f0, double* out, double F43,
                        int fsf, hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tile_grid(g, g.nx, g.ny, g.nz, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_reini_stage, grid, dim3(256), 0, s, g, nbx, nby, nbtot, stage, fs, f0, out, F43, fsf);
}

__global__ void __launch_bounds__(256) k_flux_acc(DevGrid g, int nbx, int nby, int nbtot, int ipol,
                                                  const double* __restrict__ q, const double* __restrict__ u,
                                                  const double* __restrict__ v, const double* __restrict__ w,
                                                  double* __restrict__ acc)
{
    int bx, by, bz;
    if(!tile_of(nbx, nby, nbtot, bx, by, bz)) return;
    const int i = bx * TX + (threadIdx.x & (TX - 1));
    const int j = by * TY + (threadIdx.x >> 6);
    const int k = bz;
    if(!comp_cell(g, ipol, i, j, k)) return;
    const long c = g.base + i + g.sx * j + g.sxy * k;
    acc[c] += flux_div(g, ipol, q, u, v, w, c, i, j, k);
}

void launch_flux_acc(const DevGrid& g, int ipol, const double* q, const double* u, const double* v, const double* w,
                     double* acc, hipStream_t s)
{
    if(ipol == 2 && g.jdir == 0) return;
    int nbx, nby, nbtot;
    dim3 grid = tile_grid(g, g.nx, g.ny, g.nz, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_flux_acc, grid, dim3(256), 0, s, g, nbx, nby, nbtot, ipol, q, u, v, w, acc);
}

// Momentum predictor of one velocity component, momentum_FC3.cpp:160-208 (stage 1),
// :262-310 (stage 2), :360-408 (stage 3): F = 0 (kepsilon_void::isource), pressure
// gradient (pjm_corr.cpp:185-204), gravity (irhs, momentum_FC3.cpp:432-472), WENO flux
// (weno_flux_nug), diffusion off (diff_void copies q), then the RK combination.
__global__ void __launch_bounds__(256) k_momentum(DevGrid g, int nbx, int nby, int nbtot, int comp, int stage,
                                                  const double* __restrict__ qin, const double* __restrict__ qold,
                                                  const double* __restrict__ u, const double* __restrict__ v,
                                                  const double* __restrict__ w, const double* __restrict__ press,
                                                  const double* __restrict__ phi, double grav, double cdt,
                                                  double* __restrict__ out, double rw, double ra)
{
    int bx, by, bz;
    if(!tile_of(nbx, nby, nbtot, bx, by, bz)) return;
    const int i = bx * TX + (threadIdx.x & (TX - 1));
    const int j = by * TY + (threadIdx.x >> 6);
    const int k = bz;
    if(!comp_cell(g, comp, i, j, k)) return;
    const long c = g.base + i + g.sx * j + g.sxy * k;
    const double F = momentum_F(g, comp, qin, u, v, w, press, phi, grav, rw, ra, c, i, j, k);
    const double df = qin[c];
    if(stage == 1)
        out[c] = df + cdt * 1.0 * F;
    else if(stage == 2)
        out[c] = 0.75 * qold[c] + 0.25 * df + cdt * 1.0 * F;
    else
        out[c] = (1.0 / 3.0) * qold[c] + (2.0 / 3.0) * df + cdt * 1.0 * F;
}

void launch_momentum(const DevGrid& g, int comp, int stage, const double* qin, const double* qold, const double* u,
                     const double* v, const double* w, const double* press, const double* phi, double grav,
                     double cdt, double* out, const double* rho_w_a, hipStream_t s)
{
    if(comp == 2 && g.jdir == 0) return;
    int nbx, nby, nbtot;
    dim3 grid = tile_grid(g, g.nx, g.ny, g.nz, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_momentum, grid, dim3(256), 0, s, g, nbx, nby, nbtot, comp, stage, qin, qold, u, v, w, press,
                       phi, grav, cdt, out, rho_w_a[0], rho_w_a[1]);
}

}  // namespace reef
