// Internal declarations of the gfx950 hot-path library: host-side grid metadata, the
// device view of a domain, the solver state and the solver's vector / operator views.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/reef3d_hip.h"

namespace reef {

constexpr int MARGE = 5;        // 1-D metadata offset (iterators1D.h: IP = i + 5)
constexpr int NG = 3;           // ghost layers of a field (grid.h:70)
constexpr double PI_REF = 3.14159265359;   // definitions.h:37

// Device view of one domain.  Element (i,j,k), i,j,k in [-3, n+3), lives at
// base + i + sx*j + sxy*k.  Rows are padded so that every i = 0 is 128-byte aligned, in
// the FP64 and in the FP32 arrays (sx and base are multiples of 32 elements).
struct DevGrid
{
    int nx, ny, nz, jdir;
    long sx, sxy;
    long base;                  // index of (0,0,0)
    long nalloc;                // doubles per field
    double ydir;
    double psi;
    // 1-D metadata, MARGE-offset (entry i + MARGE)
    const double *DXP, *DYP, *DZP, *DXN, *DYN, *DZN, *XP, *YP, *ZP;
    // WENO tables per axis: qf [IP][uf][12], cf [IP][uf][6], isf [IP][uf][18]
    const double *qf[3], *cf[3], *isf[3];
    // the x tables entry-major: qfT [uf * 12 + e][IP], cfT [uf * 6 + e][IP], isfT [uf * 18 + e][IP]
    // with nT = IP entries per row, so a wavefront's 64 cells of a row read one table entry
    // as one coalesced 512-byte access
    const double *qfT, *cfT, *isfT;
    int nT;
};

// One ghost operation: f[dst] = (src < 0) ? 0 : (scaled ? 0.0 + w*f[src] : f[src])
struct GhostOp
{
    int32_t dst;
    int32_t src;     // bit 30 set: scaled; -1: zero
    double w;
};

// Host copy of the metadata and the compiled ghost programs
struct HostGrid
{
    int nx, ny, nz, jdir;
    int group[7];
    std::vector<double> X[3], P[3], DN[3], DP[3];
    double DXM, DYM, DZM, psi;
    std::vector<double> qf[3], cf[3], isf[3];
    long sx, sy, sz, sxy, base, nalloc;
    int last[4];
    std::vector<int> flag[5];   // flag1..4 on the compact (nx+6)(ny+6)(nz+6) layout
    long cid(int i, int j, int k) const
    {
        return (long)(i + NG) + (long)(nx + 2 * NG) * ((long)(j + NG) + (long)(ny + 2 * NG) * (long)(k + NG));
    }
    long did(int i, int j, int k) const { return base + i + sx * j + sxy * k; }
};

// builds metadata and lists; returns empty string or an error message
std::string build_grid(HostGrid& g, const reef_grid_desc& d, double F45);
// compiled ghost program of ghostcell::start<kind>(f, gcv)
std::string build_ghost_program(const HostGrid& g, int kind, int gcv, std::vector<GhostOp>& ops);
// the programs of n (kind, gcv) pairs of one grid, ops[q] for pair q (shares the boundary
// lists and mark array between them)
std::string build_ghost_programs(const HostGrid& g, int n, const int (*kind_gcv)[2], std::vector<GhostOp>* ops);

// ------------------------------------------------------------------ BiCGStab state
struct BiState
{
    double r_j, norm_r0, alpha, w, beta, residual, stop, sigma, norm_vj, norm_rj, norm_sj;
    double w1, w2, r_j1;
    int iter, maxiter, done, mode;     // mode: 0 run, 1 restart pending, 2 fresh p,
                                       // 3 reliable update pending (mixed precision)
    int branch;                        // 1: norm_sj > stop
    int mixed;                         // 1: FP32 iteration with FP64 reliable updates
    // mixed precision: r_j before the last x phase, largest ||r|| since the last
    // reliable update, the update threshold, x32 waiting to be added to x64, count of
    // reliable updates
    double r_j_prev, rmax, delta;
    double res0;                       // ||b - A x0||_2 (FP64) of a mixed solve
    double rr;                         // |r|^2 of the current residual, summed where r was
                                       // formed (initial residual, restart, x phase)
    int pending, nru;
    int buf;                           // mixed precision: current half of the FP32 p / v buffers
    long ncell;
};

// operator of the solve: FP64 recomputes the couplings of poisson_pcorr from the face
// densities; FP32 (mixed-precision inner solve) reads FP32 face conductances 1/(rho_f DP)
// and scales them by the FP32 1/DN tables (MARGE-offset)
struct SolverOp64
{
    const double *rfx, *rfy, *rfz;
};
struct SolverOp32
{
    const float *kx, *ky, *kz;
    const float *iDXN, *iDYN, *iDZN;
};

// vectors of one BiCGStab solve (bicgstab_ijk.cpp:30-40); ph = aii p, v = A ph, t = A sh
template<class V>
struct BiVec
{
    V* x;
    const V* rhs;
    V *r, *r0, *p, *ph, *v, *t;
    const V* aii;
    double* partials;
    BiState* st;
};

}  // namespace reef
