// Internal declarations of the gfx950 hot-path library: host-side grid metadata, the
// device view passed to kernels, and the kernel launchers.
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
// base + i + sx*j + sxy*k.  Rows are padded so that every i = 0 is 128-byte aligned.
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
    int pending, nru;
    long ncell;
};

// ------------------------------------------------------------------ launchers
void launch_ghost(const GhostOp* ops, int nops, double* f, hipStream_t s);

// HJ-WENO stage: out = combo(phi_n, phi_s, cdt * (0 + L_hj(phi_s; u, v, w)))
//   stage 1: out = phi_s + cdt*L ; stage 2: 0.75 phi_n + 0.25 phi_s + cdt*L ;
//   stage 3: (1/3) phi_n + (2/3) phi_s + cdt*L ; stage 0: out = L (accumulate: out += L)
void launch_hj_stage(const DevGrid& g, int stage, const double* phi_s, const double* phi_n, const double* u,
                     const double* v, const double* w, double* out, double cdt, hipStream_t s);

// reinit sub-stage q (1..3): out = combo(f0, fs, dtf * L_reini(fs)); fsf = 1 for
// reinidisc_fsf (F49 = 0), 0 for reinidisc_f (F49 = 1)
void launch_reini_stage(const DevGrid& g, int stage, const double* fs, const double* f0, double* out, double F43,
                        int fsf, hipStream_t s);

// WENO flux: acc += flux divergence of q (component ipol) with face velocities of (u,v,w)
void launch_flux_acc(const DevGrid& g, int ipol, const double* q, const double* u, const double* v, const double* w,
                     double* acc, hipStream_t s);

// velocity predictor of one component (momentum_FC3 stages): F = 0 - dp/(DP*rho_f) + g + conv;
// stage 1: out = qin + (dt)*F ; stage 2: 0.75 qold + 0.25 qin + (0.25 dt) F ;
// stage 3: (1/3) qold + (2/3) qin + ((2/3) dt) F   (written to out)
void launch_momentum(const DevGrid& g, int comp, int stage, const double* qin, const double* qold, const double* u,
                     const double* v, const double* w, const double* press, const double* phi, double grav,
                     double cdt, double* out, const double* rho_w_a, hipStream_t s);

void launch_heaviside(const DevGrid& g, const double* phi, double* ro, double* visc, const double* prm4,
                      double* partials, double* vol_out, unsigned int* ticket, hipStream_t s);

void launch_copy_interior(const DevGrid& g, const double* src, double* dst, hipStream_t s);
void launch_fill(double* f, long n, double v, hipStream_t s);

// projection
void launch_poisson_rhs(const DevGrid& g, const double* u, const double* v, const double* w, double adt, double* rhs,
                        hipStream_t s);
void launch_face_rho(const DevGrid& g, const double* phi, double rw, double ra, double* rfx, double* rfy,
                     double* rfz, double* aii, hipStream_t s);
void launch_correct(const DevGrid& g, const double* pcorr, const double* phi, double rw, double ra, double adt,
                    double* u, double* v, double* w, double* press, hipStream_t s);

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

void launch_bicg_init(const DevGrid& g, const SolverOp64& op, const BiVec<double>& b, double stop, int maxiter,
                      hipStream_t s);
void launch_bicg_init(const DevGrid& g, const SolverOp32& op, const BiVec<float>& b, double stop, int maxiter,
                      hipStream_t s);
// one phase of a BiCGStab iteration: 0 restart or p update, 1 v = A ph, 2 t = A sh, 3 x/r update;
// part 0 is the grid kernel, part 1 the one-block finalize kernel that must follow it
void launch_bicg_phase(const DevGrid& g, const SolverOp64& op, const BiVec<double>& b, int phase, int part,
                       hipStream_t s);
void launch_bicg_phase(const DevGrid& g, const SolverOp32& op, const BiVec<float>& b, int phase, int part,
                       hipStream_t s);

// mixed precision: FP32 operator and preconditioner from the FP64 face densities;
// ||b - A x||_2 in FP64 (norm_out on the device); x64 += x32, x32 = 0
void launch_coef32(const DevGrid& g, const double* rfx, const double* rfy, const double* rfz, float* kx, float* ky,
                   float* kz, float* aii, hipStream_t s);
void launch_defect(const DevGrid& g, const SolverOp64& op, const double* x, const double* rhs, double* partials,
                   double* norm_out, hipStream_t s);
void launch_add32(const DevGrid& g, double* x, float* d, hipStream_t s);

// FP32 BiCGStab with reliable updates: the iteration of launch_bicg_phase on the FP32
// vectors of b32 (operator op32) while x64 accumulates the solution and every reliable
// update recomputes the residual as b64 - A64 (x64 + x32) in FP64
struct MixedVec
{
    double* x64;
    const double* b64;
};
void launch_mixed_init(const DevGrid& g, const SolverOp64& op64, const MixedVec& m, const BiVec<float>& b32,
                       double stop, int maxiter, double delta, hipStream_t s);
// phase 0..3 as launch_bicg_phase, 4 = reliable update / restart (a no-op unless the state
// asks for one); part 0 grid kernel, part 1 finalize
void launch_mixed_phase(const DevGrid& g, const SolverOp64& op64, const SolverOp32& op32, const MixedVec& m,
                        const BiVec<float>& b32, int phase, int part, hipStream_t s);

// prescribed vortex field (benchmark_vortex3D.cpp:67-101) and the void momentum_RK3 update
void launch_vortex_velocity(const DevGrid& g, int comp, double* q, double simtime, double DXM, hipStream_t s);
void launch_void_rk3(const DevGrid& g, int comp, double* q, double dt, double Fv, hipStream_t s);

}  // namespace reef
