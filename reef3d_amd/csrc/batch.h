// Batched tanks: many independent box domains ("tanks") advanced by the same kernel
// launches.  Each tank keeps its own fields, metadata, work decomposition, time step and
// solver state in device memory (TankDev); a launch covers a set of tanks (TankSet).  A
// single domain is a batch of one.
#pragma once

#include "reef_internal.h"
#include "tiling.h"
#include "wave.h"

namespace reef {

// FP64 arrays of a tank: the public REEF_* fields (0..15) and the step / solver work arrays
enum Arr64
{
    A_LS = REEF_NFIELDS, A_FRK1, A_FRK2, A_RK1, A_RK2, A_URK1, A_VRK1, A_WRK1, A_URK2, A_VRK2, A_WRK2, A_RHS,
    A_R, A_R0, A_P, A_PH, A_V, A_T, A_AII, A_RFX, A_RFY, A_RFZ,
    NARR64
};
// FP32 arrays of the mixed-precision solve; p and v are double buffered (S_P / S_P2,
// S_V / S_V2, BiState::buf selects the current half)
enum Arr32
{
    S_R, S_R0, S_P, S_PH, S_V, S_T, S_X, S_KX, S_KY, S_KZ, S_AII, S_P2, S_V2,
    NARR32
};

// ghost programs a step uses: (kind, gcv)
constexpr int NPROG = 13;
constexpr int PROG_KIND_GCV[NPROG][2] = {{4, 52}, {4, 1}, {4, 40}, {1, 10}, {2, 11}, {3, 12}, {1, 7},
                                         {2, 8},  {3, 9}, {2, 12}, {4, 50}, {1, 14}, {3, 16}};

struct TankDev
{
    DevGrid g;                     // geometry and metadata (base relative to each array)
    double* f[NARR64];
    float* s[NARR32];
    const float *iDXN32, *iDYN32, *iDZN32;   // FP32 1/DN of the mixed solve (MARGE-offset)
    TileShape shape;               // tiles, chunk length, items
    int item0;                     // first reduction slot (work item) of the tank
    double* tstep;                 // device: [0] dt of the next step, [1..3] umax, vmax, wmax of the
                                   // last time-step pass, [4] smallest cell width (ietimestep)
    const GhostOp* gops[NPROG];    // compiled ghost programs
    int gnops[NPROG];
    BiState* st;                   // solver state (FP64 or mixed)
    double* vol;                   // volume1, volume2 of the last fluid update
    // explicit 7-diagonal operator of the matrix solve (solver::start with the caller's
    // a->M) in field layout: M.p, M.n, M.s, M.w, M.e, M.t, M.b, the right-hand side, and the
    // private x of the solver of var 1..3 (m[8]) and of var 4, 5 (m[9]); null until the
    // tank first needs them
    double* m[10];
    // relaxation-zone wave generator and numerical beach (iowave); wave.col null: none.
    // Its simulation time is tstep[6].
    WaveDev wave;
};

// The tanks one launch covers, passed by value (it lives in the kernel-argument segment,
// read with scalar loads): their slots in the tank table and the prefix sums of their
// tiles and work items.  A block finds its tank by a search over the prefix sums.
constexpr int SET_MAX = 192;
struct TankSet
{
    const TankDev* tanks;          // device tank table (every slot)
    double* partials;              // reduction partials: [TankDev::item0 + item] * 4
    double* partials_ru;           // partials of the mixed solve's reliable-update phase:
                                   // [TankDev::item0 + item] * 2, read by the next v phase
    int n;                         // tanks in the set
    int ntiles, nitems;            // tiles and work items of the set
    int has_pw, has_direct;        // tanks walked by the 3-D plane walk / by the direct stencil
    int pw_items;                  // work items of the plane-walked tanks, which come first
    int slot[SET_MAX];             // tank slots
    int tile0[SET_MAX + 1];        // first tile of each tank in the set's tile range
    int item0[SET_MAX + 1];        // first work item of each tank in the set's item range
};

// ---- launchers (kernels_batch.hip); field arguments are Arr64 ids; every launch covers
// the tanks of the set.  Stencil and elementwise kernels run one 256-cell tile per block;
// reductions run a persistent grid over the set's work items.
void bl_copy(const TankSet& b, int src, int dst, hipStream_t s);
void bl_fill(const TankSet& b, int id, double v, hipStream_t s);
void bl_ghost(const TankSet& b, int prog, int id, hipStream_t s);
void bl_hj_stage(const TankSet& b, int stage, int phi_s, int phi_n, int u, int v, int w, int out,
                 double cfac, hipStream_t s);
void bl_reini_stage(const TankSet& b, int stage, int fs, int f0, int out, double F43, int fsf,
                    hipStream_t s);
void bl_flux_acc(const TankSet& b, int ipol, int q, int u, int v, int w, int acc, hipStream_t s);
// diff = 1: with the explicit diffusion of ediff2 (D20 = 1); 2: F only into out (D20 = 2)
void bl_momentum(const TankSet& b, int comp, int stage, int qin, int qold, int u, int v, int w,
                 double grav, double cfac, int out, double rw, double ra, int diff, hipStream_t s);
void bl_heaviside(const TankSet& b, const double* prm4, hipStream_t s);
// iowave of the tanks with a wave generator (the others are skipped): the surface
// elevation of wavegen_precalc at each tank's simtime (tstep[6]); the relaxation of the
// velocities u, v, w (u_ >= 0), of the pressure (p_ >= 0) and of the level-set field f
// (f_ >= 0), with phi's band and rho of the tank; simtime += dt
void bl_wave_eta(const TankSet& b, hipStream_t s);
void bl_wave_relax(const TankSet& b, int u, int v, int w, int p, int f, hipStream_t s);
void bl_wave_time(const TankSet& b, hipStream_t s);
// allocation-free fluid update: the partials and the volumes in the device array `scratch`
// of kernels_batch.hip (NOALLOC_DOUBLES doubles); scratch_dev is its address on the current
// device from noalloc_scratch.  The set's tanks need item0 + nitems <= NOALLOC_MAX_ITEMS and
// at most NOALLOC_MAX_TANKS tanks; volumes land at scratch_dev + NOALLOC_VOL0.
constexpr int NOALLOC_MAX_ITEMS = 65536;
constexpr int NOALLOC_MAX_TANKS = 64;
constexpr int NOALLOC_VOL0 = 2 * NOALLOC_MAX_ITEMS;
constexpr int NOALLOC_DOUBLES = NOALLOC_VOL0 + 2 * NOALLOC_MAX_TANKS;
constexpr int NOALLOC_MAX_DEVICES = 64;
int noalloc_scratch(double** out);
void bl_heaviside_noalloc(const TankSet& b, const double* prm4, double* scratch_dev, hipStream_t s);
void bl_poisson_rhs(const TankSet& b, int u, int v, int w, double afac, hipStream_t s);
void bl_face_rho(const TankSet& b, double rw, double ra, int with_aii, hipStream_t s);
void bl_correct(const TankSet& b, double rw, double ra, double afac, int u, int v, int w,
                hipStream_t s);
// BiCGStab on every active tank: x, rhs array ids; stop / maxiter per call
void bl_bicg_init(const TankSet& b, int x, int rhs, double stop, int maxiter, hipStream_t s);
void bl_bicg_phase(const TankSet& b, int x, int rhs, int phase, int part, hipStream_t s);
// mixed precision (FP32 iteration, FP64 reliable updates)
void bl_coef32(const TankSet& b, hipStream_t s);
void bl_mixed_init(const TankSet& b, int x, int rhs, double stop, int maxiter, double delta,
                   hipStream_t s);
void bl_mixed_phase(const TankSet& b, int x, int rhs, int phase, int part, hipStream_t s);
// after the last iteration of a batch: stores the outcome of a reliable-update phase that
// no v phase has taken yet (phase 3 has no finalize of its own)
void bl_mixed_settle(const TankSet& b, hipStream_t s);
void bl_add32(const TankSet& b, int x, hipStream_t s);
// ietimestep::start (ini = 0) / ::ini (ini = 1): velocity maxima, then dt into tstep[0]
void bl_timestep(const TankSet& b, int u, int v, int w, double N47, int ini, hipStream_t s);
// level-set benchmark: prescribed vortex velocity (benchmark_vortex3D) and the void
// momentum_RK3 update of one component
void bl_vortex_velocity(const TankSet& b, int comp, int id, double simtime, double DXM, hipStream_t s);
void bl_void_rk3(const TankSet& b, int comp, int id, double Fv, hipStream_t s);
// bicgstab_ijk / bicgstab_ijk_2D on each tank's explicit 7-diagonal operator (TankDev::m)
// over the FLEXLOOP cells of var (1..3: the u / v / w grids, 4, 5: every cell), iterating
// on the private x m[8]; stop / maxiter per call.  bl_mat_copy_x: fillxvec (dir 0, field
// f into the private x) or finalize (dir 1) on those cells.
void bl_mat_copy_x(const TankSet& b, int f, int var, int dir, hipStream_t s);
// implicit diffusion (D20 = 2) of component comp in a momentum_FC3 stage: the idiff2_FS
// operator and the private x (= qin on the loop); after the var-comp solve, the stage
// combination from the solved diff and F; bl_df_set marks the direct-forcing flags set
void bl_idiff(const TankSet& b, int comp, double alpha, int qin, int u, int v, int w, hipStream_t s);
void bl_idiff_combine(const TankSet& b, int comp, int stage, int qold, int F, double cfac, int out, hipStream_t s);
void bl_df_set(const TankSet& b, hipStream_t s);
void bl_mat_init(const TankSet& b, int var, double stop, int maxiter, hipStream_t s);
void bl_mat_phase(const TankSet& b, int var, int phase, int part, hipStream_t s);
// solver states of the set's tanks into dst[slot]
void bl_gather_states(const TankSet& b, BiState* dst, hipStream_t s);
// ||rhs - A x||_2 per tank into norms[slot]
void bl_residual(const TankSet& b, int x, int rhs, double* norms, hipStream_t s);

}  // namespace reef
