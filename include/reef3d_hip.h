/* reef3d_hip.h - C-ABI of the MI355X (gfx950) implementation of REEF3D's CFD
 * free-surface hot path: WENO5 flux convection of u, v, w, WENO5 Hamilton-Jacobi
 * convection of the level set phi, TVD-RK3 reinitialisation, ghost-cell fill, the
 * variable-density pressure projection with Jacobi-preconditioned BiCGStab, and the
 * momentum_FC3 / levelset_RK3 time steps that orchestrate them.
 *
 * Conventions: every call returns 0 or a negative REEF_E* status and never exits.
 * Host pointers are borrowed for the duration of the call.  Device fields belong to
 * the context; all work is ordered on the context's stream.  One context is used by
 * one host thread at a time.
 *
 * Field layouts on the host side of reef_field_put/get:
 *   REEF_LAYOUT_IFAST: double[(nz+6)*(ny+6)*(nx+6)], i fastest, 3 ghost layers
 *                      (element (i,j,k) at (i+3) + (nx+6)*((j+3) + (ny+6)*(k+3)))
 *   REEF_LAYOUT_REF:   the reference's field_base storage (field_base.h:47),
 *                      V[(i+3)*jmax*kmax + (j+3)*kmax + (k+3)], k fastest,
 *                      jmax = ny+6, kmax = nz+6.
 */
#ifndef REEF3D_HIP_H
#define REEF3D_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum reef_status
{
    REEF_OK = 0,
    REEF_E_ARG = -1,        /* invalid argument (null pointer, bad id, bad size) */
    REEF_E_UNSUPPORTED = -2,/* geometry or option outside the implemented subset */
    REEF_E_HIP = -3,        /* HIP runtime error (message via reef_last_error) */
    REEF_E_NOMEM = -4,      /* device allocation failed */
    REEF_E_STATE = -5       /* call out of order (an allocation-free call on stale tables) */
};

enum reef_field_id
{
    REEF_U = 0, REEF_V, REEF_W, REEF_PRESS, REEF_PHI, REEF_RO, REEF_VISC,
    REEF_F, REEF_G, REEF_H, REEF_L,     /* convection accumulators (fdm::F,G,H,L) */
    REEF_PCORR,                          /* pressure correction (pjm_corr::pcorr)  */
    REEF_WORK0, REEF_WORK1, REEF_WORK2, REEF_WORK3,   /* caller scratch fields     */
    REEF_NFIELDS
};

enum reef_layout { REEF_LAYOUT_IFAST = 0, REEF_LAYOUT_REF = 1 };

enum reef_solver_precision
{
    REEF_SOLVER_FP64 = 0,   /* bicgstab_ijk in FP64 (the reference solver) */
    REEF_SOLVER_MIXED = 1   /* FP32 inner BiCGStab inside FP64 defect correction */
};

/* Grid of one box domain: what lexer holds after read_grid + gridini
 * (read_grid.cpp:75-432, grid.cpp:54-150). */
typedef struct reef_grid_desc
{
    int knox, knoy, knoz;       /* interior cells */
    int j_dir;                  /* 1: 3-D, 0: 2-D x-z tank (knoy must be 1) */
    const double* XN;           /* node coordinates incl. 5 ghost nodes per side: knox+11 */
    const double* YN;           /* knoy+11 */
    const double* ZN;           /* knoz+11 */
    double DYM, DZM;            /* grid-file header spacings (used by the psi width) */
    int group[7];               /* boundary group per side 1..6 (-x,+y,-y,+x,-z,+z):
                                   21 wall, 3 free-surface top; index 0 unused */
} reef_grid_desc;

/* Control values of the hot path (ini.cpp defaults unless the case sets them). */
typedef struct reef_params
{
    double W1, W2, W3, W4;      /* water/air density and kinematic viscosity */
    double W20, W21, W22;       /* gravity components */
    double F43, F45;            /* reinit pseudo time step factor, interface width factor */
    double dt;                  /* fixed time step (N48=0, N49) */
    double N44;                 /* pressure stop criterion on ||r||_2 / cells */
    int N46;                    /* maximum BiCGStab iterations */
} reef_params;

/* Wave generation by relaxation zones and the numerical beach (iowave with B98 = 2, B99 =
 * 1, 2; iowave.h:44, iowave_relax.cpp:28-395, iowave_precalc_relax.cpp:27-160) for linear
 * waves (wave_lib_linear, B92 = 2): the ctrl.txt keys it reads.  The zones are the
 * reference's defaults (B106 = B107 = B108 = 0): a generation band of half width B96_1
 * around x = xmin and a beach band of half width B96_2 around x = xmax, the wave's origin
 * at the domain's lower corner (B105 = 0) with direction B105_1. */
typedef struct reef_wave_desc
{
    int B92;                    /* wave theory: 2 (linear) */
    int B91;                    /* 1: wave height B91_1 and length B91_2 given */
    double B91_1, B91_2;
    int B93;                    /* 1: wave height B93_1 and period B93_2 given */
    double B93_1, B93_2;
    double B96_1, B96_2;        /* half widths of the generation and the beach zone */
    int B98;                    /* 2: relaxation generation, 0: none */
    int B99;                    /* beach: 0 none, 1 u, v, w, p, 2 also phi */
    int B101;                   /* ramp: 0 none, 1 over B102 periods, 2 over B102 seconds */
    double B102;
    double B105_1;              /* wave direction in degrees (0: along +x) */
    double B120;                /* phase shift in degrees (ini.cpp: -90) */
    double F60;                 /* still-water level (p->phimean) */
    double W50;                 /* air inflow velocity (0) */
} reef_wave_desc;

typedef struct reef_ctx reef_ctx;

/* ---- REEF3D case files (host only; no device work) */
typedef struct reef_grid_file reef_grid_file;
/* lexer::read_grid (read_grid.cpp:29-560) of a single-rank box domain: the DIVEMesh grid
 * file's header, flag4, node coordinates and boundary entries fill *desc (its node arrays
 * live in *file until reef_grid_file_free); solids, topography, parallel or periodic
 * grids and sides with mixed boundary groups give REEF_E_UNSUPPORTED */
int reef_grid_read(const char* path, reef_grid_file** file, reef_grid_desc* desc);
void reef_grid_file_free(reef_grid_file* file);
/* the hot-path keys of a ctrl.txt (control::read_control, read_control.cpp:28-70) over the
 * defaults of ini.cpp: prm gets W1-W4, W20-W22, F43, F45, N44, N46 and N49 as dt; ctrl5
 * (may be null) gets F49, N48, N47, N40, D20.  Keys choosing another scheme than the hot
 * path's (D10, D20 other than 0, 1 or 2, D30, F30, F35, F40, N10, N40, M10, and D20 = 1
 * with N48 = 1) give REEF_E_UNSUPPORTED. */
int reef_ctrl_read(const char* path, reef_params* prm, double* ctrl5);
/* the values of one ctrl.txt key (letter, number; e.g. 'N', 43 or 'B', 96) from the last
 * line that sets it (read_control.cpp:28-70): up to capacity values into values, *count
 * = the number of values on that line, 0 when no line sets the key */
int reef_ctrl_key(const char* path, char letter, int num, double* values, int capacity, int* count);
/* message of the last failed reef_grid_read / reef_ctrl_read of this thread */
const char* reef_case_last_error(void);

/* lexer + fdm + ghostcell set-up on the device (replaces driver.cpp:33-118 and
 * driver_makegrid.cpp:29-62 for a box domain); allocates all fields, zeroed. */
int reef_ctx_create(const reef_grid_desc* grid, const reef_params* prm, reef_ctx** out);

/* Control keys of lexer (ctrl.txt, lexer_read.cpp) that select a hot-path variant after
 * creation.  "F49": reinitialisation discretisation, 1 = reinidisc_f (default, ini.cpp),
 * 0 = reinidisc_fsf (reini_RK3.cpp:68-72).  "graphs" (no counterpart in the reference):
 * 1 = replay the launch sequences of a momentum_FC3 step and the BiCGStab iteration
 * batches as hipGraphs captured on their first run (re-captured after tanks or controls
 * change), 0 = launch by launch (default); results are identical either way.
 * "N48": time stepping, 0 = the fixed dt of reef_params (fixtimestep, default), 1 = every
 * reef_fc3_step first computes its own dt on the device as reef_timestep does (ietimestep
 * ::start, selected for N48 = 1 with D20 = 0 by driver_logic_cfd.cpp:59-66); "N47": its
 * factor (ini.cpp:593, default 0.3).
 * "D20": momentum diffusion, 0 = diff_void (default here), 1 = ediff2, the explicit viscous
 * terms of ediff2::diff_u/v/w (ediff2.cpp:78-190) added to F/G/H after the convection in
 * every momentum_FC3 predictor (momentum_FC3.cpp:170, 187, 204, 272-306, 370-404), laminar
 * (eddyv 0); not with N48 = 1, which the reference steps with another time-step class;
 * 2 = idiff2_FS, the implicit viscous solves (idiff2_FSu.cpp:40-278, var 1..3, stop "N43",
 * default 1e-5, ini.cpp:589).  "simtime": the simulation time of the wave generator
 * (reef_wave_set; p->simtime).  Other keys: REEF_E_UNSUPPORTED. */
int reef_set_control(reef_ctx* ctx, const char* key, double value);
void reef_ctx_destroy(reef_ctx* ctx);
const char* reef_last_error(void);

/* host <-> device field transfer (fdm field storage, field_base.h:28-54) */
int reef_field_put(reef_ctx* ctx, int field_id, const double* host, int layout);
int reef_field_get(reef_ctx* ctx, int field_id, double* host, int layout);
int reef_field_zero(reef_ctx* ctx, int field_id);

/* ghostcell::start1..4(f, gcv) (gc_start.cpp:27-182): kind 1..4 = u, v, w, scalar */
int reef_ghost_fill(reef_ctx* ctx, int field_id, int kind, int gcv);

/* convection::start(p, a, q, ipol, u, v, w) (convection.h:37): ipol 1..3 accumulate the
 * WENO flux divergence (weno_flux_nug.cpp:84-134) into REEF_F/G/H, ipol 4 the HJ-WENO
 * term (weno_hj_df_nug.cpp:58-108) into REEF_L; nothing is zeroed first. */
int reef_convection(reef_ctx* ctx, int ipol, int q_id, int u_id, int v_id, int w_id);

/* reini::start(a, p, phi, pgc, pflow) (reini.h:39, reini_RK3.cpp:77-145) with
 * p->count > 0: iters TVD-RK3 pseudo steps of reinidisc_f, in place, ghosts refilled. */
int reef_reini(reef_ctx* ctx, int phi_id, int iters);

/* reini::start with p->count == 0, the initial reinitialisation of driver_ini_cfd
 * (reini_RK3.cpp:88-101): start4(phi, gcv), then iters TVD-RK3 pseudo steps (the caller
 * passes 2*int(p->maxlength/(p->F43*p->DXM))) with the ghost rule gcv = gcval_iniphi
 * (50; 52 when F50 = 2 and F60 or F62 is set, reini_RK3.cpp:44-57) after every stage. */
int reef_reini_ini(reef_ctx* ctx, int phi_id, int iters, int gcv);

/* ioflow = iowave (B90 = 1, driver_logic_cfd.cpp:512-513) with relaxation-zone generation
 * and the numerical beach (iowave::iowave + ini, iowave.cpp:32-263, iowave_ini.cpp:27-46):
 * the static zones and factors of the context's tank; params5 (may be null) gets wT, wL,
 * wk, ww and the amplitude.  From then on reef_fc3_step runs, as driver::loop_cfd does
 * around momentum_FC3 (driver_loop_cfd.cpp:78-100), wavegen_precalc at the tank's simtime,
 * u/v/w_relax of the fields and the fluid update before the step, the relaxations
 * momentum_FC3 calls inside it (phi_relax after each phi stage, u/v/w/p_relax after each
 * projection, momentum_FC3.cpp:145, 218-221, 244, 320-323, 346, 418-421), and
 * simtime += dt after it.  The simtime starts at 0; reef_set_control(ctx, "simtime", t)
 * sets it.  Outside the zones nothing changes. */
int reef_wave_set(reef_ctx* ctx, const reef_wave_desc* w, double* params5);
/* wavegen_precalc at simtime (iowave_precalc_relax.cpp:27-160), then iowave's relaxations
 * (iowave_relax.cpp:28-395): u_relax, v_relax, w_relax of fields u_id, v_id, w_id with the
 * band of REEF_PHI (u_id < 0: skipped), p_relax of press_id with REEF_RO (< 0: skipped),
 * phi_relax of f_id (< 0: skipped); the tank's simtime becomes simtime. */
int reef_wave_relax(reef_ctx* ctx, double simtime, int u_id, int v_id, int w_id, int press_id, int f_id);

/* fluid_update_fsf::start (fluid_update_fsf.cpp:39-104): REEF_RO/REEF_VISC from
 * REEF_PHI, ghosts refilled; vol2 (may be NULL) receives volume1, volume2. */
int reef_fluid_update(reef_ctx* ctx, double* vol2);

/* reef_fluid_update without any device allocation.  No counterpart in the reference: this
 * entry point is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same kernels and the
 * same results as reef_fluid_update (bit-identical REEF_RO, REEF_VISC, volume1, volume2),
 * with the reduction partials and the volumes in a static device array of the library
 * instead of library-allocated memory: 2 doubles per work item, at most 65536 work items
 * (1 MiB; about 16.7M cells of 256-cell tiles); a larger tank gives REEF_E_ARG.  The array
 * is resolved per GPU by reef_ctx_create and held by each GPU until the process ends.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (two host
 * threads, two contexts) race and must be serialised by the caller.  The context's device
 * tables must be current: they are after reef_ctx_create and after any other call;
 * reef_set_control("D20", 2) or a first reef_solver_start rebuilds them on their next call,
 * and this call returns REEF_E_STATE until then. */
int reef_fluid_update_noalloc(reef_ctx* ctx, double* vol2);

/* poisson::start + solver::start (poisson_pcorr.cpp:78-231, bicgstab_ijk.cpp:51-278):
 * assembles the 7-point operator on REEF_PHI and solves A x = rhs for field x_id.
 * rhs_id holds the right-hand side in field layout (interior cells). */
int reef_poisson_solve(reef_ctx* ctx, int x_id, int rhs_id, double stop, int maxiter,
                       int* iters, double* residual);

/* The same solve in mixed precision; no counterpart in the reference, which solves in
 * FP64 only (it replaces bicgstab_ijk::start, bicgstab_ijk.cpp:51-123, where the caller
 * accepts a result at the same FP64 residual).  bicgstab_ijk's iteration runs on FP32
 * vectors with the FP32 operator (face conductances 1/(rho_f DXP) and the Jacobi
 * preconditioner rounded once; dot products and scalars in FP64); the solution is
 * accumulated in FP64, and reliable updates replace the iterated residual by the true
 * FP64 residual b - A x of the reference operator whenever the iterated one has dropped by
 * the factor delta (reef_solver_config, default 0.1) and before any exit.  The solve ends
 * when the true residual meets ||r||_2 / N < stop (stop >= 0, the reference criterion) or
 * ||r||_2 < |stop| * ||b - A x0||_2 (stop < 0), or at maxiter.  iters: iterations
 * (restarts counted as solveriter counts them); updates: reliable updates;
 * rel_residual: final ||b - A x||_2 / ||b - A x0||_2 in FP64.  x_id, rhs_id must differ. */
int reef_poisson_solve_mixed(reef_ctx* ctx, int x_id, int rhs_id, double stop, int maxiter,
                             int* iters, int* updates, double* rel_residual);

/* solver::start(p, a, pgc, f, rhsvec, var) (solver.h:42) of bicgstab_ijk, or of
 * bicgstab_ijk_2D on a 2-D context (driver_logic_cfd.cpp:464-468), on the caller's operator
 * (bicgstab_ijk.cpp:51-123, bicgstab_ijk_2D.cpp:49-94): M7 = {M.p, M.n, M.s, M.w, M.e, M.t,
 * M.b} of matrix_diag and rhs, each reef_solver_veclen(ctx, var) doubles in the reference's
 * vec order (the FLEXLOOP cells of var, i outermost and k innermost; var 1, 2, 3 leave out
 * the last i, j, k row of the u, v, w grid, var 4, 5 take every cell).  Field x_id holds
 * the initial guess (fillxvec) and receives the solution on those cells (finalize); its
 * other cells are not touched.  stop: the caller's p->N43 (var 1..3) or p->N44 (var 4, 5)
 * on ||r||_2 / cells (negative: |stop| * ||b - A x0||_2 / cells); maxiter: p->N46; iters
 * receives p->solveriter, residual the final ||r||_2 / cells.  The operator's arrays are
 * allocated on the device at the first call. */
int reef_solver_start(reef_ctx* ctx, int var, const double* const* M7, const double* rhs, int x_id, double stop,
                      int maxiter, int* iters, double* residual);
/* entries of the vec of var (1..5) on the context's grid, -1 on a bad argument */
long reef_solver_veclen(reef_ctx* ctx, int var);

/* ||rhs - A x||_2 in FP64 with the operator assembled from REEF_PHI: res_calc
 * (bicgstab_ijk.cpp:314-338) without its 1/N. */
int reef_poisson_residual(reef_ctx* ctx, int x_id, int rhs_id, double* norm);

/* Solver of the projection inside reef_pjm_corr / reef_fc3_step: precision
 * REEF_SOLVER_FP64 (default) or REEF_SOLVER_MIXED, and the reliable-update threshold
 * delta of the mixed form, 0 < delta < 1. */
int reef_solver_config(reef_ctx* ctx, int precision, double delta);
/* counts of the last mixed-precision solve */
int reef_solver_last(reef_ctx* ctx, long* iters, long* updates, double* rel_residual);

/* pressure::start(a, p, ppois, psolv, pgc, pflow, u, v, w, alpha) with pjm_corr
 * (pjm_corr.cpp:90-121): projection of the three velocity fields, REEF_PRESS updated. */
int reef_pjm_corr(reef_ctx* ctx, int u_id, int v_id, int w_id, double alpha, int* iters);

/* one momentum_FC3 time step (momentum_FC3.cpp:122-430) of the two-phase state held in
 * REEF_U/V/W/PRESS/PHI/RO/VISC; iters receives the summed BiCGStab iterations. */
int reef_fc3_step(reef_ctx* ctx, int* iters, double* vol2);

/* ietimestep::start (ini = 0; ietimestep.cpp:38-207) or ietimestep::ini (ini = 1; :209-304)
 * with N50 = 1, the adaptive time step of N48 = 1 (driver_logic_cfd.cpp:59-66): from the
 * largest |u|, |v|, |w| over the staggered loops, dt = N47 * 2 min(dx) / |u|max (::ini: |u|max
 * at least 2, a quarter of that, at least 1e-6), computed on the device and used by the
 * following steps.  Inflow, floating-body and EARSM terms are zero on the supported box
 * domains.  out4 (may be null): dt, umax, vmax, wmax. */
int reef_timestep(reef_ctx* c, double N47, int ini, double* out4);
/* FP64 checkpoint of a context (no counterpart in the reference, whose restart files are
 * FP32, cfd_state_write.cpp:34-146): grid size, the device dt and every public field with
 * its ghost layers.  reef_checkpoint_read restores them into a context of the same grid;
 * the steps that follow are bit-identical to those of the uninterrupted run. */
int reef_checkpoint_write(reef_ctx* c, const char* path);
int reef_checkpoint_read(reef_ctx* c, const char* path);
/* The reference's FP32 state file (cfd_state, REEF3D_CFD_STATE/REEF3D-CFD-State-<num>-<rank>.r3d,
 * cfd_state_filename.cpp:27) of a single-rank box, for Paraview and restart interop.
 * meta8 = count, printcount, simtime, printtime, sedprinttime, fsfprinttime, probeprinttime,
 * stateprinttime (the two counts as whole numbers).
 * reef_cfd_state_write replaces cfd_state::write_result (cfd_state_write.cpp:34-146) for
 * the context's u, v, w, press and phi; the fields the hot path does not hold are written
 * as the reference holds them on such a case: topo 1 (no topography), kin, eps, eddyv 0
 * (laminar), qbe 0 (no sediment), conc 0.
 * reef_cfd_state_read replaces cfd_state::read (cfd_state_read.cpp:46-192): meta8 and the
 * loop cells of u, v, w, press, phi as double(float), then the reader's ghost fills (u 10,
 * v 11, w 12, press 40, phi 52 for F50 = 2). */
int reef_cfd_state_write(reef_ctx* c, const char* path, const double* meta8);
int reef_cfd_state_read(reef_ctx* c, const char* path, double* meta8);
/* the same two on host fields (REEF_LAYOUT_IFAST of an nx x ny x nz box; fields = u, v, w,
 * press, phi); no device work; errors in reef_case_last_error */
int reef_cfd_state_write_host(int nx, int ny, int nz, const char* path, const double* const* fields,
                              const double* meta8);
int reef_cfd_state_read_host(int nx, int ny, int nz, const char* path, double* const* fields, double* meta8);
/* the dt of the following steps (the p->dt that fixtimestep::start, ietimestep::start or the
 * caller sets before momentum::start, fixtimestep.cpp:38-114) */
int reef_set_timestep(reef_ctx* c, double dt);

/* one level-set benchmark step (levelset_RK3.cpp:103-178 + benchmark_vortex3D.cpp:67-101,
 * F150=3) at time simtime before the increment. */
int reef_vortex_step(reef_ctx* ctx, double simtime, double* vol2);

/* levelset_RK3 transport stages only (levelset_RK3.cpp:109-160) on REEF_PHI */
int reef_phi_rk3(reef_ctx* ctx);

int reef_sync(reef_ctx* ctx);

/* ---- Batches: many independent tanks advanced by the same launches ------------------
 * No counterpart in the reference, which runs one domain per process; this is the
 * serving form of reef_fc3_step for many small tanks.  Every tank keeps its own fields,
 * grid, time step (reef_params.dt) and solver state; the other reef_params values, the
 * F49 control and the solver configuration are shared by the batch.  Each tank's work
 * is cut into chunks that depend on the tank alone, and every reduction of a tank sums
 * that tank's chunks in a fixed order, so a tank's results are the same alone and in
 * any batch.  Tank ids are slots: a removed tank's slot is reused by a later add. */
typedef struct reef_batch reef_batch;
/* hbm_budget_bytes: device bytes the batch's tanks may hold (0 = no limit);
 * reef_batch_add fails with REEF_E_NOMEM beyond it */
int reef_batch_create(double hbm_budget_bytes, reef_batch** out);
void reef_batch_destroy(reef_batch* b);
/* device bytes one tank of this grid needs (fields, solver arrays, metadata) */
long reef_batch_tank_bytes(const reef_grid_desc* grid);
int reef_batch_add(reef_batch* b, const reef_grid_desc* grid, const reef_params* prm, int* tank);
int reef_batch_remove(reef_batch* b, int tank);
/* reef_wave_set for one tank of a batch (its simtime, advanced by its own dt) */
int reef_batch_wave_set(reef_batch* b, int tank, const reef_wave_desc* w, double* params5);
/* tanks the next operations advance (n = 0: every tank) */
int reef_batch_set_active(reef_batch* b, const int* tanks, int n);
int reef_batch_info(reef_batch* b, int* slots, int* active, double* bytes_in_use);
int reef_batch_put(reef_batch* b, int tank, int field_id, const double* host, int layout);
int reef_batch_get(reef_batch* b, int tank, int field_id, double* host, int layout);
int reef_batch_ghost_fill(reef_batch* b, int field_id, int kind, int gcv);
int reef_batch_fluid_update(reef_batch* b);
/* the keys of reef_set_control ("F49", "graphs", "N48", "N47", "D20") for every tank of the batch */
int reef_batch_control(reef_batch* b, const char* key, double value);
int reef_batch_solver_config(reef_batch* b, int precision, double delta);
/* Tile plans per size bucket (no counterpart in the reference).  A bucket is (jdir,
 * size_log2 = floor(log2(nx*ny*nz))); its plan is the tile's rows ty x tz (64 x 4 x 1,
 * 64 x 2 x 2 or 64 x 1 x 4; 2-D tanks 64 x 1 x 4) and the chunk length cpt (tiles per work
 * item, 1..64).  Buckets without a plan use the default (64x4x1 in 3-D, 64x1x4 in 2-D,
 * about 2048 chunks per tank).  A tank takes its bucket's plan when added and when that
 * plan is set; its results depend on its plan (the chunks fix the reduction order), not on
 * the other tanks.
 * reef_batch_autotune: for each bucket of the active tanks that has no plan, probes every
 * candidate plan on that bucket's tanks (reps timed runs after a warm-up; only step-internal
 * arrays are written, the tanks' state is unchanged) and keeps the fastest; nbuckets
 * (may be NULL) receives the number of buckets tuned.
 * reef_batch_plan_set: sets (cpt = 0: removes) a bucket's plan, e.g. one saved earlier.
 * reef_batch_plans: the table, 5 ints per entry (jdir, size_log2, ty, tz, cpt); n receives
 * the number of entries, at most capacity are written. */
int reef_batch_autotune(reef_batch* b, int reps, int* nbuckets);
int reef_batch_plan_set(reef_batch* b, int jdir, int size_log2, int ty, int tz, int cpt);
int reef_batch_plans(reef_batch* b, int* entries, int capacity, int* n);
/* one momentum_FC3 step of every active tank; iters[slot] (may be NULL) receives the
 * summed BiCGStab iterations, vol[2*slot..] (may be NULL) volume1, volume2 */
int reef_batch_fc3_step(reef_batch* b, long* iters, double* vol);
/* reef_timestep for every active tank (each its own dt); out (may be null) receives
 * dt, umax, vmax, wmax per slot */
int reef_batch_timestep(reef_batch* b, double N47, int ini, double* out);

/* nsteps momentum_FC3 steps of every active tank with iteration-level scheduling: each
 * tank moves through its own steps, so a tank whose pressure solve has converged goes on
 * while others still iterate (no lockstep at the solves); iters[slot] (may be NULL)
 * receives the BiCGStab iterations of each tank's last step.  Results are those of
 * nsteps calls of reef_batch_fc3_step. */
int reef_batch_run(reef_batch* b, int nsteps, long* iters);
int reef_batch_sync(reef_batch* b);
int reef_batch_timing_enable(reef_batch* b, int on);
int reef_batch_timing_get(reef_batch* b, const char* name, double* ms, long* launches);

/* Kernel timing: the context records HIP events around its launches when enabled. */
int reef_timing_enable(reef_ctx* ctx, int on);
/* name: kernel family ("bicgstab", "hj", "reini", "flux", ...); returns total ms and launches */
int reef_timing_get(reef_ctx* ctx, const char* name, double* ms, long* launches);
int reef_timing_reset(reef_ctx* ctx);

/* sizes: out[0..4] = interior cells, field doubles allocated per field, row stride,
 * plane stride, offset of element (-3,-3,-3) */
int reef_ctx_sizes(reef_ctx* ctx, long* out5);

/* Host-only views of the set-up a context would upload (no device needed): the grid
 * arrays and WENO tables (axis 0..2; P/DN/DP n+10 values, tables (n+2) rows of
 * 24/12/36), scalars4 = DXM, DYM, DZM, psi; and the compiled ghostcell::start<kind>(gcv)
 * program applied to a REEF_LAYOUT_IFAST field on the host.  These are diagnostics with no
 * counterpart in the reference; tests use them to check the host set-up on machines
 * without a GPU. */
int reef_host_grid_arrays(const reef_grid_desc* grid, double F45, int axis, double* P, double* DN,
                          double* DP, double* qf, double* cf, double* isf, double* scalars4);
int reef_host_ghost_apply(const reef_grid_desc* grid, double F45, int kind, int gcv, double* field_ifast);

#ifdef __cplusplus
}
#endif
#endif
