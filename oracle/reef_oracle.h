/* CPU oracle for the REEF3D free-surface hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++ restatement of the reference's CPU
 * algorithm (REEF3D src/, cited file:line in reef_oracle.cpp), used by tests/ as the
 * parity checker and by bench.py as the timed CPU baseline.  Nothing in the product
 * (reef3d_amd/) includes, links or calls it.
 *
 * Field layout (same as the GPU code): double[(nz+6)*(ny+6)*(nx+6)], i fastest,
 * 3 ghost layers: f[(i+3) + (nx+6)*((j+3) + (ny+6)*(k+3))] = field(i,j,k).
 */
#ifndef REEF_ORACLE_H
#define REEF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_domain orc_domain;

typedef struct orc_params
{
    double rho_water, visc_water;   /* W1, W2 (kinematic) */
    double rho_air, visc_air;       /* W3, W4 */
    double gi, gj, gk;              /* W20, W21, W22 */
    double F43, F45;                /* reinit time step factor, epsilon factor */
    double dt;                      /* fixed time step (N49) */
    double N44;                     /* pressure stop criterion on ||r||_2 / cells */
    int N46;                        /* max BiCGStab iterations */
} orc_params;

/* Box domain: nx x ny x nz cells, node coordinates XN/YN/ZN with 5 ghost nodes per side
 * (n+11 values each), DYM/DZM as stored in the grid file header, group[s] = boundary
 * group of side s (index 1..6: -x, +y, -y, +x, -z, +z; 21 wall, 3 free-surface top). */
orc_domain* orc_create(int nx, int ny, int nz, int jdir, const double* XN, const double* YN,
                       const double* ZN, double DYM, double DZM, const int* group7,
                       const orc_params* prm);
void orc_destroy(orc_domain*);

/* grid metadata: which = 0 XP,1 DXN,2 DXP (axis 0..2; n+10 values),
 * tables: orc_tables(d, axis, qf[(n+2)*24], cf[(n+2)*12], isf[(n+2)*36]) for IP=4..n+5 */
void orc_grid_array(const orc_domain*, int which, int axis, double* out);
void orc_tables(const orc_domain*, int axis, double* qf, double* cf, double* isf);
/* [DXM, DYM, DZM, psi, cellnum] */
void orc_scalars(const orc_domain*, double* out5);

/* ghostcell::start1..4(f, gcv) */
void orc_ghost(orc_domain*, double* f, int kind, int gcv);

/* convection: L += HJ-WENO (ipol 4); acc += WENO flux divergence (ipol 1,2,3) */
void orc_hj(orc_domain*, const double* phi, const double* u, const double* v, const double* w,
            double* L);
void orc_flux(orc_domain*, int ipol, const double* q, const double* u, const double* v,
              const double* w, double* acc);

/* reini_RK3::start with count>0, iters pseudo steps (phi in place, ghosts refilled) */
void orc_reini(orc_domain*, double* phi, int iters);
/* the same pseudo steps with the level-set ghost rule gcv after every stage (50 = gcval_iniphi
 * of the initial reinitialisation, 52 = gcval_phi) */
void orc_reini_gcv(orc_domain*, double* phi, int iters, int gcv);

/* ietimestep::start (ini = 0) / ::ini (ini = 1), N50 = 1: out4 = dt, umax, vmax, wmax */
void orc_timestep(orc_domain*, const double* u, const double* v, const double* w, double N47, int ini,
                  double* out4);

/* fluid_update_fsf: rho, visc from phi; volumes out2[0..1] */
void orc_fluid_update(orc_domain*, const double* phi, double* ro, double* visc, double* vol2);

/* poisson_pcorr::start on (phi) -> 7 coefficient arrays + rhs (LOOP order, cellnum each) */
void orc_poisson(orc_domain*, const double* phi, const double* pcorr, double* M7, double* rhs);
/* bicgstab_ijk::start(var=5): x (field layout, in/out), M7 + rhs as from orc_poisson */
/* control key of the hot path: "F49" (1 reinidisc_f, 0 reinidisc_fsf); or of the oracle
 * itself: "dot_ext" (1: BiCGStab dot products accumulated in long double, a rounding
 * perturbation for sensitivity checks); returns 0 or -1 */
int orc_set_control(orc_domain*, const char* key, double value);
/* r = rhs - A x (res_calc, bicgstab_ijk.cpp:314-338) per cell into r (may be NULL); returns ||r||_2 */
double orc_residual(orc_domain*, const double* x, const double* M7, const double* rhs, double* r);
int orc_bicgstab(orc_domain*, double* x, const double* M7, const double* rhs, double stop,
                 int maxiter, double* residual);
/* bicgstab_ijk::start with var 1..5 (FLEXLOOP of flag1..4, bicgstab_ijk.cpp:51-123): M7 rows
 * and rhs of orc_veclen(var) entries in vec order; x in field layout, solved in place */
int orc_bicgstab_var(orc_domain*, int var, double* x, const double* M7, const double* rhs, double stop,
                     int maxiter, double* residual);
long orc_veclen(orc_domain*, int var);
/* idiff2_FS / idiff2_FS_2D diff_u/v/w of component comp (1..3): the operator (M7 rows and
 * rhs in vec order, rhs zeroed by the caller), and the whole call (diff = vin, ghosts, the
 * var-comp solve with stop N43, ghosts) returning the iterations; au, av are the fdm's u,
 * v that diff_w reads at k+1 */
void orc_idiff_assemble(orc_domain*, int comp, double alpha, const double* vin, const double* u, const double* v,
                        const double* w, const double* au, const double* av, const double* visc, double* M7,
                        double* rhs);
int orc_idiff(orc_domain*, int comp, double alpha, double* diff, const double* vin, const double* u, const double* v,
              const double* w, const double* au, const double* av, const double* visc);

/* pjm_corr::start with alpha: u,v,w corrected in place, press updated, pcorr out */
int orc_pjm(orc_domain*, const double* phi, double* u, double* v, double* w, double* press,
            double* pcorr, double alpha);

/* one momentum_FC3 time step of the two-phase state (all fields in/out) */
int orc_fc3_step(orc_domain*, double* phi, double* u, double* v, double* w, double* press,
                 double* ro, double* visc, double* vol2);

/* one level-set benchmark step (levelset_RK3 + momentum_RK3 with void operators +
 * benchmark_vortex3D velocity update at simtime); returns volumes */
void orc_vortex_step(orc_domain*, double* phi, double* u, double* v, double* w, double* ro,
                     double* visc, double simtime, double* vol2);

/* HJ-WENO phi transport TVD-RK3 only (levelset_RK3.cpp:109-160, no reinit) */
void orc_phi_rk3(orc_domain*, double* phi, const double* u, const double* v, const double* w);

/* iowave wave generation by relaxation zones (B98 = 2) and numerical beach (B99 = 1, 2)
 * for linear waves (B92 = 2) with B91 (height, length) or B93 (height, period); the
 * default zones along x = xmin and x = xmax (B96_1, B96_2 long); with it set,
 * orc_fc3_step runs the driver loop's wave prologue (precalc at the domain's simtime,
 * relaxation of u, v, w, fluid update) before the step and advances simtime by dt */
typedef struct orc_wave_desc
{
    int B92;
    int B91;
    double B91_1, B91_2;
    int B93;
    double B93_1, B93_2;
    double B96_1, B96_2;
    int B98, B99, B101;
    double B102, B105_1, B120, F60, W50;
} orc_wave_desc;
/* also sets psi from phi as initialize::inipsi does for wave runs (B90 > 0) */
void orc_wave_set(orc_domain*, const orc_wave_desc*, const double* phi);
/* wT, wL, wk, ww, wA */
void orc_wave_params(const orc_domain*, double* out5);
void orc_set_simtime(orc_domain*, double t);
/* wavegen_precalc at simtime, then u/v/w_relax (face band of phi), p_relax (ro) and
 * phi_relax of f, in place */
void orc_wave_relax(orc_domain*, double simtime, const double* phi, const double* ro, double* u, double* v, double* w,
                    double* press, double* f);

/* number of cells of the field layout */
long orc_field_size(const orc_domain*);

#ifdef __cplusplus
}
#endif
#endif
