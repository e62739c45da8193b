// Adapters a REEF3D maintainer adds to src/ to run the hot path on an MI355X through
// libreef3d_hip.so.  They derive from the reference's own strategy interfaces
// (convection.h:33-39, reini.h:35-40, momentum.h:45-50) so driver::logic_cfd can select
// them like any other scheme, and they talk to the device only through the C-ABI of
// include/reef3d_hip.h (plain pointers, no device types in the signatures).
//
// Two ways to use them:
//  * per call (hip_weno_flux, hip_weno_hj, hip_reini): the fields are copied to the
//    device, the operator runs there and the result is copied back; a drop-in for
//    one scheme, bounded by PCIe traffic;
//  * per time step (hip_momentum_FC3): the whole momentum_FC3 step runs on the device
//    and the fdm fields are copied in and out once per step (or left on the device
//    between steps when nothing on the host reads them).
#ifndef REEF3D_HIP_ADAPTERS_H_
#define REEF3D_HIP_ADAPTERS_H_

#include <cstdio>
#include <cstdlib>

#include "convection.h"
#include "density_f.h"
#include "fdm.h"
#include "field.h"
#include "ghostcell.h"
#include "increment.h"
#include "lexer.h"
#include "looping.h"
#include "matrix_diag.h"
#include "momentum.h"
#include "pressure.h"
#include "reini.h"
#include "solver.h"
#include "timestep.h"
#include "vec.h"
#include "reef3d_hip.h"

namespace reef_adapt {

inline void check(int rc, const char* what)
{
    if(rc != 0)
    {
        std::fprintf(stderr, "reef3d_hip: %s failed (%d): %s\n", what, rc, reef_last_error());
        std::exit(1);     // the reference's own convention for fatal errors
    }
}

// reef_grid_desc from the lexer of a single-rank box run (read_grid.cpp:75-432): node
// coordinates incl. 5 ghost nodes, header spacings, and the boundary group of each side
// taken from the gcb4 list (every entry of a side must share one group)
inline reef_grid_desc grid_from_lexer(lexer* p)
{
    reef_grid_desc d;
    d.knox = p->knox;
    d.knoy = p->knoy;
    d.knoz = p->knoz;
    d.j_dir = p->j_dir;
    d.XN = p->XN;
    d.YN = p->YN;
    d.ZN = p->ZN;
    d.DYM = p->DYM;
    d.DZM = p->DZM;
    for(int s = 0; s < 7; ++s) d.group[s] = 0;
    for(int q = 0; q < p->gcb4_count; ++q)
    {
        const int side = p->gcb4[q][3], group = p->gcb4[q][4];
        if(side < 1 || side > 6) continue;
        if(d.group[side] != 0 && d.group[side] != group)
        {
            std::fprintf(stderr, "reef3d_hip: side %d mixes boundary groups\n", side);
            std::exit(1);
        }
        d.group[side] = group;
    }
    return d;
}

inline reef_params params_from_lexer(lexer* p)
{
    reef_params r;
    r.W1 = p->W1;
    r.W2 = p->W2;
    r.W3 = p->W3;
    r.W4 = p->W4;
    r.W20 = p->W20;
    r.W21 = p->W21;
    r.W22 = p->W22;
    r.F43 = p->F43;
    r.F45 = p->F45;
    r.dt = p->dt;
    r.N44 = p->N44;
    r.N46 = p->N46;
    return r;
}

// one device context per process (the reference is single-threaded per rank)
inline reef_ctx* context(lexer* p)
{
    static reef_ctx* ctx = nullptr;
    if(!ctx)
    {
        reef_grid_desc d = grid_from_lexer(p);
        reef_params r = params_from_lexer(p);
        check(reef_ctx_create(&d, &r, &ctx), "reef_ctx_create");
        // the scheme controls the case chose (reini_RK3.cpp:68-72, driver_logic_cfd.cpp:332-335)
        check(reef_set_control(ctx, "F49", p->F49 == 0 ? 0.0 : 1.0), "F49");
        if(p->D20 == 1 || p->D20 == 2) check(reef_set_control(ctx, "D20", double(p->D20)), "D20");
        if(p->D20 == 2) check(reef_set_control(ctx, "N43", p->N43), "N43");
    }
    return ctx;
}

inline void put(reef_ctx* c, int id, field& f) { check(reef_field_put(c, id, f.V, REEF_LAYOUT_REF), "put"); }
inline void get(reef_ctx* c, int id, field& f) { check(reef_field_get(c, id, f.V, REEF_LAYOUT_REF), "get"); }

// convection::start for D10 = 4 (weno_flux_nug): accumulates into a->F/G/H
class hip_weno_flux : public convection
{
public:
    explicit hip_weno_flux(lexer* p) { context(p); }
    void start(lexer* p, fdm* a, field& q, int ipol, field& u, field& v, field& w) override
    {
        reef_ctx* c = context(p);
        field* accp = &a->L;
        if(ipol == 1) accp = &a->F;
        if(ipol == 2) accp = &a->G;
        if(ipol == 3) accp = &a->H;
        field& acc = *accp;
        const int acc_id = ipol == 1 ? REEF_F : ipol == 2 ? REEF_G : ipol == 3 ? REEF_H : REEF_L;
        put(c, REEF_WORK0, q);
        put(c, REEF_U, u);
        put(c, REEF_V, v);
        put(c, REEF_W, w);
        put(c, acc_id, acc);
        check(reef_convection(c, ipol, REEF_WORK0, REEF_U, REEF_V, REEF_W), "reef_convection");
        get(c, acc_id, acc);
    }
};

// convection::start for F35 = 5 (weno_hj_df_nug): accumulates into a->L
class hip_weno_hj : public convection
{
public:
    explicit hip_weno_hj(lexer* p) { context(p); }
    void start(lexer* p, fdm* a, field& q, int ipol, field& u, field& v, field& w) override
    {
        reef_ctx* c = context(p);
        put(c, REEF_WORK0, q);
        put(c, REEF_U, u);
        put(c, REEF_V, v);
        put(c, REEF_W, w);
        put(c, REEF_L, a->L);
        check(reef_convection(c, 4, REEF_WORK0, REEF_U, REEF_V, REEF_W), "reef_convection");
        get(c, REEF_L, a->L);
    }
};

// reini::start for F40 = 3 (reini_RK3, reinidisc_f or reinidisc_fsf by F49): at
// p->count == 0 the initial reinitialisation of driver_ini_cfd (2*int(maxlength/(F43 DXM))
// pseudo steps with gcval_iniphi, reini_RK3.cpp:88-101), afterwards p->F44 steps with
// gcval_phi
class hip_reini : public reini
{
public:
    explicit hip_reini(lexer* p) { context(p); }
    void start(fdm* a, lexer* p, field& f, ghostcell* pgc, ioflow* pflow) override
    {
        (void)a;
        (void)pgc;
        (void)pflow;
        reef_ctx* c = context(p);
        if(p->F50 != 2)
        {
            std::fprintf(stderr, "reef3d_hip: hip_reini supports F50 = 2 only\n");
            std::exit(1);
        }
        put(c, REEF_WORK1, f);
        if(p->count == 0)
        {
            // gcval_iniphi of reini_RK3::reini_RK3 (reini_RK3.cpp:44-57) for F50 = 2
            const int gcv = (p->F62 > 1.0e-20 || p->F60 > 1.0e-20) ? 52 : 50;
            const int iters = 2 * int(p->maxlength / (p->F43 * p->DXM));
            check(reef_reini_ini(c, REEF_WORK1, iters, gcv), "reef_reini_ini");
        }
        else
            check(reef_reini(c, REEF_WORK1, p->F44), "reef_reini");
        get(c, REEF_WORK1, f);
    }
};

// solver::start for N10 = 1 (bicgstab_ijk, or bicgstab_ijk_2D on a 2-D case; driver_logic_cfd.cpp:
// 464-478) on the caller's a->M and rhsvec: the seven diagonals and the right-hand side
// go to the device over the vec of var, the field is the initial guess and receives the
// solution on the FLEXLOOP cells of var (bicgstab_ijk.cpp:51-123).  The driver keeps two
// solver objects, psolv (var 1..3) and ppoissonsolv (var 4, 5); the device keeps one
// private x for each of the two roles, so two instances of this class behave like them.
// The other entry points are not on the hot path (solver.h:43-46): the reference's own
// startf / startF / startM are empty, startV (sigma grids) is refused.
class hip_bicgstab : public solver
{
public:
    explicit hip_bicgstab(lexer* p) { context(p); }
    void start(lexer* p, fdm* a, ghostcell* pgc, field& f, vec& rhsvec, int var) override
    {
        (void)pgc;
        reef_ctx* c = context(p);
        p->preconiter = 0;
        const double* M7[7] = {a->M.p, a->M.n, a->M.s, a->M.w, a->M.e, a->M.t, a->M.b};
        const double stop = var <= 3 ? p->N43 : p->N44;
        put(c, REEF_WORK2, f);
        int iters = 0;
        double res = 0.0;
        check(reef_solver_start(c, var, M7, rhsvec.V, REEF_WORK2, stop, p->N46, &iters, &res), "reef_solver_start");
        get(c, REEF_WORK2, f);
        p->solveriter = iters;
    }
    void startf(lexer*, ghostcell*, field&, vec&, matrix_diag&, int) override {}
    void startF(lexer*, ghostcell*, double*, vec&, matrix_diag&, int) override {}
    void startV(lexer*, ghostcell*, double*, vec&, matrix_diag&, int) override
    {
        std::fprintf(stderr, "reef3d_hip: hip_bicgstab::startV (sigma grids) is not supported\n");
        std::exit(1);
    }
    void startM(lexer*, ghostcell*, double*, double*, double*, int) override {}
};

// pressure::start for D30 = 1 (pjm_corr with poisson_pcorr, driver_logic_cfd.cpp:385-395):
// the projection of u, v, w with the face densities of a->phi, assembled, solved and
// applied on the device (pjm_corr.cpp:90-121); a->press is updated and p->solveriter /
// p->poissoniter receive the iterations.  The ppois and psolv the driver passes are not
// used: the assembly and the solve are part of the device projection.  upgrad / vpgrad /
// wpgrad add the pressure gradient to the host's a->F, a->G, a->H, which momentum_FC3's
// irhs/jrhs/krhs assemble on the host (momentum_FC3.cpp:167, :184, :201): they run here, on
// the host, with the caller's own density_f::roface, as pjm_corr::upgrad..wpgrad do
// (pjm_corr.cpp:185-201).  ucorr / vcorr / wcorr are only called from inside
// pjm_corr::start and are refused.
class hip_pjm_corr : public pressure, public increment
{
public:
    explicit hip_pjm_corr(lexer* p) : pd_(p) { context(p); }
    void start(fdm* a, lexer* p, poisson* ppois, solver* psolv, ghostcell* pgc, ioflow* pflow, field& u, field& v,
               field& w, double alpha) override
    {
        (void)ppois;
        (void)psolv;
        (void)pgc;
        (void)pflow;
        reef_ctx* c = context(p);
        check(reef_set_timestep(c, p->dt), "reef_set_timestep");
        put(c, REEF_U, u);
        put(c, REEF_V, v);
        put(c, REEF_W, w);
        put(c, REEF_PRESS, a->press);
        put(c, REEF_PHI, a->phi);
        int iters = 0;
        check(reef_pjm_corr(c, REEF_U, REEF_V, REEF_W, alpha, &iters), "reef_pjm_corr");
        get(c, REEF_U, u);
        get(c, REEF_V, v);
        get(c, REEF_W, w);
        get(c, REEF_PRESS, a->press);
        p->solveriter = iters;
        p->poissoniter = iters;
    }
    // pjm_corr::ini sets up the pressure reference of B30 (pressure_reference_ini.cpp:28);
    // the device projection has none (B30 = 0, reference_start a no-op)
    void ini(lexer* p, fdm*, ghostcell*) override
    {
        if(p->B30 != 0) refuse("ini with B30 != 0");
    }
    void upgrad(lexer* p, fdm* a, slice&, slice&) override
    {
        ULOOP
        a->F(i, j, k) -= PORVAL1 * (a->press(i + 1, j, k) - a->press(i, j, k)) / (p->DXP[IP] * pd_.roface(p, a, 1, 0, 0));
    }
    void vpgrad(lexer* p, fdm* a, slice&, slice&) override
    {
        if(p->j_dir == 1)
            VLOOP
            a->G(i, j, k) -= PORVAL2 * (a->press(i, j + 1, k) - a->press(i, j, k)) / (p->DYP[JP] * pd_.roface(p, a, 0, 1, 0));
    }
    void wpgrad(lexer* p, fdm* a, slice&, slice&) override
    {
        WLOOP
        a->H(i, j, k) -= PORVAL3 * (a->press(i, j, k + 1) - a->press(i, j, k)) / (p->DZP[KP] * pd_.roface(p, a, 0, 0, 1));
    }
    void ucorr(lexer*, fdm*, field&, double) override { refuse("ucorr"); }
    void vcorr(lexer*, fdm*, field&, double) override { refuse("vcorr"); }
    void wcorr(lexer*, fdm*, field&, double) override { refuse("wcorr"); }

private:
    static void refuse(const char* what)
    {
        std::fprintf(stderr, "reef3d_hip: hip_pjm_corr::%s is not on the supported path\n", what);
        std::exit(1);
    }
    density_f pd_;
};

// timestep for N48 = 1 with D20 = 0 or >= 2 (ietimestep, driver_logic_cfd.cpp:59-66):
// p->dt, p->umax, p->vmax, p->wmax from the device (reef_timestep).  The diagnostics the
// reference also gathers there (viscmax, kinmax, epsmax, pressmax/min, fbdt) are not set.
class hip_ietimestep : public timestep
{
public:
    explicit hip_ietimestep(lexer* p) { context(p); }
    void start(fdm* a, lexer* p, ghostcell* pgc, turbulence* pturb) override
    {
        (void)pgc;
        (void)pturb;
        run(a, p, 0);
    }
    void ini(fdm* a, lexer* p, ghostcell* pgc) override
    {
        (void)pgc;
        run(a, p, 1);
        p->dt_old = p->dt;
    }

private:
    static void run(fdm* a, lexer* p, int ini)
    {
        reef_ctx* c = context(p);
        put(c, REEF_U, a->u);
        put(c, REEF_V, a->v);
        put(c, REEF_W, a->w);
        double out[4];
        check(reef_timestep(c, p->N47, ini, out), "reef_timestep");
        p->dt = out[0];
        p->umax = out[1];
        p->vmax = out[2];
        p->wmax = out[3];
    }
};

// momentum::start for N40 = 3 (momentum_FC3 with weno_flux_nug, weno_hj_df_nug,
// reini_RK3, diff_void, pjm_corr, bicgstab_ijk): the whole step on the device
class hip_momentum_FC3 : public momentum
{
public:
    explicit hip_momentum_FC3(lexer* p) { context(p); }
    void start(lexer* p, fdm* a, ghostcell* pgc, vrans* pvrans, sixdof* p6dof) override
    {
        (void)pgc;
        (void)pvrans;
        (void)p6dof;
        reef_ctx* c = context(p);
        check(reef_set_timestep(c, p->dt), "reef_set_timestep");   // this step's p->dt
        put(c, REEF_U, a->u);
        put(c, REEF_V, a->v);
        put(c, REEF_W, a->w);
        put(c, REEF_PRESS, a->press);
        put(c, REEF_PHI, a->phi);
        int iters = 0;
        double vol[2];
        check(reef_fc3_step(c, &iters, vol), "reef_fc3_step");
        get(c, REEF_U, a->u);
        get(c, REEF_V, a->v);
        get(c, REEF_W, a->w);
        get(c, REEF_PRESS, a->press);
        get(c, REEF_PHI, a->phi);
        get(c, REEF_RO, a->ro);
        get(c, REEF_VISC, a->visc);
        p->solveriter = iters;
        p->volume1 = vol[0];
        p->volume2 = vol[1];
    }
};

}  // namespace reef_adapt

#endif
