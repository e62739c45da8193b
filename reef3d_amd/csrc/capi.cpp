// Context, host <-> device transfers, time-step orchestration and the C-ABI of the
// gfx950 hot-path library (declared in include/reef3d_hip.h).
#include "reef_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <utility>

using namespace reef;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do                                                                                        \
    {                                                                                         \
        hipError_t e_ = (expr);                                                               \
        if(e_ != hipSuccess)                                                                  \
            return fail(REEF_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while(0)

struct TimedRange
{
    std::string name;
    hipEvent_t a, b;
};

}  // namespace

struct reef_ctx
{
    HostGrid hg;
    DevGrid dg;
    reef_params prm;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    double* field[REEF_NFIELDS] = {};
    // work buffers of the time steps
    double *ls = nullptr, *frk1 = nullptr, *frk2 = nullptr, *rk1 = nullptr, *rk2 = nullptr;
    double *urk1 = nullptr, *vrk1 = nullptr, *wrk1 = nullptr, *urk2 = nullptr, *vrk2 = nullptr, *wrk2 = nullptr;
    double* rhs = nullptr;
    BiVec<double> bi{};
    SolverOp64 op64{};
    double *rfx = nullptr, *rfy = nullptr, *rfz = nullptr, *aii64 = nullptr;
    // mixed-precision projection: FP32 inner vectors and operator, allocated on first use
    int F49 = 1;                         // reinit discretisation (ini.cpp default 1)
    int precision = REEF_SOLVER_FP64;
    double delta = 0.1;                  // reliable-update threshold of the mixed solve
    BiVec<float> bf{};
    SolverOp32 op32{};
    float *kx32 = nullptr, *ky32 = nullptr, *kz32 = nullptr, *aii32 = nullptr, *r32 = nullptr;
    double* norm_dev = nullptr;
    BiState* st_host = nullptr;          // pinned, two status slots of the batched iteration
    hipEvent_t st_ev[2] = {nullptr, nullptr};
    long last_inner = 0, last_outer = 0;
    double last_rel = 0.0;
    double* vol_dev = nullptr;
    unsigned int* vol_ticket = nullptr;
    double* partials = nullptr;
    std::map<std::pair<int, int>, std::pair<GhostOp*, int>> ghost_progs;
    bool timing = false;
    std::vector<TimedRange> ranges;
    std::map<std::string, std::pair<double, long>> totals;
    std::vector<double> staging;

    int dalloc(double** p, size_t n)
    {
        void* q = nullptr;
        if(hipMalloc(&q, n * sizeof(double)) != hipSuccess) return fail(REEF_E_NOMEM, "hipMalloc failed");
        allocs.push_back(q);
        if(hipMemsetAsync(q, 0, n * sizeof(double), stream) != hipSuccess) return fail(REEF_E_HIP, "hipMemset failed");
        *p = (double*)q;
        return 0;
    }
    int falloc(float** p, size_t n)
    {
        void* q = nullptr;
        if(hipMalloc(&q, n * sizeof(float)) != hipSuccess) return fail(REEF_E_NOMEM, "hipMalloc failed");
        allocs.push_back(q);
        if(hipMemsetAsync(q, 0, n * sizeof(float), stream) != hipSuccess) return fail(REEF_E_HIP, "hipMemset failed");
        *p = (float*)q;
        return 0;
    }
    template<class T> int dupload(const T* host, size_t n, const T** out)
    {
        void* q = nullptr;
        if(hipMalloc(&q, n * sizeof(T)) != hipSuccess) return fail(REEF_E_NOMEM, "hipMalloc failed");
        allocs.push_back(q);
        if(hipMemcpy(q, host, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return fail(REEF_E_HIP, "hipMemcpy failed");
        *out = (const T*)q;
        return 0;
    }

    // timing of one launch family: HIP events from a reusable pool, recorded on the
    // context stream around the launches
    std::vector<hipEvent_t> pool;
    hipEvent_t take()
    {
        if(pool.empty())
        {
            hipEvent_t e;
            hipEventCreate(&e);
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    void tbegin(const char* name)
    {
        if(!timing) return;
        TimedRange r;
        r.name = name;
        r.a = take();
        r.b = take();
        hipEventRecord(r.a, stream);
        ranges.push_back(r);
    }
    void tend()
    {
        if(!timing || ranges.empty()) return;
        hipEventRecord(ranges.back().b, stream);
    }
    void tcollect()
    {
        hipStreamSynchronize(stream);
        for(auto& r : ranges)
        {
            float ms = 0.f;
            hipEventElapsedTime(&ms, r.a, r.b);
            auto& t = totals[r.name];
            t.first += ms;
            t.second += 1;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        ranges.clear();
    }

    int ghost(double* f, int kind, int gcv)
    {
        auto key = std::make_pair(kind, gcv);
        auto it = ghost_progs.find(key);
        if(it == ghost_progs.end())
        {
            std::vector<GhostOp> ops;
            std::string err = build_ghost_program(hg, kind, gcv, ops);
            if(!err.empty()) return fail(REEF_E_UNSUPPORTED, "ghost program: " + err);
            const GhostOp* d = nullptr;
            if(!ops.empty())
            {
                int rc = dupload(ops.data(), ops.size(), &d);
                if(rc) return rc;
            }
            it = ghost_progs.emplace(key, std::make_pair((GhostOp*)d, (int)ops.size())).first;
        }
        tbegin("ghost");
        launch_ghost(it->second.first, it->second.second, f, stream);
        tend();
        return 0;
    }
};

namespace {

int check_ctx(reef_ctx* c) { return c ? 0 : fail(REEF_E_ARG, "null context"); }

double* field_ptr(reef_ctx* c, int id)
{
    if(id < 0 || id >= REEF_NFIELDS) return nullptr;
    return c->field[id];
}

int do_reini(reef_ctx* c, double* f, int iters)
{
    const DevGrid& g = c->dg;
    int rc;
    for(int q = 0; q < iters; ++q)
    {
        c->tbegin("reini");
        launch_reini_stage(g, 1, f, f, c->rk1, c->prm.F43, c->F49 == 0, c->stream);
        c->tend();
        if((rc = c->ghost(c->rk1, 4, 52))) return rc;
        c->tbegin("reini");
        launch_reini_stage(g, 2, c->rk1, f, c->rk2, c->prm.F43, c->F49 == 0, c->stream);
        c->tend();
        if((rc = c->ghost(c->rk2, 4, 52))) return rc;
        c->tbegin("reini");
        launch_reini_stage(g, 3, c->rk2, f, f, c->prm.F43, c->F49 == 0, c->stream);
        c->tend();
        if((rc = c->ghost(f, 4, 52))) return rc;
    }
    return 0;
}

int do_fluid_update(reef_ctx* c, double* vol2)
{
    const double prm4[4] = {c->prm.W1, c->prm.W2, c->prm.W3, c->prm.W4};
    c->tbegin("heaviside");
    launch_heaviside(c->dg, c->field[REEF_PHI], c->field[REEF_RO], c->field[REEF_VISC], prm4, c->partials,
                     c->vol_dev, c->vol_ticket, c->stream);
    c->tend();
    int rc;
    if((rc = c->ghost(c->field[REEF_RO], 4, 1))) return rc;
    if((rc = c->ghost(c->field[REEF_VISC], 4, 1))) return rc;
    if(vol2)
    {
        HIPCHK(hipMemcpyAsync(vol2, c->vol_dev, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return 0;
}

const char* const BICG_NAMES[5] = {"bicg_p", "bicg_v", "bicg_t", "bicg_x", ""};
const char* const MIX_NAMES[5] = {"mix_p", "bicg32_v", "bicg32_t", "bicg32_x", "mix_ru"};

// BiCGStab iterations after the solver's init kernels, enqueued in batches; the device
// state says when the solve ended.  Two batches are in flight: batch k+1 is enqueued
// before the host waits for the status copied out after batch k, so the device does not
// idle while the host polls.  Iterations launched after the end are no-ops (every kernel
// tests done).  launch(phase, part) enqueues part 0 (grid kernel) or 1 (finalize) of a
// phase; nphase phases make one iteration.
template<class Launch>
int run_iterations(reef_ctx* c, const BiState* st_dev, int nphase, const char* const* names, Launch launch,
                   int maxiter, BiState& hs)
{
    const int batch = 8;
    auto enqueue = [&](long k) -> int {
        for(int q = 0; q < batch; ++q)
            for(int ph = 0; ph < nphase; ++ph)
            {
                c->tbegin(names[ph]);
                launch(ph, 0);
                c->tend();
                launch(ph, 1);
            }
        HIPCHK(hipMemcpyAsync(&c->st_host[k & 1], st_dev, sizeof(BiState), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipEventRecord(c->st_ev[k & 1], c->stream));
        return 0;
    };
    int rc;
    if((rc = enqueue(0))) return rc;
    for(long k = 1;; ++k)
    {
        if((rc = enqueue(k))) return rc;
        HIPCHK(hipEventSynchronize(c->st_ev[(k - 1) & 1]));
        hs = c->st_host[(k - 1) & 1];
        if(hs.done) break;
        if(c->timing && c->ranges.size() > 2048) c->tcollect();
        // every launched iteration advances iter, restarts or updates; each of these
        // counts toward maxiter, so the solve ends within a bounded number of batches
        if(k * batch > 4L * (maxiter + 16)) return fail(REEF_E_HIP, "BiCGStab did not terminate");
    }
    if(c->timing) c->tcollect();
    return 0;
}

// BiCGStab on the operator assembled from REEF_PHI; x holds the initial guess
int do_solve(reef_ctx* c, double* x, const double* rhs, double stop, int maxiter, int* iters, double* residual)
{
    const DevGrid& g = c->dg;
    BiVec<double> b = c->bi;
    b.x = x;
    b.rhs = rhs;
    c->tbegin("face_rho");
    launch_face_rho(g, c->field[REEF_PHI], c->prm.W1, c->prm.W3, c->rfx, c->rfy, c->rfz, c->aii64, c->stream);
    c->tend();
    c->tbegin("bicg_init");
    launch_bicg_init(g, c->op64, b, stop, maxiter, c->stream);
    c->tend();
    BiState hs;
    int rc;
    auto launch = [&](int ph, int part) { launch_bicg_phase(g, c->op64, b, ph, part, c->stream); };
    if((rc = run_iterations(c, b.st, 4, BICG_NAMES, launch, maxiter, hs))) return rc;
    if(iters) *iters = hs.iter;
    if(residual) *residual = hs.residual;
    return 0;
}

int alloc_mixed(reef_ctx* c)
{
    if(c->r32) return 0;
    const size_t n = (size_t)c->dg.nalloc;
    BiVec<float>& b = c->bf;
    int rc;
    float* fx = nullptr;
    float** bw[] = {&b.r, &b.r0, &b.p, &b.ph, &b.v, &b.t, &fx, &c->kx32, &c->ky32, &c->kz32, &c->aii32, &c->r32};
    for(float** p : bw)
        if((rc = c->falloc(p, n))) return rc;
    b.x = fx;
    b.rhs = nullptr;
    b.aii = c->aii32;
    b.partials = c->bi.partials;
    double* stmem = nullptr;
    if((rc = c->dalloc(&stmem, (sizeof(BiState) + 7) / 8 + 8))) return rc;
    b.st = (BiState*)stmem;
    if(!c->norm_dev && (rc = c->dalloc(&c->norm_dev, 8))) return rc;
    const HostGrid& h = c->hg;
    const float** tab[3] = {&c->op32.iDXN, &c->op32.iDYN, &c->op32.iDZN};
    for(int a = 0; a < 3; ++a)
    {
        std::vector<float> inv(h.DN[a].size());
        for(size_t q = 0; q < inv.size(); ++q) inv[q] = h.DN[a][q] != 0.0 ? (float)(1.0 / h.DN[a][q]) : 0.0f;
        if((rc = c->dupload(inv.data(), inv.size(), tab[a]))) return rc;
    }
    c->op32.kx = c->kx32;
    c->op32.ky = c->ky32;
    c->op32.kz = c->kz32;
    return 0;
}

// Mixed-precision projection solve: BiCGStab iterating in FP32 (FP32 face conductances and
// Jacobi preconditioner, FP64 dot products) with FP64 reliable updates of the residual
// (launch_mixed_phase); stops when the true FP64 residual meets the criterion (stop >= 0:
// ||r||_2 / N < stop as bicgstab_ijk.cpp:264; stop < 0: ||r||_2 < |stop| ||b - A x0||_2).
// iters receives the iterations (restarts counted as the reference counts them), nru the
// reliable updates, rel_residual the final ||b - A x||_2 / ||b - A x0||_2.
int do_solve_mixed(reef_ctx* c, double* x, const double* rhs, double stop, int maxiter, int* iters, int* nru,
                   double* rel_residual)
{
    int rc;
    if((rc = alloc_mixed(c))) return rc;
    const DevGrid& g = c->dg;
    c->tbegin("face_rho");
    launch_face_rho(g, c->field[REEF_PHI], c->prm.W1, c->prm.W3, c->rfx, c->rfy, c->rfz, nullptr, c->stream);
    c->tend();
    c->tbegin("coef32");
    launch_coef32(g, c->rfx, c->rfy, c->rfz, c->kx32, c->ky32, c->kz32, c->aii32, c->stream);
    c->tend();
    const MixedVec m{x, rhs};
    c->tbegin("mix_init");
    launch_mixed_init(g, c->op64, m, c->bf, stop, maxiter, c->delta, c->stream);
    c->tend();
    BiState hs;
    auto launch = [&](int ph, int part) { launch_mixed_phase(g, c->op64, c->op32, m, c->bf, ph, part, c->stream); };
    if((rc = run_iterations(c, c->bf.st, 5, MIX_NAMES, launch, maxiter, hs))) return rc;
    c->tbegin("add32");
    launch_add32(g, x, c->bf.x, c->stream);
    c->tend();
    c->last_inner = hs.iter;
    c->last_outer = hs.nru;
    c->last_rel = hs.res0 > 0.0 ? hs.residual * double(hs.ncell) / hs.res0 : 0.0;
    if(iters) *iters = hs.iter;
    if(nru) *nru = hs.nru;
    if(rel_residual) *rel_residual = c->last_rel;
    return 0;
}

int do_pjm(reef_ctx* c, double* u, double* v, double* w, double alpha, int* iters)
{
    const DevGrid& g = c->dg;
    int rc;
    if((rc = c->ghost(u, 1, 7))) return rc;
    if((rc = c->ghost(v, 2, 8))) return rc;
    if((rc = c->ghost(w, 3, 9))) return rc;
    const double adt = alpha * c->prm.dt;
    c->tbegin("poisson_rhs");
    launch_poisson_rhs(g, u, v, w, adt, c->rhs, c->stream);
    c->tend();
    double* pc = c->field[REEF_PCORR];
    launch_fill(pc, g.nalloc, 0.0, c->stream);
    int it = 0;
    if(c->precision == REEF_SOLVER_MIXED)
        rc = do_solve_mixed(c, pc, c->rhs, c->prm.N44, c->prm.N46, &it, nullptr, nullptr);
    else
        rc = do_solve(c, pc, c->rhs, c->prm.N44, c->prm.N46, &it, nullptr);
    if(rc) return rc;
    if((rc = c->ghost(pc, 4, 40))) return rc;
    c->tbegin("correct");
    launch_correct(g, pc, c->field[REEF_PHI], c->prm.W1, c->prm.W3, adt, u, v, w, c->field[REEF_PRESS], c->stream);
    c->tend();
    if((rc = c->ghost(c->field[REEF_PRESS], 4, 40))) return rc;
    if(iters) *iters = it;
    return 0;
}

int do_momentum(reef_ctx* c, int stage, double* const qin[3], double* const qold[3], double* const vel[3],
                double* const out[3])
{
    const double rwa[2] = {c->prm.W1, c->prm.W3};
    const double grav[3] = {c->prm.W20, c->prm.W21, c->prm.W22};
    const double dt = c->prm.dt;
    const double cdt = stage == 1 ? dt : stage == 2 ? 0.25 * dt : (2.0 / 3.0) * dt;
    for(int comp = 1; comp <= 3; ++comp)
    {
        c->tbegin("momentum");
        launch_momentum(c->dg, comp, stage, qin[comp - 1], qold[comp - 1], vel[0], vel[1], vel[2],
                        c->field[REEF_PRESS], c->field[REEF_PHI], grav[comp - 1], cdt, out[comp - 1], rwa, c->stream);
        c->tend();
    }
    return 0;
}

int do_fc3_step(reef_ctx* c, int* iters_out, double* vol2)
{
    const DevGrid& g = c->dg;
    const double dt = c->prm.dt;
    double* phi = c->field[REEF_PHI];
    double *u = c->field[REEF_U], *v = c->field[REEF_V], *w = c->field[REEF_W];
    int rc, it, itsum = 0;
    // ---- stage 1 (momentum_FC3.cpp:132-230)
    c->tbegin("copy");
    launch_copy_interior(g, phi, c->ls, c->stream);   // ls ghosts stay zero, as in the reference
    c->tend();
    c->tbegin("hj");
    launch_hj_stage(g, 1, c->ls, nullptr, u, v, w, c->frk1, dt, c->stream);
    c->tend();
    if((rc = c->ghost(c->frk1, 4, 52))) return rc;
    if((rc = do_reini(c, c->frk1, 2))) return rc;
    c->tbegin("copy");
    launch_copy_interior(g, c->frk1, phi, c->stream);
    c->tend();
    if((rc = c->ghost(phi, 4, 52))) return rc;
    {
        double* qin[3] = {u, v, w};
        double* vel[3] = {u, v, w};
        double* out[3] = {c->urk1, c->vrk1, c->wrk1};
        if((rc = do_momentum(c, 1, qin, qin, vel, out))) return rc;
    }
    if((rc = do_pjm(c, c->urk1, c->vrk1, c->wrk1, 1.0, &it))) return rc;
    itsum += it;
    if((rc = c->ghost(c->urk1, 1, 10)) || (rc = c->ghost(c->vrk1, 2, 11)) || (rc = c->ghost(c->wrk1, 3, 12))) return rc;
    // ---- stage 2 (momentum_FC3.cpp:232-332)
    c->tbegin("hj");
    launch_hj_stage(g, 2, c->frk1, c->ls, c->urk1, c->vrk1, c->wrk1, c->frk2, 0.25 * dt, c->stream);
    c->tend();
    if((rc = c->ghost(c->frk2, 4, 52))) return rc;
    if((rc = do_reini(c, c->frk2, 2))) return rc;
    c->tbegin("copy");
    launch_copy_interior(g, c->frk2, phi, c->stream);
    c->tend();
    if((rc = c->ghost(phi, 4, 52))) return rc;
    if((rc = do_reini(c, phi, 2))) return rc;
    {
        double* qin[3] = {c->urk1, c->vrk1, c->wrk1};
        double* qold[3] = {u, v, w};
        double* out[3] = {c->urk2, c->vrk2, c->wrk2};
        if((rc = do_momentum(c, 2, qin, qold, qin, out))) return rc;
    }
    if((rc = do_pjm(c, c->urk2, c->vrk2, c->wrk2, 0.25, &it))) return rc;
    itsum += it;
    if((rc = c->ghost(c->urk2, 1, 10)) || (rc = c->ghost(c->vrk2, 2, 11)) || (rc = c->ghost(c->wrk2, 3, 12))) return rc;
    // ---- stage 3 (momentum_FC3.cpp:334-430)
    c->tbegin("hj");
    launch_hj_stage(g, 3, c->frk2, c->ls, c->urk2, c->vrk2, c->wrk2, phi, (2.0 / 3.0) * dt, c->stream);
    c->tend();
    if((rc = c->ghost(phi, 4, 52))) return rc;
    if((rc = do_reini(c, phi, 3))) return rc;
    {
        double* qin[3] = {c->urk2, c->vrk2, c->wrk2};
        double* qold[3] = {u, v, w};
        double* out[3] = {u, v, w};
        if((rc = do_momentum(c, 3, qin, qold, qin, out))) return rc;
    }
    if((rc = do_pjm(c, u, v, w, 2.0 / 3.0, &it))) return rc;
    itsum += it;
    if((rc = c->ghost(u, 1, 10)) || (rc = c->ghost(v, 2, 11)) || (rc = c->ghost(w, 3, 12))) return rc;
    // rho and nu depend on phi only: the last fluid_update of the step is the one that stays
    if((rc = do_fluid_update(c, vol2))) return rc;
    if(iters_out) *iters_out = itsum;
    return 0;
}

int do_phi_rk3(reef_ctx* c)
{
    const DevGrid& g = c->dg;
    const double dt = c->prm.dt;
    double *phi = c->field[REEF_PHI], *u = c->field[REEF_U], *v = c->field[REEF_V], *w = c->field[REEF_W];
    int rc;
    c->tbegin("hj");
    launch_hj_stage(g, 1, phi, nullptr, u, v, w, c->frk1, dt, c->stream);
    c->tend();
    if((rc = c->ghost(c->frk1, 4, 52))) return rc;
    c->tbegin("hj");
    launch_hj_stage(g, 2, c->frk1, phi, u, v, w, c->frk2, 0.25 * dt, c->stream);
    c->tend();
    if((rc = c->ghost(c->frk2, 4, 52))) return rc;
    c->tbegin("hj");
    launch_hj_stage(g, 3, c->frk2, phi, u, v, w, phi, (2.0 / 3.0) * dt, c->stream);
    c->tend();
    return c->ghost(phi, 4, 52);
}

int do_vortex_step(reef_ctx* c, double simtime, double* vol2)
{
    int rc;
    if((rc = do_phi_rk3(c))) return rc;
    if((rc = do_reini(c, c->field[REEF_PHI], 3))) return rc;
    // momentum_RK3 with void operators, then benchmark_vortex3D::start
    const double grav[3] = {c->prm.W20, c->prm.W21, c->prm.W22};
    for(int comp = 1; comp <= 3; ++comp)
    {
        if(comp == 2 && c->dg.jdir == 0) continue;
        const double Fv = 0.0 + (0.0 + grav[comp - 1] + 0.0 + 0.0) * 1.0;
        launch_void_rk3(c->dg, comp, c->field[REEF_U + comp - 1], c->prm.dt, Fv, c->stream);
        if((rc = c->ghost(c->field[REEF_U + comp - 1], comp, 9 + comp))) return rc;
    }
    for(int comp = 1; comp <= 3; ++comp)
        launch_vortex_velocity(c->dg, comp, c->field[REEF_U + comp - 1], simtime, c->hg.DXM, c->stream);
    if((rc = c->ghost(c->field[REEF_U], 1, 10))) return rc;
    if((rc = c->ghost(c->field[REEF_V], 2, 11))) return rc;
    if((rc = c->ghost(c->field[REEF_V], 2, 12))) return rc;
    return do_fluid_update(c, vol2);
}

}  // namespace

// ------------------------------------------------------------------ C-ABI
extern "C" {

const char* reef_last_error(void) { return g_err.c_str(); }

int reef_ctx_create(const reef_grid_desc* grid, const reef_params* prm, reef_ctx** out)
{
    if(!grid || !prm || !out || !grid->XN || !grid->YN || !grid->ZN) return fail(REEF_E_ARG, "null argument");
    *out = nullptr;
    std::unique_ptr<reef_ctx> c(new reef_ctx());
    c->prm = *prm;
    std::string err = build_grid(c->hg, *grid, prm->F45);
    if(!err.empty()) return fail(REEF_E_UNSUPPORTED, err);
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HostGrid& h = c->hg;
    DevGrid& g = c->dg;
    g.nx = h.nx; g.ny = h.ny; g.nz = h.nz; g.jdir = h.jdir;
    g.sx = h.sx; g.sxy = h.sxy; g.base = h.base; g.nalloc = h.nalloc;
    g.ydir = h.jdir == 1 ? 1.0 : 0.0;
    g.psi = h.psi;
    int rc;
    const double** d1[3][3] = {{&g.XP, &g.DXN, &g.DXP}, {&g.YP, &g.DYN, &g.DYP}, {&g.ZP, &g.DZN, &g.DZP}};
    for(int a = 0; a < 3; ++a)
    {
        if((rc = c->dupload(h.P[a].data(), h.P[a].size(), d1[a][0]))) return rc;
        if((rc = c->dupload(h.DN[a].data(), h.DN[a].size(), d1[a][1]))) return rc;
        if((rc = c->dupload(h.DP[a].data(), h.DP[a].size(), d1[a][2]))) return rc;
        if((rc = c->dupload(h.qf[a].data(), h.qf[a].size(), &g.qf[a]))) return rc;
        if((rc = c->dupload(h.cf[a].data(), h.cf[a].size(), &g.cf[a]))) return rc;
        if((rc = c->dupload(h.isf[a].data(), h.isf[a].size(), &g.isf[a]))) return rc;
    }
    const size_t n = (size_t)g.nalloc;
    for(int f = 0; f < REEF_NFIELDS; ++f)
        if((rc = c->dalloc(&c->field[f], n))) return rc;
    double** work[] = {&c->ls, &c->frk1, &c->frk2, &c->rk1, &c->rk2, &c->urk1, &c->vrk1, &c->wrk1,
                       &c->urk2, &c->vrk2, &c->wrk2, &c->rhs};
    for(double** p : work)
        if((rc = c->dalloc(p, n))) return rc;
    BiVec<double>& b = c->bi;
    double** bw[] = {&b.r, &b.r0, &b.p, &b.ph, &b.v, &b.t, &c->aii64, &c->rfx, &c->rfy, &c->rfz};
    for(double** p : bw)
        if((rc = c->dalloc(p, n))) return rc;
    b.aii = c->aii64;
    c->op64.rfx = c->rfx;
    c->op64.rfy = c->rfy;
    c->op64.rfz = c->rfz;
    const long nblocks = ((long)((g.nx + 63) / 64) * ((g.ny + 3) / 4) * g.nz + 7) / 8 * 8;
    if((rc = c->dalloc(&c->partials, (size_t)nblocks * 4 + 64))) return rc;
    if((rc = c->dalloc(&b.partials, (size_t)nblocks * 4 + 64))) return rc;
    double* stmem = nullptr;
    if((rc = c->dalloc(&stmem, (sizeof(BiState) + 7) / 8 + 8))) return rc;
    b.st = (BiState*)stmem;
    if((rc = c->dalloc(&c->vol_dev, 4))) return rc;
    HIPCHK(hipHostMalloc((void**)&c->st_host, 2 * sizeof(BiState), hipHostMallocDefault));
    std::memset(c->st_host, 0, 2 * sizeof(BiState));
    for(hipEvent_t& e : c->st_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    double* tk = nullptr;
    if((rc = c->dalloc(&tk, 4))) return rc;
    c->vol_ticket = (unsigned int*)tk;
    HIPCHK(hipStreamSynchronize(c->stream));
    *out = c.release();
    return 0;
}

void reef_ctx_destroy(reef_ctx* c)
{
    if(!c) return;
    if(c->stream) hipStreamSynchronize(c->stream);
    for(auto& r : c->ranges)
    {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    for(hipEvent_t e : c->pool) hipEventDestroy(e);
    for(hipEvent_t e : c->st_ev)
        if(e) hipEventDestroy(e);
    if(c->st_host) hipHostFree(c->st_host);
    for(void* p : c->allocs) hipFree(p);
    if(c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int reef_host_grid_arrays(const reef_grid_desc* grid, double F45, int axis, double* P, double* DN, double* DP,
                          double* qf, double* cf, double* isf, double* scalars4)
{
    if(!grid || axis < 0 || axis > 2) return fail(REEF_E_ARG, "bad argument");
    HostGrid h;
    std::string err = build_grid(h, *grid, F45);
    if(!err.empty()) return fail(REEF_E_UNSUPPORTED, err);
    const int n = axis == 0 ? h.nx : axis == 1 ? h.ny : h.nz;
    for(int q = 0; q < n + 10; ++q)
    {
        if(P) P[q] = h.P[axis][q];
        if(DN) DN[q] = h.DN[axis][q];
        if(DP) DP[q] = h.DP[axis][q];
    }
    for(int ip = MARGE - 1; ip < n + MARGE + 1; ++ip)
    {
        const int r = ip - (MARGE - 1);
        for(int m = 0; m < 24; ++m) if(qf) qf[r * 24 + m] = h.qf[axis][(size_t)ip * 24 + m];
        for(int m = 0; m < 12; ++m) if(cf) cf[r * 12 + m] = h.cf[axis][(size_t)ip * 12 + m];
        for(int m = 0; m < 36; ++m) if(isf) isf[r * 36 + m] = h.isf[axis][(size_t)ip * 36 + m];
    }
    if(scalars4)
    {
        scalars4[0] = h.DXM;
        scalars4[1] = h.DYM;
        scalars4[2] = h.DZM;
        scalars4[3] = h.psi;
    }
    return 0;
}

int reef_host_ghost_apply(const reef_grid_desc* grid, double F45, int kind, int gcv, double* f)
{
    if(!grid || !f || kind < 1 || kind > 4) return fail(REEF_E_ARG, "bad argument");
    HostGrid h;
    std::string err = build_grid(h, *grid, F45);
    if(!err.empty()) return fail(REEF_E_UNSUPPORTED, err);
    std::vector<GhostOp> ops;
    err = build_ghost_program(h, kind, gcv, ops);
    if(!err.empty()) return fail(REEF_E_UNSUPPORTED, err);
    const long origin = h.base - 3 - 3 * h.sx - 3 * h.sxy;
    auto compact = [&](long d) {
        const long r = d - origin;
        const long k = r / h.sxy, rem = r - k * h.sxy;
        const long j = rem / h.sx, i = rem - j * h.sx;
        return i + (long)(h.nx + 6) * (j + (long)(h.ny + 6) * k);
    };
    for(const GhostOp& op : ops)
    {
        double val;
        if(op.src < 0)
            val = 0.0;
        else if(op.src & (1 << 30))
            val = 0.0 + op.w * f[compact(op.src & ~(1 << 30))];
        else
            val = f[compact(op.src)];
        f[compact(op.dst)] = val;
    }
    return 0;
}

int reef_ctx_sizes(reef_ctx* c, long* out5)
{
    if(check_ctx(c) || !out5) return fail(REEF_E_ARG, "null argument");
    out5[0] = (long)c->dg.nx * c->dg.ny * c->dg.nz;
    out5[1] = c->dg.nalloc;
    out5[2] = c->dg.sx;
    out5[3] = c->dg.sxy;
    out5[4] = c->dg.base - 3 - 3 * c->dg.sx - 3 * c->dg.sxy;
    return 0;
}

int reef_field_put(reef_ctx* c, int id, const double* host, int layout)
{
    if(check_ctx(c) || !host) return fail(REEF_E_ARG, "null argument");
    double* d = field_ptr(c, id);
    if(!d) return fail(REEF_E_ARG, "bad field id");
    const HostGrid& h = c->hg;
    const long X = h.nx + 6, Y = h.ny + 6, Z = h.nz + 6;
    const double* src = host;
    if(layout == REEF_LAYOUT_REF)
    {
        c->staging.resize((size_t)X * Y * Z);
        for(long i = 0; i < X; ++i)
        for(long j = 0; j < Y; ++j)
        for(long k = 0; k < Z; ++k)
            c->staging[(size_t)(i + X * (j + Y * k))] = host[(i * Y + j) * Z + k];
        src = c->staging.data();
    }
    else if(layout != REEF_LAYOUT_IFAST)
        return fail(REEF_E_ARG, "bad layout");
    double* dst0 = d + (h.base - 3 - 3 * h.sx - 3 * h.sxy);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy2D(dst0, h.sx * sizeof(double), src, X * sizeof(double), X * sizeof(double), Y * Z,
                       hipMemcpyHostToDevice));
    return 0;
}

int reef_field_get(reef_ctx* c, int id, double* host, int layout)
{
    if(check_ctx(c) || !host) return fail(REEF_E_ARG, "null argument");
    double* d = field_ptr(c, id);
    if(!d) return fail(REEF_E_ARG, "bad field id");
    const HostGrid& h = c->hg;
    const long X = h.nx + 6, Y = h.ny + 6, Z = h.nz + 6;
    double* dst = host;
    if(layout == REEF_LAYOUT_REF)
    {
        c->staging.resize((size_t)X * Y * Z);
        dst = c->staging.data();
    }
    else if(layout != REEF_LAYOUT_IFAST)
        return fail(REEF_E_ARG, "bad layout");
    const double* src0 = d + (h.base - 3 - 3 * h.sx - 3 * h.sxy);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy2D(dst, X * sizeof(double), src0, h.sx * sizeof(double), X * sizeof(double), Y * Z,
                       hipMemcpyDeviceToHost));
    if(layout == REEF_LAYOUT_REF)
        for(long i = 0; i < X; ++i)
        for(long j = 0; j < Y; ++j)
        for(long k = 0; k < Z; ++k)
            host[(i * Y + j) * Z + k] = c->staging[(size_t)(i + X * (j + Y * k))];
    return 0;
}

int reef_field_zero(reef_ctx* c, int id)
{
    if(check_ctx(c)) return REEF_E_ARG;
    double* d = field_ptr(c, id);
    if(!d) return fail(REEF_E_ARG, "bad field id");
    launch_fill(d, c->dg.nalloc, 0.0, c->stream);
    return 0;
}

int reef_ghost_fill(reef_ctx* c, int id, int kind, int gcv)
{
    if(check_ctx(c)) return REEF_E_ARG;
    double* d = field_ptr(c, id);
    if(!d || kind < 1 || kind > 4) return fail(REEF_E_ARG, "bad field id or kind");
    return c->ghost(d, kind, gcv);
}

int reef_convection(reef_ctx* c, int ipol, int q_id, int u_id, int v_id, int w_id)
{
    if(check_ctx(c)) return REEF_E_ARG;
    double *q = field_ptr(c, q_id), *u = field_ptr(c, u_id), *v = field_ptr(c, v_id), *w = field_ptr(c, w_id);
    if(!q || !u || !v || !w || ipol < 1 || ipol > 4) return fail(REEF_E_ARG, "bad field id or ipol");
    if(ipol == 4)
    {
        c->tbegin("hj");
        launch_hj_stage(c->dg, 0, q, nullptr, u, v, w, c->field[REEF_L], 0.0, c->stream);
        c->tend();
        return 0;
    }
    double* acc = c->field[REEF_F + ipol - 1];
    c->tbegin("flux");
    launch_flux_acc(c->dg, ipol, q, u, v, w, acc, c->stream);
    c->tend();
    return 0;
}

int reef_reini(reef_ctx* c, int phi_id, int iters)
{
    if(check_ctx(c)) return REEF_E_ARG;
    double* f = field_ptr(c, phi_id);
    if(!f || iters < 0) return fail(REEF_E_ARG, "bad field id or iteration count");
    return do_reini(c, f, iters);
}

int reef_fluid_update(reef_ctx* c, double* vol2)
{
    if(check_ctx(c)) return REEF_E_ARG;
    return do_fluid_update(c, vol2);
}

int reef_poisson_solve(reef_ctx* c, int x_id, int rhs_id, double stop, int maxiter, int* iters, double* residual)
{
    if(check_ctx(c)) return REEF_E_ARG;
    double *x = field_ptr(c, x_id), *rhs = field_ptr(c, rhs_id);
    if(!x || !rhs || maxiter < 0) return fail(REEF_E_ARG, "bad field id or maxiter");
    return do_solve(c, x, rhs, stop, maxiter, iters, residual);
}

int reef_poisson_solve_mixed(reef_ctx* c, int x_id, int rhs_id, double stop, int maxiter, int* iters,
                             int* updates, double* rel_residual)
{
    if(check_ctx(c)) return REEF_E_ARG;
    double *x = field_ptr(c, x_id), *rhs = field_ptr(c, rhs_id);
    if(!x || !rhs || maxiter < 0) return fail(REEF_E_ARG, "bad field id or maxiter");
    if(x == rhs) return fail(REEF_E_ARG, "x and rhs must be different fields");
    return do_solve_mixed(c, x, rhs, stop, maxiter, iters, updates, rel_residual);
}

int reef_poisson_residual(reef_ctx* c, int x_id, int rhs_id, double* norm)
{
    if(check_ctx(c)) return REEF_E_ARG;
    double *x = field_ptr(c, x_id), *rhs = field_ptr(c, rhs_id);
    if(!x || !rhs || !norm) return fail(REEF_E_ARG, "bad field id or null norm");
    int rc;
    if(!c->norm_dev && (rc = c->dalloc(&c->norm_dev, 8))) return rc;
    const DevGrid& g = c->dg;
    launch_face_rho(g, c->field[REEF_PHI], c->prm.W1, c->prm.W3, c->rfx, c->rfy, c->rfz, nullptr, c->stream);
    launch_defect(g, c->op64, x, rhs, c->bi.partials, c->norm_dev, c->stream);
    HIPCHK(hipMemcpyAsync(norm, c->norm_dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int reef_set_control(reef_ctx* c, const char* key, double value)
{
    if(check_ctx(c) || !key) return fail(REEF_E_ARG, "null argument");
    if(std::strcmp(key, "F49") == 0)
    {
        if(value != 0.0 && value != 1.0) return fail(REEF_E_UNSUPPORTED, "F49 must be 0 or 1");
        c->F49 = (int)value;
        return 0;
    }
    return fail(REEF_E_UNSUPPORTED, std::string("control key not supported: ") + key);
}

int reef_solver_config(reef_ctx* c, int precision, double delta)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(precision != REEF_SOLVER_FP64 && precision != REEF_SOLVER_MIXED) return fail(REEF_E_ARG, "bad precision");
    if(!(delta > 0.0 && delta < 1.0)) return fail(REEF_E_ARG, "delta must lie in (0, 1)");
    c->precision = precision;
    c->delta = delta;
    return 0;
}

int reef_solver_last(reef_ctx* c, long* iters, long* updates, double* rel_residual)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(iters) *iters = c->last_inner;
    if(updates) *updates = c->last_outer;
    if(rel_residual) *rel_residual = c->last_rel;
    return 0;
}

int reef_pjm_corr(reef_ctx* c, int u_id, int v_id, int w_id, double alpha, int* iters)
{
    if(check_ctx(c)) return REEF_E_ARG;
    double *u = field_ptr(c, u_id), *v = field_ptr(c, v_id), *w = field_ptr(c, w_id);
    if(!u || !v || !w) return fail(REEF_E_ARG, "bad field id");
    return do_pjm(c, u, v, w, alpha, iters);
}

int reef_fc3_step(reef_ctx* c, int* iters, double* vol2)
{
    if(check_ctx(c)) return REEF_E_ARG;
    return do_fc3_step(c, iters, vol2);
}

int reef_vortex_step(reef_ctx* c, double simtime, double* vol2)
{
    if(check_ctx(c)) return REEF_E_ARG;
    return do_vortex_step(c, simtime, vol2);
}

int reef_phi_rk3(reef_ctx* c)
{
    if(check_ctx(c)) return REEF_E_ARG;
    return do_phi_rk3(c);
}

int reef_sync(reef_ctx* c)
{
    if(check_ctx(c)) return REEF_E_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipGetLastError());
    return 0;
}

int reef_timing_enable(reef_ctx* c, int on)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!on) c->tcollect();
    c->timing = on != 0;
    return 0;
}

int reef_timing_get(reef_ctx* c, const char* name, double* ms, long* launches)
{
    if(check_ctx(c) || !name) return fail(REEF_E_ARG, "null argument");
    c->tcollect();
    auto it = c->totals.find(name);
    if(ms) *ms = it == c->totals.end() ? 0.0 : it->second.first;
    if(launches) *launches = it == c->totals.end() ? 0 : it->second.second;
    return 0;
}

int reef_timing_reset(reef_ctx* c)
{
    if(check_ctx(c)) return REEF_E_ARG;
    c->tcollect();
    c->totals.clear();
    return 0;
}

}  // extern "C"
