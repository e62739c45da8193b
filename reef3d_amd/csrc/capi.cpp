// C-ABI of the gfx950 hot-path library (include/reef3d_hip.h).  A reef_ctx is an engine
// (engine.h) with one tank; a reef_batch is an engine with any number of tanks that the
// same launches advance together.
#include "engine.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>

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

// engine status -> C-ABI status with the engine's message
int eng(Engine& e, int rc)
{
    if(rc) g_err = e.err;
    return rc;
}

bool field_ok(int id) { return id >= 0 && id < REEF_NFIELDS; }

}  // namespace

struct reef_batch
{
    hipStream_t stream = nullptr;
    std::unique_ptr<Engine> e;
};

struct reef_ctx
{
    reef_batch b;
    int tank = 0;
};

namespace {

int make_engine(reef_batch& b)
{
    HIPCHK(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
    b.e.reset(new Engine(b.stream));
    return 0;
}

void drop_engine(reef_batch& b)
{
    if(b.stream) (void)hipStreamSynchronize(b.stream);
    b.e.reset();
    if(b.stream) (void)hipStreamDestroy(b.stream);
    b.stream = nullptr;
}

int check_ctx(reef_ctx* c) { return c && c->b.e ? 0 : fail(REEF_E_ARG, "null context"); }
int check_batch(reef_batch* b) { return b && b->e ? 0 : fail(REEF_E_ARG, "null batch"); }

int control(Engine& e, const char* key, double value)
{
    if(!key) return fail(REEF_E_ARG, "null argument");
    if(std::strcmp(key, "F49") == 0)
    {
        if(value != 0.0 && value != 1.0) return fail(REEF_E_UNSUPPORTED, "F49 must be 0 or 1");
        e.drop_graphs();
        e.fsf = value == 0.0 ? 1 : 0;
        return 0;
    }
    if(std::strcmp(key, "simtime") == 0)
        return eng(e, e.set_simtime(value));
    if(std::strcmp(key, "graphs") == 0)
    {
        if(value != 0.0 && value != 1.0) return fail(REEF_E_ARG, "graphs must be 0 or 1");
        e.drop_graphs();
        e.graphs = value != 0.0;
        return 0;
    }
    if(std::strcmp(key, "D20") == 0)
    {
        if(value != 0.0 && value != 1.0 && value != 2.0)
            return fail(REEF_E_UNSUPPORTED, "D20 must be 0 (diff_void), 1 (ediff2) or 2 (idiff2_FS)");
        if(value == 1.0 && e.n48)
            return fail(REEF_E_UNSUPPORTED, "D20 = 1 with N48 = 1: the reference steps it with another time-step class");
        return eng(e, e.set_d20((int)value));
    }
    if(std::strcmp(key, "N43") == 0)
    {
        if(!(value > 0.0)) return fail(REEF_E_ARG, "N43 must be > 0");
        e.drop_graphs();
        e.n43 = value;
        return 0;
    }
    if(std::strcmp(key, "N48") == 0)
    {
        if(value != 0.0 && value != 1.0) return fail(REEF_E_UNSUPPORTED, "N48 must be 0 (fixed) or 1 (adaptive)");
        if(value == 1.0 && e.d20 == 1)
            return fail(REEF_E_UNSUPPORTED, "N48 = 1 with D20 = 1: the reference steps it with another time-step class");
        e.drop_graphs();
        e.n48 = value != 0.0;
        return 0;
    }
    if(std::strcmp(key, "N47") == 0)
    {
        if(!(value > 0.0)) return fail(REEF_E_ARG, "N47 must be > 0");
        e.drop_graphs();
        e.n47 = value;
        return 0;
    }
    return fail(REEF_E_UNSUPPORTED, std::string("control key not supported: ") + key);
}

int solver_config(Engine& e, int precision, double delta)
{
    if(precision != REEF_SOLVER_FP64 && precision != REEF_SOLVER_MIXED) return fail(REEF_E_ARG, "bad precision");
    if(!(delta > 0.0 && delta < 1.0)) return fail(REEF_E_ARG, "delta must lie in (0, 1)");
    e.drop_graphs();
    e.precision = precision;
    e.delta = delta;
    return 0;
}

int timing_get(Engine& e, const char* name, double* ms, long* launches)
{
    if(!name) return fail(REEF_E_ARG, "null argument");
    e.tcollect();
    auto it = e.totals.find(name);
    if(ms) *ms = it == e.totals.end() ? 0.0 : it->second.first;
    if(launches) *launches = it == e.totals.end() ? 0 : it->second.second;
    return 0;
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
    int rc;
    if((rc = make_engine(c->b))) return rc;
    // one tank: the pool holds exactly its arrays
    TankBuild tb;
    std::string err;
    if((rc = Engine::build_tank(*grid, *prm, tb, err)))
    {
        drop_engine(c->b);
        return fail(rc, err);
    }
    c->b.e->pool.chunk_bytes = Engine::built_bytes(tb) + ((size_t)1 << 20);
    if((rc = c->b.e->add_built(tb, &c->tank)))
    {
        eng(*c->b.e, rc);
        drop_engine(c->b);
        return rc;
    }
    // set-up: the device tables, and the address of the allocation-free scratch on this GPU
    // (reef_fluid_update_noalloc)
    double* scratch_dev = nullptr;
    if((rc = c->b.e->prepare()) || (rc = noalloc_scratch(&scratch_dev)))
    {
        if(rc == REEF_E_HIP && c->b.e->err.empty()) c->b.e->err = "hipGetSymbolAddress of the scratch array failed";
        eng(*c->b.e, rc);
        drop_engine(c->b);
        return rc;
    }
    HIPCHK(hipStreamSynchronize(c->b.stream));
    *out = c.release();
    return 0;
}

void reef_ctx_destroy(reef_ctx* c)
{
    if(!c) return;
    drop_engine(c->b);
    delete c;
}

int reef_set_control(reef_ctx* c, const char* key, double value)
{
    if(check_ctx(c)) return REEF_E_ARG;
    return control(*c->b.e, key, value);
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
    const HostGrid& h = c->b.e->tank(c->tank)->hg;
    out5[0] = (long)h.nx * h.ny * h.nz;
    out5[1] = h.nalloc;
    out5[2] = h.sx;
    out5[3] = h.sxy;
    out5[4] = h.base - 3 - 3 * h.sx - 3 * h.sxy;
    return 0;
}

int reef_field_put(reef_ctx* c, int id, const double* host, int layout)
{
    if(check_ctx(c) || !host) return fail(REEF_E_ARG, "null argument");
    if(!field_ok(id)) return fail(REEF_E_ARG, "bad field id");
    return eng(*c->b.e, c->b.e->put(c->tank, id, host, layout));
}

int reef_field_get(reef_ctx* c, int id, double* host, int layout)
{
    if(check_ctx(c) || !host) return fail(REEF_E_ARG, "null argument");
    if(!field_ok(id)) return fail(REEF_E_ARG, "bad field id");
    return eng(*c->b.e, c->b.e->get(c->tank, id, host, layout));
}

int reef_field_zero(reef_ctx* c, int id)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!field_ok(id)) return fail(REEF_E_ARG, "bad field id");
    return eng(*c->b.e, c->b.e->zero(id));
}

int reef_ghost_fill(reef_ctx* c, int id, int kind, int gcv)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!field_ok(id) || kind < 1 || kind > 4) return fail(REEF_E_ARG, "bad field id or kind");
    return eng(*c->b.e, c->b.e->ghost(id, kind, gcv));
}

int reef_convection(reef_ctx* c, int ipol, int q_id, int u_id, int v_id, int w_id)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!field_ok(q_id) || !field_ok(u_id) || !field_ok(v_id) || !field_ok(w_id) || ipol < 1 || ipol > 4)
        return fail(REEF_E_ARG, "bad field id or ipol");
    return eng(*c->b.e, c->b.e->convection(ipol, q_id, u_id, v_id, w_id));
}

int reef_reini(reef_ctx* c, int phi_id, int iters)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!field_ok(phi_id) || iters < 0) return fail(REEF_E_ARG, "bad field id or iteration count");
    return eng(*c->b.e, c->b.e->reini(phi_id, iters));
}

int reef_reini_ini(reef_ctx* c, int phi_id, int iters, int gcv)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!field_ok(phi_id) || iters < 0) return fail(REEF_E_ARG, "bad field id or iteration count");
    if(gcv != 50 && gcv != 52) return fail(REEF_E_UNSUPPORTED, "initial reinitialisation ghost rule must be 50 or 52");
    int rc;
    if((rc = eng(*c->b.e, c->b.e->ghost(phi_id, 4, gcv)))) return rc;
    return eng(*c->b.e, c->b.e->reini(phi_id, iters, gcv));
}

int reef_fluid_update(reef_ctx* c, double* vol2)
{
    if(check_ctx(c)) return REEF_E_ARG;
    Engine& e = *c->b.e;
    int rc;
    if((rc = e.fluid_update())) return eng(e, rc);
    if(vol2)
    {
        if((rc = e.read_volumes())) return eng(e, rc);
        vol2[0] = e.vol[2 * c->tank];
        vol2[1] = e.vol[2 * c->tank + 1];
    }
    return 0;
}

int reef_fluid_update_noalloc(reef_ctx* c, double* vol2)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!vol2) return fail(REEF_E_ARG, "null vol2");
    Engine& e = *c->b.e;
    int rc;
    if((rc = e.fluid_update_noalloc(c->tank, vol2))) return eng(e, rc);
    return 0;
}

int reef_wave_set(reef_ctx* c, const reef_wave_desc* w, double* params5)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!w) return fail(REEF_E_ARG, "null argument");
    return eng(*c->b.e, c->b.e->wave_set(c->tank, *w, params5));
}

int reef_wave_relax(reef_ctx* c, double simtime, int u_id, int v_id, int w_id, int press_id, int f_id)
{
    if(check_ctx(c)) return REEF_E_ARG;
    const bool vel = u_id >= 0;
    if((vel && (!field_ok(u_id) || !field_ok(v_id) || !field_ok(w_id))) || press_id >= REEF_NFIELDS ||
       f_id >= REEF_NFIELDS)
        return fail(REEF_E_ARG, "bad field id");
    Engine& e = *c->b.e;
    const TankHost* t = e.tank(c->tank);
    if(!t || !t->dev.wave.col) return fail(REEF_E_ARG, "no wave generator on this context (reef_wave_set)");
    int rc;
    if((rc = e.set_simtime(simtime))) return eng(e, rc);
    return eng(e, e.wave_relax(vel ? u_id : -1, vel ? v_id : -1, vel ? w_id : -1, press_id < 0 ? -1 : press_id,
                               f_id < 0 ? -1 : f_id));
}

int reef_poisson_solve(reef_ctx* c, int x_id, int rhs_id, double stop, int maxiter, int* iters, double* residual)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!field_ok(x_id) || !field_ok(rhs_id) || maxiter < 0) return fail(REEF_E_ARG, "bad field id or maxiter");
    Engine& e = *c->b.e;
    int rc;
    if((rc = e.solve(x_id, rhs_id, stop, maxiter, false))) return eng(e, rc);
    if(iters) *iters = e.last_state[c->tank].iter;
    if(residual) *residual = e.last_state[c->tank].residual;
    return 0;
}

int reef_poisson_solve_mixed(reef_ctx* c, int x_id, int rhs_id, double stop, int maxiter, int* iters, int* updates,
                             double* rel_residual)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!field_ok(x_id) || !field_ok(rhs_id) || maxiter < 0) return fail(REEF_E_ARG, "bad field id or maxiter");
    if(x_id == rhs_id) return fail(REEF_E_ARG, "x and rhs must be different fields");
    Engine& e = *c->b.e;
    int rc;
    if((rc = e.solve(x_id, rhs_id, stop, maxiter, true))) return eng(e, rc);
    const BiState& s = e.last_state[c->tank];
    if(iters) *iters = s.iter;
    if(updates) *updates = s.nru;
    if(rel_residual) *rel_residual = s.res0 > 0.0 ? s.residual * double(s.ncell) / s.res0 : 0.0;
    return 0;
}

long reef_solver_veclen(reef_ctx* c, int var)
{
    if(check_ctx(c) || var < 1 || var > 5) return -1;
    const HostGrid& h = c->b.e->tank(c->tank)->hg;
    return (long)(h.nx - (var == 1)) * (h.ny - (var == 2)) * (h.nz - (var == 3));
}

int reef_solver_start(reef_ctx* c, int var, const double* const* M7, const double* rhs, int x_id, double stop,
                      int maxiter, int* iters, double* residual)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(var < 1 || var > 5 || !field_ok(x_id) || !M7 || !rhs || maxiter < 0) return fail(REEF_E_ARG, "bad argument");
    for(int q = 0; q < 7; ++q)
        if(!M7[q]) return fail(REEF_E_ARG, "null diagonal");
    Engine& E = *c->b.e;
    int rc;
    if((rc = eng(E, E.ensure_matrix(c->tank)))) return rc;
    const TankHost* t = E.tank(c->tank);
    const HostGrid& h = t->hg;
    const long X = h.nx + 6, Y = h.ny + 6, Z = h.nz + 6;
    const int ie = h.nx - (var == 1), je = h.ny - (var == 2), ke = h.nz - (var == 3);
    // vec order (fillxvec, bicgstab_ijk.cpp:366-378): i outermost, k innermost
    std::vector<double> box((size_t)(X * Y * Z));
    const double* src[8] = {M7[0], M7[1], M7[2], M7[3], M7[4], M7[5], M7[6], rhs};
    for(int q = 0; q < 8; ++q)
    {
        std::fill(box.begin(), box.end(), 0.0);
        long n = 0;
        for(int i = 0; i < ie; ++i)
        for(int j = 0; j < je; ++j)
        for(int k = 0; k < ke; ++k)
            box[(size_t)((i + 3) + X * ((j + 3) + Y * (k + 3)))] = src[q][n++];
        if((rc = eng(E, E.put_array(c->tank, t->dev.m[q], box.data(), REEF_LAYOUT_IFAST)))) return rc;
    }
    if((rc = eng(E, E.mat_solve(var, x_id, stop, maxiter)))) return rc;
    const BiState& st = E.last_state[c->tank];
    if(iters) *iters = st.iter;
    if(residual) *residual = st.residual;
    return 0;
}

int reef_poisson_residual(reef_ctx* c, int x_id, int rhs_id, double* norm)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!field_ok(x_id) || !field_ok(rhs_id) || !norm) return fail(REEF_E_ARG, "bad field id or null norm");
    Engine& e = *c->b.e;
    std::vector<double> norms;
    int rc;
    if((rc = e.residual(x_id, rhs_id, norms))) return eng(e, rc);
    *norm = norms[c->tank];
    return 0;
}

int reef_solver_config(reef_ctx* c, int precision, double delta)
{
    if(check_ctx(c)) return REEF_E_ARG;
    return solver_config(*c->b.e, precision, delta);
}

int reef_solver_last(reef_ctx* c, long* iters, long* updates, double* rel_residual)
{
    if(check_ctx(c)) return REEF_E_ARG;
    Engine& e = *c->b.e;
    if((size_t)c->tank >= e.last_state.size()) return fail(REEF_E_ARG, "no solve yet");
    const BiState& s = e.last_state[c->tank];
    if(iters) *iters = s.iter;
    if(updates) *updates = s.nru;
    if(rel_residual) *rel_residual = s.res0 > 0.0 ? s.residual * double(s.ncell) / s.res0 : 0.0;
    return 0;
}

int reef_pjm_corr(reef_ctx* c, int u_id, int v_id, int w_id, double alpha, int* iters)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!field_ok(u_id) || !field_ok(v_id) || !field_ok(w_id)) return fail(REEF_E_ARG, "bad field id");
    Engine& e = *c->b.e;
    int rc;
    if((rc = e.prepare())) return eng(e, rc);
    e.step_iters[c->tank] = 0;
    if((rc = e.pjm(u_id, v_id, w_id, alpha))) return eng(e, rc);
    if(iters) *iters = (int)e.step_iters[c->tank];
    return 0;
}

int reef_fc3_step(reef_ctx* c, int* iters, double* vol2)
{
    if(check_ctx(c)) return REEF_E_ARG;
    Engine& e = *c->b.e;
    int rc;
    if((rc = e.fc3_step())) return eng(e, rc);
    if(iters) *iters = (int)e.step_iters[c->tank];
    if(vol2)
    {
        if((rc = e.read_volumes())) return eng(e, rc);
        vol2[0] = e.vol[2 * c->tank];
        vol2[1] = e.vol[2 * c->tank + 1];
    }
    return 0;
}

int reef_set_timestep(reef_ctx* c, double dt)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!(dt > 0.0)) return fail(REEF_E_ARG, "dt must be > 0");
    Engine& e = *c->b.e;
    const int rc = e.set_dt(dt);
    return rc ? eng(e, rc) : 0;
}

// FP64 checkpoint: "REEF3DCK", version, nx, ny, nz, j_dir, the device dt, the number of
// fields, then each public field (id, (nx+6)(ny+6)(nz+6) doubles, i fastest with ghosts)
int reef_checkpoint_write(reef_ctx* c, const char* path)
{
    if(check_ctx(c) || !path) return fail(REEF_E_ARG, "bad argument");
    Engine& e = *c->b.e;
    TankHost* t = e.tank(c->tank);
    if(!t) return fail(REEF_E_ARG, "no tank");
    const HostGrid& h = t->hg;
    const size_t n = (size_t)(h.nx + 6) * (h.ny + 6) * (h.nz + 6);
    std::vector<double> buf(n);
    double dt = 0.0;
    int rc;
    if((rc = e.get(c->tank, REEF_U, buf.data(), REEF_LAYOUT_IFAST))) return eng(e, rc);   // synchronises
    HIPCHK(hipMemcpy(&dt, t->dev.tstep, sizeof(double), hipMemcpyDeviceToHost));
    FILE* f = std::fopen(path, "wb");
    if(!f) return fail(REEF_E_ARG, std::string("cannot write ") + path);
    const int head[6] = {1, h.nx, h.ny, h.nz, h.jdir, REEF_NFIELDS};
    bool ok = std::fwrite("REEF3DCK", 1, 8, f) == 8 && std::fwrite(head, sizeof(int), 5, f) == 5 &&
              std::fwrite(&dt, sizeof(double), 1, f) == 1 && std::fwrite(head + 5, sizeof(int), 1, f) == 1;
    for(int id = 0; ok && id < REEF_NFIELDS; ++id)
    {
        if((rc = e.get(c->tank, id, buf.data(), REEF_LAYOUT_IFAST)))
        {
            std::fclose(f);
            return eng(e, rc);
        }
        ok = std::fwrite(&id, sizeof(int), 1, f) == 1 && std::fwrite(buf.data(), sizeof(double), n, f) == n;
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : fail(REEF_E_ARG, std::string("write failed: ") + path);
}

int reef_checkpoint_read(reef_ctx* c, const char* path)
{
    if(check_ctx(c) || !path) return fail(REEF_E_ARG, "bad argument");
    Engine& e = *c->b.e;
    TankHost* t = e.tank(c->tank);
    if(!t) return fail(REEF_E_ARG, "no tank");
    const HostGrid& h = t->hg;
    const size_t n = (size_t)(h.nx + 6) * (h.ny + 6) * (h.nz + 6);
    FILE* f = std::fopen(path, "rb");
    if(!f) return fail(REEF_E_ARG, std::string("cannot open ") + path);
    char magic[8];
    int head[6];
    double dt = 0.0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "REEF3DCK", 8) == 0 &&
              std::fread(head, sizeof(int), 5, f) == 5 && std::fread(&dt, sizeof(double), 1, f) == 1 &&
              std::fread(head + 5, sizeof(int), 1, f) == 1;
    if(!ok || head[0] != 1)
    {
        std::fclose(f);
        return fail(REEF_E_ARG, "not a checkpoint of this library");
    }
    if(head[1] != h.nx || head[2] != h.ny || head[3] != h.nz || head[4] != h.jdir || head[5] != REEF_NFIELDS)
    {
        std::fclose(f);
        return fail(REEF_E_ARG, "checkpoint of another grid");
    }
    std::vector<std::vector<double>> fields(REEF_NFIELDS, std::vector<double>(n));
    for(int q = 0; ok && q < REEF_NFIELDS; ++q)
    {
        int id = -1;
        ok = std::fread(&id, sizeof(int), 1, f) == 1 && id == q && std::fread(fields[q].data(), sizeof(double), n, f) == n;
    }
    std::fclose(f);
    if(!ok) return fail(REEF_E_ARG, "checkpoint is truncated");
    int rc;
    for(int q = 0; q < REEF_NFIELDS; ++q)
        if((rc = e.put(c->tank, q, fields[q].data(), REEF_LAYOUT_IFAST))) return eng(e, rc);
    HIPCHK(hipMemcpy(t->dev.tstep, &dt, sizeof(double), hipMemcpyHostToDevice));
    t->prm.dt = dt;
    return 0;
}

// cfd_state::write_result of the context's u, v, w, press, phi (the host writer does the
// format; the fields come back from the device first)
int reef_cfd_state_write(reef_ctx* c, const char* path, const double* meta8)
{
    if(check_ctx(c) || !path || !meta8) return fail(REEF_E_ARG, "bad argument");
    Engine& e = *c->b.e;
    TankHost* t = e.tank(c->tank);
    if(!t) return fail(REEF_E_ARG, "no tank");
    const HostGrid& h = t->hg;
    const size_t n = (size_t)(h.nx + 6) * (h.ny + 6) * (h.nz + 6);
    std::vector<std::vector<double>> f(5, std::vector<double>(n));
    const int ids[5] = {REEF_U, REEF_V, REEF_W, REEF_PRESS, REEF_PHI};
    const double* fp[5];
    int rc;
    for(int q = 0; q < 5; ++q)
    {
        if((rc = e.get(c->tank, ids[q], f[q].data(), REEF_LAYOUT_IFAST))) return eng(e, rc);
        fp[q] = f[q].data();
    }
    if((rc = reef_cfd_state_write_host(h.nx, h.ny, h.nz, path, fp, meta8)))
        return fail(rc, reef_case_last_error());
    return 0;
}

// cfd_state::read (cfd_state_read.cpp:46-192): the loop cells of u, v, w, press, phi from the
// file, then the reader's ghost fills: start1 u 10, start2 v 11, start3 w 12, start4 press
// 40, start4 phi 52 (F50 = 2, ini.cpp:416)
int reef_cfd_state_read(reef_ctx* c, const char* path, double* meta8)
{
    if(check_ctx(c) || !path || !meta8) return fail(REEF_E_ARG, "bad argument");
    Engine& e = *c->b.e;
    TankHost* t = e.tank(c->tank);
    if(!t) return fail(REEF_E_ARG, "no tank");
    const HostGrid& h = t->hg;
    const size_t n = (size_t)(h.nx + 6) * (h.ny + 6) * (h.nz + 6);
    std::vector<std::vector<double>> f(5, std::vector<double>(n));
    const int ids[5] = {REEF_U, REEF_V, REEF_W, REEF_PRESS, REEF_PHI};
    const int kind_gcv[5][2] = {{1, 10}, {2, 11}, {3, 12}, {4, 40}, {4, 52}};
    double* fp[5];
    int rc;
    for(int q = 0; q < 5; ++q)
    {
        if((rc = e.get(c->tank, ids[q], f[q].data(), REEF_LAYOUT_IFAST))) return eng(e, rc);
        fp[q] = f[q].data();
    }
    if((rc = reef_cfd_state_read_host(h.nx, h.ny, h.nz, path, fp, meta8))) return fail(rc, reef_case_last_error());
    for(int q = 0; q < 5; ++q)
        if((rc = e.put(c->tank, ids[q], fp[q], REEF_LAYOUT_IFAST))) return eng(e, rc);
    for(int q = 0; q < 5; ++q)
        if((rc = e.ghost(ids[q], kind_gcv[q][0], kind_gcv[q][1]))) return eng(e, rc);
    return 0;
}

int reef_timestep(reef_ctx* c, double N47, int ini, double* out4)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!(N47 > 0.0) || (ini != 0 && ini != 1)) return fail(REEF_E_ARG, "N47 must be > 0 and ini 0 or 1");
    Engine& e = *c->b.e;
    std::vector<double> out;
    int rc;
    if((rc = e.timestep(N47, ini, out4 ? &out : nullptr))) return eng(e, rc);
    if(out4)
        for(int q = 0; q < 4; ++q) out4[q] = out[4 * c->tank + q];
    return 0;
}

int reef_vortex_step(reef_ctx* c, double simtime, double* vol2)
{
    if(check_ctx(c)) return REEF_E_ARG;
    Engine& e = *c->b.e;
    int rc;
    if((rc = e.vortex_step(simtime))) return eng(e, rc);
    if(vol2)
    {
        if((rc = e.read_volumes())) return eng(e, rc);
        vol2[0] = e.vol[2 * c->tank];
        vol2[1] = e.vol[2 * c->tank + 1];
    }
    return 0;
}

int reef_phi_rk3(reef_ctx* c)
{
    if(check_ctx(c)) return REEF_E_ARG;
    return eng(*c->b.e, c->b.e->phi_rk3());
}

int reef_sync(reef_ctx* c)
{
    if(check_ctx(c)) return REEF_E_ARG;
    HIPCHK(hipStreamSynchronize(c->b.stream));
    HIPCHK(hipGetLastError());
    return 0;
}

int reef_timing_enable(reef_ctx* c, int on)
{
    if(check_ctx(c)) return REEF_E_ARG;
    if(!on) c->b.e->tcollect();
    c->b.e->timing = on != 0;
    return 0;
}

int reef_timing_get(reef_ctx* c, const char* name, double* ms, long* launches)
{
    if(check_ctx(c)) return REEF_E_ARG;
    return timing_get(*c->b.e, name, ms, launches);
}

int reef_timing_reset(reef_ctx* c)
{
    if(check_ctx(c)) return REEF_E_ARG;
    c->b.e->tcollect();
    c->b.e->totals.clear();
    return 0;
}

// ------------------------------------------------------------------ batches
int reef_batch_create(double hbm_budget_bytes, reef_batch** out)
{
    if(!out || hbm_budget_bytes < 0.0) return fail(REEF_E_ARG, "bad argument");
    *out = nullptr;
    std::unique_ptr<reef_batch> b(new reef_batch());
    int rc;
    if((rc = make_engine(*b))) return rc;
    b->e->pool.budget = (size_t)hbm_budget_bytes;
    *out = b.release();
    return 0;
}

void reef_batch_destroy(reef_batch* b)
{
    if(!b) return;
    drop_engine(*b);
    delete b;
}

long reef_batch_tank_bytes(const reef_grid_desc* grid)
{
    if(!grid || !grid->XN || !grid->YN || !grid->ZN) return -1;
    return (long)Engine::tank_bytes(*grid);
}

int reef_batch_add(reef_batch* b, const reef_grid_desc* grid, const reef_params* prm, int* tank)
{
    if(check_batch(b)) return REEF_E_ARG;
    if(!grid || !prm || !grid->XN || !grid->YN || !grid->ZN) return fail(REEF_E_ARG, "null argument");
    return eng(*b->e, b->e->add_tank(*grid, *prm, tank));
}

int reef_batch_wave_set(reef_batch* b, int tank, const reef_wave_desc* w, double* params5)
{
    if(check_batch(b)) return REEF_E_ARG;
    if(!w) return fail(REEF_E_ARG, "null argument");
    return eng(*b->e, b->e->wave_set(tank, *w, params5));
}

int reef_batch_remove(reef_batch* b, int tank)
{
    if(check_batch(b)) return REEF_E_ARG;
    return eng(*b->e, b->e->remove_tank(tank));
}

int reef_batch_set_active(reef_batch* b, const int* tanks, int n)
{
    if(check_batch(b) || (n > 0 && !tanks)) return fail(REEF_E_ARG, "bad argument");
    std::vector<int> ids(tanks, tanks + (n > 0 ? n : 0));
    return eng(*b->e, b->e->set_active(ids));
}

int reef_batch_info(reef_batch* b, int* slots, int* active, double* bytes_in_use)
{
    if(check_batch(b)) return REEF_E_ARG;
    if(slots) *slots = (int)b->e->tanks.size();
    if(active) *active = b->e->active_count();
    if(bytes_in_use) *bytes_in_use = (double)b->e->pool.in_use();
    return 0;
}

int reef_batch_put(reef_batch* b, int tank, int field, const double* host, int layout)
{
    if(check_batch(b) || !host) return fail(REEF_E_ARG, "null argument");
    if(!field_ok(field)) return fail(REEF_E_ARG, "bad field id");
    return eng(*b->e, b->e->put(tank, field, host, layout));
}

int reef_batch_get(reef_batch* b, int tank, int field, double* host, int layout)
{
    if(check_batch(b) || !host) return fail(REEF_E_ARG, "null argument");
    if(!field_ok(field)) return fail(REEF_E_ARG, "bad field id");
    return eng(*b->e, b->e->get(tank, field, host, layout));
}

int reef_batch_ghost_fill(reef_batch* b, int field, int kind, int gcv)
{
    if(check_batch(b)) return REEF_E_ARG;
    if(!field_ok(field) || kind < 1 || kind > 4) return fail(REEF_E_ARG, "bad field id or kind");
    return eng(*b->e, b->e->ghost(field, kind, gcv));
}

int reef_batch_fluid_update(reef_batch* b)
{
    if(check_batch(b)) return REEF_E_ARG;
    return eng(*b->e, b->e->fluid_update());
}

int reef_batch_control(reef_batch* b, const char* key, double value)
{
    if(check_batch(b)) return REEF_E_ARG;
    return control(*b->e, key, value);
}

int reef_batch_solver_config(reef_batch* b, int precision, double delta)
{
    if(check_batch(b)) return REEF_E_ARG;
    return solver_config(*b->e, precision, delta);
}

int reef_batch_autotune(reef_batch* b, int reps, int* nbuckets)
{
    if(check_batch(b)) return REEF_E_ARG;
    std::vector<Engine::TuneRecord> log;
    const int rc = b->e->autotune(reps, &log);
    if(nbuckets) *nbuckets = (int)log.size();
    return eng(*b->e, rc);
}

int reef_batch_plan_set(reef_batch* b, int jdir, int size_log2, int ty, int tz, int cpt)
{
    if(check_batch(b)) return REEF_E_ARG;
    Engine::Plan p;
    p.ty = ty;
    p.tz = tz;
    p.cpt = cpt;
    return eng(*b->e, b->e->set_plan(jdir, size_log2, p));
}

int reef_batch_plans(reef_batch* b, int* entries, int capacity, int* n)
{
    if(check_batch(b) || !n || (capacity > 0 && !entries)) return fail(REEF_E_ARG, "bad argument");
    int q = 0;
    for(const auto& e : b->e->plans)
    {
        if(q < capacity)
        {
            int* o = entries + 5 * q;
            o[0] = e.first.first;
            o[1] = e.first.second;
            o[2] = e.second.ty;
            o[3] = e.second.tz;
            o[4] = e.second.cpt;
        }
        ++q;
    }
    *n = q;
    return 0;
}

int reef_batch_fc3_step(reef_batch* b, long* iters, double* vol)
{
    if(check_batch(b)) return REEF_E_ARG;
    Engine& e = *b->e;
    int rc;
    if((rc = e.fc3_step())) return eng(e, rc);
    if(iters)
        for(size_t q = 0; q < e.tanks.size(); ++q) iters[q] = e.step_iters[q];
    if(vol)
    {
        if((rc = e.read_volumes())) return eng(e, rc);
        for(size_t q = 0; q < 2 * e.tanks.size(); ++q) vol[q] = e.vol[q];
    }
    return 0;
}

int reef_batch_timestep(reef_batch* b, double N47, int ini, double* out)
{
    if(check_batch(b)) return REEF_E_ARG;
    if(!(N47 > 0.0) || (ini != 0 && ini != 1)) return fail(REEF_E_ARG, "N47 must be > 0 and ini 0 or 1");
    Engine& e = *b->e;
    std::vector<double> v;
    int rc;
    if((rc = e.timestep(N47, ini, out ? &v : nullptr))) return eng(e, rc);
    if(out) std::copy(v.begin(), v.end(), out);
    return 0;
}

int reef_batch_run(reef_batch* b, int nsteps, long* iters)
{
    if(check_batch(b) || nsteps < 0) return fail(REEF_E_ARG, "bad argument");
    Engine& e = *b->e;
    int rc;
    if((rc = e.prepare())) return eng(e, rc);
    for(size_t q = 0; q < e.tanks.size(); ++q)
        if(e.tanks[q].used && e.tanks[q].active && (rc = e.start_steps((int)q, nsteps))) return eng(e, rc);
    if((rc = e.run_scheduled())) return eng(e, rc);
    if(iters)
        for(size_t q = 0; q < e.tanks.size(); ++q) iters[q] = e.tanks[q].used ? e.tanks[q].iters_last : 0;
    return 0;
}

int reef_batch_sync(reef_batch* b)
{
    if(check_batch(b)) return REEF_E_ARG;
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipGetLastError());
    return 0;
}

int reef_batch_timing_enable(reef_batch* b, int on)
{
    if(check_batch(b)) return REEF_E_ARG;
    if(!on) b->e->tcollect();
    b->e->timing = on != 0;
    return 0;
}

int reef_batch_timing_get(reef_batch* b, const char* name, double* ms, long* launches)
{
    if(check_batch(b)) return REEF_E_ARG;
    return timing_get(*b->e, name, ms, launches);
}

}  // extern "C"
