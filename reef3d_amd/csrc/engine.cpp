// Host side of the batched engine (engine.h): device memory pool, tank set-up, the device
// tank / work-item tables, and the orchestration of the reference's operations on every
// active tank (momentum_FC3, levelset_RK3, reini_RK3, pjm_corr + bicgstab_ijk).
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace reef {

namespace {

#define EHIP(expr)                                                                            \
    do                                                                                        \
    {                                                                                         \
        hipError_t e_ = (expr);                                                               \
        if(e_ != hipSuccess) return fail(REEF_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while(0)

constexpr size_t GRAN = 256;
inline size_t up(size_t b) { return (b + GRAN - 1) / GRAN * GRAN; }

int prog_index(int kind, int gcv)
{
    for(int q = 0; q < NPROG; ++q)
        if(PROG_KIND_GCV[q][0] == kind && PROG_KIND_GCV[q][1] == gcv) return q;
    return -1;
}

}  // namespace

// ------------------------------------------------------------------ device pool
DevicePool::~DevicePool()
{
    for(Chunk& c : chunks_) (void)hipFree(c.base);
}

int DevicePool::alloc(size_t bytes, void** out)
{
    bytes = up(bytes ? bytes : GRAN);
    if(budget && used_ + bytes > budget) return REEF_E_NOMEM;
    for(size_t ci = 0; ci < chunks_.size(); ++ci)
    {
        Chunk& c = chunks_[ci];
        for(auto it = c.free_.begin(); it != c.free_.end(); ++it)
        {
            if(it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            c.free_.erase(it);
            if(len > bytes) c.free_[off + bytes] = len - bytes;
            void* p = c.base + off;
            live_[p] = {ci, bytes};
            used_ += bytes;
            *out = p;
            return 0;
        }
    }
    const size_t size = std::max(chunk_bytes, bytes);
    void* base = nullptr;
    if(hipMalloc(&base, size) != hipSuccess) return REEF_E_NOMEM;
    Chunk c;
    c.base = (char*)base;
    c.size = size;
    if(size > bytes) c.free_[bytes] = size - bytes;
    chunks_.push_back(c);
    live_[base] = {chunks_.size() - 1, bytes};
    used_ += bytes;
    *out = base;
    return 0;
}

void DevicePool::free(void* p)
{
    auto it = live_.find(p);
    if(it == live_.end()) return;
    Chunk& c = chunks_[it->second.first];
    size_t off = (size_t)((char*)p - c.base), len = it->second.second;
    used_ -= len;
    live_.erase(it);
    // coalesce with the neighbouring free ranges
    auto nx = c.free_.lower_bound(off);
    if(nx != c.free_.end() && off + len == nx->first)
    {
        len += nx->second;
        nx = c.free_.erase(nx);
    }
    if(nx != c.free_.begin())
    {
        auto pv = std::prev(nx);
        if(pv->first + pv->second == off)
        {
            pv->second += len;
            return;
        }
    }
    c.free_[off] = len;
}

// ------------------------------------------------------------------ plan
// 64 x 4 x 1 tiles for 3-D tanks (plane walk of the solver), 64 x 1 x 4 for 2-D x-z tanks;
// chunks of up to 16 tiles so that a large tank has about 2048 work items.  The plan is a
// function of the tank alone, so its reductions are the same in any batch.
TileShape plan_shape(const HostGrid& h, int cpt_override)
{
    const int ty = h.jdir == 1 ? 4 : 1, tz = h.jdir == 1 ? 1 : 4;
    TileShape s = make_shape(h.nx, h.ny, h.nz, ty, tz, 1);
    int cpt = cpt_override > 0 ? cpt_override : (s.ntiles + 2047) / 2048;
    cpt = std::max(1, std::min(cpt, 64));
    return make_shape(h.nx, h.ny, h.nz, ty, tz, cpt);
}

// ------------------------------------------------------------------ tanks
Engine::~Engine()
{
    if(stream) (void)hipStreamSynchronize(stream);
    for(auto& r : ranges_)
    {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for(hipEvent_t e : evpool_) (void)hipEventDestroy(e);
    for(hipEvent_t e : st_ev_)
        if(e) (void)hipEventDestroy(e);
    if(h_status_) (void)hipHostFree(h_status_);
    if(d_tanks_) (void)hipFree(d_tanks_);
    if(d_items_) (void)hipFree(d_items_);
    if(d_partials_) (void)hipFree(d_partials_);
    if(d_norms_) (void)hipFree(d_norms_);
}

size_t Engine::tank_bytes(const reef_grid_desc& gd)
{
    HostGrid h;
    if(!build_grid(h, gd, 2.1).empty()) return 0;
    size_t b = (size_t)NARR64 * up((size_t)h.nalloc * 8) + (size_t)NARR32 * up((size_t)h.nalloc * 4);
    size_t meta = 0;
    for(int a = 0; a < 3; ++a)
        meta += (h.P[a].size() * 3 + h.qf[a].size() + h.cf[a].size() + h.isf[a].size()) * 8 + h.DN[a].size() * 4;
    b += up(meta) + (size_t)2 * 8 * (size_t)(h.nx + h.ny + h.nz) * 64;   // ghost programs (estimate)
    return b;
}

int Engine::add_tank(const reef_grid_desc& gd, const reef_params& p, int* id)
{
    TankHost t;
    t.used = true;
    t.prm = p;
    std::string e = build_grid(t.hg, gd, p.F45);
    if(!e.empty()) return fail(REEF_E_UNSUPPORTED, e);
    const HostGrid& h = t.hg;
    // ghost programs first (may refuse the geometry)
    std::vector<GhostOp> ops;
    std::vector<int> op0(NPROG), nop(NPROG);
    for(int q = 0; q < NPROG; ++q)
    {
        std::vector<GhostOp> one;
        e = build_ghost_program(h, PROG_KIND_GCV[q][0], PROG_KIND_GCV[q][1], one);
        if(!e.empty()) return fail(REEF_E_UNSUPPORTED, "ghost program: " + e);
        op0[q] = (int)ops.size();
        nop[q] = (int)one.size();
        ops.insert(ops.end(), one.begin(), one.end());
    }
    TankDev& D = t.dev;
    std::memset(&D, 0, sizeof(D));
    DevGrid& g = D.g;
    g.nx = h.nx;
    g.ny = h.ny;
    g.nz = h.nz;
    g.jdir = h.jdir;
    g.sx = h.sx;
    g.sxy = h.sxy;
    g.base = h.base;
    g.nalloc = h.nalloc;
    g.ydir = h.jdir == 1 ? 1.0 : 0.0;
    g.psi = h.psi;
    int rc;
    auto take = [&](size_t bytes, void** out) -> int {
        int r = pool.alloc(bytes, out);
        if(r) return fail(r, "device pool: out of memory (budget)");
        t.mem.push_back(*out);
        t.bytes += up(bytes);
        return 0;
    };
    auto release = [&]() {
        for(void* m : t.mem) pool.free(m);
    };
    for(int a = 0; a < NARR64; ++a)
    {
        void* m;
        if((rc = take((size_t)h.nalloc * 8, &m))) return release(), rc;
        D.f[a] = (double*)m;
        EHIP(hipMemsetAsync(m, 0, (size_t)h.nalloc * 8, stream));
    }
    for(int a = 0; a < NARR32; ++a)
    {
        void* m;
        if((rc = take((size_t)h.nalloc * 4, &m))) return release(), rc;
        D.s[a] = (float*)m;
        EHIP(hipMemsetAsync(m, 0, (size_t)h.nalloc * 4, stream));
    }
    // metadata: P, DN, DP per axis, WENO tables, FP32 1/DN, in one block
    std::vector<double> meta;
    size_t offP[3], offDN[3], offDP[3], offqf[3], offcf[3], offisf[3];
    for(int a = 0; a < 3; ++a)
    {
        offP[a] = meta.size();
        meta.insert(meta.end(), h.P[a].begin(), h.P[a].end());
        offDN[a] = meta.size();
        meta.insert(meta.end(), h.DN[a].begin(), h.DN[a].end());
        offDP[a] = meta.size();
        meta.insert(meta.end(), h.DP[a].begin(), h.DP[a].end());
        offqf[a] = meta.size();
        meta.insert(meta.end(), h.qf[a].begin(), h.qf[a].end());
        offcf[a] = meta.size();
        meta.insert(meta.end(), h.cf[a].begin(), h.cf[a].end());
        offisf[a] = meta.size();
        meta.insert(meta.end(), h.isf[a].begin(), h.isf[a].end());
    }
    size_t offinv[3];
    std::vector<float> inv;
    for(int a = 0; a < 3; ++a)
    {
        offinv[a] = inv.size();
        for(double d : h.DN[a]) inv.push_back(d != 0.0 ? (float)(1.0 / d) : 0.0f);
    }
    const size_t metabytes = up(meta.size() * 8);
    void* mb;
    if((rc = take(metabytes + inv.size() * 4 + ops.size() * sizeof(GhostOp) + 64, &mb))) return release(), rc;
    EHIP(hipMemcpy(mb, meta.data(), meta.size() * 8, hipMemcpyHostToDevice));
    float* invd = (float*)((char*)mb + metabytes);
    EHIP(hipMemcpy(invd, inv.data(), inv.size() * 4, hipMemcpyHostToDevice));
    GhostOp* opd = (GhostOp*)((char*)invd + up(inv.size() * 4));
    if(!ops.empty()) EHIP(hipMemcpy(opd, ops.data(), ops.size() * sizeof(GhostOp), hipMemcpyHostToDevice));
    const double* md = (const double*)mb;
    const double** d1[3][3] = {{&g.XP, &g.DXN, &g.DXP}, {&g.YP, &g.DYN, &g.DYP}, {&g.ZP, &g.DZN, &g.DZP}};
    for(int a = 0; a < 3; ++a)
    {
        *d1[a][0] = md + offP[a];
        *d1[a][1] = md + offDN[a];
        *d1[a][2] = md + offDP[a];
        g.qf[a] = md + offqf[a];
        g.cf[a] = md + offcf[a];
        g.isf[a] = md + offisf[a];
    }
    D.iDXN32 = invd + offinv[0];
    D.iDYN32 = invd + offinv[1];
    D.iDZN32 = invd + offinv[2];
    for(int q = 0; q < NPROG; ++q)
    {
        D.gops[q] = opd + op0[q];
        D.gnops[q] = nop[q];
    }
    D.shape = plan_shape(h, plan_cpt_override);
    D.active = 1;
    D.dt = p.dt;
    // slot
    int slot = -1;
    for(size_t q = 0; q < tanks.size(); ++q)
        if(!tanks[q].used)
        {
            slot = (int)q;
            break;
        }
    if(slot < 0)
    {
        slot = (int)tanks.size();
        tanks.emplace_back();
    }
    tanks[slot] = std::move(t);
    dirty_ = true;
    if(id) *id = slot;
    return 0;
}

int Engine::remove_tank(int id)
{
    TankHost* t = tank(id);
    if(!t) return fail(REEF_E_ARG, "no such tank");
    EHIP(hipStreamSynchronize(stream));
    for(void* m : t->mem) pool.free(m);
    *t = TankHost();
    dirty_ = true;
    return 0;
}

int Engine::active_count() const
{
    int n = 0;
    for(const TankHost& t : tanks) n += t.used && t.dev.active;
    return n;
}

int Engine::set_active(const std::vector<int>& ids)
{
    for(size_t q = 0; q < tanks.size(); ++q)
        if(tanks[q].used) tanks[q].dev.active = ids.empty() ? 1 : 0;
    for(int id : ids)
    {
        TankHost* t = tank(id);
        if(!t) return fail(REEF_E_ARG, "no such tank");
        t->dev.active = 1;
    }
    dirty_ = true;
    return 0;
}

// device tank table (every slot; unused or inactive slots have active = 0), work items
// of the active tanks in slot order, per-slot solver states, launch sizes
int Engine::sync_tables()
{
    if(!dirty_) return 0;
    EHIP(hipStreamSynchronize(stream));
    const int nslot = (int)tanks.size();
    if(nslot > cap_tanks_ || !d_tanks_)
    {
        const int cap = std::max(nslot, 8);
        if(d_tanks_) hipFree(d_tanks_);
        if(d_norms_) hipFree(d_norms_);
        if(h_status_) hipHostFree(h_status_);
        EHIP(hipMalloc(&d_tanks_, sizeof(TankDev) * cap));
        // solver states, volumes and norms of every slot in one array
        EHIP(hipMalloc(&d_norms_, (sizeof(BiState) + 4 * sizeof(double)) * cap));
        EHIP(hipMemset(d_norms_, 0, (sizeof(BiState) + 4 * sizeof(double)) * cap));
        EHIP(hipHostMalloc((void**)&h_status_, 2 * sizeof(BiState) * cap, hipHostMallocDefault));
        std::memset(h_status_, 0, 2 * sizeof(BiState) * cap);
        cap_tanks_ = cap;
        cap_status_ = cap;
    }
    BiState* states = (BiState*)((char*)d_norms_ + 4 * sizeof(double) * cap_tanks_);
    double* vols = d_norms_ + cap_tanks_;
    std::vector<TankDev> host(nslot);
    std::vector<int> items;
    bool pw = false, direct = false;
    for(int q = 0; q < nslot; ++q)
    {
        TankHost& t = tanks[q];
        TankDev D{};
        if(t.used)
        {
            D = t.dev;
            D.st = states + q;
            D.vol = vols + 2 * q;
            D.item0 = (int)items.size();
            if(D.active)
            {
                for(int it = 0; it < D.shape.nitems; ++it) items.push_back(q);
                if(D.g.jdir == 1 && D.shape.ty == 4 && D.shape.tz == 1)
                    pw = true;
                else
                    direct = true;
            }
            t.dev = D;
        }
        host[q] = D;
    }
    if(!host.empty()) EHIP(hipMemcpy(d_tanks_, host.data(), sizeof(TankDev) * nslot, hipMemcpyHostToDevice));
    nitems_ = (int)items.size();
    if(nitems_ > cap_items_ || !d_items_)
    {
        const int cap = std::max(nitems_, 64);
        if(d_items_) hipFree(d_items_);
        if(d_partials_) hipFree(d_partials_);
        EHIP(hipMalloc(&d_items_, sizeof(int) * cap));
        EHIP(hipMalloc(&d_partials_, sizeof(double) * 4 * cap));
        cap_items_ = cap;
    }
    if(nitems_) EHIP(hipMemcpy(d_items_, items.data(), sizeof(int) * nitems_, hipMemcpyHostToDevice));
    if(!st_ev_[0])
        for(hipEvent_t& e : st_ev_) EHIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    view_.tanks = d_tanks_;
    view_.item_tank = d_items_;
    view_.ntank = nslot;
    view_.nitems = nitems_;
    view_.partials = d_partials_;
    view_.has_pw = pw ? 1 : 0;
    view_.has_direct = direct ? 1 : 0;
    auto grid_of = [&](int cap) {
        int gsz = std::min(cap, (nitems_ + 7) / 8 * 8);
        return std::max(gsz, 8);
    };
    grid_solver_ = grid_of(2048);
    grid_simple_ = grid_of(2048);
    grid_stencil_ = grid_of(1024);
    last_state.assign(nslot, BiState{});
    step_iters.assign(nslot, 0);
    vol.assign(2 * (size_t)nslot, 0.0);
    dirty_ = false;
    return 0;
}

// ------------------------------------------------------------------ fields
int Engine::put(int id, int field, const double* host, int layout)
{
    TankHost* t = tank(id);
    if(!t || field < 0 || field >= NARR64) return fail(REEF_E_ARG, "bad tank or field id");
    const HostGrid& h = t->hg;
    const long X = h.nx + 6, Y = h.ny + 6, Z = h.nz + 6;
    const double* src = host;
    if(layout == REEF_LAYOUT_REF)
    {
        t->staging.resize((size_t)X * Y * Z);
        for(long i = 0; i < X; ++i)
        for(long j = 0; j < Y; ++j)
        for(long k = 0; k < Z; ++k)
            t->staging[(size_t)(i + X * (j + Y * k))] = host[(i * Y + j) * Z + k];
        src = t->staging.data();
    }
    else if(layout != REEF_LAYOUT_IFAST)
        return fail(REEF_E_ARG, "bad layout");
    double* dst0 = t->dev.f[field] + (h.base - 3 - 3 * h.sx - 3 * h.sxy);
    EHIP(hipStreamSynchronize(stream));
    EHIP(hipMemcpy2D(dst0, h.sx * sizeof(double), src, X * sizeof(double), X * sizeof(double), Y * Z,
                     hipMemcpyHostToDevice));
    return 0;
}

int Engine::get(int id, int field, double* host, int layout)
{
    TankHost* t = tank(id);
    if(!t || field < 0 || field >= NARR64) return fail(REEF_E_ARG, "bad tank or field id");
    const HostGrid& h = t->hg;
    const long X = h.nx + 6, Y = h.ny + 6, Z = h.nz + 6;
    const double* src0 = t->dev.f[field] + (h.base - 3 - 3 * h.sx - 3 * h.sxy);
    double* dst = host;
    if(layout == REEF_LAYOUT_REF)
    {
        t->staging.resize((size_t)X * Y * Z);
        dst = t->staging.data();
    }
    else if(layout != REEF_LAYOUT_IFAST)
        return fail(REEF_E_ARG, "bad layout");
    EHIP(hipStreamSynchronize(stream));
    EHIP(hipMemcpy2D(dst, X * sizeof(double), src0, h.sx * sizeof(double), X * sizeof(double), Y * Z,
                     hipMemcpyDeviceToHost));
    if(layout == REEF_LAYOUT_REF)
        for(long i = 0; i < X; ++i)
        for(long j = 0; j < Y; ++j)
        for(long k = 0; k < Z; ++k)
            host[(i * Y + j) * Z + k] = t->staging[(size_t)(i + X * (j + Y * k))];
    return 0;
}

int Engine::zero(int field)
{
    int rc;
    if((rc = sync_tables())) return rc;
    bl_fill(view_, grid_simple_, field, 0.0, stream);
    return 0;
}

// ------------------------------------------------------------------ timing
hipEvent_t Engine::take()
{
    if(evpool_.empty())
    {
        hipEvent_t e;
        hipEventCreate(&e);
        return e;
    }
    hipEvent_t e = evpool_.back();
    evpool_.pop_back();
    return e;
}

void Engine::tbegin(const char* name)
{
    if(!timing) return;
    TimedRange r;
    r.name = name;
    r.a = take();
    r.b = take();
    hipEventRecord(r.a, stream);
    ranges_.push_back(r);
}

void Engine::tend()
{
    if(!timing || ranges_.empty()) return;
    hipEventRecord(ranges_.back().b, stream);
}

void Engine::tcollect()
{
    hipStreamSynchronize(stream);
    for(auto& r : ranges_)
    {
        float ms = 0.f;
        hipEventElapsedTime(&ms, r.a, r.b);
        auto& t = totals[r.name];
        t.first += ms;
        t.second += 1;
        evpool_.push_back(r.a);
        evpool_.push_back(r.b);
    }
    ranges_.clear();
}

// ------------------------------------------------------------------ operations
int Engine::ghost(int field, int kind, int gcv)
{
    const int q = prog_index(kind, gcv);
    if(q < 0) return fail(REEF_E_UNSUPPORTED, "ghost program (kind, gcv) not compiled");
    int rc;
    if((rc = sync_tables())) return rc;
    tbegin("ghost");
    bl_ghost(view_, 256, q, field, stream);
    tend();
    return 0;
}

int Engine::convection(int ipol, int q, int u, int v, int w)
{
    int rc;
    if((rc = sync_tables())) return rc;
    if(ipol == 4)
    {
        tbegin("hj");
        bl_hj_stage(view_, grid_stencil_, 0, q, -1, u, v, w, REEF_L, 0.0, stream);
        tend();
        return 0;
    }
    tbegin("flux");
    bl_flux_acc(view_, grid_stencil_, ipol, q, u, v, w, ipol == 1 ? REEF_F : ipol == 2 ? REEF_G : REEF_H, stream);
    tend();
    return 0;
}

int Engine::reini(int f, int iters)
{
    int rc;
    if((rc = sync_tables())) return rc;
    // F43 is a control value shared by the batch (the first used tank's)
    double f43 = 0.55;
    for(const TankHost& t : tanks)
        if(t.used)
        {
            f43 = t.prm.F43;
            break;
        }
    for(int q = 0; q < iters; ++q)
    {
        tbegin("reini");
        bl_reini_stage(view_, grid_stencil_, 1, f, f, A_RK1, f43, fsf, stream);
        tend();
        if((rc = ghost(A_RK1, 4, 52))) return rc;
        tbegin("reini");
        bl_reini_stage(view_, grid_stencil_, 2, A_RK1, f, A_RK2, f43, fsf, stream);
        tend();
        if((rc = ghost(A_RK2, 4, 52))) return rc;
        tbegin("reini");
        bl_reini_stage(view_, grid_stencil_, 3, A_RK2, f, f, f43, fsf, stream);
        tend();
        if((rc = ghost(f, 4, 52))) return rc;
    }
    return 0;
}

int Engine::fluid_update()
{
    int rc;
    if((rc = sync_tables())) return rc;
    const reef_params* p = nullptr;
    for(const TankHost& t : tanks)
        if(t.used)
        {
            p = &t.prm;
            break;
        }
    if(!p) return 0;
    const double prm4[4] = {p->W1, p->W2, p->W3, p->W4};
    tbegin("heaviside");
    bl_heaviside(view_, grid_simple_, prm4, stream);
    tend();
    if((rc = ghost(REEF_RO, 4, 1))) return rc;
    return ghost(REEF_VISC, 4, 1);
}

int Engine::read_volumes()
{
    int rc;
    if((rc = sync_tables())) return rc;
    const int nslot = (int)tanks.size();
    EHIP(hipMemcpyAsync(vol.data(), d_norms_ + cap_tanks_, sizeof(double) * 2 * nslot, hipMemcpyDeviceToHost, stream));
    EHIP(hipStreamSynchronize(stream));
    return 0;
}

// BiCGStab iterations after the init kernels, enqueued in batches; the device states say
// when every active tank has finished.  Two batches are in flight: batch k+1 is enqueued
// before the host waits for the states copied out after batch k.  Iterations launched
// after a tank's end are no-ops for it.
template<class Launch>
int Engine::run_iterations(int nphase, const char* const* names, Launch launch, int maxiter)
{
    const int batch = 8;
    const int nslot = (int)tanks.size();
    BiState* states = (BiState*)((char*)d_norms_ + 4 * sizeof(double) * cap_tanks_);
    auto enqueue = [&](long k) -> int {
        for(int q = 0; q < batch; ++q)
            for(int ph = 0; ph < nphase; ++ph)
            {
                tbegin(names[ph]);
                launch(ph, 0);
                tend();
                launch(ph, 1);
            }
        EHIP(hipMemcpyAsync(h_status_ + (k & 1) * cap_status_, states, sizeof(BiState) * nslot,
                            hipMemcpyDeviceToHost, stream));
        EHIP(hipEventRecord(st_ev_[k & 1], stream));
        return 0;
    };
    int rc;
    if((rc = enqueue(0))) return rc;
    for(long k = 1;; ++k)
    {
        if((rc = enqueue(k))) return rc;
        EHIP(hipEventSynchronize(st_ev_[(k - 1) & 1]));
        const BiState* hs = h_status_ + ((k - 1) & 1) * cap_status_;
        bool all = true;
        for(int q = 0; q < nslot; ++q)
            if(tanks[q].used && tanks[q].dev.active && !hs[q].done) all = false;
        if(all)
        {
            for(int q = 0; q < nslot; ++q)
                if(tanks[q].used && tanks[q].dev.active) last_state[q] = hs[q];
            break;
        }
        if(timing && ranges_.size() > 2048) tcollect();
        if(k * batch > 4L * (maxiter + 16)) return fail(REEF_E_HIP, "BiCGStab did not terminate");
    }
    if(timing) tcollect();
    return 0;
}

static const char* const BICG_NAMES[5] = {"bicg_p", "bicg_v", "bicg_t", "bicg_x", ""};
static const char* const MIX_NAMES[5] = {"mix_p", "bicg32_v", "bicg32_t", "bicg32_x", "mix_ru"};

// poisson::start + solver::start on every active tank: FP64 bicgstab_ijk, or the mixed
// form (FP32 iteration with FP64 reliable updates) ending on the true FP64 residual
int Engine::solve(int x, int rhs, double stop, int maxiter, bool mixed)
{
    int rc;
    if((rc = sync_tables())) return rc;
    double rw = 1000.0, ra = 1.205;
    for(const TankHost& t : tanks)
        if(t.used)
        {
            rw = t.prm.W1;
            ra = t.prm.W3;
            break;
        }
    tbegin("face_rho");
    bl_face_rho(view_, grid_simple_, rw, ra, mixed ? 0 : 1, stream);
    tend();
    if(!mixed)
    {
        tbegin("bicg_init");
        bl_bicg_init(view_, grid_solver_, x, rhs, stop, maxiter, stream);
        tend();
        auto launch = [&](int ph, int part) { bl_bicg_phase(view_, grid_solver_, x, rhs, ph, part, stream); };
        return run_iterations(4, BICG_NAMES, launch, maxiter);
    }
    tbegin("coef32");
    bl_coef32(view_, grid_simple_, stream);
    tend();
    tbegin("mix_init");
    bl_mixed_init(view_, grid_solver_, x, rhs, stop, maxiter, delta, stream);
    tend();
    auto launch = [&](int ph, int part) { bl_mixed_phase(view_, grid_solver_, x, rhs, ph, part, stream); };
    if((rc = run_iterations(5, MIX_NAMES, launch, maxiter))) return rc;
    tbegin("add32");
    bl_add32(view_, grid_simple_, x, stream);
    tend();
    return 0;
}

int Engine::residual(int x, int rhs, std::vector<double>& norms)
{
    int rc;
    if((rc = sync_tables())) return rc;
    double rw = 1000.0, ra = 1.205;
    for(const TankHost& t : tanks)
        if(t.used)
        {
            rw = t.prm.W1;
            ra = t.prm.W3;
            break;
        }
    bl_face_rho(view_, grid_simple_, rw, ra, 0, stream);
    bl_residual(view_, grid_solver_, x, rhs, d_norms_, stream);
    norms.assign(tanks.size(), 0.0);
    EHIP(hipMemcpyAsync(norms.data(), d_norms_, sizeof(double) * tanks.size(), hipMemcpyDeviceToHost, stream));
    EHIP(hipStreamSynchronize(stream));
    return 0;
}

// pjm_corr::start (pjm_corr.cpp:90-121) on every active tank
int Engine::pjm(int u, int v, int w, double alpha)
{
    int rc;
    if((rc = ghost(u, 1, 7)) || (rc = ghost(v, 2, 8)) || (rc = ghost(w, 3, 9))) return rc;
    const TankHost* t0 = nullptr;
    for(const TankHost& t : tanks)
        if(t.used)
        {
            t0 = &t;
            break;
        }
    if(!t0) return 0;
    tbegin("poisson_rhs");
    bl_poisson_rhs(view_, grid_simple_, u, v, w, alpha, stream);
    tend();
    bl_fill(view_, grid_simple_, REEF_PCORR, 0.0, stream);
    if((rc = solve(REEF_PCORR, A_RHS, t0->prm.N44, t0->prm.N46, precision == REEF_SOLVER_MIXED))) return rc;
    for(size_t q = 0; q < tanks.size(); ++q)
        if(tanks[q].used && tanks[q].dev.active) step_iters[q] += last_state[q].iter;
    if((rc = ghost(REEF_PCORR, 4, 40))) return rc;
    tbegin("correct");
    bl_correct(view_, grid_simple_, t0->prm.W1, t0->prm.W3, alpha, u, v, w, stream);
    tend();
    return ghost(REEF_PRESS, 4, 40);
}

// velocity predictor of the three components (momentum_FC3.cpp:160-208 and the later stages)
int Engine::momentum_stage(int stage, const int qin[3], const int qold[3], const int vel[3], const int out[3])
{
    const TankHost* t0 = nullptr;
    for(const TankHost& t : tanks)
        if(t.used)
        {
            t0 = &t;
            break;
        }
    if(!t0) return 0;
    const double grav[3] = {t0->prm.W20, t0->prm.W21, t0->prm.W22};
    const double cfac = stage == 1 ? 1.0 : stage == 2 ? 0.25 : (2.0 / 3.0);
    int rc;
    for(int comp = 1; comp <= 3; ++comp)
    {
        tbegin("momentum");
        bl_momentum(view_, grid_stencil_, comp, stage, qin[comp - 1], qold[comp - 1], vel[0], vel[1], vel[2],
                    grav[comp - 1], cfac, out[comp - 1], t0->prm.W1, t0->prm.W3, stream);
        tend();
    }
    const int kinds[3] = {1, 2, 3}, gcvs[3] = {10, 11, 12};
    for(int comp = 0; comp < 3; ++comp)
        if((rc = ghost(out[comp], kinds[comp], gcvs[comp]))) return rc;
    return 0;
}

// momentum_FC3::start (momentum_FC3.cpp:122-430): three TVD-RK3 stages of phi transport,
// reinitialisation, velocity predictor and projection; rho / nu from the final phi
int Engine::fc3_step()
{
    int rc;
    if((rc = sync_tables())) return rc;
    std::fill(step_iters.begin(), step_iters.end(), 0L);
    // ---- stage 1 (momentum_FC3.cpp:132-230)
    tbegin("copy");
    bl_copy(view_, grid_simple_, REEF_PHI, A_LS, stream);   // ls ghosts stay zero, as in the reference
    tend();
    tbegin("hj");
    bl_hj_stage(view_, grid_stencil_, 1, A_LS, -1, REEF_U, REEF_V, REEF_W, A_FRK1, 1.0, stream);
    tend();
    if((rc = ghost(A_FRK1, 4, 52))) return rc;
    if((rc = reini(A_FRK1, 2))) return rc;
    tbegin("copy");
    bl_copy(view_, grid_simple_, A_FRK1, REEF_PHI, stream);
    tend();
    if((rc = ghost(REEF_PHI, 4, 52))) return rc;
    {
        const int qin[3] = {REEF_U, REEF_V, REEF_W}, out[3] = {A_URK1, A_VRK1, A_WRK1};
        if((rc = momentum_stage(1, qin, qin, qin, out))) return rc;
    }
    if((rc = pjm(A_URK1, A_VRK1, A_WRK1, 1.0))) return rc;
    if((rc = ghost(A_URK1, 1, 10)) || (rc = ghost(A_VRK1, 2, 11)) || (rc = ghost(A_WRK1, 3, 12))) return rc;
    // ---- stage 2 (momentum_FC3.cpp:232-332)
    tbegin("hj");
    bl_hj_stage(view_, grid_stencil_, 2, A_FRK1, A_LS, A_URK1, A_VRK1, A_WRK1, A_FRK2, 0.25, stream);
    tend();
    if((rc = ghost(A_FRK2, 4, 52))) return rc;
    if((rc = reini(A_FRK2, 2))) return rc;
    tbegin("copy");
    bl_copy(view_, grid_simple_, A_FRK2, REEF_PHI, stream);
    tend();
    if((rc = ghost(REEF_PHI, 4, 52))) return rc;
    if((rc = reini(REEF_PHI, 2))) return rc;
    {
        const int qin[3] = {A_URK1, A_VRK1, A_WRK1}, qold[3] = {REEF_U, REEF_V, REEF_W},
                  out[3] = {A_URK2, A_VRK2, A_WRK2};
        if((rc = momentum_stage(2, qin, qold, qin, out))) return rc;
    }
    if((rc = pjm(A_URK2, A_VRK2, A_WRK2, 0.25))) return rc;
    if((rc = ghost(A_URK2, 1, 10)) || (rc = ghost(A_VRK2, 2, 11)) || (rc = ghost(A_WRK2, 3, 12))) return rc;
    // ---- stage 3 (momentum_FC3.cpp:334-430)
    tbegin("hj");
    bl_hj_stage(view_, grid_stencil_, 3, A_FRK2, A_LS, A_URK2, A_VRK2, A_WRK2, REEF_PHI, 2.0 / 3.0, stream);
    tend();
    if((rc = ghost(REEF_PHI, 4, 52))) return rc;
    if((rc = reini(REEF_PHI, 3))) return rc;
    {
        const int qin[3] = {A_URK2, A_VRK2, A_WRK2}, qold[3] = {REEF_U, REEF_V, REEF_W},
                  out[3] = {REEF_U, REEF_V, REEF_W};
        if((rc = momentum_stage(3, qin, qold, qin, out))) return rc;
    }
    if((rc = pjm(REEF_U, REEF_V, REEF_W, 2.0 / 3.0))) return rc;
    if((rc = ghost(REEF_U, 1, 10)) || (rc = ghost(REEF_V, 2, 11)) || (rc = ghost(REEF_W, 3, 12))) return rc;
    // rho and nu depend on phi only: the last fluid_update of the step is the one that stays
    return fluid_update();
}

// levelset_RK3 transport stages (levelset_RK3.cpp:109-160)
int Engine::phi_rk3()
{
    int rc;
    if((rc = sync_tables())) return rc;
    tbegin("hj");
    bl_hj_stage(view_, grid_stencil_, 1, REEF_PHI, -1, REEF_U, REEF_V, REEF_W, A_FRK1, 1.0, stream);
    tend();
    if((rc = ghost(A_FRK1, 4, 52))) return rc;
    tbegin("hj");
    bl_hj_stage(view_, grid_stencil_, 2, A_FRK1, REEF_PHI, REEF_U, REEF_V, REEF_W, A_FRK2, 0.25, stream);
    tend();
    if((rc = ghost(A_FRK2, 4, 52))) return rc;
    tbegin("hj");
    bl_hj_stage(view_, grid_stencil_, 3, A_FRK2, REEF_PHI, REEF_U, REEF_V, REEF_W, REEF_PHI, 2.0 / 3.0, stream);
    tend();
    return ghost(REEF_PHI, 4, 52);
}

// one level-set benchmark step (levelset_RK3 + reini + momentum_RK3 void + benchmark_vortex3D)
int Engine::vortex_step(double simtime)
{
    int rc;
    if((rc = phi_rk3())) return rc;
    if((rc = reini(REEF_PHI, 3))) return rc;
    const TankHost* t0 = nullptr;
    for(const TankHost& t : tanks)
        if(t.used)
        {
            t0 = &t;
            break;
        }
    if(!t0) return 0;
    const double grav[3] = {t0->prm.W20, t0->prm.W21, t0->prm.W22};
    for(int comp = 1; comp <= 3; ++comp)
    {
        if(comp == 2 && t0->hg.jdir == 0) continue;
        const double Fv = 0.0 + (0.0 + grav[comp - 1] + 0.0 + 0.0) * 1.0;
        bl_void_rk3(view_, grid_simple_, comp, REEF_U + comp - 1, Fv, stream);
        if((rc = ghost(REEF_U + comp - 1, comp, 9 + comp))) return rc;
    }
    for(int comp = 1; comp <= 3; ++comp)
        bl_vortex_velocity(view_, grid_simple_, comp, REEF_U + comp - 1, simtime, t0->hg.DXM, stream);
    if((rc = ghost(REEF_U, 1, 10))) return rc;
    if((rc = ghost(REEF_V, 2, 11))) return rc;
    if((rc = ghost(REEF_V, 2, 12))) return rc;
    return fluid_update();
}

}  // namespace reef
