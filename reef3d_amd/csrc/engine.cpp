// Host side of the batched engine (engine.h): device memory pool, tank set-up, the device
// tank table and the tank sets of the launches, and the orchestration of the reference's
// operations (momentum_FC3, levelset_RK3, reini_RK3, pjm_corr + bicgstab_ijk) in lockstep
// or by the iteration-level scheduler.
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
// 64 x 4 x 1 tiles for 3-D tanks (plane walk of the solver), 64 x 1 x 4 for 2-D x-z tanks.
// A 3-D tank gets about one work item per block of the plane-walk grid (a large tank's
// items are columns of several planes); a 2-D tank about 2048 items.  The plan is a
// function of the tank alone, so its reductions are the same in any batch.
TileShape plan_shape(const HostGrid& h, int cpt_override)
{
    const int ty = h.jdir == 1 ? 4 : 1, tz = h.jdir == 1 ? 1 : 4;
    TileShape s = make_shape(h.nx, h.ny, h.nz, ty, tz, 1);
    const int target = h.jdir == 1 ? PW_BLOCKS : 2048;
    int cpt = cpt_override > 0 ? cpt_override : (s.ntiles + target - 1) / target;
    cpt = std::max(1, std::min(cpt, 64));
    return make_shape(h.nx, h.ny, h.nz, ty, tz, cpt);
}

Engine::Bucket Engine::bucket_of(const HostGrid& h)
{
    const double cells = (double)h.nx * h.ny * h.nz;
    return {h.jdir, (int)std::floor(std::log2(std::max(cells, 1.0)))};
}

TileShape Engine::shape_of(const HostGrid& h) const
{
    auto it = plans.find(bucket_of(h));
    if(it == plans.end() || it->second.cpt <= 0) return plan_shape(h, 0);
    const Plan& p = it->second;
    return make_shape(h.nx, h.ny, h.nz, p.ty, p.tz, p.cpt);
}

static bool plan_valid(int jdir, const Engine::Plan& p)
{
    const bool shape = (p.ty == 4 && p.tz == 1) || (p.ty == 2 && p.tz == 2) || (p.ty == 1 && p.tz == 4);
    return shape && (jdir == 1 || p.ty == 1) && p.cpt >= 1 && p.cpt <= 64;
}

int Engine::set_plan(int jdir, int lg, const Plan& p)
{
    if((jdir != 0 && jdir != 1) || lg < 0 || lg > 40) return fail(REEF_E_ARG, "bad size bucket");
    if(p.cpt == 0)
        plans.erase({jdir, lg});
    else if(!plan_valid(jdir, p))
        return fail(REEF_E_ARG, "bad plan: tile rows (4,1), (2,2) or (1,4) (2-D tanks: (1,4)), 1 <= cpt <= 64");
    else
        plans[{jdir, lg}] = p;
    for(TankHost& t : tanks)
        if(t.used && bucket_of(t.hg) == Bucket{jdir, lg}) t.dev.shape = shape_of(t.hg);
    dirty_ = true;
    return 0;
}

// ------------------------------------------------------------------ tanks
Engine::~Engine()
{
    if(stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    for(auto& r : ranges_)
    {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for(hipEvent_t e : evpool_) (void)hipEventDestroy(e);
    for(hipEvent_t e : st_ev_)
        if(e) (void)hipEventDestroy(e);
    if(h_status_) (void)hipHostFree(h_status_);
    if(d_status_) (void)hipFree(d_status_);
    if(d_tanks_) (void)hipFree(d_tanks_);
    if(d_partials_) (void)hipFree(d_partials_);
    if(d_norms_) (void)hipFree(d_norms_);
}

// device bytes add_built takes for a build: the pool blocks of the FP64 and FP32 arrays,
// the metadata block (tables, FP32 1/DN, ghost programs), the solver state and the
// time-step block, each rounded to the pool's granule
static size_t meta_block_bytes(const TankBuild& b)
{
    return up(b.meta.size() * 8) + up(b.inv.size() * 4) + b.ops.size() * sizeof(GhostOp);
}

size_t Engine::built_bytes(const TankBuild& b)
{
    const size_t n = (size_t)b.hg.nalloc;
    return (size_t)NARR64 * up(n * 8) + (size_t)NARR32 * up(n * 4) + up(meta_block_bytes(b)) + up(sizeof(BiState)) +
           up(8 * sizeof(double));
}

size_t Engine::tank_bytes(const reef_grid_desc& gd)
{
    reef_params p{};
    p.F45 = 2.1;
    TankBuild b;
    std::string e;
    if(build_tank(gd, p, b, e)) return 0;
    return built_bytes(b);
}

size_t Engine::tank_bytes_min(const reef_grid_desc& gd)
{
    HostGrid h;
    if(!build_grid(h, gd, 2.1).empty()) return 0;
    return (size_t)NARR64 * up((size_t)h.nalloc * 8) + (size_t)NARR32 * up((size_t)h.nalloc * 4);
}

int Engine::build_tank(const reef_grid_desc& gd, const reef_params& p, TankBuild& b, std::string& err)
{
    b.prm = p;
    std::string e = build_grid(b.hg, gd, p.F45);
    if(!e.empty()) return err = e, REEF_E_UNSUPPORTED;
    const HostGrid& h = b.hg;
    // ghost programs first (may refuse the geometry)
    std::vector<GhostOp> progs[NPROG];
    e = build_ghost_programs(h, NPROG, PROG_KIND_GCV, progs);
    if(!e.empty()) return err = "ghost program: " + e, REEF_E_UNSUPPORTED;
    b.ops.clear();
    for(int q = 0; q < NPROG; ++q)
    {
        b.op0[q] = (int)b.ops.size();
        b.nop[q] = (int)progs[q].size();
        b.ops.insert(b.ops.end(), progs[q].begin(), progs[q].end());
    }
    // metadata: P, DN, DP per axis, WENO tables, FP32 1/DN, in one block
    b.meta.clear();
    b.inv.clear();
    for(int a = 0; a < 3; ++a)
    {
        b.offP[a] = b.meta.size();
        b.meta.insert(b.meta.end(), h.P[a].begin(), h.P[a].end());
        b.offDN[a] = b.meta.size();
        b.meta.insert(b.meta.end(), h.DN[a].begin(), h.DN[a].end());
        b.offDP[a] = b.meta.size();
        b.meta.insert(b.meta.end(), h.DP[a].begin(), h.DP[a].end());
        b.offqf[a] = b.meta.size();
        b.meta.insert(b.meta.end(), h.qf[a].begin(), h.qf[a].end());
        b.offcf[a] = b.meta.size();
        b.meta.insert(b.meta.end(), h.cf[a].begin(), h.cf[a].end());
        b.offisf[a] = b.meta.size();
        b.meta.insert(b.meta.end(), h.isf[a].begin(), h.isf[a].end());
        if(a == 0)
        {
            // the x tables again, entry-major: entry e of every IP in a row
            auto transpose = [&](const std::vector<double>& t, int w, size_t& off) {
                const size_t n = t.size() / (size_t)w;
                off = b.meta.size();
                for(int e = 0; e < w; ++e)
                    for(size_t ip = 0; ip < n; ++ip) b.meta.push_back(t[ip * (size_t)w + (size_t)e]);
                b.nT = (int)n;
            };
            transpose(h.qf[0], 24, b.offqfT);
            transpose(h.cf[0], 12, b.offcfT);
            transpose(h.isf[0], 36, b.offisfT);
        }
    }
    for(int a = 0; a < 3; ++a)
    {
        b.offinv[a] = b.inv.size();
        for(double d : h.DN[a]) b.inv.push_back(d != 0.0 ? (float)(1.0 / d) : 0.0f);
    }
    return 0;
}

int Engine::add_tank(const reef_grid_desc& gd, const reef_params& p, int* id)
{
    TankBuild b;
    std::string e;
    const int rc = build_tank(gd, p, b, e);
    if(rc) return fail(rc, e);
    return add_built(b, id);
}

// the reef_params values a batch shares (everything except the per-tank dt and F45)
static bool same_shared(const reef_params& a, const reef_params& b)
{
    return a.W1 == b.W1 && a.W2 == b.W2 && a.W3 == b.W3 && a.W4 == b.W4 && a.W20 == b.W20 && a.W21 == b.W21 &&
           a.W22 == b.W22 && a.F43 == b.F43 && a.N44 == b.N44 && a.N46 == b.N46;
}

int Engine::add_built(TankBuild& b, int* id)
{
    if(!first_used()) has_shared_ = false;
    if(has_shared_ && !same_shared(shared_, b.prm))
        return fail(REEF_E_ARG, "tank parameters differ from the batch's shared ones (W1-W4, W20-W22, F43, N44, N46; "
                                "only dt and F45 may differ per tank)");
    TankHost t;
    t.used = true;
    t.prm = b.prm;
    t.hg = std::move(b.hg);
    const HostGrid& h = t.hg;
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
        b.hg = std::move(t.hg);   // the build stays usable for another attempt
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
    const std::vector<double>& meta = b.meta;
    const std::vector<float>& inv = b.inv;
    const std::vector<GhostOp>& ops = b.ops;
    const size_t metabytes = up(meta.size() * 8);
    void* mb;
    if((rc = take(meta_block_bytes(b), &mb))) return release(), rc;
    // the solver state stays with the tank (iterations in flight survive table rebuilds)
    void* stb;
    if((rc = take(sizeof(BiState), &stb))) return release(), rc;
    EHIP(hipMemsetAsync(stb, 0, sizeof(BiState), stream));
    D.st = (BiState*)stb;
    // the time-step block: dt, the velocity maxima of the last time-step pass, and the
    // smallest cell width MIN over the cells of MIN3(DXN, DYN, DZN) (ietimestep.cpp:180-184)
    void* tsb;
    if((rc = take(8 * sizeof(double), &tsb))) return release(), rc;
    D.tstep = (double*)tsb;
    {
        double mindx = 1.0e300;
        for(int a = 0; a < 3; ++a)
        {
            const int n = a == 0 ? h.nx : a == 1 ? h.ny : h.nz;
            for(int q = 0; q < n; ++q) mindx = std::min(mindx, h.DN[a][q + MARGE]);
        }
        const double init[8] = {t.prm.dt, 0.0, 0.0, 0.0, mindx, 0.0, 0.0, 0.0};
        EHIP(hipMemcpy(tsb, init, sizeof(init), hipMemcpyHostToDevice));
    }
    EHIP(hipMemcpy(mb, meta.data(), meta.size() * 8, hipMemcpyHostToDevice));
    float* invd = (float*)((char*)mb + metabytes);
    EHIP(hipMemcpy(invd, inv.data(), inv.size() * 4, hipMemcpyHostToDevice));
    GhostOp* opd = (GhostOp*)((char*)invd + up(inv.size() * 4));
    if(!ops.empty()) EHIP(hipMemcpy(opd, ops.data(), ops.size() * sizeof(GhostOp), hipMemcpyHostToDevice));
    const double* md = (const double*)mb;
    const double** d1[3][3] = {{&g.XP, &g.DXN, &g.DXP}, {&g.YP, &g.DYN, &g.DYP}, {&g.ZP, &g.DZN, &g.DZP}};
    for(int a = 0; a < 3; ++a)
    {
        *d1[a][0] = md + b.offP[a];
        *d1[a][1] = md + b.offDN[a];
        *d1[a][2] = md + b.offDP[a];
        g.qf[a] = md + b.offqf[a];
        g.cf[a] = md + b.offcf[a];
        g.isf[a] = md + b.offisf[a];
    }
    g.qfT = md + b.offqfT;
    g.cfT = md + b.offcfT;
    g.isfT = md + b.offisfT;
    g.nT = b.nT;
    D.iDXN32 = invd + b.offinv[0];
    D.iDYN32 = invd + b.offinv[1];
    D.iDZN32 = invd + b.offinv[2];
    for(int q = 0; q < NPROG; ++q)
    {
        D.gops[q] = opd + b.op0[q];
        D.gnops[q] = b.nop[q];
    }
    D.shape = shape_of(h);
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
    if(!has_shared_)
    {
        shared_ = tanks[slot].prm;
        has_shared_ = true;
    }
    dirty_ = true;
    if(id) *id = slot;
    if(d20 == 2)
    {
        const int rc2 = ensure_matrix(slot);
        if(rc2)
        {
            const std::string keep = err;
            remove_tank(slot);
            return fail(rc2, keep);
        }
    }
    return 0;
}

int Engine::remove_tank(int id)
{
    TankHost* t = tank(id);
    if(!t) return fail(REEF_E_ARG, "no such tank");
    EHIP(hipStreamSynchronize(stream));
    int rc;
    if((rc = drain())) return rc;
    for(void* m : t->mem) pool.free(m);
    *t = TankHost();
    if(!first_used()) has_shared_ = false;
    dirty_ = true;
    return 0;
}

int Engine::active_count() const
{
    int n = 0;
    for(const TankHost& t : tanks) n += t.used && t.active;
    return n;
}

int Engine::set_active(const std::vector<int>& ids)
{
    for(size_t q = 0; q < tanks.size(); ++q)
        if(tanks[q].used) tanks[q].active = ids.empty();
    for(int id : ids)
    {
        TankHost* t = tank(id);
        if(!t) return fail(REEF_E_ARG, "no such tank");
        t->active = true;
    }
    dirty_ = true;
    return 0;
}

// device tank table (every slot), reduction slots of the used tanks, per-slot norms,
// volumes and gathered solver states, and the tank sets of the active tanks
int Engine::sync_tables()
{
    if(!dirty_) return 0;
    EHIP(hipStreamSynchronize(stream));
    int rc;
    if((rc = drain())) return rc;
    drop_graphs();   // their launches carry the old tables, sets and parameters
    const int nslot = (int)tanks.size();
    if(nslot > cap_tanks_ || !d_tanks_)
    {
        const int cap = std::max(2 * nslot, 8);
        if(d_tanks_) (void)hipFree(d_tanks_);
        if(d_norms_) (void)hipFree(d_norms_);
        if(d_status_) (void)hipFree(d_status_);
        if(h_status_) (void)hipHostFree(h_status_);
        d_tanks_ = nullptr;
        d_norms_ = nullptr;
        d_status_ = nullptr;
        h_status_ = nullptr;
        EHIP(hipMalloc(&d_tanks_, sizeof(TankDev) * cap));
        // norms (1 per slot) and volumes (2 per slot)
        EHIP(hipMalloc(&d_norms_, 3 * sizeof(double) * cap));
        EHIP(hipMemset(d_norms_, 0, 3 * sizeof(double) * cap));
        EHIP(hipMalloc(&d_status_, 2 * sizeof(BiState) * cap));
        EHIP(hipMemset(d_status_, 0, 2 * sizeof(BiState) * cap));
        EHIP(hipHostMalloc((void**)&h_status_, 2 * sizeof(BiState) * cap, hipHostMallocDefault));
        std::memset(h_status_, 0, 2 * sizeof(BiState) * cap);
        cap_tanks_ = cap;
    }
    double* vols = d_norms_ + cap_tanks_;
    std::vector<TankDev> host(nslot);
    size_t items = 0;
    for(int q = 0; q < nslot; ++q)
    {
        TankHost& t = tanks[q];
        if(!t.used) continue;
        t.dev.item0 = (int)items;
        t.dev.vol = vols + 2 * q;
        items += (size_t)t.dev.shape.nitems;
        host[q] = t.dev;
    }
    if(!host.empty()) EHIP(hipMemcpy(d_tanks_, host.data(), sizeof(TankDev) * nslot, hipMemcpyHostToDevice));
    // two halves: 4 partials per item, then the reliable-update phase's 2 per item
    if(8 * std::max(items, (size_t)64) > cap_partials_)
    {
        if(d_partials_) (void)hipFree(d_partials_);
        d_partials_ = nullptr;
        cap_partials_ = 16 * std::max(items, (size_t)64);
        EHIP(hipMalloc(&d_partials_, sizeof(double) * cap_partials_));
    }
    if(!st_ev_[0])
        for(hipEvent_t& e : st_ev_) EHIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    std::vector<int> act;
    for(int q = 0; q < nslot; ++q)
        if(tanks[q].used && tanks[q].active) act.push_back(q);
    all_ = make_sets(act);
    cur_ = &all_;
    solving_key_.clear();
    solving_sets_.clear();
    last_state.resize(nslot, BiState{});
    step_iters.resize(nslot, 0);
    vol.resize(2 * (size_t)nslot, 0.0);
    dirty_ = false;
    return 0;
}

// the launch sets of the given slots, at most SET_MAX tanks each
std::vector<TankSet> Engine::make_sets(const std::vector<int>& slots_in) const
{
    // plane-walked tanks first: their items are a prefix of each set's items, so the
    // launches that take one kind of tank spread that kind's items over every XCD group
    auto pw = [&](int q) {
        const TankHost& t = tanks[q];
        return t.hg.jdir == 1 && t.dev.shape.ty == 4 && t.dev.shape.tz == 1;
    };
    std::vector<int> slots(slots_in);
    std::stable_partition(slots.begin(), slots.end(), pw);
    std::vector<TankSet> out;
    for(size_t a = 0; a < slots.size(); a += SET_MAX)
    {
        TankSet S;
        std::memset(&S, 0, sizeof(S));
        S.tanks = d_tanks_;
        S.partials = d_partials_;
        S.partials_ru = d_partials_ + cap_partials_ / 2;
        const size_t b = std::min(slots.size(), a + (size_t)SET_MAX);
        int nt = 0, ni = 0;
        for(size_t q = a; q < b; ++q)
        {
            const TankHost& t = tanks[slots[q]];
            const int p = (int)(q - a);
            S.slot[p] = slots[q];
            S.tile0[p] = nt;
            S.item0[p] = ni;
            nt += t.dev.shape.ntiles;
            ni += t.dev.shape.nitems;
            if(pw(slots[q]))
            {
                S.has_pw = 1;
                S.pw_items = ni;
            }
            else
                S.has_direct = 1;
        }
        S.n = (int)(b - a);
        S.tile0[S.n] = nt;
        S.item0[S.n] = ni;
        S.ntiles = nt;
        S.nitems = ni;
        out.push_back(S);
    }
    return out;
}

// ------------------------------------------------------------------ fields
// the ghosted box of a field (X x Y x Z with X = nx + 6, ...) is one span of the device
// layout: rows sx apart, planes sx * Y apart, so a field moves in one contiguous copy and
// the host gathers or scatters the rows (a pitched copy of Y * Z short rows is slow)
static size_t box_span(const HostGrid& h)
{
    return (size_t)((long)(h.ny + 6) * (h.nz + 6) - 1) * h.sx + (size_t)(h.nx + 6);
}

int Engine::put(int id, int field, const double* host, int layout)
{
    TankHost* t = tank(id);
    if(!t || field < 0 || field >= NARR64) return fail(REEF_E_ARG, "bad tank or field id");
    return put_array(id, t->dev.f[field], host, layout);
}

int Engine::put_array(int id, double* dev, const double* host, int layout)
{
    TankHost* t = tank(id);
    if(!t || !dev) return fail(REEF_E_ARG, "bad tank or array");
    if(layout != REEF_LAYOUT_IFAST && layout != REEF_LAYOUT_REF) return fail(REEF_E_ARG, "bad layout");
    const HostGrid& h = t->hg;
    const long X = h.nx + 6, Y = h.ny + 6, Z = h.nz + 6, sx = h.sx;
    const size_t span = box_span(h);
    t->staging.assign(span, 0.0);   // the row padding beyond i = nx + 2 stays zero
    double* st = t->staging.data();
    if(layout == REEF_LAYOUT_REF)
    {
        for(long i = 0; i < X; ++i)
        for(long j = 0; j < Y; ++j)
        for(long k = 0; k < Z; ++k)
            st[(size_t)(i + sx * (j + Y * k))] = host[(i * Y + j) * Z + k];
    }
    else
        for(long r = 0; r < Y * Z; ++r) std::memcpy(st + (size_t)r * sx, host + (size_t)r * X, X * sizeof(double));
    double* dst0 = dev + (h.base - 3 - 3 * h.sx - 3 * h.sxy);
    EHIP(hipStreamSynchronize(stream));
    EHIP(hipMemcpy(dst0, st, span * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int Engine::get(int id, int field, double* host, int layout)
{
    TankHost* t = tank(id);
    if(!t || field < 0 || field >= NARR64) return fail(REEF_E_ARG, "bad tank or field id");
    if(layout != REEF_LAYOUT_IFAST && layout != REEF_LAYOUT_REF) return fail(REEF_E_ARG, "bad layout");
    const HostGrid& h = t->hg;
    const long X = h.nx + 6, Y = h.ny + 6, Z = h.nz + 6, sx = h.sx;
    const size_t span = box_span(h);
    const double* src0 = t->dev.f[field] + (h.base - 3 - 3 * h.sx - 3 * h.sxy);
    t->staging.resize(span);
    double* st = t->staging.data();
    EHIP(hipStreamSynchronize(stream));
    EHIP(hipMemcpy(st, src0, span * sizeof(double), hipMemcpyDeviceToHost));
    if(layout == REEF_LAYOUT_REF)
    {
        for(long i = 0; i < X; ++i)
        for(long j = 0; j < Y; ++j)
        for(long k = 0; k < Z; ++k)
            host[(i * Y + j) * Z + k] = st[(size_t)(i + sx * (j + Y * k))];
    }
    else
        for(long r = 0; r < Y * Z; ++r) std::memcpy(host + (size_t)r * X, st + (size_t)r * sx, X * sizeof(double));
    return 0;
}

int Engine::zero(int field)
{
    int rc;
    if((rc = sync_tables())) return rc;
    each([&](const TankSet& S) { bl_fill(S, field, 0.0, stream); });
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
const TankHost* Engine::first_used() const
{
    for(const TankHost& t : tanks)
        if(t.used) return &t;
    return nullptr;
}

int Engine::ghost(int field, int kind, int gcv)
{
    const int q = prog_index(kind, gcv);
    if(q < 0) return fail(REEF_E_UNSUPPORTED, "ghost program (kind, gcv) not compiled");
    int rc;
    if((rc = sync_tables())) return rc;
    tbegin("ghost");
    each([&](const TankSet& S) { bl_ghost(S, q, field, stream); });
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
        each([&](const TankSet& S) { bl_hj_stage(S, 0, q, -1, u, v, w, REEF_L, 0.0, stream); });
        tend();
        return 0;
    }
    const int acc = ipol == 1 ? REEF_F : ipol == 2 ? REEF_G : REEF_H;
    tbegin("flux");
    each([&](const TankSet& S) { bl_flux_acc(S, ipol, q, u, v, w, acc, stream); });
    tend();
    return 0;
}

int Engine::reini(int f, int iters, int gcv)
{
    int rc;
    if((rc = sync_tables())) return rc;
    // F43 is a control value shared by the batch
    const reef_params* sp = shared();
    const double f43 = sp ? sp->F43 : 0.55;
    for(int q = 0; q < iters; ++q)
    {
        tbegin("reini");
        each([&](const TankSet& S) { bl_reini_stage(S, 1, f, f, A_RK1, f43, fsf, stream); });
        tend();
        if((rc = ghost(A_RK1, 4, gcv))) return rc;
        tbegin("reini");
        each([&](const TankSet& S) { bl_reini_stage(S, 2, A_RK1, f, A_RK2, f43, fsf, stream); });
        tend();
        if((rc = ghost(A_RK2, 4, gcv))) return rc;
        tbegin("reini");
        each([&](const TankSet& S) { bl_reini_stage(S, 3, A_RK2, f, f, f43, fsf, stream); });
        tend();
        if((rc = ghost(f, 4, gcv))) return rc;
    }
    return 0;
}

int Engine::fluid_update()
{
    int rc;
    if((rc = sync_tables())) return rc;
    const reef_params* p = shared();
    if(!p) return 0;
    const double prm4[4] = {p->W1, p->W2, p->W3, p->W4};
    tbegin("heaviside");
    each([&](const TankSet& S) { bl_heaviside(S, prm4, stream); });
    tend();
    if((rc = ghost(REEF_RO, 4, 1))) return rc;
    return ghost(REEF_VISC, 4, 1);
}

// fluid_update of one active tank with its reduction memory in the allocation-free scratch
// (bl_heaviside_noalloc): no device allocation, no table rebuild, no timing events.  The
// tables must be current (the context syncs them at creation); vol2 receives volume1,
// volume2 of the tank.
int Engine::fluid_update_noalloc(int slot, double* vol2)
{
    if(dirty_)
        return fail(REEF_E_STATE, "allocation-free fluid update: the device tables are out of date (tanks, plans or "
                                  "operators changed since the last call); call any other entry point once first");
    const reef_params* p = shared();
    const TankHost* t = tank(slot);
    if(!p || !t || !t->active) return fail(REEF_E_ARG, "allocation-free fluid update: no such active tank");
    int set_of = -1, idx = -1;
    for(size_t q = 0; q < cur_->size(); ++q)
    {
        const TankSet& S = (*cur_)[q];
        if(S.n > NOALLOC_MAX_TANKS)
            return fail(REEF_E_ARG, "allocation-free fluid update: more than " + std::to_string(NOALLOC_MAX_TANKS) +
                                        " tanks in one launch set");
        for(int m = 0; m < S.n; ++m)
        {
            const TankDev& d = tanks[S.slot[m]].dev;
            if((long)d.item0 + d.shape.nitems > NOALLOC_MAX_ITEMS)
                return fail(REEF_E_ARG, "allocation-free fluid update: the tanks need " +
                                            std::to_string((long)d.item0 + d.shape.nitems) +
                                            " reduction slots, the scratch holds " + std::to_string(NOALLOC_MAX_ITEMS));
            if(S.slot[m] == slot) set_of = (int)q, idx = m;
        }
    }
    if(set_of < 0) return fail(REEF_E_ARG, "allocation-free fluid update: tank not in the active sets");
    double* sp = nullptr;
    if(noalloc_scratch(&sp)) return fail(REEF_E_HIP, "hipGetSymbolAddress of the scratch array failed");
    const double prm4[4] = {p->W1, p->W2, p->W3, p->W4};
    for(size_t q = 0; q < cur_->size(); ++q)
    {
        bl_heaviside_noalloc((*cur_)[q], prm4, sp, stream);
        // the next set's launch reuses the volume slots: copy this set's first
        if((int)q == set_of)
            EHIP(hipMemcpyAsync(vol2, sp + NOALLOC_VOL0 + 2 * idx, 2 * sizeof(double), hipMemcpyDeviceToHost, stream));
    }
    const int q_ro = prog_index(4, 1);
    if(q_ro < 0) return fail(REEF_E_UNSUPPORTED, "ghost program (4, 1) not compiled");
    each([&](const TankSet& S) {
        bl_ghost(S, q_ro, REEF_RO, stream);
        bl_ghost(S, q_ro, REEF_VISC, stream);
    });
    EHIP(hipStreamSynchronize(stream));
    return 0;
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

int Engine::prepare_tanks(const std::vector<int>& ids)
{
    int rc;
    if((rc = sync_tables())) return rc;
    for(int id : ids)
        if(!tank(id)) return fail(REEF_E_ARG, "no such tank");
    const std::vector<TankSet> sets = make_sets(ids);
    cur_ = &sets;
    if(!(rc = ghost(REEF_PHI, 4, 52)) && !(rc = ghost(REEF_U, 1, 10)) && !(rc = ghost(REEF_V, 2, 11)) &&
       !(rc = ghost(REEF_W, 3, 12)) && !(rc = ghost(REEF_PRESS, 4, 40)))
        rc = fluid_update();
    cur_ = &all_;
    return rc;
}

// ------------------------------------------------------------------ graphs
void Engine::drop_graphs()
{
    if(graphs_.empty() && set_graphs_.empty()) return;
    if(stream) (void)hipStreamSynchronize(stream);
    for(auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
    graphs_.clear();
    for(auto& g : set_graphs_)
        if(g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
    set_graphs_.clear();
}

// captures body() (a launch sequence on the stream) into an executable graph
template<class Body>
int Engine::capture(Body body, hipGraphExec_t* out)
{
    EHIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    const int rc = body();
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(stream, &g);
    if(rc || ec != hipSuccess)
    {
        if(g) (void)hipGraphDestroy(g);
        return rc ? rc : fail(REEF_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if(ei != hipSuccess) return fail(REEF_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    *out = ex;
    return 0;
}

// runs body() (a launch sequence on the stream), or replays it as a graph captured from
// its first run under this key
template<class Body>
int Engine::replay(const GraphKey& key, Body body)
{
    if(!graphs || timing) return body();
    auto it = graphs_.find(key);
    if(it == graphs_.end())
    {
        hipGraphExec_t ex = nullptr;
        const int rc = capture(body, &ex);
        if(rc) return rc;
        it = graphs_.emplace(key, ex).first;
    }
    EHIP(hipGraphLaunch(it->second, stream));
    return 0;
}

// ------------------------------------------------------------------ BiCGStab iterations
static const char* const BICG_NAMES[5] = {"bicg_p", "bicg_v", "bicg_t", "bicg_x", ""};
static const char* const MIX_NAMES[4] = {"mix_v", "bicg32_t", "bicg32_x", "mix_ru"};
static const char* const MAT_NAMES[4] = {"mat_p", "mat_v", "mat_t", "mat_x"};

// 8 iterations of the tanks of the current sets, then their states gathered into ring
// slot `ring` and copied to the host.  Iterations after a tank's end are no-ops for it.
// kind: SOLVE_FP64 (the Poisson operator), SOLVE_MIXED, or SOLVE_MAT (the explicit
// 7-diagonal operator; rhs is then the FLEXLOOP var)
int Engine::enqueue_batch(int kind, int x, int rhs, int ring, int iters)
{
    const int nph = 4;
    const char* const* names = kind == SOLVE_MIXED ? MIX_NAMES : kind == SOLVE_MAT ? MAT_NAMES : BICG_NAMES;
    auto body = [&]() -> int {
        for(int q = 0; q < iters; ++q)
            for(int ph = 0; ph < nph; ++ph)
                for(int part = 0; part < 2; ++part)
                {
                    if(part == 0) tbegin(names[ph]);
                    each([&](const TankSet& S) {
                        if(kind == SOLVE_MIXED)
                            bl_mixed_phase(S, x, rhs, ph, part, stream);
                        else if(kind == SOLVE_MAT)
                            bl_mat_phase(S, rhs, ph, part, stream);
                        else
                            bl_bicg_phase(S, x, rhs, ph, part, stream);
                    });
                    if(part == 0) tend();
                }
        if(kind == SOLVE_MIXED) each([&](const TankSet& S) { bl_mixed_settle(S, stream); });
        return 0;
    };
    int rc;
    // the lockstep sets only change with the tables; the scheduler's solving sets change
    // from round to round, and a set's batch graph is captured when the set comes back
    if(cur_ == &all_)
        rc = replay({1, kind, x, rhs * 64 + iters}, body);
    else if(graphs && !timing && cur_ == &solving_sets_)
    {
        std::vector<int> key(solving_key_);
        key.insert(key.end(), {-1, kind, x, rhs, iters});
        // a full cache captures no more sets (their batches launch directly) rather than
        // destroying executable graphs in the middle of a run: graph execs are destroyed only
        // when the tables change (drop_graphs, after a stream synchronisation) and at teardown
        auto it = set_graphs_.find(key);
        if(it == set_graphs_.end() && set_graphs_.size() < SET_GRAPHS_MAX) it = set_graphs_.emplace(key, SetGraph{}).first;
        if(it != set_graphs_.end())
        {
            SetGraph& sg = it->second;
            if(!sg.exec && ++sg.seen >= 2 && (rc = capture(body, &sg.exec))) return rc;
            if(sg.exec)
            {
                EHIP(hipGraphLaunch(sg.exec, stream));
                rc = 0;
            }
            else
                rc = body();
        }
        else
            rc = body();
    }
    else
        rc = body();
    if(rc) return rc;
    BiState* dst = d_status_ + (size_t)ring * cap_tanks_;
    each([&](const TankSet& S) { bl_gather_states(S, dst, stream); });
    EHIP(hipMemcpyAsync(h_status_ + (size_t)ring * cap_tanks_, dst, sizeof(BiState) * tanks.size(),
                        hipMemcpyDeviceToHost, stream));
    EHIP(hipEventRecord(st_ev_[ring], stream));
    return 0;
}

// BiCGStab iterations of the current sets until every tank of them has finished.  Two
// batches are in flight: batch k+1 is enqueued before the host waits for the states of
// batch k.
int Engine::run_iterations(int kind, int x, int rhs, int maxiter)
{
    std::vector<int> members;
    for(const TankSet& S : *cur_)
        for(int p = 0; p < S.n; ++p) members.push_back(S.slot[p]);
    int rc;
    if((rc = enqueue_batch(kind, x, rhs, 0, 8))) return rc;
    for(long k = 1;; ++k)
    {
        if((rc = enqueue_batch(kind, x, rhs, (int)(k & 1), 8))) return rc;
        EHIP(hipEventSynchronize(st_ev_[(k - 1) & 1]));
        const BiState* hs = h_status_ + (size_t)((k - 1) & 1) * cap_tanks_;
        bool all = true;
        for(int q : members)
            if(!hs[q].done) all = false;
        if(all)
        {
            for(int q : members) last_state[q] = hs[q];
            break;
        }
        if(timing && ranges_.size() > 2048) tcollect();
        if(k * 8 > 4L * (maxiter + 16)) return fail(REEF_E_HIP, "BiCGStab did not terminate");
    }
    if(timing) tcollect();
    return 0;
}

// face densities and the solver's init kernels (FP64, or FP32 coefficients and the mixed
// solver's init)
int Engine::solve_begin(int x, int rhs, double stop, int maxiter, bool mixed)
{
    const reef_params* sp = shared();
    const double rw = sp ? sp->W1 : 1000.0, ra = sp ? sp->W3 : 1.205;
    tbegin("face_rho");
    each([&](const TankSet& S) { bl_face_rho(S, rw, ra, mixed ? 0 : 1, stream); });
    tend();
    if(!mixed)
    {
        tbegin("bicg_init");
        each([&](const TankSet& S) { bl_bicg_init(S, x, rhs, stop, maxiter, stream); });
        tend();
        return 0;
    }
    tbegin("coef32");
    each([&](const TankSet& S) { bl_coef32(S, stream); });
    tend();
    tbegin("mix_init");
    each([&](const TankSet& S) { bl_mixed_init(S, x, rhs, stop, maxiter, delta, stream); });
    tend();
    return 0;
}

int Engine::solve_iterate(int x, int rhs, int maxiter, bool mixed)
{
    return run_iterations(mixed ? SOLVE_MIXED : SOLVE_FP64, x, rhs, maxiter);
}

// the mixed solve keeps its FP64 iterate split until the end: x64 += x32
int Engine::solve_end(int x, bool mixed)
{
    if(!mixed) return 0;
    tbegin("add32");
    each([&](const TankSet& S) { bl_add32(S, x, stream); });
    tend();
    return 0;
}

// poisson::start + solver::start on every active tank: FP64 bicgstab_ijk, or the mixed
// form (FP32 iteration with FP64 reliable updates) ending on the true FP64 residual
int Engine::solve(int x, int rhs, double stop, int maxiter, bool mixed)
{
    int rc;
    if((rc = sync_tables()) || (rc = drain())) return rc;
    if((rc = solve_begin(x, rhs, stop, maxiter, mixed))) return rc;
    if((rc = solve_iterate(x, rhs, maxiter, mixed))) return rc;
    return solve_end(x, mixed);
}

// the explicit-operator arrays of a tank (M.p .. M.b, the right-hand side and the solver's
// private x), zeroed, from the pool the first time a tank needs them
int Engine::ensure_matrix(int id)
{
    TankHost* t = tank(id);
    if(!t) return fail(REEF_E_ARG, "no such tank");
    if(t->dev.m[0]) return 0;
    const size_t bytes = (size_t)t->hg.nalloc * 8;
    for(int q = 0; q < 10; ++q)
    {
        void* m = nullptr;
        const int rc = pool.alloc(bytes, &m);
        if(rc) return fail(rc, "device pool: out of memory (budget) for the explicit operator");
        t->mem.push_back(m);
        t->bytes += up(bytes);
        t->dev.m[q] = (double*)m;
        EHIP(hipMemsetAsync(m, 0, bytes, stream));
    }
    dirty_ = true;
    return 0;
}

int Engine::wave_set(int id, const reef_wave_desc& w, double* params5)
{
    TankHost* t = tank(id);
    if(!t) return fail(REEF_E_ARG, "no such tank");
    if(t->dev.wave.col) return fail(REEF_E_ARG, "the tank already has a wave generator");
    WaveSetup ws;
    const std::string e = build_wave(t->hg, w, ws);
    if(!e.empty()) return fail(REEF_E_UNSUPPORTED, e);
    HostGrid& h = t->hg;
    const size_t ncol = (size_t)h.nx * h.ny, nzn = h.X[2].size();
    const reef_params* sp = shared();
    ws.dev.gabs = std::fabs(sp ? sp->W22 : t->prm.W22);
    void* m[3] = {nullptr, nullptr, nullptr};
    const size_t bytes[3] = {WC_N * ncol * 8, ncol * 8, nzn * 8};
    int rc;
    if((rc = drain())) return rc;
    EHIP(hipStreamSynchronize(stream));
    // initialize::inipsi with an ioflow of waves (B90 > 0, ini_psi.cpp:40-56): the interface
    // half width from the mean DZN of the cells within 5 DZM of the current phi, summed in
    // LOOP order
    {
        std::vector<double> phi((size_t)(h.nx + 6) * (h.ny + 6) * (h.nz + 6));
        if((rc = get(id, REEF_PHI, phi.data(), REEF_LAYOUT_IFAST))) return rc;
        double psim = 0.0;
        int count = 0;
        for(int i = 0; i < h.nx; ++i)
            for(int j = 0; j < h.ny; ++j)
                for(int k = 0; k < h.nz; ++k)
                    if(std::fabs(phi[(size_t)h.cid(i, j, k)]) < 5.0 * h.DZM)
                    {
                        psim += h.DN[2][k + MARGE];
                        ++count;
                    }
        if(count == 0) return fail(REEF_E_ARG, "wave set-up: no cell within 5 DZM of the interface (set phi first)");
        h.psi = t->prm.F45 * psim / double(count);
        t->dev.g.psi = h.psi;
    }
    for(int q = 0; q < 3; ++q)
    {
        if((rc = pool.alloc(bytes[q], &m[q])))
        {
            for(int r = 0; r < q; ++r) pool.free(m[r]);
            return fail(rc, "device pool: out of memory (budget) for the wave generator");
        }
    }
    const std::vector<double> zero(ncol, 0.0);
    EHIP(hipMemcpy(m[0], ws.col.data(), bytes[0], hipMemcpyHostToDevice));
    EHIP(hipMemcpy(m[1], zero.data(), bytes[1], hipMemcpyHostToDevice));
    EHIP(hipMemcpy(m[2], h.X[2].data(), bytes[2], hipMemcpyHostToDevice));
    for(int q = 0; q < 3; ++q)
    {
        t->mem.push_back(m[q]);
        t->bytes += up(bytes[q]);
    }
    ws.dev.col = (const double*)m[0];
    ws.dev.eta = (double*)m[1];
    ws.dev.zn = (const double*)m[2];
    t->dev.wave = ws.dev;
    if(params5)
    {
        params5[0] = ws.wT;
        params5[1] = ws.wL;
        params5[2] = ws.dev.wk;
        params5[3] = ws.dev.ww;
        params5[4] = ws.dev.wa;
    }
    any_wave_ = true;
    drop_graphs();      // captured steps have no wave launches
    dirty_ = true;
    return 0;
}

int Engine::set_simtime(double tm)
{
    int rc;
    if((rc = sync_tables()) || (rc = drain())) return rc;
    EHIP(hipStreamSynchronize(stream));
    for(TankHost& t : tanks)
        if(t.used && t.active && t.dev.wave.col) EHIP(hipMemcpy(t.dev.tstep + 6, &tm, sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int Engine::wave_relax(int u, int v, int w, int p, int f)
{
    int rc;
    if((rc = sync_tables()) || (rc = drain())) return rc;
    tbegin("wave");
    each([&](const TankSet& S) {
        bl_wave_eta(S, stream);
        bl_wave_relax(S, u, v, w, p, f, stream);
    });
    tend();
    return 0;
}

// solver::start (bicgstab_ijk / bicgstab_ijk_2D, bicgstab_ijk.cpp:51-123) on the explicit
// operator of every active tank: var 1..3 over the u / v / w grids (stop = N43), 4, 5 over
// every cell (N44); the solution in place in array x
int Engine::mat_solve(int var, int x, double stop, int maxiter)
{
    if(var < 1 || var > 5) return fail(REEF_E_ARG, "var must be 1..5");
    int rc;
    for(const TankHost& t : tanks)
        if(t.used && t.active && !t.dev.m[0]) return fail(REEF_E_ARG, "tank without an explicit operator");
    if((rc = sync_tables()) || (rc = drain())) return rc;
    tbegin("mat_init");
    each([&](const TankSet& S) {
        bl_mat_copy_x(S, x, var, 0, stream);
        bl_mat_init(S, var, stop, maxiter, stream);
    });
    tend();
    if((rc = run_iterations(SOLVE_MAT, -1, var, maxiter))) return rc;
    each([&](const TankSet& S) { bl_mat_copy_x(S, x, var, 1, stream); });
    return 0;
}

int Engine::residual(int x, int rhs, std::vector<double>& norms)
{
    int rc;
    if((rc = sync_tables())) return rc;
    const reef_params* sp = shared();
    const double rw = sp ? sp->W1 : 1000.0, ra = sp ? sp->W3 : 1.205;
    each([&](const TankSet& S) {
        bl_face_rho(S, rw, ra, 0, stream);
        bl_residual(S, x, rhs, d_norms_, stream);
    });
    norms.assign(tanks.size(), 0.0);
    EHIP(hipMemcpyAsync(norms.data(), d_norms_, sizeof(double) * tanks.size(), hipMemcpyDeviceToHost, stream));
    EHIP(hipStreamSynchronize(stream));
    return 0;
}

// pjm_corr::start (pjm_corr.cpp:90-121) up to the solve: velocity ghosts, rhs, p' = 0,
// solver set-up
int Engine::pjm_pre(int u, int v, int w, double alpha, bool mixed)
{
    int rc;
    if((rc = ghost(u, 1, 7)) || (rc = ghost(v, 2, 8)) || (rc = ghost(w, 3, 9))) return rc;
    const reef_params* sp = shared();
    if(!sp) return 0;
    tbegin("poisson_rhs");
    each([&](const TankSet& S) {
        bl_poisson_rhs(S, u, v, w, alpha, stream);
        bl_fill(S, REEF_PCORR, 0.0, stream);
    });
    tend();
    return solve_begin(REEF_PCORR, A_RHS, sp->N44, sp->N46, mixed);
}

// pjm_corr::start after the solve: p' ghosts, velocity and pressure correction
int Engine::pjm_post(int u, int v, int w, double alpha, bool mixed)
{
    int rc;
    const reef_params* sp = shared();
    if(!sp) return 0;
    if((rc = solve_end(REEF_PCORR, mixed))) return rc;
    if((rc = ghost(REEF_PCORR, 4, 40))) return rc;
    tbegin("correct");
    each([&](const TankSet& S) { bl_correct(S, sp->W1, sp->W3, alpha, u, v, w, stream); });
    tend();
    return ghost(REEF_PRESS, 4, 40);
}

// pjm_corr::start (pjm_corr.cpp:90-121) on every active tank
int Engine::pjm(int u, int v, int w, double alpha)
{
    int rc;
    if((rc = sync_tables()) || (rc = drain())) return rc;
    const reef_params* sp = shared();
    const bool mixed = precision == REEF_SOLVER_MIXED;
    if((rc = pjm_pre(u, v, w, alpha, mixed))) return rc;
    if(!sp) return 0;
    if((rc = solve_iterate(REEF_PCORR, A_RHS, sp->N46, mixed))) return rc;
    for(const TankSet& S : all_)
        for(int p = 0; p < S.n; ++p) step_iters[S.slot[p]] += last_state[S.slot[p]].iter;
    return pjm_post(u, v, w, alpha, mixed);
}

// velocity predictor of the three components (momentum_FC3.cpp:160-208 and the later stages)
int Engine::momentum_stage(int stage, const int qin[3], const int qold[3], const int vel[3], const int out[3])
{
    const reef_params* sp = shared();
    if(!sp) return 0;
    const double grav[3] = {sp->W20, sp->W21, sp->W22};
    const double cfac = stage == 1 ? 1.0 : stage == 2 ? 0.25 : (2.0 / 3.0);
    int rc;
    if(d20 == 2)
    {
        // idiff2_FS: per component F (irhs + convection) into F / G / H, the implicit
        // viscous solve (var = comp, stop N43) on the private x, then the stage combination
        static const int FGH[3] = {REEF_F, REEF_G, REEF_H};
        for(int comp = 1; comp <= 3; ++comp)
        {
            const int q = comp - 1;
            tbegin("momentum");
            each([&](const TankSet& S) {
                bl_momentum(S, comp, stage, qin[q], qold[q], vel[0], vel[1], vel[2], grav[q], cfac, FGH[q], sp->W1,
                            sp->W3, 2, stream);
            });
            tend();
            tbegin("idiff");
            each([&](const TankSet& S) {
                bl_idiff(S, comp, cfac, qin[q], vel[0], vel[1], vel[2], stream);
                bl_mat_init(S, comp, n43, sp->N46, stream);
            });
            tend();
            if((rc = run_iterations(SOLVE_MAT, -1, comp, sp->N46))) return rc;
            tbegin("idiff");
            each([&](const TankSet& S) { bl_idiff_combine(S, comp, stage, qold[q], FGH[q], cfac, out[q], stream); });
            tend();
        }
        if(stage == 1) each([&](const TankSet& S) { bl_df_set(S, stream); });
    }
    else
        for(int comp = 1; comp <= 3; ++comp)
        {
            tbegin("momentum");
            each([&](const TankSet& S) {
                bl_momentum(S, comp, stage, qin[comp - 1], qold[comp - 1], vel[0], vel[1], vel[2], grav[comp - 1], cfac,
                            out[comp - 1], sp->W1, sp->W3, d20, stream);
            });
            tend();
        }
    const int kinds[3] = {1, 2, 3}, gcvs[3] = {10, 11, 12};
    for(int comp = 0; comp < 3; ++comp)
        if((rc = ghost(out[comp], kinds[comp], gcvs[comp]))) return rc;
    return 0;
}

// D20 = 2 needs each tank's explicit-operator arrays
int Engine::set_d20(int v)
{
    if(v < 0 || v > 2) return fail(REEF_E_ARG, "D20 must be 0, 1 or 2");
    drop_graphs();
    if(v == 2)
        for(size_t q = 0; q < tanks.size(); ++q)
            if(tanks[q].used)
            {
                const int rc = ensure_matrix((int)q);
                if(rc) return rc;
            }
    d20 = v;
    return 0;
}

// The launches of momentum_FC3::start (momentum_FC3.cpp:122-430) between its pressure
// solves.  Segment 0: stage-1 phi transport, reinit, predictor and the stage-1 projection
// up to its solve; segments 1 and 2: the end of the previous projection, velocity ghosts,
// the next stage up to its solve; segment 3: the end of the stage-3 projection, velocity
// ghosts and the fluid update (rho and nu depend on phi only, so the last fluid_update of
// the step is the one that stays).
int Engine::fc3_segment(int s, bool mixed)
{
    static const int VEL[3][3] = {{A_URK1, A_VRK1, A_WRK1}, {A_URK2, A_VRK2, A_WRK2}, {REEF_U, REEF_V, REEF_W}};
    static const double ALPHA[3] = {1.0, 0.25, 2.0 / 3.0};
    int rc;
    if(s > 0)
    {
        const int* vel = VEL[s - 1];
        if((rc = pjm_post(vel[0], vel[1], vel[2], ALPHA[s - 1], mixed))) return rc;
        // pflow->u_relax, v_relax, w_relax, p_relax (momentum_FC3.cpp:218-221, 320-323, 418-421)
        if(any_wave_)
        {
            tbegin("wave");
            each([&](const TankSet& S) { bl_wave_relax(S, vel[0], vel[1], vel[2], REEF_PRESS, -1, stream); });
            tend();
        }
        if((rc = ghost(vel[0], 1, 10)) || (rc = ghost(vel[1], 2, 11)) || (rc = ghost(vel[2], 3, 12))) return rc;
    }
    if(s == 0 && n48)
    {
        // adaptive time step at the start of the step (driver_loop_cfd.cpp:103-105)
        tbegin("timestep");
        each([&](const TankSet& S) { bl_timestep(S, REEF_U, REEF_V, REEF_W, n47, 0, stream); });
        tend();
    }
    if(s == 0 && any_wave_)
    {
        // the driver loop before momentum (driver_loop_cfd.cpp:78-92): wavegen_precalc at the
        // step's simtime, u_relax, v_relax, w_relax of the fdm's velocities, then
        // levelset_void::update, the fluid update (levelset_void.cpp:86-89)
        tbegin("wave");
        each([&](const TankSet& S) {
            bl_wave_eta(S, stream);
            bl_wave_relax(S, REEF_U, REEF_V, REEF_W, -1, -1, stream);
        });
        tend();
        if((rc = fluid_update())) return rc;
    }
    if(s == 0)
    {
        // stage 1 (momentum_FC3.cpp:132-230)
        tbegin("copy");
        each([&](const TankSet& S) { bl_copy(S, REEF_PHI, A_LS, stream); });   // ls ghosts stay zero, as in the reference
        tend();
        tbegin("hj");
        each([&](const TankSet& S) { bl_hj_stage(S, 1, A_LS, -1, REEF_U, REEF_V, REEF_W, A_FRK1, 1.0, stream); });
        tend();
        if(any_wave_) wave_phi(A_FRK1);         // pflow->phi_relax (momentum_FC3.cpp:145)
        if((rc = ghost(A_FRK1, 4, 52))) return rc;
        if((rc = reini(A_FRK1, 2))) return rc;
        tbegin("copy");
        each([&](const TankSet& S) { bl_copy(S, A_FRK1, REEF_PHI, stream); });
        tend();
        if((rc = ghost(REEF_PHI, 4, 52))) return rc;
        // the stage's fluid update (pupdate, momentum_FC3.cpp:157-159): the diffusion reads the
        // mixture viscosity of this stage's phi, the beach's pressure relaxation its density;
        // otherwise nothing in the step reads rho or nu before the final update (face
        // densities come from phi)
        if((d20 || any_wave_) && (rc = fluid_update())) return rc;
        const int qin[3] = {REEF_U, REEF_V, REEF_W};
        if((rc = momentum_stage(1, qin, qin, qin, VEL[0]))) return rc;
    }
    else if(s == 1)
    {
        // stage 2 (momentum_FC3.cpp:232-332)
        tbegin("hj");
        each([&](const TankSet& S) { bl_hj_stage(S, 2, A_FRK1, A_LS, A_URK1, A_VRK1, A_WRK1, A_FRK2, 0.25, stream); });
        tend();
        if(any_wave_) wave_phi(A_FRK2);         // momentum_FC3.cpp:244
        if((rc = ghost(A_FRK2, 4, 52))) return rc;
        if((rc = reini(A_FRK2, 2))) return rc;
        tbegin("copy");
        each([&](const TankSet& S) { bl_copy(S, A_FRK2, REEF_PHI, stream); });
        tend();
        if((rc = ghost(REEF_PHI, 4, 52))) return rc;
        if((rc = reini(REEF_PHI, 2))) return rc;
        if((d20 || any_wave_) && (rc = fluid_update())) return rc;   // momentum_FC3.cpp:260
        const int qold[3] = {REEF_U, REEF_V, REEF_W};
        if((rc = momentum_stage(2, VEL[0], qold, VEL[0], VEL[1]))) return rc;
    }
    else if(s == 2)
    {
        // stage 3 (momentum_FC3.cpp:334-430)
        tbegin("hj");
        each([&](const TankSet& S) {
            bl_hj_stage(S, 3, A_FRK2, A_LS, A_URK2, A_VRK2, A_WRK2, REEF_PHI, 2.0 / 3.0, stream);
        });
        tend();
        if(any_wave_) wave_phi(REEF_PHI);       // momentum_FC3.cpp:346
        if((rc = ghost(REEF_PHI, 4, 52))) return rc;
        if((rc = reini(REEF_PHI, 3))) return rc;
        if((d20 || any_wave_) && (rc = fluid_update())) return rc;   // momentum_FC3.cpp:358
        const int qold[3] = {REEF_U, REEF_V, REEF_W};
        if((rc = momentum_stage(3, VEL[1], qold, VEL[1], VEL[2]))) return rc;
    }
    else
    {
        if((rc = fluid_update())) return rc;
        // p->simtime += p->dt after the step (driver_loop_cfd.cpp:100)
        if(any_wave_) each([&](const TankSet& S) { bl_wave_time(S, stream); });
        return 0;
    }
    return pjm_pre(VEL[s][0], VEL[s][1], VEL[s][2], ALPHA[s], mixed);
}

// iowave::phi_relax (iowave_relax.cpp:358-395) of field f on the active wave tanks
void Engine::wave_phi(int f)
{
    tbegin("wave");
    each([&](const TankSet& S) { bl_wave_relax(S, -1, -1, -1, -1, f, stream); });
    tend();
}

// momentum_FC3::start (momentum_FC3.cpp:122-430): three TVD-RK3 stages of phi transport,
// reinitialisation, velocity predictor and projection; rho / nu from the final phi
int Engine::fc3_step()
{
    int rc;
    if((rc = sync_tables()) || (rc = drain())) return rc;
    std::fill(step_iters.begin(), step_iters.end(), 0L);
    const reef_params* sp = shared();
    if(!sp) return 0;
    const bool mixed = precision == REEF_SOLVER_MIXED;
    for(int s = 0; s < 4; ++s)
    {
        // the implicit diffusion (D20 = 2) synchronises with the host inside its segments,
        // which are then launched directly
        if(d20 == 2)
            rc = fc3_segment(s, mixed);
        else
            rc = replay({0, s, mixed ? 1 : 0, fsf}, [&] { return fc3_segment(s, mixed); });
        if(rc) return rc;
        if(s == 3) break;
        if((rc = solve_iterate(REEF_PCORR, A_RHS, sp->N46, mixed))) return rc;
        for(const TankSet& S : all_)
            for(int p = 0; p < S.n; ++p) step_iters[S.slot[p]] += last_state[S.slot[p]].iter;
    }
    return 0;
}

// the step size of the active tanks' following steps (fixtimestep / the reference's own
// time stepping setting p->dt)
int Engine::set_dt(double dt)
{
    int rc;
    if((rc = sync_tables()) || (rc = drain())) return rc;
    EHIP(hipStreamSynchronize(stream));
    for(TankHost& t : tanks)
        if(t.used && t.active)
        {
            t.prm.dt = dt;
            EHIP(hipMemcpy(t.dev.tstep, &dt, sizeof(double), hipMemcpyHostToDevice));
        }
    return 0;
}

// ietimestep::start / ::ini on the active tanks; out (optional) receives dt, umax, vmax,
// wmax per slot
int Engine::timestep(double N47, int ini, std::vector<double>* out)
{
    int rc;
    if((rc = sync_tables()) || (rc = drain())) return rc;
    tbegin("timestep");
    each([&](const TankSet& S) { bl_timestep(S, REEF_U, REEF_V, REEF_W, N47, ini, stream); });
    tend();
    if(!out) return 0;
    out->assign(4 * tanks.size(), 0.0);
    EHIP(hipStreamSynchronize(stream));
    for(size_t q = 0; q < tanks.size(); ++q)
        if(tanks[q].used && tanks[q].active)
            EHIP(hipMemcpy(out->data() + 4 * q, tanks[q].dev.tstep, 4 * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

// ------------------------------------------------------------------ iteration-level scheduling
int Engine::start_steps(int id, int n)
{
    TankHost* t = tank(id);
    if(!t || n < 0) return fail(REEF_E_ARG, "bad tank or step count");
    if(d20 == 2)
        return fail(REEF_E_UNSUPPORTED, "D20 = 2 (implicit diffusion) runs in lockstep steps only (reef_batch_fc3_step)");
    t->steps_left += n;
    if(t->pc < 0 && t->steps_left > 0) t->pc = 0;
    return 0;
}

// pc 1, 3, 5: in the pressure solve of stage 1, 2, 3 (pc -1, idle, is not odd here)
static inline bool solving_pc(int pc) { return pc == 1 || pc == 3 || pc == 5; }

bool Engine::busy() const
{
    if(!inflight_.empty()) return true;
    for(const TankHost& t : tanks)
        if(t.used && t.pc >= 0) return true;
    return false;
}

// states of an iteration batch: a tank whose solve (the one the batch ran) has ended moves
// on to its next segment
int Engine::read_batch(const Inflight& f)
{
    EHIP(hipEventSynchronize(st_ev_[f.ring]));
    const BiState* hs = h_status_ + (size_t)f.ring * cap_tanks_;
    for(const auto& m : f.members)
    {
        TankHost& t = tanks[m.first];
        if(!t.used || t.seq != m.second || !solving_pc(t.pc)) continue;
        const BiState& st = hs[m.first];
        if(!st.done) continue;
        t.iters_step += st.iter;
        last_state[m.first] = st;
        t.pc += 1;
    }
    return 0;
}

int Engine::drain()
{
    int rc;
    while(!inflight_.empty())
    {
        const Inflight f = std::move(inflight_.front());
        inflight_.pop_front();
        if((rc = read_batch(f))) return rc;
    }
    return 0;
}

// One round: the segments the tanks have reached (grouped by segment, in the order 3, 0,
// 1, 2 so that a tank that ends a step starts the next one in the same round), then one
// batch of 8 iterations for the tanks that are solving.  The states of the batch before
// are read while this one runs.
int Engine::pump(std::vector<int>& stepped)
{
    stepped.clear();
    int rc;
    if((rc = sync_tables())) return rc;
    if(!first_used()) return 0;
    const bool mixed = precision == REEF_SOLVER_MIXED;
    static const int ORDER[4] = {3, 0, 1, 2};
    for(int s : ORDER)
    {
        std::vector<int> group;
        for(size_t q = 0; q < tanks.size(); ++q)
            if(tanks[q].used && tanks[q].pc == 2 * s) group.push_back((int)q);
        if(group.empty()) continue;
        const std::vector<TankSet> sets = make_sets(group);
        cur_ = &sets;
        rc = fc3_segment(s, mixed);
        cur_ = &all_;
        if(rc) return rc;
        for(int q : group)
        {
            TankHost& t = tanks[q];
            if(s < 3)
            {
                t.pc = 2 * s + 1;
                t.seq += 1;
                continue;
            }
            t.steps_done += 1;
            t.steps_left -= 1;
            t.iters_last = t.iters_step;
            t.iters_step = 0;
            step_iters[q] = t.iters_last;
            stepped.push_back(q);
            t.pc = t.steps_left > 0 ? 0 : -1;
        }
    }
    std::vector<int> solving;
    for(size_t q = 0; q < tanks.size(); ++q)
        if(tanks[q].used && solving_pc(tanks[q].pc)) solving.push_back((int)q);
    if(!solving.empty())
    {
        if(solving != solving_key_)
        {
            solving_sets_ = make_sets(solving);
            solving_key_ = solving;
        }
        const int ring = (int)(batch_no_++ & 1);
        cur_ = &solving_sets_;
        // many tanks solving: a deeper batch per round amortises the round's segment
        // launches and host synchronisation (iterations after a tank's end are no-ops)
        rc = enqueue_batch(mixed ? SOLVE_MIXED : SOLVE_FP64, REEF_PCORR, A_RHS, ring, (int)solving.size() >= 8 ? 32 : 8);
        cur_ = &all_;
        if(rc) return rc;
        Inflight f;
        f.ring = ring;
        for(int q : solving) f.members.push_back({q, tanks[q].seq});
        inflight_.push_back(std::move(f));
    }
    if(inflight_.size() > 1 || (solving.empty() && !inflight_.empty()))
    {
        const Inflight f = std::move(inflight_.front());
        inflight_.pop_front();
        if((rc = read_batch(f))) return rc;
    }
    if(timing && ranges_.size() > 2048) tcollect();
    return 0;
}

int Engine::run_scheduled()
{
    std::vector<int> stepped;
    int rc;
    while(busy())
        if((rc = pump(stepped))) return rc;
    if(timing) tcollect();
    return 0;
}

// levelset_RK3 transport stages (levelset_RK3.cpp:109-160)
int Engine::phi_rk3()
{
    int rc;
    if((rc = sync_tables())) return rc;
    tbegin("hj");
    each([&](const TankSet& S) { bl_hj_stage(S, 1, REEF_PHI, -1, REEF_U, REEF_V, REEF_W, A_FRK1, 1.0, stream); });
    tend();
    if((rc = ghost(A_FRK1, 4, 52))) return rc;
    tbegin("hj");
    each([&](const TankSet& S) { bl_hj_stage(S, 2, A_FRK1, REEF_PHI, REEF_U, REEF_V, REEF_W, A_FRK2, 0.25, stream); });
    tend();
    if((rc = ghost(A_FRK2, 4, 52))) return rc;
    tbegin("hj");
    each([&](const TankSet& S) {
        bl_hj_stage(S, 3, A_FRK2, REEF_PHI, REEF_U, REEF_V, REEF_W, REEF_PHI, 2.0 / 3.0, stream);
    });
    tend();
    return ghost(REEF_PHI, 4, 52);
}

// one level-set benchmark step (levelset_RK3 + reini + momentum_RK3 void + benchmark_vortex3D)
int Engine::vortex_step(double simtime)
{
    int rc;
    if((rc = phi_rk3())) return rc;
    if((rc = reini(REEF_PHI, 3))) return rc;
    const TankHost* t0 = first_used();
    const reef_params* sp = shared();
    if(!t0 || !sp) return 0;
    const double grav[3] = {sp->W20, sp->W21, sp->W22};
    for(int comp = 1; comp <= 3; ++comp)
    {
        if(comp == 2 && t0->hg.jdir == 0) continue;
        const double Fv = 0.0 + (0.0 + grav[comp - 1] + 0.0 + 0.0) * 1.0;
        each([&](const TankSet& S) { bl_void_rk3(S, comp, REEF_U + comp - 1, Fv, stream); });
        if((rc = ghost(REEF_U + comp - 1, comp, 9 + comp))) return rc;
    }
    for(int comp = 1; comp <= 3; ++comp)
        each([&](const TankSet& S) { bl_vortex_velocity(S, comp, REEF_U + comp - 1, simtime, t0->hg.DXM, stream); });
    // benchmark_vortex3D.cpp:98-100 fills u (10), v (11) and then v again with gcv 12; w's
    // ghosts keep the values of the void RK3 fill above.  Kept as the reference has it.
    if((rc = ghost(REEF_U, 1, 10))) return rc;
    if((rc = ghost(REEF_V, 2, 11))) return rc;
    if((rc = ghost(REEF_V, 2, 12))) return rc;
    return fluid_update();
}

// ------------------------------------------------------------------ autotuner
// One sample of the step's kernels on step-internal arrays only: phi copied to ls, an HJ
// stage into frk1, one reinit iteration (3 stages) of frk1, the stage-1 predictor into
// urk1..wrk1, then the projection set-up on those and 16 iterations of the configured
// solver on p'.  phi, u, v, w, press, ro and visc are read, never written; ls, frk1,
// urk1..wrk1, the rhs, p' and the solver arrays are rewritten by every step before use.
int Engine::probe(double& stencil_ms, double& solver_ms)
{
    int rc;
    stencil_ms = solver_ms = 0.0;
    if((rc = sync_tables())) return rc;
    if(!first_used()) return 0;
    struct Ev
    {
        hipEvent_t e = nullptr;
        ~Ev()
        {
            if(e) (void)hipEventDestroy(e);
        }
    } ev[3];
    for(Ev& x : ev) EHIP(hipEventCreate(&x.e));
    const bool mixed = precision == REEF_SOLVER_MIXED;
    EHIP(hipEventRecord(ev[0].e, stream));
    each([&](const TankSet& S) {
        bl_copy(S, REEF_PHI, A_LS, stream);
        bl_hj_stage(S, 1, A_LS, -1, REEF_U, REEF_V, REEF_W, A_FRK1, 1.0, stream);
    });
    if((rc = ghost(A_FRK1, 4, 52)) || (rc = reini(A_FRK1, 1))) return rc;
    const int qin[3] = {REEF_U, REEF_V, REEF_W}, out[3] = {A_URK1, A_VRK1, A_WRK1};
    // the predictor without the implicit diffusion, whose solves would move the private x
    const int keep_d20 = d20;
    d20 = d20 == 2 ? 0 : d20;
    rc = momentum_stage(1, qin, qin, qin, out);
    d20 = keep_d20;
    if(rc) return rc;
    EHIP(hipEventRecord(ev[1].e, stream));
    if((rc = ghost(A_URK1, 1, 7)) || (rc = ghost(A_VRK1, 2, 8)) || (rc = ghost(A_WRK1, 3, 9))) return rc;
    each([&](const TankSet& S) {
        bl_poisson_rhs(S, A_URK1, A_VRK1, A_WRK1, 1.0, stream);
        bl_fill(S, REEF_PCORR, 0.0, stream);
    });
    if((rc = solve_begin(REEF_PCORR, A_RHS, 1e-300, 16, mixed))) return rc;
    for(int it = 0; it < 16; ++it)
        for(int ph = 0; ph < 4; ++ph)
            for(int part = 0; part < 2; ++part)
                each([&](const TankSet& S) {
                    if(mixed)
                        bl_mixed_phase(S, REEF_PCORR, A_RHS, ph, part, stream);
                    else
                        bl_bicg_phase(S, REEF_PCORR, A_RHS, ph, part, stream);
                });
    if(mixed) each([&](const TankSet& S) { bl_mixed_settle(S, stream); });
    if((rc = solve_end(REEF_PCORR, mixed))) return rc;
    EHIP(hipEventRecord(ev[2].e, stream));
    EHIP(hipEventSynchronize(ev[2].e));
    float a = 0.0f, b = 0.0f;
    EHIP(hipEventElapsedTime(&a, ev[0].e, ev[1].e));
    EHIP(hipEventElapsedTime(&b, ev[1].e, ev[2].e));
    stencil_ms = a;
    solver_ms = b;
    return 0;
}

// For each bucket of the active tanks without a plan: every tile shape (3-D: 64x4x1,
// 64x2x2, 64x1x4; 2-D: 64x1x4) and chunk length 1..64 (powers of two) is probed on that
// bucket's tanks alone, launch by launch, reps times after one warm-up; the score is the
// best probe time weighted to about one momentum_FC3 step (9 x the stencil sample, which
// holds a third of a stage's HJ / reinit / predictor work, plus 300 / 16 x the 16-iteration
// solver sample).  The fastest plan is kept.
int Engine::autotune(int reps, std::vector<TuneRecord>* log)
{
    int rc;
    if((rc = sync_tables())) return rc;
    std::map<Bucket, std::vector<int>> members;
    std::vector<int> was_active;
    for(size_t q = 0; q < tanks.size(); ++q)
        if(tanks[q].used && tanks[q].active)
        {
            was_active.push_back((int)q);
            if(!plans.count(bucket_of(tanks[q].hg))) members[bucket_of(tanks[q].hg)].push_back((int)q);
        }
    if(members.empty()) return 0;
    const bool keep_graphs = graphs, keep_timing = timing;
    graphs = false;
    timing = false;
    reps = std::max(1, reps);
    static const int SHAPES[3][2] = {{4, 1}, {2, 2}, {1, 4}};
    auto tune_bucket = [&](const Bucket& bk, const std::vector<int>& ids) -> int {
        int r;
        if((r = set_active(ids))) return r;
        int max_tiles = 1;
        for(int id : ids) max_tiles = std::max(max_tiles, plan_shape(tanks[id].hg, 1).ntiles);
        Plan best{};
        double best_score = 1e300;
        for(const auto& sh : SHAPES)
        {
            if(bk.first == 0 && sh[0] != 1) continue;
            for(int cpt = 1; cpt <= 64; cpt *= 2)
            {
                if(cpt / 2 >= max_tiles) break;   // one item per tank already
                const Plan p{sh[0], sh[1], cpt};
                for(int id : ids)
                {
                    const HostGrid& h = tanks[id].hg;
                    tanks[id].dev.shape = make_shape(h.nx, h.ny, h.nz, p.ty, p.tz, p.cpt);
                }
                dirty_ = true;
                double score = 1e300;
                for(int k = 0; k <= reps; ++k)
                {
                    double a, b;
                    if((r = probe(a, b))) return r;
                    if(k > 0) score = std::min(score, 9.0 * a + (300.0 / 16.0) * b);
                }
                if(score < best_score)
                {
                    best_score = score;
                    best = p;
                }
            }
        }
        plans[bk] = best;
        if(log) log->push_back({bk, best, best_score});
        return 0;
    };
    for(const auto& m : members)
        if((rc = tune_bucket(m.first, m.second))) break;
    // every tank back on its bucket's plan, the active set as before
    for(TankHost& t : tanks)
        if(t.used) t.dev.shape = shape_of(t.hg);
    dirty_ = true;
    const int rc2 = set_active(was_active);
    graphs = keep_graphs;
    timing = keep_timing;
    return rc ? rc : rc2;
}

}  // namespace reef
