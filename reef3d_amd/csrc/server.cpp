// Serving loop over a batch engine (engine.h).  Requests (one tank: grid, initial fields,
// time step and step count) are submitted from any host thread.  Builder threads make the
// host half of each tank (grid metadata and compiled ghost programs, Engine::build_tank)
// as requests arrive, while the GPU keeps working; one worker thread owns the engine and
// its stream.  Each round the worker
//   1. admits built requests in arrival order while their device bytes fit in the HBM
//      budget and the live count allows (the first that does not fit stops admission, so
//      a large request is not overtaken for ever), prepares the admitted tanks alone
//      (ghost fills and fluid update of those tanks) and schedules their steps,
//   2. runs one round of the engine's iteration-level scheduler: every live tank moves on
//      through its own momentum_FC3 steps (a tank whose pressure solve has converged
//      goes on to its next segment while the others keep iterating),
//   3. retires the tanks that have done all their steps: fields copied to the request,
//      tank removed, its device memory back in the pool for the next admission.
// A tank's arithmetic does not depend on the other tanks of a launch, so a request's
// result is the same whether it was served alone or among others (given the same plan
// table and controls).
#include "engine.h"

#include "../../include/reef3d_serve.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

using namespace reef;

namespace {

thread_local std::string s_err;

int sfail(int code, const std::string& m)
{
    s_err = m;
    return code;
}

enum
{
    RQ_QUEUED = 0,
    RQ_LIVE = 1,
    RQ_DONE = 2,
    RQ_FAILED = 3
};

constexpr int NOUT = REEF_VISC + 1;   // u, v, w, press, phi, ro, visc

struct Request
{
    long id = 0;
    reef_grid_desc gd{};
    std::vector<double> XN, YN, ZN;
    double dt = 0.0;
    int steps = 0;
    size_t bytes = 0;
    std::vector<double> init[5];      // phi, u, v, w, press (i fastest, 3 ghost layers)
    std::unique_ptr<TankBuild> built; // host half of the tank, once a builder made it
    bool ready = false;               // built, or failed to build (code != 0)
    std::vector<double> out[NOUT];    // every field id up to REEF_VISC after the last step
    int state = RQ_QUEUED;
    int tank = -1;
    int steps_done = 0;
    long iters = 0;
    double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
    int code = 0;
    std::string err;
};

using RequestPtr = std::shared_ptr<Request>;

}  // namespace

struct reef_server
{
    hipStream_t stream = nullptr;
    std::unique_ptr<Engine> e;
    reef_server_config cfg{};
    std::vector<int> plans;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_build;
    std::deque<RequestPtr> queue;     // admission order
    std::deque<RequestPtr> to_build;  // not yet taken by a builder
    std::unordered_map<long, RequestPtr> all;
    std::vector<RequestPtr> live;     // worker only
    long next_id = 1;
    bool stop = false;
    int fatal = 0;
    std::string fatal_msg;
    long steps = 0;
    size_t bytes_in_use = 0;
    size_t peak_bytes = 0;            // largest bytes_in_use since creation
    std::thread worker;
    std::vector<std::thread> builders;
    int device = 0;
    std::chrono::steady_clock::time_point t0;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    void run();
    void build_loop();
    int round(std::vector<RequestPtr>& admitted);
};

namespace {

const int INIT_FIELDS[5] = {REEF_PHI, REEF_U, REEF_V, REEF_W, REEF_PRESS};

}  // namespace

// one admission + step + retirement round of the worker; a non-zero status is fatal
int reef_server::round(std::vector<RequestPtr>& admitted)
{
    Engine& E = *e;
    int rc;
    std::vector<int> new_ids;
    for(size_t q = 0; q < admitted.size(); ++q)
    {
        RequestPtr& r = admitted[q];
        int tank = -1;
        rc = E.add_built(*r->built, &tank);
        if(rc == REEF_E_NOMEM && !(live.empty() && new_ids.empty()))
        {
            // the pool is fuller than the estimate said: back to the head of the queue
            std::lock_guard<std::mutex> lk(mu);
            for(size_t b = admitted.size(); b-- > q;) queue.push_front(admitted[b]);
            admitted.resize(q);
            break;
        }
        if(rc)
        {
            std::lock_guard<std::mutex> lk(mu);
            r->state = RQ_FAILED;
            r->code = rc;
            r->err = E.err;
            admitted.erase(admitted.begin() + q--);
            cv_done.notify_all();
            continue;
        }
        r->tank = tank;
        r->built.reset();
        for(int f = 0; f < 5; ++f)
        {
            if((rc = E.put(tank, INIT_FIELDS[f], r->init[f].data(), REEF_LAYOUT_IFAST))) return rc;
            std::vector<double>().swap(r->init[f]);
        }
        new_ids.push_back(tank);
    }
    if(!new_ids.empty() && (rc = E.prepare_tanks(new_ids))) return rc;
    {
        std::lock_guard<std::mutex> lk(mu);
        const double t = now();
        for(RequestPtr& r : admitted)
        {
            r->state = RQ_LIVE;
            r->t_admit = t;
            live.push_back(r);
        }
        bytes_in_use = E.pool.in_use();
        peak_bytes = std::max(peak_bytes, bytes_in_use);
    }
    for(RequestPtr& r : admitted)
        if((rc = E.start_steps(r->tank, r->steps))) return rc;
    admitted.clear();
    if(live.empty()) return 0;
    std::vector<int> stepped;
    if((rc = E.pump(stepped))) return rc;
    ++steps;
    if(stepped.empty()) return 0;
    std::vector<RequestPtr> still;
    std::vector<RequestPtr> finished;
    {
        std::lock_guard<std::mutex> lk(mu);
        for(RequestPtr& r : live)
        {
            const TankHost* t = E.tank(r->tank);
            if(std::find(stepped.begin(), stepped.end(), r->tank) != stepped.end())
            {
                r->steps_done += 1;
                r->iters += t->iters_last;
            }
            (t->steps_left == 0 && t->pc < 0 ? finished : still).push_back(r);
        }
    }
    for(RequestPtr& r : finished)
    {
        const TankHost* t = E.tank(r->tank);
        const size_t n = (size_t)(t->hg.nx + 6) * (t->hg.ny + 6) * (t->hg.nz + 6);
        for(int f = 0; f < NOUT; ++f)
        {
            r->out[f].resize(n);
            if((rc = E.get(r->tank, f, r->out[f].data(), REEF_LAYOUT_IFAST))) return rc;
        }
        if((rc = E.remove_tank(r->tank))) return rc;
    }
    live.swap(still);
    if(!finished.empty())
    {
        std::lock_guard<std::mutex> lk(mu);
        const double t = now();
        for(RequestPtr& r : finished)
        {
            r->state = RQ_DONE;
            r->t_done = t;
        }
        bytes_in_use = E.pool.in_use();
        cv_done.notify_all();
    }
    return 0;
}

// builder thread: the host half of each submitted tank, in submission order
void reef_server::build_loop()
{
    for(;;)
    {
        RequestPtr r;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_build.wait(lk, [&] { return stop || !to_build.empty(); });
            if(stop) return;
            r = to_build.front();
            to_build.pop_front();
        }
        reef_params p = cfg.shared;
        p.dt = r->dt;
        std::unique_ptr<TankBuild> b(new TankBuild());
        std::string err;
        int rc = Engine::build_tank(r->gd, p, *b, err);
        const size_t exact = rc ? 0 : Engine::built_bytes(*b);
        if(!rc && cfg.hbm_budget_bytes > 0.0 && (double)exact > cfg.hbm_budget_bytes)
        {
            rc = REEF_E_NOMEM;
            err = "tank larger than the HBM budget";
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            if(rc)
            {
                r->code = rc;
                r->err = err;
            }
            else
            {
                r->built = std::move(b);
                r->bytes = exact;
            }
            r->ready = true;
        }
        cv_work.notify_all();
    }
}

void reef_server::run()
{
    (void)hipSetDevice(device);   // the device of the creating thread
    std::vector<RequestPtr> admitted;
    for(;;)
    {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || (!queue.empty() && queue.front()->ready) || !live.empty(); });
            if(stop) break;
            // admission in arrival order by the HBM budget and the live limit; a request
            // still being built holds back the ones behind it
            size_t in_use = e->pool.in_use();
            while(!queue.empty() && queue.front()->ready)
            {
                RequestPtr& r = queue.front();
                if(r->code)
                {
                    r->state = RQ_FAILED;
                    queue.pop_front();
                    cv_done.notify_all();
                    continue;
                }
                if(cfg.max_live > 0 && (int)(live.size() + admitted.size()) >= cfg.max_live) break;
                if(cfg.hbm_budget_bytes > 0.0 && (double)(in_use + r->bytes) > cfg.hbm_budget_bytes) break;
                in_use += r->bytes;
                admitted.push_back(r);
                queue.pop_front();
            }
        }
        const int rc = round(admitted);
        if(rc)
        {
            std::lock_guard<std::mutex> lk(mu);
            fatal = rc;
            fatal_msg = e->err;
            break;
        }
    }
    // stopped or failed: whatever is not done fails
    std::lock_guard<std::mutex> lk(mu);
    for(auto& kv : all)
        if(kv.second->state == RQ_QUEUED || kv.second->state == RQ_LIVE)
        {
            kv.second->state = RQ_FAILED;
            kv.second->code = fatal ? fatal : REEF_E_ARG;
            kv.second->err = fatal ? fatal_msg : "server stopped";
        }
    queue.clear();
    to_build.clear();
    cv_done.notify_all();
}

extern "C" {

const char* reef_server_last_error(void) { return s_err.c_str(); }

int reef_server_create(const reef_server_config* cfg, reef_server** out)
{
    if(!cfg || !out) return sfail(REEF_E_ARG, "null argument");
    *out = nullptr;
    if(cfg->hbm_budget_bytes < 0.0 || cfg->max_live < 0 || (cfg->nplans > 0 && !cfg->plans))
        return sfail(REEF_E_ARG, "bad configuration");
    if(cfg->precision != REEF_SOLVER_FP64 && cfg->precision != REEF_SOLVER_MIXED)
        return sfail(REEF_E_ARG, "bad precision");
    if(cfg->precision == REEF_SOLVER_MIXED && !(cfg->delta > 0.0 && cfg->delta < 1.0))
        return sfail(REEF_E_ARG, "delta must lie in (0, 1)");
    if(cfg->F49 != 0 && cfg->F49 != 1) return sfail(REEF_E_UNSUPPORTED, "F49 must be 0 or 1");
    std::unique_ptr<reef_server> s(new reef_server());
    s->cfg = *cfg;
    if(cfg->nplans > 0) s->plans.assign(cfg->plans, cfg->plans + 5 * cfg->nplans);
    s->cfg.plans = nullptr;
    hipError_t he = hipGetDevice(&s->device);
    if(he == hipSuccess) he = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if(he != hipSuccess) return sfail(REEF_E_HIP, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(he));
    s->e.reset(new Engine(s->stream));
    Engine& E = *s->e;
    E.pool.budget = 0;   // admission keeps the budget; the pool serves what was admitted
    E.fsf = cfg->F49 == 0 ? 1 : 0;
    E.precision = cfg->precision;
    if(cfg->precision == REEF_SOLVER_MIXED) E.delta = cfg->delta;
    E.graphs = cfg->graphs != 0;
    for(int q = 0; q < cfg->nplans; ++q)
    {
        const int* p = s->plans.data() + 5 * q;
        Engine::Plan pl;
        pl.ty = p[2];
        pl.tz = p[3];
        pl.cpt = p[4];
        if(E.set_plan(p[0], p[1], pl))
        {
            const std::string m = E.err;
            s->e.reset();
            (void)hipStreamDestroy(s->stream);
            return sfail(REEF_E_ARG, "plan entry " + std::to_string(q) + ": " + m);
        }
    }
    s->t0 = std::chrono::steady_clock::now();
    reef_server* raw = s.release();
    raw->worker = std::thread([raw] { raw->run(); });
    const int nb = (int)std::max(1u, std::min(4u, std::thread::hardware_concurrency()));
    for(int q = 0; q < nb; ++q) raw->builders.emplace_back([raw] { raw->build_loop(); });
    *out = raw;
    return 0;
}

void reef_server_destroy(reef_server* s)
{
    if(!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_build.notify_all();
    if(s->worker.joinable()) s->worker.join();
    for(std::thread& b : s->builders)
        if(b.joinable()) b.join();
    s->e.reset();
    if(s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int reef_server_submit(reef_server* s, const reef_grid_desc* grid, double dt, int steps, const double* phi,
                       const double* u, const double* v, const double* w, const double* press, long* id)
{
    if(!s || !grid || !grid->XN || !grid->YN || !grid->ZN || !phi || !u || !v || !w || !press || !id)
        return sfail(REEF_E_ARG, "null argument");
    if(steps < 1 || !(dt > 0.0)) return sfail(REEF_E_ARG, "steps must be >= 1 and dt > 0");
    // the field arrays alone (a lower bound, cheap); the builder replaces it by the exact
    // bytes, ghost programs included, before the request can be admitted
    const size_t bytes = Engine::tank_bytes_min(*grid);
    if(bytes == 0) return sfail(REEF_E_UNSUPPORTED, "grid outside the implemented subset");
    if(s->cfg.hbm_budget_bytes > 0.0 && (double)bytes > s->cfg.hbm_budget_bytes)
        return sfail(REEF_E_NOMEM, "tank larger than the HBM budget");
    RequestPtr r = std::make_shared<Request>();
    r->gd = *grid;
    r->XN.assign(grid->XN, grid->XN + grid->knox + 11);
    r->YN.assign(grid->YN, grid->YN + grid->knoy + 11);
    r->ZN.assign(grid->ZN, grid->ZN + grid->knoz + 11);
    r->gd.XN = r->XN.data();
    r->gd.YN = r->YN.data();
    r->gd.ZN = r->ZN.data();
    r->dt = dt;
    r->steps = steps;
    r->bytes = bytes;
    const size_t n = (size_t)(grid->knox + 6) * (grid->knoy + 6) * (grid->knoz + 6);
    const double* src[5] = {phi, u, v, w, press};
    for(int f = 0; f < 5; ++f) r->init[f].assign(src[f], src[f] + n);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if(s->stop || s->fatal) return sfail(s->fatal ? s->fatal : REEF_E_ARG, "server is not running: " + s->fatal_msg);
        r->id = s->next_id++;
        r->t_submit = s->now();
        s->all[r->id] = r;
        s->queue.push_back(r);
        s->to_build.push_back(r);
        *id = r->id;
    }
    s->cv_build.notify_one();
    return 0;
}

int reef_server_wait(reef_server* s, long id, double timeout_s, int* state)
{
    if(!s) return sfail(REEF_E_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if(it == s->all.end()) return sfail(REEF_E_ARG, "no such request");
    RequestPtr r = it->second;
    auto ended = [&] { return r->state == RQ_DONE || r->state == RQ_FAILED; };
    if(timeout_s < 0.0)
        s->cv_done.wait(lk, ended);
    else
        s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ended);
    if(state) *state = r->state;
    if(r->state == RQ_FAILED) return sfail(r->code ? r->code : REEF_E_ARG, r->err);
    return 0;
}

int reef_server_result(reef_server* s, long id, int field, double* host)
{
    if(!s || !host) return sfail(REEF_E_ARG, "null argument");
    if(field < 0 || field >= NOUT) return sfail(REEF_E_ARG, "field id must be REEF_U .. REEF_VISC");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if(it == s->all.end()) return sfail(REEF_E_ARG, "no such request");
    const Request& r = *it->second;
    if(r.state != RQ_DONE) return sfail(REEF_E_ARG, "request not done");
    std::memcpy(host, r.out[field].data(), r.out[field].size() * sizeof(double));
    return 0;
}

int reef_server_request_info(reef_server* s, long id, double* times, int* steps_done, long* iters)
{
    if(!s) return sfail(REEF_E_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if(it == s->all.end()) return sfail(REEF_E_ARG, "no such request");
    const Request& r = *it->second;
    if(times)
    {
        times[0] = r.t_submit;
        times[1] = r.t_admit;
        times[2] = r.t_done;
    }
    if(steps_done) *steps_done = r.steps_done;
    if(iters) *iters = r.iters;
    return 0;
}

int reef_server_release(reef_server* s, long id)
{
    if(!s) return sfail(REEF_E_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if(it == s->all.end()) return sfail(REEF_E_ARG, "no such request");
    if(it->second->state != RQ_DONE && it->second->state != RQ_FAILED)
        return sfail(REEF_E_ARG, "request still queued or live");
    s->all.erase(it);
    return 0;
}

int reef_server_stats(reef_server* s, int* queued, int* live, long* steps, double* bytes_in_use)
{
    if(!s) return sfail(REEF_E_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if(queued) *queued = (int)s->queue.size();
    if(live)
    {
        int n = 0;
        for(auto& kv : s->all) n += kv.second->state == RQ_LIVE;
        *live = n;
    }
    if(steps) *steps = s->steps;
    if(bytes_in_use) *bytes_in_use = (double)s->bytes_in_use;
    if(s->fatal) return sfail(s->fatal, s->fatal_msg);
    return 0;
}

int reef_server_peak_bytes(reef_server* s, double* peak)
{
    if(!s || !peak) return sfail(REEF_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *peak = (double)s->peak_bytes;
    return 0;
}

}  // extern "C"
