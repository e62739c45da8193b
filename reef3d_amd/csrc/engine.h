// Host side of the batched engine: tanks (box domains) live in a device memory pool, the
// engine keeps the device tank table and work-item table, plans the launches and runs
// the reference's operations (ghost fills, convection, reinit, fluid update, projection
// and BiCGStab, momentum_FC3 and level-set steps) on all active tanks at once.  The
// single-domain C-ABI context is an engine with one tank.
#pragma once

#include "batch.h"

#include <array>
#include <map>
#include <string>
#include <vector>

namespace reef {

// first-fit suballocator over large device chunks (256-byte granules)
class DevicePool
{
public:
    ~DevicePool();
    // returns 0 or a REEF_E status; budget (bytes, 0 = none) caps the bytes in use
    int alloc(size_t bytes, void** out);
    void free(void* p);
    size_t in_use() const { return used_; }
    size_t budget = 0;
    size_t chunk_bytes = (size_t)256 << 20;

private:
    struct Chunk
    {
        char* base;
        size_t size;
        std::map<size_t, size_t> free_;   // offset -> length
    };
    std::vector<Chunk> chunks_;
    std::map<void*, std::pair<size_t, size_t>> live_;   // ptr -> (chunk, length)
    size_t used_ = 0;
};

struct TankHost
{
    bool used = false;
    HostGrid hg;
    reef_params prm{};
    TankDev dev{};                  // host copy of the table entry
    std::vector<void*> mem;         // pool blocks
    size_t bytes = 0;               // device bytes of the tank
    std::vector<double> staging;
};

class Engine
{
public:
    explicit Engine(hipStream_t s) : stream(s) {}
    ~Engine();

    hipStream_t stream;
    DevicePool pool;
    std::vector<TankHost> tanks;    // slots
    int fsf = 0;                    // reinidisc_fsf (F49 = 0) for every tank
    int precision = REEF_SOLVER_FP64;
    double delta = 0.1;             // reliable-update threshold of the mixed solve
    // tile plans per size bucket (jdir, floor(log2(nx ny nz))): tile rows ty x tz and
    // chunk length cpt; a bucket without an entry uses the default plan.  A tank takes the
    // plan of its bucket when it is added (and when its bucket's plan is set), so its
    // chunks, and with them its reduction order, are the same alone and in any batch that
    // shares the plan table.
    struct Plan
    {
        int ty = 0, tz = 0, cpt = 0;
    };
    using Bucket = std::pair<int, int>;
    std::map<Bucket, Plan> plans;
    static Bucket bucket_of(const HostGrid& h);
    TileShape shape_of(const HostGrid& h) const;
    int set_plan(int jdir, int lg, const Plan& p);
    // times every candidate plan of each bucket of the active tanks that has none and
    // keeps the fastest; leaves the tanks' fields and results unchanged
    struct TuneRecord
    {
        Bucket bucket;
        Plan plan;
        double score_ms;
    };
    int autotune(int reps, std::vector<TuneRecord>* log);
    // replay fc3_step's launch sequences (the four segments between its pressure solves)
    // and the BiCGStab iteration batches as hipGraphs, captured on first use; the graphs
    // are dropped whenever the tank tables or the controls change.  Launch-by-launch while
    // timing is on.
    bool graphs = false;
    void drop_graphs();

    // ---- tanks
    static size_t tank_bytes(const reef_grid_desc& g);
    int add_tank(const reef_grid_desc& g, const reef_params& p, int* id);
    int remove_tank(int id);
    int active_count() const;
    TankHost* tank(int id) { return id >= 0 && id < (int)tanks.size() && tanks[id].used ? &tanks[id] : nullptr; }
    // marks a subset active (others skipped by every launch); empty = all used tanks
    int set_active(const std::vector<int>& ids);

    // device tables up to date (after tanks were added, removed or (de)activated)
    int prepare() { return sync_tables(); }

    // ---- host <-> device fields of one tank (REEF_LAYOUT_IFAST / REEF_LAYOUT_REF)
    int put(int id, int field, const double* host, int layout);
    int get(int id, int field, double* host, int layout);
    int zero(int field);

    // ---- operations on every active tank
    int ghost(int field, int kind, int gcv);
    int convection(int ipol, int q, int u, int v, int w);
    int reini(int field, int iters);
    int fluid_update();
    int solve(int x, int rhs, double stop, int maxiter, bool mixed);
    int residual(int x, int rhs, std::vector<double>& norms);
    int pjm(int u, int v, int w, double alpha);
    int fc3_step();
    int phi_rk3();
    int vortex_step(double simtime);

    // results of the last operation per tank slot
    std::vector<BiState> last_state;       // solver state after the last solve
    std::vector<long> step_iters;          // summed iterations of the last fc3 step
    std::vector<double> vol;               // volume1, volume2 per slot (2 each)
    int read_volumes();

    // ---- timing of launch families (HIP events around each launch on the stream)
    bool timing = false;
    void tbegin(const char* name);
    void tend();
    void tcollect();
    std::map<std::string, std::pair<double, long>> totals;

    std::string err;

private:
    struct TimedRange
    {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<TimedRange> ranges_;
    std::vector<hipEvent_t> evpool_;
    hipEvent_t take();

    // device tables
    TankDev* d_tanks_ = nullptr;
    int cap_tanks_ = 0;
    int* d_items_ = nullptr;
    int cap_items_ = 0;
    double* d_partials_ = nullptr;
    double* d_norms_ = nullptr;
    BiState* h_status_ = nullptr;          // pinned, 2 slots x tanks
    int cap_status_ = 0;
    hipEvent_t st_ev_[2] = {nullptr, nullptr};
    bool dirty_ = true;
    BatchView view_{};
    int nitems_ = 0;
    int grid_solver_ = 8, grid_stencil_ = 8, grid_simple_ = 8;

    int sync_tables();
    int fail(int code, const std::string& m)
    {
        err = m;
        return code;
    }
    using GraphKey = std::array<int, 4>;
    std::map<GraphKey, hipGraphExec_t> graphs_;
    template<class Body>
    int replay(const GraphKey& key, Body body);
    template<class Launch>
    int run_iterations(const GraphKey& key, int nphase, const char* const* names, Launch launch, int maxiter);
    const TankHost* first_used() const;
    int solve_begin(int x, int rhs, double stop, int maxiter, bool mixed);
    int solve_iterate(int x, int rhs, int maxiter, bool mixed);
    int solve_end(int x, bool mixed);
    int pjm_pre(int u, int v, int w, double alpha, bool mixed);
    int pjm_post(int u, int v, int w, double alpha, bool mixed);
    int fc3_segment(int s, bool mixed);
    int probe(double& stencil_ms, double& solver_ms);
    int momentum_stage(int stage, const int qin[3], const int qold[3], const int vel[3], const int out[3]);
};

// default tile shape and chunk length of a tank
TileShape plan_shape(const HostGrid& h, int cpt_override);

}  // namespace reef
