// Host side of the batched engine: tanks (box domains) live in a device memory pool, the
// engine keeps the device tank table, builds the tank sets its launches cover and runs
// the reference's operations (ghost fills, convection, reinit, fluid update, projection
// and BiCGStab, momentum_FC3 and level-set steps) on the active tanks.  The single-domain
// C-ABI context is an engine with one tank.
//
// Two ways to advance momentum_FC3 steps:
//   fc3_step(): every active tank by one step in lockstep (each pressure solve runs until
//     the slowest tank has converged); its launch sequences can be replayed as hipGraphs;
//   start_steps() + pump(): iteration-level scheduling.  Each tank moves through its own
//     steps: a round launches the step segment each tank has reached (tanks grouped by
//     segment) and one batch of BiCGStab iterations for the tanks that are solving, so a
//     tank whose solve has converged goes on without waiting for the others.
// A tank's arithmetic is the same either way, alone or with any other tanks.
#pragma once

#include "batch.h"

#include <array>
#include <deque>
#include <map>
#include <string>
#include <vector>

namespace reef {

// the three iterations of the solver kernels
enum SolveKind
{
    SOLVE_FP64 = 0,    // bicgstab_ijk on the Poisson operator (face densities)
    SOLVE_MIXED = 1,   // FP32 iteration with FP64 reliable updates, Poisson operator
    SOLVE_MAT = 2      // bicgstab_ijk on an explicit 7-diagonal operator (TankDev::m)
};

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
    bool active = true;             // covered by the launches of the whole batch
    HostGrid hg;
    reef_params prm{};
    TankDev dev{};                  // host copy of the table entry
    std::vector<void*> mem;         // pool blocks
    size_t bytes = 0;               // device bytes of the tank
    std::vector<double> staging;
    // iteration-level scheduling (start_steps / pump)
    int pc = -1;                    // -1 idle; 0, 2, 4, 6: segment pc / 2 next; 1, 3, 5: solving
    int seq = 0;                    // solves started
    int steps_left = 0;
    int steps_done = 0;
    long iters_step = 0;            // iterations of the step in progress
    long iters_last = 0;            // iterations of the last completed step
};

// host half of a tank (grid metadata, ghost programs and the tables its device block
// holds); Engine::build_tank makes it without touching the engine, on any thread
struct TankBuild
{
    HostGrid hg;
    reef_params prm{};
    std::vector<GhostOp> ops;           // the NPROG ghost programs, concatenated
    int op0[NPROG] = {}, nop[NPROG] = {};
    std::vector<double> meta;           // P, DN, DP and WENO tables per axis
    std::vector<float> inv;             // FP32 1/DN per axis
    size_t offP[3] = {}, offDN[3] = {}, offDP[3] = {}, offqf[3] = {}, offcf[3] = {}, offisf[3] = {}, offinv[3] = {};
    size_t offqfT = 0, offcfT = 0, offisfT = 0;   // the x tables entry-major (DevGrid::qfT)
    int nT = 0;
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
    // N48 = 1: every momentum_FC3 step starts with the adaptive time step (ietimestep::start
    // with the factor N47) computed on the device; 0: the fixed dt of each tank
    int n48 = 0;
    double n47 = 0.3;
    // D20 = 1: the momentum predictors add ediff2's explicit viscous terms; 2: idiff2_FS,
    // the implicit viscous solve of each component (stop N43); 0: diff_void.  Set with
    // set_d20 (D20 = 2 gives every tank its explicit-operator arrays).
    int d20 = 0;
    double n43 = 1.0e-5;
    int set_d20(int v);
    double delta = 0.1;             // reliable-update threshold of the mixed solve
    // replay fc3_step's launch sequences (the four segments between its pressure solves)
    // and the BiCGStab iteration batches as hipGraphs, captured on first use; the graphs
    // are dropped whenever the tank tables or the controls change.  Launch-by-launch while
    // timing is on.
    bool graphs = false;
    void drop_graphs();
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

    // ---- tanks
    // device bytes of one tank of this grid: exactly what add_built takes from the pool
    // (builds the tank's host half, ghost programs included); 0 if the grid is refused
    static size_t tank_bytes(const reef_grid_desc& g);
    // the same for a build already made
    static size_t built_bytes(const TankBuild& b);
    // a lower bound without the ghost programs (the field arrays only), cheap
    static size_t tank_bytes_min(const reef_grid_desc& g);
    int add_tank(const reef_grid_desc& g, const reef_params& p, int* id);
    // add_tank in two halves: the host build (thread-safe, no engine state) and the device
    // part; add_built leaves b usable again when it fails
    static int build_tank(const reef_grid_desc& g, const reef_params& p, TankBuild& b, std::string& err);
    int add_built(TankBuild& b, int* id);
    int remove_tank(int id);
    int active_count() const;
    TankHost* tank(int id) { return id >= 0 && id < (int)tanks.size() && tanks[id].used ? &tanks[id] : nullptr; }
    // marks a subset active (others skipped by every launch); empty = all used tanks
    int set_active(const std::vector<int>& ids);

    // device tables up to date (after tanks were added, removed or (de)activated)
    int prepare() { return sync_tables(); }

    // ---- host <-> device fields of one tank (REEF_LAYOUT_IFAST / REEF_LAYOUT_REF)
    int put(int id, int field, const double* host, int layout);
    // the same into any FP64 array of the tank's layout (e.g. TankDev::m)
    int put_array(int id, double* dev, const double* host, int layout);
    int get(int id, int field, double* host, int layout);
    int zero(int field);

    // ---- operations on every active tank
    int ghost(int field, int kind, int gcv);
    int convection(int ipol, int q, int u, int v, int w);
    // reini_RK3 pseudo steps with the ghost rule gcv of the level set (52 = gcval_phi for
    // F50 = 2; 50 = gcval_iniphi of the initial reinitialisation)
    int reini(int field, int iters, int gcv = 52);
    int fluid_update();
    // fluid_update of tank slot with the reduction memory in the allocation-free scratch
    // array (no device allocation); vol2 receives its volume1, volume2
    int fluid_update_noalloc(int slot, double* vol2);
    int solve(int x, int rhs, double stop, int maxiter, bool mixed);
    // the tank's explicit-operator arrays (TankDev::m), from the pool on first use
    int ensure_matrix(int id);
    // relaxation-zone wave generation and numerical beach on tank id (iowave, B98 = 2): the
    // static set-up from the pool; params5 (may be null) receives wT, wL, wk, ww, wA.  From
    // then on every fc3 step of the tank runs the driver loop's wave prologue and the
    // relaxations inside momentum_FC3, and advances the tank's simtime by its dt.
    int wave_set(int id, const reef_wave_desc& w, double* params5);
    // the simulation time of the active wave tanks (p->simtime)
    int set_simtime(double t);
    // wavegen_precalc at the tanks' simtime, then the relaxations of iowave on the given
    // fields of every active wave tank (u >= 0: u, v, w; p >= 0: pressure; f >= 0: phi_relax)
    int wave_relax(int u, int v, int w, int p, int f);
    // solver::start on the explicit operators of the active tanks (bicgstab_ijk with var
    // 1..5), the solution in place in array x
    int mat_solve(int var, int x, double stop, int maxiter);
    int residual(int x, int rhs, std::vector<double>& norms);
    int pjm(int u, int v, int w, double alpha);
    int fc3_step();
    // ietimestep::start (ini = 0) or ::ini (ini = 1): the tanks' next dt on the device; out
    // (optional) receives dt, umax, vmax, wmax per slot
    int timestep(double N47, int ini, std::vector<double>* out);
    // the dt of the active tanks' following steps
    int set_dt(double dt);
    int phi_rk3();
    int vortex_step(double simtime);
    // the ghost fills and fluid update a run starts with, on the given tanks only
    int prepare_tanks(const std::vector<int>& ids);

    // ---- iteration-level scheduling of momentum_FC3 steps
    int start_steps(int id, int n);
    // one round; stepped receives the slots that completed a step in it
    int pump(std::vector<int>& stepped);
    bool busy() const;
    // pump() until no tank has steps left
    int run_scheduled();

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
    double* d_partials_ = nullptr;
    size_t cap_partials_ = 0;
    double* d_norms_ = nullptr;            // per slot
    BiState* d_status_ = nullptr;          // gathered states, 2 ring slots x cap_tanks_
    BiState* h_status_ = nullptr;          // pinned mirror
    hipEvent_t st_ev_[2] = {nullptr, nullptr};
    long batch_no_ = 0;
    bool dirty_ = true;
    bool any_wave_ = false;                // some tank has a wave generator
    std::vector<TankSet> all_;             // the active tanks (chunks of SET_MAX)
    const std::vector<TankSet>* cur_ = &all_;

    // iteration batches in flight (pump)
    struct Inflight
    {
        int ring;
        std::vector<std::pair<int, int>> members;   // (slot, seq)
    };
    std::deque<Inflight> inflight_;
    std::vector<int> solving_key_;
    std::vector<TankSet> solving_sets_;

    int sync_tables();
    std::vector<TankSet> make_sets(const std::vector<int>& slots) const;
    int fail(int code, const std::string& m)
    {
        err = m;
        return code;
    }
    template<class F>
    void each(F f)
    {
        for(const TankSet& S : *cur_) f(S);
    }
    using GraphKey = std::array<int, 4>;
    std::map<GraphKey, hipGraphExec_t> graphs_;
    // the scheduler's iteration batches, keyed by the solving slots and the batch's
    // arguments: captured when a key comes back (the tanks solving together at the end of
    // a stream repeat their batch many times), launched directly the first time
    struct SetGraph
    {
        int seen = 0;
        hipGraphExec_t exec = nullptr;
    };
    std::map<std::vector<int>, SetGraph> set_graphs_;
    static constexpr size_t SET_GRAPHS_MAX = 4096;   // solving sets with a graph entry
    template<class Body>
    int capture(Body body, hipGraphExec_t* out);
    template<class Body>
    int replay(const GraphKey& key, Body body);
    int enqueue_batch(int kind, int x, int rhs, int ring, int iters);
    int drain();
    int read_batch(const Inflight& b);
    int run_iterations(int kind, int x, int rhs, int maxiter);
    const TankHost* first_used() const;
    // the parameters every tank of the engine shares (all of reef_params except dt and
    // F45): those of the first tank added to an empty engine; add_built refuses a tank
    // whose values differ.  Null while the engine holds no tank.
    reef_params shared_{};
    bool has_shared_ = false;
    const reef_params* shared() const { return has_shared_ ? &shared_ : nullptr; }
    int solve_begin(int x, int rhs, double stop, int maxiter, bool mixed);
    int solve_iterate(int x, int rhs, int maxiter, bool mixed);
    int solve_end(int x, bool mixed);
    int pjm_pre(int u, int v, int w, double alpha, bool mixed);
    int pjm_post(int u, int v, int w, double alpha, bool mixed);
    int fc3_segment(int s, bool mixed);
    void wave_phi(int f);
    int probe(double& stencil_ms, double& solver_ms);
    int momentum_stage(int stage, const int qin[3], const int qold[3], const int vel[3], const int out[3]);
};

// default tile shape and chunk length of a tank
TileShape plan_shape(const HostGrid& h, int cpt_override);

}  // namespace reef
