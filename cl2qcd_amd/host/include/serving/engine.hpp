// Inversion serving on one MI355X: gauge configurations stay resident in HBM, a
// FIFO queue of inversion requests is admitted against an HBM budget, requests
// that share an operator (configuration, action and mass parameters) are merged
// into one multi-right-hand-side solve, and a pool of workers -- each with its own
// HIP stream -- runs admitted batches concurrently, so requests arriving while
// others run are started as soon as a worker and memory are free (continuous
// batching).  No counterpart in the reference, whose inverter executable solves
// one source after the other (physics/algorithms/inversion.cpp:78-242).
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../physics/algorithms/solvers/solvers.hpp"

namespace serving {

enum class Action { twistedmass = 0, staggered = 1 };

struct Request {
    int config = 0;  // id returned by Engine::load_config
    Action action = Action::twistedmass;
    double kappa = 0.16, mu = 0.005;  // twisted mass: QplusQminus_eo with mubar = 2 kappa mu
    double mass  = 0.01;              // staggered: MdagM_eo
    // right-hand sides as host eo fields in the reference's AoS order: Veo x 24 doubles (Wilson) or
    // Veo x 6 doubles (staggered)
    std::vector<std::vector<double>> sources;
    // or, when sources is empty, one Gaussian right-hand side per seed, generated on the device
    std::vector<uint64_t> gaussian_seeds;
    double rel_tol        = 1e-10;
    bool return_solutions = true;  // false: only residuals and timings come back
    int nrhs() const { return sources.empty() ? (int)gaussian_seeds.size() : (int)sources.size(); }
};

struct Result {
    long id = -1;
    bool ok = false;
    std::string error;
    std::vector<std::vector<double>> solutions;  // same layout as the sources
    int width = 0;                               // doubles per eo site: 24 (Wilson) or 6 (staggered)
    std::vector<double> true_rel_residual;
    int iterations = 0;           // inner iterations of the solve that carried this request
    int batch_requests = 0;       // requests merged into that solve
    int batch_rhs      = 0;       // right-hand sides in that solve
    bool joined_running = false;  // taken into free slots of a lockstep solve that was already running
    long solve = -1;              // sequence number of the solve that carried this request
    double arrival = 0., admitted = 0., finished = 0.;  // seconds on the engine clock
};

/* scheduler counters since the engine started */
struct EngineStats {
    long solves_started      = 0;  // solves a worker started (each may carry several requests)
    long requests_merged     = 0;  // requests merged into another request's solve when it started
    long requests_joined     = 0;  // requests taken into free slots of a running lockstep solve
    long max_queue_depth     = 0;  // most requests waiting at once
    long queue_depth         = 0;  // requests waiting now
};

struct EngineOptions {
    size_t hbm_budget_bytes = (size_t)64 << 30;  // device memory for solver work fields
    int workers             = 4;                 // concurrent solves, one HIP stream each
    int inner_precision     = CL2_F32;           // inner loop of the mixed Wilson solvers (CL2_F32 or CL2_F16;
                                                 // staggered solves run FP32 inner)
    int max_batch_rhs       = 12;                // right-hand sides merged into one Wilson solve
    bool mixed              = true;              // false: the reference's FP64 solvers, one request at a time
    // twisted-mass lockstep solves take queued compatible requests into slots that are free (padding, or
    // converged) at each reliable update; such requests wait for those slots instead of starting a new solve
    bool join_running = true;
};

class Engine {
  public:
    explicit Engine(const EngineOptions& options);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    /* heatbath configuration of ns^3 x nt at beta (cold start, sweeps x (heatbath + overrelaxation),
     * antiperiodic fermion boundary in time), resident until the engine is destroyed; returns its id */
    int load_config(int ns, int nt, double beta, int sweeps, uint64_t seed);
    /* the configuration of an ILDG/LIME file of an ns^3 x nt lattice, resident until the engine is
     * destroyed; returns its id */
    int load_config_file(int ns, int nt, const std::string& path);
    /* checkpoint: resident configuration `config` to an ILDG/LIME file (with its plaquette and checksum) */
    void save_config(int config, const std::string& path) const;
    double plaquette(int config) const;
    long eo_sites(int config) const;
    size_t resident_bytes() const;

    /* queue a request (arrival time = now); returns its id */
    long submit(Request r);
    /* block until request id has finished */
    Result wait(long id);
    /* block until every submitted request has finished */
    void drain();
    double now() const;
    /* create every worker's device context and operators for every loaded lattice size now, so the
     * first requests do not pay for them */
    void warmup();
    /* largest HBM use of work fields seen so far */
    size_t peak_work_bytes() const;
    EngineStats stats() const;
    /* restart max_queue_depth from the current depth (for measurements over a window) */
    void reset_peaks();

  private:
    struct Config;
    struct Pending;
    struct Batch;
    struct Worker;
    struct OpenBlock;
    int add_config(std::unique_ptr<Config> c);
    bool claimable_locked(const Pending& p, std::map<const OpenBlock*, int>& claimed) const;
    void publish_locked(Pending& p, Result r);
    void run_tm_block(Worker& wk, Batch& b, Config& cfg);
    size_t work_bytes(const Pending& p) const;
    bool compatible(const Pending& a, const Pending& b) const;
    void schedule_locked();
    void worker_loop(int w);
    void run_batch(Worker& wk, Batch& b);

    EngineOptions opt;
    std::vector<std::unique_ptr<Config>> configs;
    mutable std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::unique_ptr<Pending>> queue;
    std::map<long, Result> done;
    std::vector<std::unique_ptr<Worker>> workers;
    std::vector<OpenBlock*> open_blocks;  // running lockstep solves that take refills
    EngineStats counters;
    std::vector<std::thread> threads;
    long next_id       = 0;
    long outstanding   = 0;
    size_t used_bytes  = 0, peak_bytes = 0;
    bool stopping      = false;
    double t0          = 0.;
};

}  // namespace serving
