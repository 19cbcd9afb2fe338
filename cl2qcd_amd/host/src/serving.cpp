// Inversion serving engine: resident configurations, FIFO admission against an
// HBM budget, merging of requests that share an operator into one multi-RHS solve,
// and a worker pool with one HIP stream per worker so admitted solves overlap.
#include "../include/serving/engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>

#include "../include/physics/algorithms/su3heatbath.hpp"
#include "../include/physics/observables/gaugeObservables.hpp"

namespace serving {

using namespace physics;
using namespace physics::lattices;
namespace fm = physics::fermionmatrix;
namespace sv = physics::algorithms::solvers;

namespace {

std::vector<std::string> lattice_args(int ns, int nt, const char* action)
{
    return {"--nSpace=" + std::to_string(ns), "--nTime=" + std::to_string(nt), std::string("--fermionAction=") + action,
            "--thetaFermionTemporal=1", "--solverMaxIterations=1000000"};
}

/* right-hand-side counts with a compiled multi-RHS operator */
int padded_rhs(int n)
{
    for (int s : {1, 2, 4, 8, 12, 16})
        if (n <= s) return s;
    return -1;
}

}  // namespace

/* one resident configuration and the objects that own it */
struct Engine::Config {
    int ns, nt;
    std::unique_ptr<meta::Inputparameters> params;
    std::unique_ptr<hardware::System> system;
    std::unique_ptr<PrngParametersImplementation> prng_params;
    std::unique_ptr<PRNG> prng;
    std::unique_ptr<InterfacesHandlerImplementation> handler;
    std::unique_ptr<Gaugefield> gf;
    double plaquette = 0.;
    size_t bytes     = 0;
};

struct Engine::Pending {
    long id;
    Request req;
    double arrival;
    double admitted  = 0.;     // when a solve took it
    long solve       = -1;     // which solve
    bool joined      = false;  // taken into a running lockstep solve
    bool published   = false;  // its result is in `done`
    int nrhs() const { return req.nrhs(); }
};

/* a running twisted-mass lockstep solve that takes refills: its configuration, tolerance and free slots */
struct Engine::OpenBlock {
    int config;
    double rel_tol;
    int free;
};

struct Engine::Batch {
    std::vector<std::unique_ptr<Pending>> items;
    size_t bytes   = 0;
    double admitted = 0.;
    long solve      = -1;  // sequence number
};

/* a worker's device context and operators for one lattice size */
struct WorkerLattice {
    std::unique_ptr<meta::Inputparameters> p_tm, p_st;
    std::unique_ptr<hardware::System> system;
    std::unique_ptr<InterfacesHandlerImplementation> h_tm, h_st;
    std::unique_ptr<fm::QplusQminus_eo> qq;
    std::unique_ptr<fm::MdagM_eo> mdagm;
};

struct Engine::Worker {
    std::map<std::pair<int, int>, std::unique_ptr<WorkerLattice>> lattices;
    std::unique_ptr<Batch> batch;  // assigned, not yet finished
    bool busy = false;
    WorkerLattice& lattice(int ns, int nt)
    {
        auto& w = lattices[{ns, nt}];
        if (!w) {
            w.reset(new WorkerLattice);
            w->p_tm.reset(new meta::Inputparameters(lattice_args(ns, nt, "twistedmass")));
            w->p_st.reset(new meta::Inputparameters(lattice_args(ns, nt, "rooted_stagg")));
            w->system.reset(new hardware::System(*w->p_tm));
            w->h_tm.reset(new InterfacesHandlerImplementation(*w->p_tm));
            w->h_st.reset(new InterfacesHandlerImplementation(*w->p_st));
            w->qq.reset(new fm::QplusQminus_eo(*w->system, w->h_tm->getInterface<fm::QplusQminus_eo>()));
            w->mdagm.reset(new fm::MdagM_eo(*w->system, w->h_st->getInterface<fm::MdagM_eo>()));
        }
        return *w;
    }
};

Engine::Engine(const EngineOptions& options) : opt(options)
{
    if (opt.workers < 1 || opt.workers > 16) throw std::invalid_argument("serving: 1 to 16 workers");
    if (opt.max_batch_rhs < 1 || opt.max_batch_rhs > 16) throw std::invalid_argument("serving: max_batch_rhs in [1, 16]");
    if (opt.inner_precision != CL2_F32 && opt.inner_precision != CL2_F16)
        throw std::invalid_argument("serving: inner precision FP32 or FP16");
    t0 = 0.;
    t0 = now();
    for (int w = 0; w < opt.workers; ++w) workers.emplace_back(new Worker);
    for (int w = 0; w < opt.workers; ++w) threads.emplace_back(&Engine::worker_loop, this, w);
}

Engine::~Engine()
{
    {
        std::lock_guard<std::mutex> lk(mu);
        stopping = true;
    }
    cv_work.notify_all();
    for (auto& t : threads) t.join();
}

double Engine::now() const
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count() - t0;
}

int Engine::load_config(int ns, int nt, double beta, int sweeps, uint64_t seed)
{
    std::unique_ptr<Config> c(new Config);
    c->ns = ns;
    c->nt = nt;
    std::vector<std::string> a = lattice_args(ns, nt, "twistedmass");
    a.push_back("--beta=" + std::to_string(beta));
    a.push_back("--hostSeed=" + std::to_string(seed));
    c->params.reset(new meta::Inputparameters(a));
    c->system.reset(new hardware::System(*c->params));
    c->prng_params.reset(new PrngParametersImplementation(*c->params));
    c->prng.reset(new PRNG(*c->system, c->prng_params.get()));
    c->handler.reset(new InterfacesHandlerImplementation(*c->params));
    c->gf.reset(new Gaugefield(*c->system, &c->handler->getInterface<Gaugefield>(), *c->prng, false));
    for (int s = 0; s < sweeps; ++s) physics::algorithms::su3heatbath(*c->gf, *c->prng, 1);
    return add_config(std::move(c));
}

int Engine::load_config_file(int ns, int nt, const std::string& path)
{
    std::unique_ptr<Config> c(new Config);
    c->ns = ns;
    c->nt = nt;
    c->params.reset(new meta::Inputparameters(lattice_args(ns, nt, "twistedmass")));
    c->system.reset(new hardware::System(*c->params));
    c->prng_params.reset(new PrngParametersImplementation(*c->params));
    c->prng.reset(new PRNG(*c->system, c->prng_params.get()));
    c->handler.reset(new InterfacesHandlerImplementation(*c->params));
    c->gf.reset(new Gaugefield(*c->system, &c->handler->getInterface<Gaugefield>(), *c->prng, path));
    return add_config(std::move(c));
}

/* measures the plaquette, makes the reduced-precision link copies (so workers only read them) and makes
 * the configuration resident */
int Engine::add_config(std::unique_ptr<Config> c)
{
    c->plaquette = physics::observables::measurePlaquette(c->gf.get());
    c->gf->dev_f32();
    if (opt.inner_precision == CL2_F16) c->gf->dev_f16();
    c->system->synchronize();
    const size_t veo = (size_t)c->system->eo_sites();
    c->bytes = veo * 8 * (144 + 48 + (opt.inner_precision == CL2_F16 ? 24 : 0));
    std::lock_guard<std::mutex> lk(mu);
    configs.push_back(std::move(c));
    return (int)configs.size() - 1;
}

void Engine::save_config(int config, const std::string& path) const
{
    const Config* c = nullptr;
    {
        std::lock_guard<std::mutex> lk(mu);
        c = configs.at(config).get();
    }
    // workers only read a resident configuration, so it can be written while they run
    c->gf->save(path, 0);
}

double Engine::plaquette(int config) const
{
    std::lock_guard<std::mutex> lk(mu);
    return configs.at(config)->plaquette;
}
long Engine::eo_sites(int config) const
{
    std::lock_guard<std::mutex> lk(mu);
    return configs.at(config)->system->eo_sites();
}
size_t Engine::resident_bytes() const
{
    std::lock_guard<std::mutex> lk(mu);
    size_t s = 0;
    for (const auto& c : configs) s += c->bytes;
    return s;
}
size_t Engine::peak_work_bytes() const
{
    std::lock_guard<std::mutex> lk(mu);
    return peak_bytes;
}
EngineStats Engine::stats() const
{
    std::lock_guard<std::mutex> lk(mu);
    EngineStats s = counters;
    s.queue_depth = (long)queue.size();
    return s;
}

void Engine::reset_peaks()
{
    std::lock_guard<std::mutex> lk(mu);
    counters.max_queue_depth = (long)queue.size();
}

/* whether an open lockstep solve will take p at its next reliable update (free slots not yet claimed by
 * requests ahead of p in this scheduling pass) */
bool Engine::claimable_locked(const Pending& p, std::map<const OpenBlock*, int>& claimed) const
{
    if (!opt.join_running || !opt.mixed || p.req.action != Action::twistedmass) return false;
    for (const OpenBlock* o : open_blocks) {
        if (o->config != p.req.config || o->rel_tol != p.req.rel_tol) continue;
        if (o->free - claimed[o] >= p.nrhs()) {
            claimed[o] += p.nrhs();
            return true;
        }
    }
    return false;
}

void Engine::publish_locked(Pending& p, Result r)
{
    r.id           = p.id;
    r.arrival      = p.arrival;
    r.admitted     = p.admitted;
    r.joined_running = p.joined;
    r.solve          = p.solve;
    done[p.id]     = std::move(r);
    p.published    = true;
    --outstanding;
}

/* device bytes a request's solve needs: the block solver's FP64 and inner-precision fields and
 * temporaries per (padded) right-hand side, plus for staggered each right-hand side's own source and
 * solution.  Called with mu held. */
size_t Engine::work_bytes(const Pending& p) const
{
    const size_t veo = (size_t)configs.at(p.req.config)->system->eo_sites();
    if (p.req.action == Action::twistedmass) {
        const int n = std::max(1, padded_rhs(p.nrhs()));
        return veo * (size_t)n * (6 * 192 + 6 * 96 + 2 * 192);
    }
    return veo * (size_t)p.nrhs() * (4 * 48 + 4 * 24 + 2 * 48 + 2 * 24 + 2 * 48);
}

bool Engine::compatible(const Pending& a, const Pending& b) const
{
    // requests of one action on one configuration: the twisted-mass block operator takes kappa and mu per
    // right-hand side, the staggered one a shift per right-hand side (m_r^2 - m^2 on MdagM_eo of mass m)
    return opt.mixed && a.req.action == b.req.action && a.req.config == b.req.config && a.req.rel_tol == b.req.rel_tol;
}

long Engine::submit(Request r)
{
    if (r.nrhs() < 1 || r.nrhs() > 16) throw std::invalid_argument("serving: 1 to 16 right-hand sides");
    if (!r.sources.empty() && !r.gaussian_seeds.empty())
        throw std::invalid_argument("serving: give sources or Gaussian seeds, not both");
    std::unique_ptr<Pending> p(new Pending);
    p->arrival = now();
    std::lock_guard<std::mutex> lk(mu);
    if (r.config < 0 || r.config >= (int)configs.size()) throw std::invalid_argument("serving: unknown configuration");
    const size_t veo = (size_t)configs[r.config]->system->eo_sites();
    const size_t per = veo * (r.action == Action::twistedmass ? 24 : 6);
    for (const auto& s : r.sources)
        if (s.size() != per) throw std::invalid_argument("serving: source of the wrong size for its lattice");
    p->req = std::move(r);
    p->id = next_id++;
    const long id = p->id;
    queue.push_back(std::move(p));
    ++outstanding;
    counters.max_queue_depth = std::max(counters.max_queue_depth, (long)queue.size());
    schedule_locked();
    return id;
}

/* FIFO admission: the head of the queue starts when a worker is idle and its work fields fit the HBM
 * budget (or nothing is running, so an oversized request cannot wait forever); queued requests that
 * share its operator are merged into the same solve while right-hand sides allow and the merged solve's
 * work fields, next to those of the running solves, fit the budget.  Called with mu held. */
void Engine::schedule_locked()
{
    while (!queue.empty()) {
        int free_w = -1;
        bool any_busy = false;
        for (int w = 0; w < (int)workers.size(); ++w) {
            if (workers[w]->busy)
                any_busy = true;
            else if (free_w < 0)
                free_w = w;
        }
        if (free_w < 0) return;
        if (!opt.mixed && any_busy) return;  // the FP64 baseline runs one request at a time
        // the first queued request that no running lockstep solve will take into its free slots
        std::map<const OpenBlock*, int> claimed;
        auto head = queue.begin();
        while (head != queue.end() && claimable_locked(**head, claimed)) ++head;
        if (head == queue.end()) return;
        const size_t need = work_bytes(**head);
        if (used_bytes + need > opt.hbm_budget_bytes && any_busy) return;
        std::unique_ptr<Batch> b(new Batch);
        b->items.push_back(std::move(*head));
        queue.erase(head);
        int rhs = b->items[0]->nrhs();
        // memory of a merged solve of k right-hand sides (its padded count)
        auto merged_bytes = [&](int k) {
            Pending m;
            m.req.config = b->items[0]->req.config;
            m.req.action = b->items[0]->req.action;
            m.req.gaussian_seeds.resize(k);
            return work_bytes(m);
        };
        std::map<const OpenBlock*, int> claimed2;
        for (auto it = queue.begin(); it != queue.end();) {
            const int k = rhs + (*it)->nrhs();
            // a candidate joins only while the merged solve still fits the budget next to the running ones
            // (only a lone head may exceed it, and only with nothing running); requests a running lockstep
            // solve will take stay for it
            if (compatible(*b->items[0], **it) && k <= opt.max_batch_rhs &&
                used_bytes + merged_bytes(k) <= opt.hbm_budget_bytes && !claimable_locked(**it, claimed2)) {
                rhs = k;
                b->items.push_back(std::move(*it));
                it = queue.erase(it);
                ++counters.requests_merged;
            } else {
                ++it;
            }
        }
        b->bytes    = merged_bytes(rhs);
        b->admitted = now();
        b->solve    = counters.solves_started++;
        for (auto& it : b->items) {
            it->admitted = b->admitted;
            it->solve    = b->solve;
        }
        used_bytes += b->bytes;
        peak_bytes = std::max(peak_bytes, used_bytes);
        workers[free_w]->busy  = true;
        workers[free_w]->batch = std::move(b);
        cv_work.notify_all();
    }
}

void Engine::worker_loop(int w)
{
    Worker& wk = *workers[w];
    for (;;) {
        std::unique_ptr<Batch> b;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stopping || wk.batch != nullptr; });
            if (!wk.batch) return;  // stopping with nothing assigned
            b = std::move(wk.batch);
        }
        run_batch(wk, *b);
        {
            std::lock_guard<std::mutex> lk(mu);
            used_bytes -= b->bytes;
            wk.busy = false;
            schedule_locked();
        }
        cv_done.notify_all();
    }
}

void Engine::run_batch(Worker& wk, Batch& b)
{
    const Request& head = b.items[0]->req;
    Config* cfgp        = nullptr;
    {
        // configurations are owned by unique_ptr: the object stays put when the vector grows
        std::lock_guard<std::mutex> lk(mu);
        cfgp = configs[head.config].get();
    }
    Config& cfg = *cfgp;
    int total_rhs = 0;
    for (const auto& it : b.items) total_rhs += it->nrhs();
    if (opt.mixed && head.action == Action::twistedmass && padded_rhs(total_rhs) >= 2) {
        run_tm_block(wk, b, cfg);
        return;
    }
    std::vector<Result> res(b.items.size());
    try {
        WorkerLattice& L              = wk.lattice(cfg.ns, cfg.nt);
        const hardware::System& sys   = *L.system;
        const Gaugefield& gf          = *cfg.gf;
        int iterations                = 0;
        std::vector<std::vector<double>> sol;
        std::vector<double> resid;
        std::vector<bool> keep;  // right-hand sides whose solutions go back to the host
        for (const auto& it : b.items)
            for (int k = 0; k < it->nrhs(); ++k) keep.push_back(it->req.return_solutions);
        if (head.action == Action::twistedmass) {
            const int n = opt.mixed ? padded_rhs(total_rhs) : total_rhs;
            // each right-hand side with its own request's kappa and mu (padding right-hand sides: the last
            // request's)
            std::vector<std::unique_ptr<WilsonAdditionalParameters>> own;
            std::vector<const physics::AdditionalParameters*> aps;
            for (const auto& it : b.items) {
                own.emplace_back(new WilsonAdditionalParameters(it->req.kappa, 2. * it->req.kappa * it->req.mu));
                for (int j = 0; j < it->nrhs(); ++j) aps.push_back(own.back().get());
            }
            while ((int)aps.size() < n) aps.push_back(own.back().get());
            std::vector<std::unique_ptr<Spinorfield_eo>> bs, xs;
            for (int r = 0; r < n; ++r) {
                bs.emplace_back(new Spinorfield_eo(sys));
                xs.emplace_back(new Spinorfield_eo(sys));
                xs.back()->zero();
            }
            int r = 0;
            for (const auto& it : b.items) {
                for (const auto& s : it->req.sources) bs[r++]->import(reinterpret_cast<const spinor*>(s.data()));
                for (uint64_t seed : it->req.gaussian_seeds) bs[r++]->gaussian_from_seed(seed);
            }
            for (; r < n; ++r) bs[r]->zero();  // padding right-hand sides are inactive from the start
            if (!opt.mixed) {
                // the reference's CG stops on its iterated residual; where the recomputed true residual
                // misses rel_tol it is entered again from the current solution (at most twice more)
                const Spinorfield_eo t(sys);
                for (int k = 0; k < n; ++k) {
                    const hmc_float b2 = squarenorm(*bs[k]);
                    for (int round = 0; round < 3; ++round) {
                        iterations += sv::cg(xs[k].get(), *L.qq, gf, *bs[k], sys, *L.h_tm,
                                             head.rel_tol * head.rel_tol * b2, *aps[k]);
                        (*L.qq)(&t, gf, *xs[k], *aps[k]);
                        saxpy(&t, hmc_complex{1., 0.}, t, *bs[k]);
                        if (squarenorm(t) <= head.rel_tol * head.rel_tol * b2) break;
                    }
                }
            } else {  // a single right-hand side (blocks of 2 and more run in run_tm_block)
                sv::ReliableUpdateOptions o;
                o.inner_precision = opt.inner_precision;
                iterations = sv::cg_reliable_update(xs[0].get(), *L.qq, gf, *bs[0], sys, head.rel_tol, *aps[0], o);
            }
            // FP64 true residual of every answer
            const Spinorfield_eo t(sys);
            for (int k = 0; k < total_rhs; ++k) {
                (*L.qq)(&t, gf, *xs[k], *aps[k]);
                saxpy(&t, hmc_complex{1., 0.}, t, *bs[k]);  // b - A x
                resid.push_back(std::sqrt(squarenorm(t) / squarenorm(*bs[k])));
                if (keep[k]) {
                    const std::vector<spinor> h = xs[k]->get_host();
                    const double* d             = reinterpret_cast<const double*>(h.data());
                    sol.emplace_back(d, d + h.size() * 24);
                } else {
                    sol.emplace_back();
                }
            }
        } else if (opt.mixed && total_rhs > 1) {
            // every right-hand side of the batch in lockstep blocks over one pass of the links per D_KS_eo; a
            // request of mass m_r is the head's MdagM_eo shifted by m_r^2 - m^2
            const StaggeredAdditionalParameters ap(head.mass);
            std::vector<std::unique_ptr<StaggeredAdditionalParameters>> own;
            std::vector<const StaggeredAdditionalParameters*> aps;
            std::vector<hmc_float> shift;
            std::vector<std::unique_ptr<Staggeredfield_eo>> bs, xs;
            for (const auto& it : b.items) {
                own.emplace_back(new StaggeredAdditionalParameters(it->req.mass));
                for (int j = 0; j < it->nrhs(); ++j) {
                    aps.push_back(own.back().get());
                    shift.push_back(it->req.mass * it->req.mass - head.mass * head.mass);
                    bs.emplace_back(new Staggeredfield_eo(sys));
                    xs.emplace_back(new Staggeredfield_eo(sys));
                    if (it->req.sources.empty())
                        bs.back()->gaussian_from_seed(it->req.gaussian_seeds[j]);
                    else
                        bs.back()->import(reinterpret_cast<const su3vec*>(it->req.sources[j].data()));
                    xs.back()->set_zero();
                }
            }
            std::vector<const Staggeredfield_eo*> xv, bv;
            for (int k = 0; k < total_rhs; ++k) {
                xv.push_back(xs[k].get());
                bv.push_back(bs[k].get());
            }
            sv::ReliableUpdateOptions o;
            o.inner_precision = CL2_F32;
            iterations = sv::cg_reliable_update_stagg_block(xv, *L.mdagm, gf, shift, bv, sys, head.rel_tol, ap, o);
            const Staggeredfield_eo t(sys);
            for (int k = 0; k < total_rhs; ++k) {  // FP64 true residual with the request's own mass
                (*L.mdagm)(&t, gf, *xs[k], aps[k]);
                saxpby(&t, 1., *bs[k], -1., t);
                resid.push_back(std::sqrt(squarenorm(t) / squarenorm(*bs[k])));
                if (keep[k]) {
                    const std::vector<su3vec> h = xs[k]->get_host();
                    const double* d             = reinterpret_cast<const double*>(h.data());
                    sol.emplace_back(d, d + h.size() * 6);
                } else {
                    sol.emplace_back();
                }
            }
        } else {
            const StaggeredAdditionalParameters ap(head.mass);
            const Staggeredfield_eo bf(sys), xf(sys), t(sys);
            int k = 0;
            for (const auto& it : b.items)
                for (int j = 0; j < it->nrhs(); ++j, ++k) {
                    if (it->req.sources.empty())
                        bf.gaussian_from_seed(it->req.gaussian_seeds[j]);
                    else
                        bf.import(reinterpret_cast<const su3vec*>(it->req.sources[j].data()));
                    xf.set_zero();
                    if (opt.mixed) {
                        sv::ReliableUpdateOptions o;
                        o.inner_precision = CL2_F32;
                        iterations += sv::cg_reliable_update_stagg(&xf, *L.mdagm, gf, 0., bf, sys, head.rel_tol, ap, o);
                    } else {
                        // the reference's CG-M (one shift 0) from zero; where the true residual misses rel_tol,
                        // CG-M again on the residual and the correction added (at most twice more)
                        const Staggeredfield_eo d(sys);
                        const hmc_float b2 = squarenorm(bf);
                        for (int round = 0; round < 3; ++round) {
                            const Staggeredfield_eo& rhs = round == 0 ? bf : t;
                            const Staggeredfield_eo& sol = round == 0 ? xf : d;
                            std::vector<std::shared_ptr<Staggeredfield_eo>> xv{std::shared_ptr<Staggeredfield_eo>(
                                const_cast<Staggeredfield_eo*>(&sol), [](Staggeredfield_eo*) {})};
                            iterations += sv::cg_m(xv, *L.mdagm, gf, {0.}, rhs, sys, *L.h_st,
                                                   head.rel_tol * head.rel_tol * b2, ap);
                            if (round > 0) saxpby(&xf, 1., xf, 1., d);
                            (*L.mdagm)(&t, gf, xf, &ap);
                            saxpby(&t, 1., bf, -1., t);
                            if (squarenorm(t) <= head.rel_tol * head.rel_tol * b2) break;
                        }
                    }
                    (*L.mdagm)(&t, gf, xf, &ap);
                    saxpby(&t, 1., bf, -1., t);
                    resid.push_back(std::sqrt(squarenorm(t) / squarenorm(bf)));
                    if (keep[k]) {
                        const std::vector<su3vec> h = xf.get_host();
                        const double* d             = reinterpret_cast<const double*>(h.data());
                        sol.emplace_back(d, d + h.size() * 6);
                    } else {
                        sol.emplace_back();
                    }
                }
        }
        const double fin = now();
        int r            = 0;
        for (size_t i = 0; i < b.items.size(); ++i) {
            Result& o        = res[i];
            o.ok             = true;
            o.iterations     = iterations;
            o.batch_requests = (int)b.items.size();
            o.batch_rhs      = total_rhs;
            o.width          = head.action == Action::twistedmass ? 24 : 6;
            for (int k = 0; k < b.items[i]->nrhs(); ++k, ++r) {
                if (b.items[i]->req.return_solutions) o.solutions.push_back(std::move(sol[r]));
                o.true_rel_residual.push_back(resid[r]);
            }
            o.finished = fin;
        }
    } catch (const std::exception& e) {
        const double fin = now();
        for (auto& o : res) {
            o.ok       = false;
            o.error    = e.what();
            o.finished = fin;
        }
    }
    std::lock_guard<std::mutex> lk(mu);
    for (size_t i = 0; i < b.items.size(); ++i) publish_locked(*b.items[i], std::move(res[i]));
}

/* A twisted-mass lockstep solve (FP32 inner, FP64 reliable updates) of 2 to 16 slots whose right-hand sides
 * each carry their request's kappa and mu.  Each request's result is published as soon as its own right-hand
 * sides have converged; while join_running is set, the slots that are free at a reliable update (padding, or
 * converged) take queued compatible requests, which start there from x = 0 while the others continue. */
void Engine::run_tm_block(Worker& wk, Batch& b, Config& cfg)
{
    const double rel_tol = b.items[0]->req.rel_tol;
    const int config     = b.items[0]->req.config;
    int total_rhs        = 0;
    for (const auto& it : b.items) total_rhs += it->nrhs();
    const int n = padded_rhs(total_rhs);
    OpenBlock open{config, rel_tol, n - total_rhs};
    struct Refill : sv::BlockRefill {
        Engine& e;
        Batch& b;
        OpenBlock& open;
        WorkerLattice* L = nullptr;
        const Gaugefield* gf = nullptr;
        std::vector<std::unique_ptr<Spinorfield_eo>> bs, xs;
        std::vector<int> item, rhs;  // per slot: request (index into b.items, -1 free) and its right-hand side
        std::vector<int> left;       // per request: right-hand sides not yet converged
        std::vector<std::vector<double>> resid;
        std::vector<std::vector<std::vector<double>>> sols;
        std::vector<int> iters;
        std::vector<hmc_float> kappa_of, mubar_of;  // per slot, for the FP64 residual of its answer
        Refill(Engine& e, Batch& b, OpenBlock& o) : e(e), b(b), open(o) {}

        /* request i into `slots`: sources, x = 0, parameters */
        void place(size_t i, const std::vector<int>& slots, std::vector<hmc_float>& kappa,
                   std::vector<hmc_float>& mubar)
        {
            const Request& q = b.items[i]->req;
            const hmc_float mb = 2. * q.kappa * q.mu;
            for (int j = 0; j < q.nrhs(); ++j) {
                const int r = slots[j];
                if (q.sources.empty())
                    bs[r]->gaussian_from_seed(q.gaussian_seeds[j]);
                else
                    bs[r]->import(reinterpret_cast<const spinor*>(q.sources[j].data()));
                xs[r]->zero();
                item[r] = (int)i;
                rhs[r]  = j;
                kappa[r] = kappa_of[r] = q.kappa;
                mubar[r] = mubar_of[r] = mb;
            }
            if (left.size() <= i) {
                left.resize(i + 1);
                resid.resize(i + 1);
                sols.resize(i + 1);
                iters.resize(i + 1);
            }
            left[i] = q.nrhs();
            resid[i].assign(q.nrhs(), 0.);
            sols[i].assign(q.nrhs(), std::vector<double>());
            iters[i] = 0;
        }

        void converged(int r, const sv::SolveStats& st) override
        {
            const int i = item[r];
            if (i < 0) return;  // padding
            const hardware::System& sys = *L->system;
            const WilsonAdditionalParameters ap(kappa_of[r], mubar_of[r]);
            const Spinorfield_eo t(sys);
            (*L->qq)(&t, *gf, *xs[r], ap);
            saxpy(&t, hmc_complex{1., 0.}, t, *bs[r]);  // b - A x in FP64 with the request's own operator
            const hmc_float b2 = squarenorm(*bs[r]);
            resid[i][rhs[r]]   = b2 > 0. ? std::sqrt(squarenorm(t) / b2) : 0.;
            if (b.items[i]->req.return_solutions) {
                const std::vector<spinor> h = xs[r]->get_host();
                const double* d             = reinterpret_cast<const double*>(h.data());
                sols[i][rhs[r]].assign(d, d + h.size() * 24);
            }
            iters[i] = std::max(iters[i], st.iterations);
            item[r]  = -1;
            if (--left[i] > 0) {
                std::lock_guard<std::mutex> lk(e.mu);
                ++open.free;
                return;
            }
            Result o;
            o.ok             = true;
            o.iterations     = iters[i];
            o.batch_rhs      = (int)item.size();
            o.width          = 24;
            o.true_rel_residual = resid[i];
            if (b.items[i]->req.return_solutions) o.solutions = std::move(sols[i]);
            o.finished = e.now();
            {
                std::lock_guard<std::mutex> lk(e.mu);
                o.batch_requests = (int)b.items.size();
                ++open.free;
                e.publish_locked(*b.items[i], std::move(o));
            }
            e.cv_done.notify_all();
        }

        std::vector<int> refill(const std::vector<int>& free_slots, std::vector<hmc_float>& kappa,
                                std::vector<hmc_float>& mubar) override
        {
            std::vector<size_t> joined;
            {
                std::lock_guard<std::mutex> lk(e.mu);
                if (!e.opt.join_running) return {};
                int room = (int)free_slots.size();
                const double t = e.now();
                for (auto it = e.queue.begin(); it != e.queue.end() && room > 0;) {
                    const Request& q = (*it)->req;
                    if (q.action == Action::twistedmass && q.config == open.config && q.rel_tol == open.rel_tol &&
                        q.nrhs() <= room) {
                        room -= q.nrhs();
                        (*it)->admitted = t;
                        (*it)->joined   = true;
                        (*it)->solve    = b.solve;
                        joined.push_back(b.items.size());
                        b.items.push_back(std::move(*it));
                        it = e.queue.erase(it);
                        ++e.counters.requests_joined;
                    } else {
                        ++it;
                    }
                }
                open.free = room;
            }
            std::vector<int> taken;
            size_t next = 0;
            for (size_t i : joined) {
                std::vector<int> slots(free_slots.begin() + next, free_slots.begin() + next + b.items[i]->nrhs());
                next += slots.size();
                place(i, slots, kappa, mubar);
                taken.insert(taken.end(), slots.begin(), slots.end());
            }
            return taken;
        }
    } rf(*this, b, open);
    {
        std::lock_guard<std::mutex> lk(mu);
        open_blocks.push_back(&open);
    }
    try {
        WorkerLattice& L = wk.lattice(cfg.ns, cfg.nt);
        rf.L             = &L;
        rf.gf            = cfg.gf.get();
        const hardware::System& sys = *L.system;
        std::vector<hmc_float> kappa(n), mubar(n);
        rf.item.assign(n, -1);
        rf.rhs.assign(n, 0);
        rf.kappa_of.assign(n, 0.);
        rf.mubar_of.assign(n, 0.);
        for (int r = 0; r < n; ++r) {
            rf.bs.emplace_back(new Spinorfield_eo(sys));
            rf.xs.emplace_back(new Spinorfield_eo(sys));
            rf.xs.back()->zero();
        }
        int next = 0;
        for (size_t i = 0; i < b.items.size(); ++i) {
            std::vector<int> slots;
            for (int j = 0; j < b.items[i]->nrhs(); ++j) slots.push_back(next++);
            rf.place(i, slots, kappa, mubar);
        }
        for (int r = next; r < n; ++r) {  // padding right-hand sides are zero: free from the start
            rf.bs[r]->zero();
            kappa[r] = kappa[0];
            mubar[r] = mubar[0];
        }
        std::vector<const Spinorfield_eo*> xv, bv;
        for (int r = 0; r < n; ++r) {
            xv.push_back(rf.xs[r].get());
            bv.push_back(rf.bs[r].get());
        }
        sv::ReliableUpdateOptions o;
        o.inner_precision = opt.inner_precision;  // FP32, or FP16 blocks with per-site norms
        sv::cg_reliable_update_block(xv, *L.qq, *cfg.gf, bv, sys, rel_tol, kappa, mubar, o, rf);
    } catch (const std::exception& ex) {
        const double fin = now();
        std::lock_guard<std::mutex> lk(mu);
        for (auto& it : b.items)
            if (!it->published) {
                Result r;
                r.ok       = false;
                r.error    = ex.what();
                r.finished = fin;
                publish_locked(*it, std::move(r));
            }
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        open_blocks.erase(std::find(open_blocks.begin(), open_blocks.end(), &open));
        for (auto& it : b.items)  // every right-hand side converged and was published by the solver's callback
            if (!it->published) {
                Result r;
                r.ok       = false;
                r.error    = "serving: lockstep solve ended before this request converged";
                r.finished = now();
                publish_locked(*it, std::move(r));
            }
    }
    cv_done.notify_all();
}

void Engine::warmup()
{
    std::lock_guard<std::mutex> lk(mu);
    for (auto& w : workers) {
        if (w->busy) throw std::logic_error("serving: warmup while requests run");
        for (const auto& c : configs) w->lattice(c->ns, c->nt);
    }
}

Result Engine::wait(long id)
{
    std::unique_lock<std::mutex> lk(mu);
    if (id < 0 || id >= next_id) throw std::invalid_argument("serving: unknown request id");
    cv_done.wait(lk, [&] { return done.count(id) != 0; });
    Result r = std::move(done[id]);
    done.erase(id);
    return r;
}

void Engine::drain()
{
    std::unique_lock<std::mutex> lk(mu);
    cv_done.wait(lk, [&] { return outstanding == 0; });
}

}  // namespace serving
