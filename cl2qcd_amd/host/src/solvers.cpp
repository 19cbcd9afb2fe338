// cg, bicgstab and cg_m of the drop-in: the reference's algorithms and stopping
// rules (physics/algorithms/solvers/cg.cpp:277-400, bicgstab.cpp:401-733,
// physics/algorithms/solver_shifted.cpp:114-475) on device-resident fields and
// device scalars; the host reads a scalar only where the reference does.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>

#include "../include/physics/algorithms/solvers/solvers.hpp"

namespace physics {
namespace algorithms {
namespace solvers {

using namespace physics::lattices;

namespace {

bool is_nan(hmc_float v)
{
    return v != v;
}

void debug(const char* fmt, int iter, double v)
{
    if (std::getenv("CL2QCD_DEBUG")) std::fprintf(stderr, fmt, iter, v);
}

}  // namespace

int cg(const Spinorfield_eo* x, const fermionmatrix::Fermionmatrix_eo& f, const Gaugefield& gf,
       const Spinorfield_eo& b, const hardware::System& system, physics::InterfacesHandler& interfacesHandler,
       hmc_float prec, const physics::AdditionalParameters& additionalParameters)
{
    const SolversParametersInterface& par = interfacesHandler.getSolversParametersInterface();
    const int cgmax = par.getCgMax(), refresh = par.getIterRefresh(), check = par.getCgIterationBlockSize();
    const int min_iter = par.getCgMinimumIterationCount();
    const bool merged  = par.getUseMergeKernelsSpinor();
    const int prec_f   = b.precision();

    const Spinorfield_eo p(system, prec_f), rn(system, prec_f), v(system, prec_f);
    const Scalar<hmc_complex> alpha(system), beta(system), omega(system), rho(system), rho_next(system),
        tmp1(system), tmp2(system), one(system), minus_one(system);
    one.store(hmc_complex_one);
    minus_one.store(hmc_complex_minusone);

    int iter = 0;
    hmc_float resid = 0.;
    for (iter = 0; iter < cgmax || iter < min_iter; iter++) {
        if (iter % refresh == 0) {
            f(&rn, gf, *x, additionalParameters);  // rn = A x
            saxpy(&rn, one, rn, b);                // rn = b - A x
            copyData(&p, rn);
            scalar_product(&omega, rn, rn);
        } else {
            copyData(&omega, rho_next);
        }
        f(&v, gf, p, additionalParameters);  // v = A p
        scalar_product(&rho, p, v);
        divide(&alpha, omega, rho);  // alpha = (r,r)/(p,Ap)
        multiply(&tmp1, minus_one, alpha);
        saxpy(x, tmp1, p, *x);  // x = x + alpha p
        if (merged) {
            saxpy_AND_squarenorm(&rn, alpha, v, rn, rho_next);
        } else {
            saxpy(&rn, alpha, v, rn);  // rn = rn - alpha v
            scalar_product(&rho_next, rn, rn);
        }
        if (iter % check == 0) {
            resid = rho_next.get().re;
            debug("CG [%d] resid %.16g\n", iter, resid);
            if (is_nan(resid)) throw SolverStuck(iter, "cg: NAN occured");
            if (resid < prec && iter >= min_iter) return iter;
        }
        divide(&beta, rho_next, omega);
        multiply(&tmp2, minus_one, beta);
        saxpy(&p, tmp2, p, rn);  // p = rn + beta p
    }
    throw SolverDidNotSolve(iter, "cg");
}

namespace {

int bicgstab_save(const Spinorfield_eo* x, const fermionmatrix::Fermionmatrix_eo& f, const Gaugefield& gf,
                  const Spinorfield_eo& b, const hardware::System& system, const SolversParametersInterface& par,
                  hmc_float prec, const physics::AdditionalParameters& ap)
{
    const int pf = b.precision();
    const Spinorfield_eo s(system, pf), t(system, pf), v(system, pf), p(system, pf), rn(system, pf),
        rhat(system, pf), aux(system, pf);
    const Scalar<hmc_complex> alpha(system), beta(system), omega(system), rho(system), rho_next(system),
        tmp1(system), tmp2(system), one(system), minus_one(system);
    one.store(hmc_complex_one);
    minus_one.store(hmc_complex_minusone);
    const int cgmax = par.getCgMax();
    hmc_float resid = 1.;
    int iter        = 0;
    for (iter = 0; iter < cgmax; iter++) {
        if (iter % par.getIterRefresh() == 0) {
            v.zero();
            p.zero();
            f(&rn, gf, *x, ap);
            saxpy(&rn, one, rn, b);
            copyData(&rhat, rn);
            copyData(&alpha, one);
            copyData(&omega, one);
            copyData(&rho, one);
        }
        scalar_product(&rho_next, rhat, rn);
        const hmc_complex test = rho_next.get();
        if (std::fabs(test.re) < 1e-25 && std::fabs(test.im) < 1e-25) throw SolverStuck(iter, "bicgstab_save");
        divide(&tmp1, rho_next, rho);
        copyData(&rho, rho_next);
        divide(&tmp2, alpha, omega);
        multiply(&beta, tmp1, tmp2);
        multiply(&tmp1, beta, omega);
        multiply(&tmp2, minus_one, tmp1);
        saxsbypz(&p, beta, p, tmp2, v, rn);  // p = beta p - beta omega v + rn
        f(&v, gf, p, ap);
        scalar_product(&tmp1, rhat, v);
        divide(&alpha, rho, tmp1);
        saxpy(&s, alpha, v, rn);  // s = rn - alpha v
        f(&t, gf, s, ap);
        scalar_product(&tmp1, t, s);
        scalar_product(&tmp2, t, t);
        divide(&omega, tmp1, tmp2);
        saxpy(&rn, omega, t, s);               // rn = s - omega t
        saxsbypz(x, alpha, p, omega, s, *x);  // x = alpha p + omega s + x
        resid = squarenorm(rn);
        debug("BiCGstab_save [%d] resid %.16g\n", iter, resid);
        if (is_nan(resid)) throw SolverStuck(iter, "bicgstab_save: NAN occured");
        if (resid < prec) {
            f(&aux, gf, *x, ap);
            saxpy(&aux, one, aux, b);
            const hmc_float trueresid = squarenorm(aux);
            if (trueresid < prec) return iter;
        }
    }
    throw SolverDidNotSolve(iter, "bicgstab_save");
}

int bicgstab_fast(const Spinorfield_eo* x, const fermionmatrix::Fermionmatrix_eo& f, const Gaugefield& gf,
                  const Spinorfield_eo& b, const hardware::System& system, const SolversParametersInterface& par,
                  hmc_float prec, const physics::AdditionalParameters& ap)
{
    const int pf = b.precision();
    const Spinorfield_eo p(system, pf), rn(system, pf), rhat(system, pf), s(system, pf), t(system, pf),
        v(system, pf);
    const Scalar<hmc_complex> alpha(system), beta(system), omega(system), rho(system), rho_next(system),
        tmp1(system), tmp2(system), one(system), minus_one(system);
    one.store(hmc_complex_one);
    minus_one.store(hmc_complex_minusone);
    hmc_float resid = NAN;
    int iter        = 0;
    for (iter = 0; iter < par.getCgMax(); iter++) {
        if (iter % par.getIterRefresh() == 0) {
            f(&rn, gf, *x, ap);
            saxpy(&p, one, rn, b);
            copyData(&rhat, p);
            copyData(&rn, p);
            scalar_product(&rho, rhat, rn);
        }
        resid = squarenorm(rn);
        debug("BiCGstab [%d] resid %.16g\n", iter, resid);
        if (is_nan(resid)) throw SolverStuck(iter, "bicgstab: NAN occured");
        if (resid < prec) return iter;
        f(&v, gf, p, ap);
        scalar_product(&tmp1, rhat, v);
        divide(&alpha, rho, tmp1);
        saxpy(&s, alpha, v, rn);
        f(&t, gf, s, ap);
        scalar_product(&tmp1, t, s);
        scalar_product(&tmp2, t, t);
        divide(&omega, tmp1, tmp2);
        saxsbypz(x, alpha, p, omega, s, *x);
        saxpy(&rn, omega, t, s);
        scalar_product(&rho_next, rhat, rn);
        const hmc_complex test = rho_next.get();
        if (std::fabs(test.re) < 1e-25 && std::fabs(test.im) < 1e-25) throw SolverStuck(iter, "bicgstab");
        divide(&tmp1, rho_next, rho);
        divide(&tmp2, alpha, omega);
        multiply(&beta, tmp1, tmp2);
        multiply(&tmp1, beta, omega);
        multiply(&tmp2, minus_one, tmp1);
        saxsbypz(&p, beta, p, tmp2, v, rn);
        copyData(&rho, rho_next);
    }
    throw SolverDidNotSolve(iter, "bicgstab");
}

}  // namespace

int bicgstab(const Spinorfield_eo* x, const fermionmatrix::Fermionmatrix_eo& A, const Gaugefield& gf,
             const Spinorfield_eo& b, const hardware::System& system, physics::InterfacesHandler& interfacesHandler,
             hmc_float prec, const physics::AdditionalParameters& additionalParameters)
{
    const SolversParametersInterface& par = interfacesHandler.getSolversParametersInterface();
    if (par.getSolver() == common::bicgstab_save)
        return bicgstab_save(x, A, gf, b, system, par, prec, additionalParameters);
    return bicgstab_fast(x, A, gf, b, system, par, prec, additionalParameters);
}

/* SolverShifted::solve (solver_shifted.cpp:428-475) */
int cg_m(std::vector<std::shared_ptr<Staggeredfield_eo>> x, const fermionmatrix::Fermionmatrix_stagg_eo& A,
         const Gaugefield& gf, const std::vector<hmc_float> sigma, const Staggeredfield_eo& b,
         const hardware::System& system, physics::InterfacesHandler& interfacesHandler, hmc_float prec,
         const physics::AdditionalParameters& additionalParameters)
{
    if (sigma.size() != x.size()) throw std::invalid_argument("Wrong size of multi-shifted inverter parameters!");
    const SolversParametersInterface& par = interfacesHandler.getSolversParametersInterface();
    const int n      = (int)sigma.size();
    const int pf     = b.precision();
    const bool merged = par.getUseMergeKernelsSpinor();
    const int min_iter = par.getCgMinimumIterationCount();
    if (squarenorm(b) == 0) {
        for (auto& xi : x) xi->set_zero();
        return 0;
    }
    const Staggeredfield_eo r(system, pf), p(system, pf), v(system, pf);
    std::vector<std::unique_ptr<Staggeredfield_eo>> ps;
    for (int k = 0; k < n; ++k) ps.emplace_back(new Staggeredfield_eo(system, pf));
    const Vector<hmc_float> zeta_prev(n, system), zeta(n, system), zeta_foll(n, system), alpha_vec(n, system),
        beta_vec(n, system), shift(n, system), single_eq_resid(n, system);
    const Scalar<hmc_float> alpha_scalar_prev(system), alpha_scalar(system), beta_scalar_prev(system),
        beta_scalar(system), zero(system), tmp1(system), tmp2(system), tmp3(system);
    std::vector<bool> converged(n, false);
    int nconv = 0;

    // setInitialConditions (solver_shifted.cpp:114-134)
    zero.store(0.0);
    zeta_prev.store(std::vector<hmc_float>(n, 1.));
    zeta.store(std::vector<hmc_float>(n, 1.));
    alpha_vec.store(std::vector<hmc_float>(n, 0.));
    shift.store(sigma);
    for (int k = 0; k < n; ++k) {
        x[k]->set_zero();
        copyData(ps[k].get(), b);
    }
    copyData(&r, b);
    copyData(&p, b);
    scalar_product_real_part(&tmp1, r, r);
    beta_scalar.store(1.0);
    alpha_scalar.store(0.0);

    int iter = 0;
    while (iter < par.getCgMax() || iter < min_iter) {
        // updateBetaScalar
        copyData(&beta_scalar_prev, beta_scalar);
        A(&v, gf, p, &additionalParameters);
        scalar_product_real_part(&tmp3, p, v);
        divide(&beta_scalar, tmp1, tmp3);
        subtract(&beta_scalar, zero, beta_scalar);
        // updateAuxiliaryFieldR: r = beta v + r
        saxpy(&r, beta_scalar, v, r);
        // updateAlphaScalar
        scalar_product_real_part(&tmp2, r, r);
        copyData(&alpha_scalar_prev, alpha_scalar);
        divide(&alpha_scalar, tmp2, tmp1);
        // updateAuxiliaryFieldP: p = alpha p + r
        saxpy(&p, alpha_scalar, p, r);
        // updateVectorQuantities
        update_zeta_cgm(&zeta_foll, zeta, zeta_prev, beta_scalar_prev, beta_scalar, alpha_scalar_prev, shift, n);
        update_beta_cgm(&beta_vec, beta_scalar, zeta_foll, zeta, n);
        update_alpha_cgm(&alpha_vec, alpha_scalar, zeta_foll, beta_vec, zeta, beta_scalar, n);
        const bool check = iter % par.getCgIterationBlockSize() == 0;
        std::vector<hmc_float> resid_host;
        if (merged) {
            sax_vec_and_squarenorm(&single_eq_resid, zeta_foll, r);
            resid_host = single_eq_resid.get();
        }
        for (int k = 0; k < n; ++k) {
            if (converged[k]) continue;
            saxpy(x[k].get(), beta_vec, k, *ps[k], *x[k]);               // x_k = beta_k ps_k + x_k
            saxpby(ps[k].get(), zeta_foll, k, r, alpha_vec, k, *ps[k]);  // ps_k = zeta_k r + alpha_k ps_k
            if (check) {
                hmc_float res_k;
                if (merged) {
                    res_k = resid_host[k];
                } else {
                    sax(&v, zeta_foll, k, r);
                    res_k = squarenorm(v);
                }
                if (res_k < prec) {
                    converged[k] = true;
                    ++nconv;
                }
            }
        }
        // updateQuantitiesForFollowingIteration
        copyData(&zeta_prev, zeta);
        copyData(&zeta, zeta_foll);
        copyData(&tmp1, tmp2);
        if (nconv == n) {
            const hmc_float resid = tmp2.get();
            debug("CG-M [%d] resid %.16g\n", iter, resid);
            if (is_nan(resid)) throw SolverStuck(iter, "cg_m: NAN occured");
            if (resid < prec && iter >= min_iter) return iter;
        }
        ++iter;
    }
    throw SolverDidNotSolve(iter, "cg_m");
}

}  // namespace solvers
}  // namespace algorithms
}  // namespace physics

namespace physics {
namespace algorithms {
namespace solvers {

/* Reliable-update CG: the inner recurrence runs on inner_precision fields with
 * the operator applied in that precision; whenever the iterated residual has
 * dropped by delta since the last update (or claims convergence), the partial
 * solution is folded into the FP64 solution and the residual is recomputed in
 * FP64 from b - A x, which also decides convergence.  Scalars (alpha, beta,
 * |r|^2) and all reductions are FP64.
 *
 * Two standard measures keep the low-precision recurrence close to exact CG:
 * beta is the Polak-Ribiere value <r_new, r_new - r_old> / |r_old|^2 (from the
 * fused residual update; the Fletcher-Reeves |r_new|^2 / |r_old|^2 when that is
 * negative), and at every reliable update the search direction is rebuilt on the
 * true residual: p = r_true + d with d = beta p_old minus its component along
 * r_true, so that <r_true, beta p_old> = 0 holds as in exact CG.  One inner
 * iteration (operator, alpha, x and r updates, beta, p update) is a fixed kernel
 * sequence with every scalar on the device; it is recorded once and replayed. */
int cg_reliable_update(const Spinorfield_eo* x, const fermionmatrix::Fermionmatrix_eo& A, const Gaugefield& gf,
                       const Spinorfield_eo& b, const hardware::System& system, hmc_float rel_tol,
                       const physics::AdditionalParameters& ap, const ReliableUpdateOptions& opt, SolveStats* stats)
{
    if (x->precision() != CL2_F64 || b.precision() != CL2_F64)
        throw std::invalid_argument("cg_reliable_update: x and b are FP64 fields");
    if (opt.delta <= 0. || opt.delta >= 1.) throw std::invalid_argument("cg_reliable_update: delta must be in (0,1)");
    const int lo = opt.inner_precision;
    const Spinorfield_eo r(system, CL2_F64), t64(system, CL2_F64);
    const Spinorfield_eo r_lo(system, lo), p(system, lo), v(system, lo), x_lo(system, lo);
    const Scalar<hmc_complex> alpha(system), malpha(system), pAp(system), r2c(system), mbeta(system),
        minus_one(system);
    const Vector<hmc_float> norms(4, system);  // (|r|^2, Re <r_new, r_old>, |r|^2, 0) of the last update
    minus_one.store(hmc_complex_minusone);
    const hmc_complex one{1., 0.};
    cl2_ctx* c = system.ctx();

    const hmc_float b2     = squarenorm(b);
    const hmc_float target = rel_tol * rel_tol * b2;
    auto true_residual = [&]() {
        A(&t64, gf, *x, ap);
        saxpy(&r, one, t64, b);  // r = b - A x
        return squarenorm(r);
    };
    hmc_float r2 = true_residual();
    SolveStats st;
    if (b2 == 0. || r2 <= target) {
        st.true_relative_residual = b2 > 0. ? std::sqrt(r2 / b2) : 0.;
        if (stats) *stats = st;
        return 0;
    }
    convert_precision(&r_lo, r);
    copyData(&p, r_lo);
    x_lo.zero();
    r2c.store(hmc_complex{r2, 0.});  // |r|^2 of the current step as a complex device scalar
    A(&v, gf, p, ap);                // makes the reduced-precision link copy before anything is recorded
    // QplusQminus_eo = Qminus^dagger Qminus: <p, A p> = |Qminus p|^2 comes out of the operator's own kernels
    const auto* builtin  = dynamic_cast<const fermionmatrix::BuiltinEoOperator*>(&A);
    const bool fused_dot = builtin != nullptr && builtin->abi_operator() == CL2_OP_QPLUSQMINUS_EO;
    // one inner iteration, every scalar on the device: v = A p with <p, v>, alpha = r2 / <p, v>,
    // r_lo -= alpha v with its norms (for QplusQminus_eo all of it inside the operator's kernels, v never
    // stored), beta (Polak-Ribiere) and r2 <- |r_lo|^2, then x_lo += alpha p and p = r_lo + beta p in one pass
    auto body = [&]() {
        if (fused_dot) {
            builtin->cg_step(&r_lo, gf, p, ap, r2c, pAp, alpha, norms);
        } else {
            A(&v, gf, p, ap);
            scalar_product(&pAp, p, v);
            divide(&alpha, r2c, pAp);
            saxpy_AND_cgnorm(&r_lo, alpha, v, r_lo, norms);
        }
        multiply(&malpha, minus_one, alpha);
        system.check(cl2_cg_beta(c, mbeta.dev(), r2c.dev(), norms.dev()));
        cg_update_xp(&x_lo, &p, r_lo, malpha, mbeta);
    };
    int graph_nodes = 0, graph_kernel_nodes = 0;  // census of the captured iteration
    struct Graph {
        cl2_graph* g = nullptr;
        ~Graph() { cl2_graph_destroy(g); }
    } graph;
    if (opt.use_graph) {
        system.check(cl2_graph_begin(c));
        try {
            body();
        } catch (...) {
            cl2_graph* g = nullptr;
            (void)cl2_graph_end(c, &g);
            cl2_graph_destroy(g);
            throw;
        }
        system.check(cl2_graph_end(c, &graph.g));
        system.check(cl2_graph_node_count(graph.g, &graph_kernel_nodes, &graph_nodes));
        // replays are checked against eager execution for kernel launches only (memset nodes have replayed
        // wrongly on this hardware): a capture that recorded anything else is a bug here
        if (graph_nodes != graph_kernel_nodes) throw std::logic_error("captured solver iteration holds non-kernel nodes");
    }
    hmc_float rmax = std::sqrt(r2);
    hmc_float rtrue = r2;  // |b - A x|^2 of the FP64 solution
    bool pending    = false;  // x_lo holds a correction not yet added to x
    int k           = 0;
    for (k = 0; k < opt.max_iterations; ++k) {
        if (graph.g)
            system.check(cl2_graph_launch(c, graph.g, 1));
        else
            body();
        pending                         = true;
        const std::vector<hmc_float> nr = norms.get();
        hmc_float rn                    = nr[0];
        if (rn != rn) throw SolverStuck(k, "cg_reliable_update: NAN occured");
        if (std::sqrt(rn) < opt.delta * rmax || rn <= target) {
            convert_precision(&t64, x_lo);
            saxpy(x, hmc_complex{-1., 0.}, t64, *x);  // x += x_lo
            x_lo.zero();
            pending = false;
            rn      = true_residual();
            rtrue   = rn;
            ++st.reliable_updates;
            if (rn <= target) {
                ++k;
                break;
            }
            // p = r_iter + beta p_old was formed with the iterated residual: rebuild it on the true one,
            // p = r_true + (beta p_old with its component along r_true removed)
            saxpy(&p, one, r_lo, p);  // beta p_old = p - r_iter
            convert_precision(&r_lo, r);
            const hmc_complex rp = scalar_product(r_lo, p);
            saxpy(&p, hmc_complex{rp.re / rn, rp.im / rn}, r_lo, p);
            saxpy(&p, hmc_complex{-1., 0.}, r_lo, p);  // + r_true
            r2c.store(hmc_complex{rn, 0.});
            rmax = std::sqrt(rn);
        }
        rmax = std::max(rmax, std::sqrt(rn));
    }
    if (pending) {  // stopped by the iteration limit: keep the correction gathered since the last update
        convert_precision(&t64, x_lo);
        saxpy(x, hmc_complex{-1., 0.}, t64, *x);
        rtrue = true_residual();
    }
    st.iterations             = k;
    st.true_relative_residual = std::sqrt(rtrue / b2);
    st.graph_nodes            = graph_nodes;
    st.graph_kernel_nodes     = graph_kernel_nodes;
    if (stats) *stats = st;
    if (st.true_relative_residual > rel_tol) throw SolverDidNotSolve(k, "cg_reliable_update");
    return k;
}

}  // namespace solvers
}  // namespace algorithms
}  // namespace physics

namespace physics {
namespace algorithms {
namespace solvers {

namespace {
/* the block operations of the lockstep solver for Wilson spinors and staggered vectors */
struct WilsonBlockOps {
    using Field = Spinorfield_eo;
    using Block = Spinorfield_eo_block;
    static constexpr bool half = true;  // FP16 blocks (per-site norm) for the inner loop
    static int squarenorm(cl2_ctx* c, double* res, const void* x, int prec, int n)
    {
        return cl2_block_squarenorm_eo(c, res, x, prec, n);
    }
    static int dot(cl2_ctx* c, double* res, const void* x, const void* y, int prec, int n)
    {
        return cl2_block_dot_eo(c, res, x, y, prec, n);
    }
    static int axpy(cl2_ctx* c, void* out, const void* x, const void* y, int prec, int n, const double* coef,
                    double sign, double* norms)
    {
        return cl2_block_axpy_eo(c, out, x, y, prec, n, coef, sign, norms);
    }
    static int convert(cl2_ctx* c, void* dst, int dprec, const void* src, int sprec, int n)
    {
        return cl2_block_convert_eo(c, dst, dprec, src, sprec, n);
    }
    static int xp(cl2_ctx* c, void* x, void* p, const void* r, int prec, int n, const double* a, const double* b)
    {
        return cl2_block_cg_update_xp_eo(c, x, p, r, prec, n, a, b);
    }
};
struct StaggBlockOps {
    using Field = Staggeredfield_eo;
    using Block = Staggeredfield_eo_block;
    static constexpr bool half = false;
    static int squarenorm(cl2_ctx* c, double* res, const void* x, int prec, int n)
    {
        return cl2_block_squarenorm_stagg_eo(c, res, x, prec, n);
    }
    static int dot(cl2_ctx* c, double* res, const void* x, const void* y, int prec, int n)
    {
        return cl2_block_dot_stagg_eo(c, res, x, y, prec, n);
    }
    static int axpy(cl2_ctx* c, void* out, const void* x, const void* y, int prec, int n, const double* coef,
                    double sign, double* norms)
    {
        return cl2_block_axpy_stagg_eo(c, out, x, y, prec, n, coef, sign, norms);
    }
    static int convert(cl2_ctx* c, void* dst, int dprec, const void* src, int sprec, int n)
    {
        return cl2_block_convert_stagg_eo(c, dst, dprec, src, sprec, n);
    }
    static int xp(cl2_ctx* c, void* x, void* p, const void* r, int prec, int n, const double* a, const double* b)
    {
        return cl2_block_cg_update_xp_stagg_eo(c, x, p, r, prec, n, a, b);
    }
};

/* out = A in for a block (tmp: two blocks of in's shape) */
template<class Ops>
using BlockApplyT = std::function<void(const typename Ops::Block*, const typename Ops::Block&, const DeviceBuffer&)>;
using BlockApply  = BlockApplyT<WilsonBlockOps>;
/* the operator half of a lockstep step (r, p, tmp, r2, active, pAp, alpha, norms) where the operator has one
 * (an empty function where not) */
template<class Ops>
using BlockStepT = std::function<bool(const typename Ops::Block*, const typename Ops::Block&, const DeviceBuffer&,
                                      const Vector<hmc_float>&, const Vector<hmc_float>&, const Vector<hmc_float>&,
                                      const Vector<hmc_float>&, const Vector<hmc_float>&)>;
using BlockStep = BlockStepT<WilsonBlockOps>;

/* slot refills of a running block (continuous batching): at a reliable update, `converged` hears of every
 * right-hand side frozen there and `refill` is offered the free slots; it returns the slots it gave new
 * problems (new b[r], x[r] = 0, new operator parameters) */
struct BlockHooks {
    std::function<void(int, const SolveStats&)> converged;
    std::function<std::vector<int>(const std::vector<int>&)> refill;
};

template<class Ops>
int block_solve(const std::vector<const typename Ops::Field*>& x, const std::vector<const typename Ops::Field*>& b,
                const hardware::System& system, hmc_float rel_tol, const BlockApplyT<Ops>& apply,
                const BlockStepT<Ops>& step, const ReliableUpdateOptions& opt, std::vector<SolveStats>* stats,
                bool shrink = false, const BlockHooks* hooks = nullptr)
{
    const int n = (int)b.size();
    if (n < 1 || x.size() != b.size()) throw std::invalid_argument("cg_reliable_update_block: x and b differ in size");
    if (!(n == 1 || n == 2 || n == 4 || n == 8 || n == 12 || n == 16))
        throw std::invalid_argument("cg_reliable_update_block: 1, 2, 4, 8, 12 or 16 right-hand sides");
    for (int r = 0; r < n; ++r)
        if (x[r]->precision() != CL2_F64 || b[r]->precision() != CL2_F64)
            throw std::invalid_argument("cg_reliable_update_block: x and b are FP64 fields");
    if (opt.delta <= 0. || opt.delta >= 1.) throw std::invalid_argument("cg_reliable_update_block: delta must be in (0,1)");
    const int lo = opt.inner_precision;
    if (lo != CL2_F32 && lo != CL2_F64 && !(lo == CL2_F16 && Ops::half))
        throw std::invalid_argument(Ops::half ? "cg_reliable_update_block: inner precision FP64, FP32 or FP16"
                                              : "cg_reliable_update_block: inner precision FP32 or FP64");
    cl2_ctx* c = system.ctx();

    const typename Ops::Block B(system, n, CL2_F64), X(system, n, CL2_F64), R(system, n, CL2_F64),
        T(system, n, CL2_F64);
    const typename Ops::Block Rlo(system, n, lo), P(system, n, lo), V(system, n, lo), Xlo(system, n, lo);
    const DeviceBuffer tmp64(system, 2 * T.bytes()), tmplo(system, 2 * P.bytes());
    // per-right-hand-side device scalars
    const Vector<hmc_float> ones(2 * n, system), alpha(2 * n, system), beta(2 * n, system), pAp(2 * n, system),
        norms(2 * n, system), r2(n, system), active(n, system), coef(2 * n, system), res(2 * n, system);
    std::vector<hmc_float> one_v(2 * n, 0.);
    for (int r = 0; r < n; ++r) one_v[2 * r] = 1.;
    ones.store(one_v);
    for (int r = 0; r < n; ++r) {
        B.insert(r, *b[r]);
        X.insert(r, *x[r]);
    }
    auto block_norms = [&](const typename Ops::Block& f) {
        system.check(Ops::squarenorm(c, res.dev(), f.dev(), f.precision(), n));
        std::vector<hmc_float> v = res.get();
        v.resize(n);
        return v;
    };
    auto true_residual = [&]() {
        apply(&T, X, tmp64);
        system.check(Ops::axpy(c, R.dev(), T.dev(), B.dev(), CL2_F64, n, ones.dev(), -1., nullptr));
        return block_norms(R);  // |b - A x|^2 per right-hand side
    };
    std::vector<hmc_float> b2 = block_norms(B);
    std::vector<hmc_float> target(n), rn = true_residual(), act(n), r2h(n), rmax(n);
    std::vector<SolveStats> st(n);
    std::vector<int> start_k(n, 0), start_rel(n, 0);  // iteration and reliable update at which a slot started
    int n_active = 0;
    for (int r = 0; r < n; ++r) {
        target[r] = rel_tol * rel_tol * b2[r];
        act[r]    = (b2[r] > 0. && rn[r] > target[r]) ? 1. : 0.;
        n_active += act[r] != 0.;
        r2h[r]  = rn[r];
        rmax[r] = std::sqrt(rn[r]);
    }
    auto finish = [&](int k) {
        for (int r = 0; r < n; ++r) {
            X.extract(r, x[r]);
            st[r].true_relative_residual = b2[r] > 0. ? std::sqrt(rn[r] / b2[r]) : 0.;
            if (act[r] != 0.) st[r].iterations = k - start_k[r];
        }
        if (stats) *stats = st;
    };
    if (hooks)  // right-hand sides that need no iteration (zero, or already solved) are answered now
        for (int r = 0; r < n; ++r)
            if (act[r] == 0.) {
                X.extract(r, x[r]);
                st[r].true_relative_residual = b2[r] > 0. ? std::sqrt(rn[r] / b2[r]) : 0.;
                hooks->converged(r, st[r]);
            }
    if (n_active == 0) {
        finish(0);
        return 0;
    }
    system.check(Ops::convert(c, Rlo.dev(), lo, R.dev(), CL2_F64, n));
    system.check(cl2_copy(c, P.dev(), Rlo.dev(), P.bytes()));
    Xlo.zero();
    r2.store(r2h);
    active.store(act);
    int k           = 0;
    int reliable    = 0;
    bool partial    = false;  // shrink: returned with some right-hand sides still active
    apply(&V, P, tmplo);  // reduced-precision links made before recording
    // whether the operator does the step's operator half itself (QplusQminus_eo: no A p stored, no dot pass)
    const bool fused = static_cast<bool>(step) && n > 1;
    // one lockstep iteration, every scalar on the device
    auto body = [&]() {
        if (!fused) {
            apply(&V, P, tmplo);
            system.check(Ops::dot(c, pAp.dev(), P.dev(), V.dev(), lo, n));
            system.check(cl2_block_cg_alpha(c, alpha.dev(), r2.dev(), pAp.dev(), active.dev(), n));
            system.check(Ops::axpy(c, Rlo.dev(), V.dev(), Rlo.dev(), lo, n, alpha.dev(), -1., norms.dev()));
        } else {
            if (!step(&Rlo, P, tmplo, r2, active, pAp, alpha, norms))
                throw std::logic_error("cg_reliable_update_block: the operator has no fused CG step");
        }
        system.check(cl2_block_cg_beta(c, beta.dev(), r2.dev(), norms.dev(), active.dev(), n));
        // x += alpha p and p = r + beta p in one pass
        system.check(Ops::xp(c, Xlo.dev(), P.dev(), Rlo.dev(), lo, n, alpha.dev(), beta.dev()));
    };
    int graph_nodes = 0, graph_kernel_nodes = 0;  // census of the captured iteration
    struct Graph {
        cl2_graph* g = nullptr;
        ~Graph() { cl2_graph_destroy(g); }
    } graph;
    // (re)record the iteration: the operator's per-right-hand-side parameters are kernel arguments, so a refill
    // that changes them records again
    auto capture = [&]() {
        cl2_graph_destroy(graph.g);
        graph.g = nullptr;
        system.check(cl2_graph_begin(c));
        try {
            body();
        } catch (...) {
            cl2_graph* g = nullptr;
            (void)cl2_graph_end(c, &g);
            cl2_graph_destroy(g);
            throw;
        }
        system.check(cl2_graph_end(c, &graph.g));
        system.check(cl2_graph_node_count(graph.g, &graph_kernel_nodes, &graph_nodes));
        // replays are checked against eager execution for kernel launches only (memset nodes have replayed
        // wrongly on this hardware): a capture that recorded anything else is a bug here
        if (graph_nodes != graph_kernel_nodes) throw std::logic_error("captured solver iteration holds non-kernel nodes");
    };
    if (opt.use_graph) capture();
    bool pending = false;  // Xlo holds corrections not yet added to X
    for (k = 0; k < opt.max_iterations && n_active > 0; ++k) {
        if (graph.g)
            system.check(cl2_graph_launch(c, graph.g, 1));
        else
            body();
        pending                         = true;
        const std::vector<hmc_float> nr = norms.get();
        bool update                     = false;
        // a right-hand side whose iterated residual has reached its target waits, still iterating in lockstep,
        // for the next update to verify it in FP64: that update comes when another one has dropped by delta,
        // when every active one has reached its target, or (a shrinking block) when half of them have.  A
        // right-hand side verified one at a time would cost one FP64 update each as the sources converge a few
        // iterations apart.  With hooks (serving) each is verified at once, so that its answer leaves early.
        int reached = 0, open = 0;
        for (int r = 0; r < n; ++r) {
            if (act[r] == 0.) continue;
            const hmc_float q = nr[2 * r];
            if (q != q) throw SolverStuck(k, "cg_reliable_update_block: NAN occured");
            if (q <= target[r]) {
                ++reached;
                if (hooks) update = true;
            } else {
                ++open;
                if (std::sqrt(q) < opt.delta * rmax[r]) update = true;
            }
            rmax[r] = std::max(rmax[r], std::sqrt(q));
        }
        if (open == 0 || (shrink && 2 * reached >= n_active)) update = true;
        if (!update) continue;
        system.check(Ops::convert(c, T.dev(), CL2_F64, Xlo.dev(), lo, n));
        system.check(Ops::axpy(c, X.dev(), T.dev(), X.dev(), CL2_F64, n, ones.dev(), 1., nullptr));
        Xlo.zero();
        pending = false;
        rn      = true_residual();
        ++reliable;
        std::vector<hmc_float> oc(2 * n, 0.);
        std::vector<int> frozen;
        for (int r = 0; r < n; ++r) {
            if (act[r] == 0.) continue;
            if (rn[r] <= target[r]) {  // converged: freeze this right-hand side
                act[r]           = 0.;
                st[r].iterations = k + 1 - start_k[r];
                --n_active;
                frozen.push_back(r);
            }
        }
        std::vector<int> taken;  // slots refilled with new problems at this update
        if (hooks) {
            for (int r : frozen) {
                X.extract(r, x[r]);
                st[r].true_relative_residual = std::sqrt(rn[r] / b2[r]);
                st[r].reliable_updates       = reliable - start_rel[r];
                hooks->converged(r, st[r]);
            }
            std::vector<int> free_slots;
            for (int r = 0; r < n; ++r)
                if (act[r] == 0.) free_slots.push_back(r);
            if (!free_slots.empty()) taken = hooks->refill(free_slots);
            if (!taken.empty()) {
                for (int r : taken) {
                    B.insert(r, *b[r]);
                    X.insert(r, *x[r]);  // zero
                }
                b2 = block_norms(B);
                rn = true_residual();  // r = b in the new slots
                for (int r : taken) {
                    if (act[r] != 0.) throw std::logic_error("cg_reliable_update_block: a refill took an active slot");
                    target[r]    = rel_tol * rel_tol * b2[r];
                    st[r]        = SolveStats();
                    start_k[r]   = k + 1;
                    start_rel[r] = reliable;
                    if (b2[r] > 0.) {
                        act[r] = 1.;
                        ++n_active;
                    } else {  // a zero right-hand side: x = 0 is its answer
                        hooks->converged(r, st[r]);
                    }
                }
            }
        }
        if (n_active == 0) {
            ++k;
            break;
        }
        if (shrink && 2 * n_active <= n) {  // at most half still active: return them for a smaller block
            ++k;
            partial = true;
            break;
        }
        // p = r_iter + beta p_old was formed with the iterated residuals: rebuild it on the true ones,
        // p = r_true + (beta p_old with its component along r_true removed)
        system.check(Ops::axpy(c, P.dev(), Rlo.dev(), P.dev(), lo, n, ones.dev(), -1., nullptr));
        if (!taken.empty()) {  // refilled slots start afresh: their p_old is zero, so p = r there
            std::vector<hmc_float> z(2 * n, 0.);
            for (int r : taken) z[2 * r] = 1.;
            coef.store(z);
            system.check(Ops::axpy(c, P.dev(), P.dev(), P.dev(), lo, n, coef.dev(), -1., nullptr));  // p_r -= p_r
        }
        system.check(Ops::convert(c, Rlo.dev(), lo, R.dev(), CL2_F64, n));
        system.check(Ops::dot(c, res.dev(), Rlo.dev(), P.dev(), lo, n));
        const std::vector<hmc_float> rp = res.get();
        for (int r = 0; r < n; ++r) {
            if (act[r] == 0.) continue;
            oc[2 * r]     = rp[2 * r] / rn[r];
            oc[2 * r + 1] = rp[2 * r + 1] / rn[r];
            r2h[r]        = rn[r];
            rmax[r]       = std::sqrt(rn[r]);
        }
        coef.store(oc);
        system.check(Ops::axpy(c, P.dev(), Rlo.dev(), P.dev(), lo, n, coef.dev(), -1., nullptr));
        system.check(Ops::axpy(c, P.dev(), Rlo.dev(), P.dev(), lo, n, ones.dev(), 1., nullptr));
        r2.store(r2h);
        active.store(act);
        if (!taken.empty() && graph.g) capture();
    }
    if (pending) {  // stopped by the iteration limit: keep the corrections gathered since the last update
        system.check(Ops::convert(c, T.dev(), CL2_F64, Xlo.dev(), lo, n));
        system.check(Ops::axpy(c, X.dev(), T.dev(), X.dev(), CL2_F64, n, ones.dev(), 1., nullptr));
        rn = true_residual();
    }
    for (int r = 0; r < n; ++r) {
        st[r].reliable_updates   = reliable;
        st[r].graph_nodes        = graph_nodes;
        st[r].graph_kernel_nodes = graph_kernel_nodes;
    }
    const bool done = n_active == 0;
    finish(k);
    if (!done && !partial) throw SolverDidNotSolve(k, "cg_reliable_update_block");
    return k;
}

}  // namespace

int cg_reliable_update_block(const std::vector<const Spinorfield_eo*>& x, const fermionmatrix::Fermionmatrix_eo& A,
                             const Gaugefield& gf, const std::vector<const Spinorfield_eo*>& b,
                             const hardware::System& system, hmc_float rel_tol, const physics::AdditionalParameters& ap,
                             const ReliableUpdateOptions& opt, std::vector<SolveStats>* stats)
{
    BlockStep step;
    if (A.abi_operator() == CL2_OP_QPLUSQMINUS_EO)
        step = [&](const Spinorfield_eo_block* r, const Spinorfield_eo_block& p, const DeviceBuffer& tmp,
                   const Vector<hmc_float>& r2, const Vector<hmc_float>& active, const Vector<hmc_float>& pAp,
                   const Vector<hmc_float>& alpha, const Vector<hmc_float>& norms) {
            return fermionmatrix::cg_step_block(A, r, gf, p, &ap, {}, {}, tmp, r2, active, pAp, alpha, norms);
        };
    return block_solve<WilsonBlockOps>(
        x, b, system, rel_tol,
        [&](const Spinorfield_eo_block* out, const Spinorfield_eo_block& in, const DeviceBuffer& tmp) {
            fermionmatrix::apply_block(A, out, gf, in, ap, tmp);
        },
        step, opt, stats);
}

int cg_reliable_update_block(const std::vector<const Spinorfield_eo*>& x, const fermionmatrix::Fermionmatrix_eo& A,
                             const Gaugefield& gf, const std::vector<const Spinorfield_eo*>& b,
                             const hardware::System& system, hmc_float rel_tol,
                             const std::vector<const physics::AdditionalParameters*>& aps,
                             const ReliableUpdateOptions& opt, std::vector<SolveStats>* stats)
{
    if (aps.size() != b.size()) throw std::invalid_argument("cg_reliable_update_block: one parameter set per right-hand side");
    const bool tm = A.action() == common::twistedmass;
    std::vector<hmc_float> kappa, mubar;
    bool uniform = true;
    for (const auto* p : aps) {
        if (!p) throw std::invalid_argument("cg_reliable_update_block: null parameters");
        kappa.push_back(p->getKappa());
        mubar.push_back(tm ? p->getMubar() : 0.);
        uniform = uniform && kappa.back() == kappa[0] && mubar.back() == mubar[0];
    }
    if (uniform) return cg_reliable_update_block(x, A, gf, b, system, rel_tol, *aps[0], opt, stats);
    BlockStep step;
    if (A.abi_operator() == CL2_OP_QPLUSQMINUS_EO)
        step = [&](const Spinorfield_eo_block* r, const Spinorfield_eo_block& p, const DeviceBuffer& tmp,
                   const Vector<hmc_float>& r2, const Vector<hmc_float>& active, const Vector<hmc_float>& pAp,
                   const Vector<hmc_float>& alpha, const Vector<hmc_float>& norms) {
            return fermionmatrix::cg_step_block(A, r, gf, p, nullptr, kappa, mubar, tmp, r2, active, pAp, alpha,
                                                norms);
        };
    return block_solve<WilsonBlockOps>(
        x, b, system, rel_tol,
        [&](const Spinorfield_eo_block* out, const Spinorfield_eo_block& in, const DeviceBuffer& tmp) {
            fermionmatrix::apply_block(A, out, gf, in, kappa, mubar, tmp);
        },
        step, opt, stats);
}

int cg_reliable_update_block(const std::vector<const Spinorfield_eo*>& x, const fermionmatrix::Fermionmatrix_eo& A,
                             const Gaugefield& gf, const std::vector<const Spinorfield_eo*>& b,
                             const hardware::System& system, hmc_float rel_tol, std::vector<hmc_float> kappa,
                             std::vector<hmc_float> mubar, const ReliableUpdateOptions& opt, BlockRefill& refill,
                             std::vector<SolveStats>* stats)
{
    if (kappa.size() != b.size() || mubar.size() != b.size())
        throw std::invalid_argument("cg_reliable_update_block: one kappa and one mubar per right-hand side");
    if (b.size() < 2) throw std::invalid_argument("cg_reliable_update_block: slot refills need a block of 2 or more");
    if (A.action() != common::twistedmass)
        for (hmc_float m : mubar)
            if (m != 0.) throw std::invalid_argument("cg_reliable_update_block: mubar is for twisted mass");
    // the operator reads kappa and mubar at every call (and at every capture), so refills that change them
    // take effect from the next iteration
    BlockStep step;
    if (A.abi_operator() == CL2_OP_QPLUSQMINUS_EO)
        step = [&](const Spinorfield_eo_block* r, const Spinorfield_eo_block& p, const DeviceBuffer& tmp,
                   const Vector<hmc_float>& r2, const Vector<hmc_float>& active, const Vector<hmc_float>& pAp,
                   const Vector<hmc_float>& alpha, const Vector<hmc_float>& norms) {
            return fermionmatrix::cg_step_block(A, r, gf, p, nullptr, kappa, mubar, tmp, r2, active, pAp, alpha,
                                                norms);
        };
    BlockHooks hooks;
    hooks.converged = [&](int r, const SolveStats& st) { refill.converged(r, st); };
    hooks.refill    = [&](const std::vector<int>& free_slots) {
        std::vector<int> taken = refill.refill(free_slots, kappa, mubar);
        for (int r : taken)
            if (std::find(free_slots.begin(), free_slots.end(), r) == free_slots.end())
                throw std::logic_error("cg_reliable_update_block: a refill took a slot that was not free");
        return taken;
    };
    return block_solve<WilsonBlockOps>(
        x, b, system, rel_tol,
        [&](const Spinorfield_eo_block* out, const Spinorfield_eo_block& in, const DeviceBuffer& tmp) {
            fermionmatrix::apply_block(A, out, gf, in, kappa, mubar, tmp);
        },
        step, opt, stats, false, &hooks);
}

}  // namespace solvers
}  // namespace algorithms
}  // namespace physics

namespace physics {
namespace algorithms {
namespace solvers {

int cg_reliable_update_stagg_block(const std::vector<const Staggeredfield_eo*>& x, const fermionmatrix::MdagM_eo& A,
                                   const Gaugefield& gf, const std::vector<hmc_float>& sigma, const Staggeredfield_eo& b,
                                   const hardware::System& system, hmc_float rel_tol,
                                   const physics::AdditionalParameters& ap, const ReliableUpdateOptions& opt,
                                   std::vector<SolveStats>* stats)
{
    return cg_reliable_update_stagg_block(x, A, gf, sigma, std::vector<const Staggeredfield_eo*>(x.size(), &b), system,
                                          rel_tol, ap, opt, stats);
}

int cg_reliable_update_stagg_block(const std::vector<const Staggeredfield_eo*>& x, const fermionmatrix::MdagM_eo& A,
                                   const Gaugefield& gf, const std::vector<hmc_float>& sigma,
                                   const std::vector<const Staggeredfield_eo*>& b, const hardware::System& system,
                                   hmc_float rel_tol, const physics::AdditionalParameters& ap,
                                   const ReliableUpdateOptions& opt, std::vector<SolveStats>* stats)
{
    if (sigma.size() != x.size() || b.size() != x.size() || x.empty())
        throw std::invalid_argument("cg_reliable_update_stagg_block: one shift and one right-hand side per solution");
    const int n = (int)x.size();
    std::vector<SolveStats> st(n);
    std::vector<int> todo(n);
    for (int r = 0; r < n; ++r) todo[r] = r;
    int total = 0;
    while (!todo.empty()) {
        // the largest supported block of pending shifts; a block returns once at most half of it is still
        // unconverged, and those continue in a smaller block
        int m = 1;
        for (int s : {16, 12, 8, 4, 2})
            if ((int)todo.size() >= s) {
                m = s;
                break;
            }
        if (m == 1) {
            SolveStats one;
            total += cg_reliable_update_stagg(x[todo[0]], A, gf, sigma[todo[0]], *b[todo[0]], system, rel_tol, ap, opt,
                                              &one);
            st[todo[0]].iterations += one.iterations;
            st[todo[0]].reliable_updates += one.reliable_updates;
            st[todo[0]].true_relative_residual = one.true_relative_residual;
            todo.erase(todo.begin());
            continue;
        }
        std::vector<const Staggeredfield_eo*> xs, bs;
        std::vector<hmc_float> sg;
        for (int j = 0; j < m; ++j) {
            xs.push_back(x[todo[j]]);
            bs.push_back(b[todo[j]]);
            sg.push_back(sigma[todo[j]]);
        }
        std::vector<SolveStats> bst;
        total += block_solve<StaggBlockOps>(
            xs, bs, system, rel_tol,
            [&](const Staggeredfield_eo_block* out, const Staggeredfield_eo_block& in, const DeviceBuffer& tmp) {
                fermionmatrix::apply_block(A, out, gf, in, ap, sg, tmp);
            },
            [&](const Staggeredfield_eo_block* r, const Staggeredfield_eo_block& p, const DeviceBuffer& tmp,
                const Vector<hmc_float>& r2, const Vector<hmc_float>& active, const Vector<hmc_float>& pAp,
                const Vector<hmc_float>& alpha, const Vector<hmc_float>& norms) {
                return fermionmatrix::cg_step_block(A, r, gf, p, ap, sg, tmp, r2, active, pAp, alpha, norms);
            },
            opt, &bst, true);
        std::vector<int> next;
        for (int j = 0; j < m; ++j) {
            SolveStats& s = st[todo[j]];
            s.iterations += bst[j].iterations;
            s.reliable_updates += bst[j].reliable_updates;
            s.true_relative_residual = bst[j].true_relative_residual;
            if (bst[j].true_relative_residual > rel_tol) next.push_back(todo[j]);
        }
        next.insert(next.end(), todo.begin() + m, todo.end());
        todo.swap(next);
    }
    if (stats) *stats = st;
    return total;
}

int cg_reliable_update_stagg(const Staggeredfield_eo* x, const fermionmatrix::Fermionmatrix_stagg_eo& A,
                             const Gaugefield& gf, hmc_float sigma, const Staggeredfield_eo& b,
                             const hardware::System& system, hmc_float rel_tol, const physics::AdditionalParameters& ap,
                             const ReliableUpdateOptions& opt, SolveStats* stats)
{
    if (x->precision() != CL2_F64 || b.precision() != CL2_F64)
        throw std::invalid_argument("cg_reliable_update_stagg: x and b are FP64 fields");
    if (opt.delta <= 0. || opt.delta >= 1.) throw std::invalid_argument("cg_reliable_update_stagg: delta must be in (0,1)");
    const int lo = opt.inner_precision;
    if (lo != CL2_F32 && lo != CL2_F64) throw std::invalid_argument("cg_reliable_update_stagg: inner precision FP32 or FP64");
    const Staggeredfield_eo r(system, CL2_F64), t64(system, CL2_F64);
    const Staggeredfield_eo r_lo(system, lo), p(system, lo), v(system, lo), x_lo(system, lo);
    const Scalar<hmc_float> alpha(system), malpha(system), pAp(system), r2(system), rn_s(system), beta(system),
        mbeta(system), zero(system);
    zero.store(0.);
    cl2_ctx* c  = system.ctx();
    auto apply = [&](const Staggeredfield_eo* out, const Staggeredfield_eo& in) {  // out = (A + sigma) in
        A(out, gf, in, &ap);
        if (sigma != 0.) saxpy(out, sigma, in, *out);
    };
    const hmc_float b2     = squarenorm(b);
    const hmc_float target = rel_tol * rel_tol * b2;
    auto true_residual = [&]() {
        apply(&t64, *x);
        saxpby(&r, 1., b, -1., t64);  // r = b - (A + sigma) x
        return squarenorm(r);
    };
    hmc_float rn = true_residual();
    SolveStats st;
    if (b2 == 0. || rn <= target) {
        st.true_relative_residual = b2 > 0. ? std::sqrt(rn / b2) : 0.;
        if (stats) *stats = st;
        return 0;
    }
    convert_precision(&r_lo, r);
    copyData(&p, r_lo);
    x_lo.set_zero();
    r2.store(rn);
    apply(&v, p);  // makes the reduced-precision link copy and MdagM temporary before anything is recorded
    // one inner iteration with every scalar on the device (beta = |r_new|^2 / |r_old|^2); on MdagM_eo the
    // operator half is the fused CG-M step (malpha = -r2 / pAp, r_lo += malpha (A + sigma) p, rn_s = |r_lo|^2)
    const auto* mdagm = dynamic_cast<const fermionmatrix::MdagM_eo*>(&A);
    auto body         = [&]() {
        if (mdagm != nullptr) {
            fermionmatrix::cgm_step(*mdagm, &r_lo, gf, p, ap, sigma, v, r2, pAp, malpha, alpha, rn_s);
        } else {
            apply(&v, p);
            scalar_product_real_part(&pAp, p, v);
            divide(&alpha, r2, pAp);
            subtract(&malpha, zero, alpha);
            saxpy(&r_lo, malpha, v, r_lo);  // r_lo -= alpha v
            squarenorm(&rn_s, r_lo);
        }
        divide(&beta, rn_s, r2);
        subtract(&mbeta, zero, beta);
        add(&r2, rn_s, zero);                       // r2 = |r_new|^2
        cg_update_xp(&x_lo, &p, r_lo, malpha, mbeta);  // x_lo += alpha p, p = r_lo + beta p
    };
    int graph_nodes = 0, graph_kernel_nodes = 0;  // census of the captured iteration
    struct Graph {
        cl2_graph* g = nullptr;
        ~Graph() { cl2_graph_destroy(g); }
    } graph;
    if (opt.use_graph) {
        system.check(cl2_graph_begin(c));
        try {
            body();
        } catch (...) {
            cl2_graph* g = nullptr;
            (void)cl2_graph_end(c, &g);
            cl2_graph_destroy(g);
            throw;
        }
        system.check(cl2_graph_end(c, &graph.g));
        system.check(cl2_graph_node_count(graph.g, &graph_kernel_nodes, &graph_nodes));
        // replays are checked against eager execution for kernel launches only (memset nodes have replayed
        // wrongly on this hardware): a capture that recorded anything else is a bug here
        if (graph_nodes != graph_kernel_nodes) throw std::logic_error("captured solver iteration holds non-kernel nodes");
    }
    hmc_float rmax  = std::sqrt(rn);
    hmc_float rtrue = rn;     // |b - (A + sigma) x|^2 of the FP64 solution
    bool pending    = false;  // x_lo holds a correction not yet added to x
    int k           = 0;
    for (k = 0; k < opt.max_iterations; ++k) {
        if (graph.g)
            system.check(cl2_graph_launch(c, graph.g, 1));
        else
            body();
        pending = true;
        rn      = rn_s.get();
        if (rn != rn) throw SolverStuck(k, "cg_reliable_update_stagg: NAN occured");
        if (std::sqrt(rn) < opt.delta * rmax || rn <= target) {
            convert_precision(&t64, x_lo);
            saxpy(x, 1., t64, *x);  // x += x_lo
            x_lo.set_zero();
            pending = false;
            rn      = true_residual();
            rtrue   = rn;
            ++st.reliable_updates;
            if (rn <= target) {
                ++k;
                break;
            }
            // p = r_iter + beta p_old was formed with the iterated residual: rebuild it on the true one,
            // p = r_true + (beta p_old with its component along r_true removed)
            saxpy(&p, -1., r_lo, p);  // beta p_old = p - r_iter
            convert_precision(&r_lo, r);
            const hmc_complex rp = scalar_product(r_lo, p);
            saxpy(&p, hmc_complex{-rp.re / rn, -rp.im / rn}, r_lo, p);
            saxpy(&p, 1., r_lo, p);  // + r_true
            r2.store(rn);
            rmax = std::sqrt(rn);
        }
        rmax = std::max(rmax, std::sqrt(rn));
    }
    if (pending) {  // stopped by the iteration limit: keep the correction gathered since the last update
        convert_precision(&t64, x_lo);
        saxpy(x, 1., t64, *x);
        rtrue = true_residual();
    }
    st.iterations             = k;
    st.true_relative_residual = std::sqrt(rtrue / b2);
    st.graph_nodes            = graph_nodes;
    st.graph_kernel_nodes     = graph_kernel_nodes;
    if (stats) *stats = st;
    if (st.true_relative_residual > rel_tol) throw SolverDidNotSolve(k, "cg_reliable_update_stagg");
    return k;
}

int cg_m_mixed(const std::vector<std::shared_ptr<Staggeredfield_eo>>& x, const fermionmatrix::Fermionmatrix_stagg_eo& A,
               const Gaugefield& gf, const std::vector<hmc_float>& sigma, const Staggeredfield_eo& b,
               const hardware::System& system, hmc_float rel_tol, const physics::AdditionalParameters& ap,
               const MultishiftOptions& opt, std::vector<SolveStats>* stats, int* multishift_iterations)
{
    const int n = (int)sigma.size();
    if (n < 1 || n > 64 || (int)x.size() != n) throw std::invalid_argument("cg_m_mixed: 1 to 64 shifts, one x per shift");
    if (b.precision() != CL2_F64) throw std::invalid_argument("cg_m_mixed: b is an FP64 field");
    for (const auto& xi : x)
        if (xi->precision() != CL2_F64) throw std::invalid_argument("cg_m_mixed: x are FP64 fields");
    const int lo = opt.inner_precision;
    const hmc_float b2 = squarenorm(b);
    std::vector<SolveStats> st(n);
    if (b2 == 0.) {
        for (auto& xi : x) xi->set_zero();
        if (stats) *stats = st;
        if (multishift_iterations) *multishift_iterations = 0;
        return 0;
    }
    // ---- phase 1: CG-M on inner-precision r and p_k, solutions accumulated in FP64, fused vector updates ----
    const Staggeredfield_eo b_lo(system, lo), r(system, lo), p(system, lo), v(system, lo);
    std::vector<std::unique_ptr<Staggeredfield_eo>> ps;
    std::vector<void*> xp(n), pp(n);
    for (int k = 0; k < n; ++k) {
        ps.emplace_back(new Staggeredfield_eo(system, lo));
        xp[k] = x[k]->dev();
        pp[k] = ps[k]->dev();
    }
    const Vector<hmc_float> zeta_prev(n, system), zeta(n, system), zeta_foll(n, system), alpha_vec(n, system),
        beta_vec(n, system), shift(n, system), active(n, system);
    const Scalar<hmc_float> alpha_scalar_prev(system), alpha_scalar(system), beta_scalar_prev(system),
        beta_scalar(system), tmp1(system), tmp2(system), tmp3(system);
    zeta_prev.store(std::vector<hmc_float>(n, 1.));
    zeta.store(std::vector<hmc_float>(n, 1.));
    alpha_vec.store(std::vector<hmc_float>(n, 0.));
    shift.store(sigma);
    std::vector<hmc_float> act(n, 1.);
    active.store(act);
    convert_precision(&b_lo, b);
    for (int k = 0; k < n; ++k) {
        x[k]->set_zero();
        copyData(ps[k].get(), b_lo);
    }
    copyData(&r, b_lo);
    copyData(&p, b_lo);
    scalar_product_real_part(&tmp1, r, r);
    beta_scalar.store(1.0);
    alpha_scalar.store(0.0);
    const hmc_float target1 = opt.sloppy_tol * opt.sloppy_tol * b2;
    cl2_ctx* c              = system.ctx();
    {  // upload the multi-shift pointer table (an update with every shift masked off) before recording
        const Vector<hmc_float> none(n, system);
        none.store(std::vector<hmc_float>(n, 0.));
        system.check(cl2_multishift_update_stagg_eo(c, xp.data(), CL2_F64, pp.data(), r.dev(), lo, n, beta_vec.dev(),
                                                    zeta_foll.dev(), alpha_vec.dev(), none.dev()));
        A(&v, gf, p, &ap);  // reduced-precision links and MdagM temporary
        system.synchronize();
    }
    // one CG-M iteration (solver_shifted.cpp:136-198), every scalar on the device: the reference's copies,
    // ratio, subtraction and zeta/beta/alpha updates run as two fused launches with the same values
    // (beta_prev = beta, beta = -tmp1/tmp3; then alpha_prev = alpha, alpha = tmp2/tmp1, the per-shift
    // recurrences, zeta_prev = zeta, zeta = zeta_foll, tmp1 = tmp2)
    const auto* mdagm = dynamic_cast<const fermionmatrix::MdagM_eo*>(&A);
    auto body         = [&]() {
        if (mdagm != nullptr) {  // the same from two D_KS_eo kernels: A p never stored, no reduction passes
            fermionmatrix::cgm_step(*mdagm, &r, gf, p, ap, 0., v, tmp1, tmp3, beta_scalar, beta_scalar_prev, tmp2);
        } else {
            A(&v, gf, p, &ap);
            scalar_product_real_part(&tmp3, p, v);
            system.check(cl2_cgm_beta_step(c, beta_scalar.dev(), beta_scalar_prev.dev(), tmp1.dev(), tmp3.dev()));
            saxpy(&r, beta_scalar, v, r);  // r = beta v + r
            scalar_product_real_part(&tmp2, r, r);
        }
        system.check(cl2_cgm_alpha_step(c, n, alpha_scalar.dev(), alpha_scalar_prev.dev(), tmp2.dev(), tmp1.dev(),
                                        beta_scalar.dev(), beta_scalar_prev.dev(), shift.dev(), zeta.dev(),
                                        zeta_prev.dev(), zeta_foll.dev(), beta_vec.dev(), alpha_vec.dev()));
        saxpy(&p, alpha_scalar, p, r);  // p = alpha p + r
        system.check(cl2_multishift_update_stagg_eo(c, xp.data(), CL2_F64, pp.data(), r.dev(), lo, n, beta_vec.dev(),
                                                    zeta_foll.dev(), alpha_vec.dev(), active.dev()));
    };
    int graph_nodes = 0, graph_kernel_nodes = 0;  // census of the captured iteration
    struct Graph {
        cl2_graph* g = nullptr;
        ~Graph() { cl2_graph_destroy(g); }
    } graph;
    if (opt.use_graph) {
        system.check(cl2_graph_begin(c));
        try {
            body();
        } catch (...) {
            cl2_graph* g = nullptr;
            (void)cl2_graph_end(c, &g);
            cl2_graph_destroy(g);
            throw;
        }
        system.check(cl2_graph_end(c, &graph.g));
        system.check(cl2_graph_node_count(graph.g, &graph_kernel_nodes, &graph_nodes));
        // replays are checked against eager execution for kernel launches only (memset nodes have replayed
        // wrongly on this hardware): a capture that recorded anything else is a bug here
        if (graph_nodes != graph_kernel_nodes) throw std::logic_error("captured solver iteration holds non-kernel nodes");
    }
    int n_active = n, iter = 0;
    while (iter < opt.max_iterations && n_active > 0) {
        // ten iterations between convergence checks, as cgIterationBlockSize
        const int steps = std::min(10, opt.max_iterations - iter);
        if (graph.g)
            system.check(cl2_graph_launch(c, graph.g, steps));
        else
            for (int s = 0; s < steps; ++s) body();
        iter += steps;
        const hmc_float rr = tmp2.get();  // shift k's residual is zeta_k r
        if (rr != rr) throw SolverStuck(iter, "cg_m_mixed: NAN occured");
        const std::vector<hmc_float> z = zeta.get();
        bool changed                   = false;
        for (int k = 0; k < n; ++k)
            if (act[k] != 0. && z[k] * z[k] * rr < target1) {
                act[k] = 0.;
                --n_active;
                changed = true;
            }
        if (changed) active.store(act);
    }
    if (multishift_iterations) *multishift_iterations = iter;
    // ---- phase 2: refine every shift in FP64 ----
    ReliableUpdateOptions ro;
    ro.inner_precision = lo;
    ro.delta           = opt.delta;
    ro.max_iterations  = opt.max_iterations;
    ro.use_graph       = opt.use_graph;
    int total          = iter;
    if (mdagm != nullptr) {
        std::vector<const Staggeredfield_eo*> xs;
        for (const auto& xi : x) xs.push_back(xi.get());
        total += cg_reliable_update_stagg_block(xs, *mdagm, gf, sigma, b, system, rel_tol, ap, ro, &st);
    } else {
        for (int k = 0; k < n; ++k)
            total += cg_reliable_update_stagg(x[k].get(), A, gf, sigma[k], b, system, rel_tol, ap, ro, &st[k]);
    }
    if (stats) *stats = st;
    return total;
}

}  // namespace solvers
}  // namespace algorithms
}  // namespace physics
