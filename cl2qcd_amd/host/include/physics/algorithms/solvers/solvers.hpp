// Krylov solvers of the fermion-inversion path: cg (solvers/cg.hpp:41-56),
// bicgstab (solvers/bicgstab.hpp), cg_m (solver_shifted.hpp:48-52) with the
// reference's semantics (absolute |r|^2 < prec, checked every
// cgIterationBlockSize iterations, restart every iterRefresh), exceptions
// (solvers/exceptions.hpp:35-64), plus the reduced-precision and batched
// variants this build adds on top.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../fermionmatrix/fermionmatrix.hpp"

namespace physics {
namespace algorithms {
namespace solvers {

class SolverException : public std::runtime_error {
  public:
    SolverException(const std::string& what, int iter) : std::runtime_error(what), iter(iter) {}
    int get_iterations() const noexcept { return iter; }

  private:
    int iter;
};
/* NaN residual or vanishing rho */
class SolverStuck : public SolverException {
  public:
    SolverStuck(int iter, const std::string& where)
        : SolverException("Solver got stuck after " + std::to_string(iter) + " iterations (" + where + ")", iter)
    {
    }
};
/* maximum iteration count reached */
class SolverDidNotSolve : public SolverException {
  public:
    SolverDidNotSolve(int iter, const std::string& where)
        : SolverException("Solver did not solve in " + std::to_string(iter) + " iterations (" + where + ")", iter)
    {
    }
};

/* CG on a hermitian eo operator; returns the iteration index at convergence */
int cg(const lattices::Spinorfield_eo* x, const fermionmatrix::Fermionmatrix_eo& A, const lattices::Gaugefield& gf,
       const lattices::Spinorfield_eo& b, const hardware::System& system, physics::InterfacesHandler& interfacesHandler,
       hmc_float prec, const physics::AdditionalParameters& additionalParameters);

/* BiCGStab (bicgstab_fast, or bicgstab_save when the solver parameter is bicgstab_save) */
int bicgstab(const lattices::Spinorfield_eo* x, const fermionmatrix::Fermionmatrix_eo& A,
             const lattices::Gaugefield& gf, const lattices::Spinorfield_eo& b, const hardware::System& system,
             physics::InterfacesHandler& interfacesHandler, hmc_float prec,
             const physics::AdditionalParameters& additionalParameters);

/* multi-shift CG for (A + sigma_k) x_k = b */
int cg_m(std::vector<std::shared_ptr<lattices::Staggeredfield_eo>> x, const fermionmatrix::Fermionmatrix_stagg_eo& A,
         const lattices::Gaugefield& gf, const std::vector<hmc_float> sigma, const lattices::Staggeredfield_eo& b,
         const hardware::System& system, physics::InterfacesHandler& interfacesHandler, hmc_float prec,
         const physics::AdditionalParameters& additionalParameters);

/* ---- reduced precision (no reference counterpart: the reference is FP64 only) ---- */

struct ReliableUpdateOptions {
    int inner_precision = CL2_F32;  // precision of the Dslash and Krylov vectors of the inner loop
    double delta        = 0.1;      // FP64 reliable update when |r| drops by this factor
    int max_iterations  = 100000;
    bool use_graph      = true;     // record one inner iteration and replay it (same arithmetic either way)
};

struct SolveStats {
    int iterations       = 0;
    int reliable_updates = 0;
    double true_relative_residual = 0.;  // |b - A x| / |b| recomputed in FP64 at exit
    int graph_nodes        = 0;  // nodes of the captured inner iteration (0: not captured)
    int graph_kernel_nodes = 0;  // of which kernel launches
};

/* CG with an FP64 reliable-update outer loop (solution accumulation, true residual,
 * reductions and scalar recurrences in FP64) around an inner CG whose operator and
 * vectors are in inner_precision.  Converged when the FP64 true residual satisfies
 * |b - A x| <= rel_tol |b|.  Returns the total number of inner iterations. */
int cg_reliable_update(const lattices::Spinorfield_eo* x, const fermionmatrix::Fermionmatrix_eo& A,
                       const lattices::Gaugefield& gf, const lattices::Spinorfield_eo& b,
                       const hardware::System& system, hmc_float rel_tol,
                       const physics::AdditionalParameters& additionalParameters, const ReliableUpdateOptions& options,
                       SolveStats* stats = nullptr);

/* Reliable-update CG for one shifted staggered system (A + sigma) x = b with A hermitian positive (MdagM_eo):
 * inner recurrence in options.inner_precision (FP32 or FP64) with the operator in that precision, FP64
 * solution accumulation and true residual; x is the initial guess and the result (FP64).  Converged when
 * |b - (A + sigma) x| <= rel_tol |b| in FP64. */
int cg_reliable_update_stagg(const lattices::Staggeredfield_eo* x, const fermionmatrix::Fermionmatrix_stagg_eo& A,
                             const lattices::Gaugefield& gf, hmc_float sigma, const lattices::Staggeredfield_eo& b,
                             const hardware::System& system, hmc_float rel_tol,
                             const physics::AdditionalParameters& additionalParameters,
                             const ReliableUpdateOptions& options, SolveStats* stats = nullptr);

/* The shifted staggered systems (A + sigma[r]) x[r] = b of one right-hand side b, A = MdagM_eo, any number of
 * shifts, solved in lockstep blocks with the reliable-update CG of cg_reliable_update_block: every inner
 * iteration applies A to all shifts of a block (2, 4, 8, 12 or 16) with one pass over the links per D_KS_eo,
 * each shift keeps its own alpha and beta and is frozen when |b - (A + sigma[r]) x[r]| <= rel_tol |b| in FP64;
 * once at most half of a block is unconverged, those shifts continue (restarted from their solutions) in the
 * next smaller block, a last single one with cg_reliable_update_stagg.  x[r] (FP64) are initial guesses and
 * results.  Returns the lockstep iterations summed over the blocks. */
int cg_reliable_update_stagg_block(const std::vector<const lattices::Staggeredfield_eo*>& x,
                                   const fermionmatrix::MdagM_eo& A, const lattices::Gaugefield& gf,
                                   const std::vector<hmc_float>& sigma, const lattices::Staggeredfield_eo& b,
                                   const hardware::System& system, hmc_float rel_tol,
                                   const physics::AdditionalParameters& additionalParameters,
                                   const ReliableUpdateOptions& options, std::vector<SolveStats>* stats = nullptr);
/* the same with a right-hand side per system: (A + sigma[r]) x[r] = b[r] */
int cg_reliable_update_stagg_block(const std::vector<const lattices::Staggeredfield_eo*>& x,
                                   const fermionmatrix::MdagM_eo& A, const lattices::Gaugefield& gf,
                                   const std::vector<hmc_float>& sigma,
                                   const std::vector<const lattices::Staggeredfield_eo*>& b,
                                   const hardware::System& system, hmc_float rel_tol,
                                   const physics::AdditionalParameters& additionalParameters,
                                   const ReliableUpdateOptions& options, std::vector<SolveStats>* stats = nullptr);

struct MultishiftOptions {
    int inner_precision = CL2_F32;  // precision of the multi-shift phase and of the refinement's inner loop
    double sloppy_tol   = 1e-7;     // relative residual of each shift at which the multi-shift phase stops
    double delta        = 0.1;      // reliable-update factor of the refinement
    int max_iterations  = 100000;
    bool use_graph      = true;     // record one CG-M iteration and replay it (same arithmetic either way)
};

/* Multi-shift CG for (A + sigma_k) x_k = b in two phases: the reference's CG-M recurrences
 * (solver_shifted.cpp:114-475) on inner-precision fields with all shifts' vector updates fused into one
 * pass per iteration, until every shift's residual zeta_k r is below sloppy_tol |b|; then the x_k are
 * refined to |b - (A + sigma_k) x_k| <= rel_tol |b| in FP64: for A = MdagM_eo in lockstep blocks of up to 16
 * shifts (cg_reliable_update_stagg_block), else one by one with cg_reliable_update_stagg.  Returns the
 * multi-shift iterations plus the refinement's operator applications (lockstep iterations of a block count
 * once); stats[k] holds shift k's refinement. */
int cg_m_mixed(const std::vector<std::shared_ptr<lattices::Staggeredfield_eo>>& x,
               const fermionmatrix::Fermionmatrix_stagg_eo& A, const lattices::Gaugefield& gf,
               const std::vector<hmc_float>& sigma, const lattices::Staggeredfield_eo& b, const hardware::System& system,
               hmc_float rel_tol, const physics::AdditionalParameters& additionalParameters,
               const MultishiftOptions& options, std::vector<SolveStats>* stats = nullptr, int* multishift_iterations = nullptr);

/* The reliable-update CG of cg_reliable_update on a block of right-hand sides b[r] (FP64 fields, nrhs = 1, 2,
 * 4, 8, 12 or 16), solved in lockstep: every inner iteration applies A once to the whole block with the
 * multi-right-hand-side operator, each right-hand side keeps its own alpha and beta (device arrays), a
 * reliable update is made for all right-hand sides when one of them has dropped by delta, and a
 * right-hand side whose FP64 true residual meets rel_tol is frozen.  Inner precision FP32, FP16 (per-site norm) or FP64.
 * Returns the number of lockstep iterations; stats (if given) receives one entry per right-hand side. */
int cg_reliable_update_block(const std::vector<const lattices::Spinorfield_eo*>& x,
                             const fermionmatrix::Fermionmatrix_eo& A, const lattices::Gaugefield& gf,
                             const std::vector<const lattices::Spinorfield_eo*>& b, const hardware::System& system,
                             hmc_float rel_tol, const physics::AdditionalParameters& additionalParameters,
                             const ReliableUpdateOptions& options, std::vector<SolveStats>* stats = nullptr);
/* the same with the operator's parameters per right-hand side (kappa and mubar of a twisted-mass or Wilson
 * operator; parameters[r] for b[r]): right-hand sides with different quark masses on one gauge field are
 * solved in one block, sharing every link load */
int cg_reliable_update_block(const std::vector<const lattices::Spinorfield_eo*>& x,
                             const fermionmatrix::Fermionmatrix_eo& A, const lattices::Gaugefield& gf,
                             const std::vector<const lattices::Spinorfield_eo*>& b, const hardware::System& system,
                             hmc_float rel_tol, const std::vector<const physics::AdditionalParameters*>& parameters,
                             const ReliableUpdateOptions& options, std::vector<SolveStats>* stats = nullptr);

/* Continuous batching of a lockstep block (no counterpart in the reference): at every reliable update the
 * block solver reports the right-hand sides that have converged since the last one and offers its free
 * slots (converged, or zero from the start) to new problems, which start from x = 0 in those slots while
 * the others continue. */
struct BlockRefill {
    virtual ~BlockRefill() = default;
    /* right-hand side r has converged: x[r] holds its final FP64 solution; st counts from the slot's start */
    virtual void converged(int r, const SolveStats& st) = 0;
    /* Offers the free slots.  For each slot it takes, the callee writes the new right-hand side to b[r],
     * zeroes x[r] and sets kappa[r], mubar[r]; returns the slots taken (empty: none). */
    virtual std::vector<int> refill(const std::vector<int>& free_slots, std::vector<hmc_float>& kappa,
                                    std::vector<hmc_float>& mubar) = 0;
};
/* cg_reliable_update_block with kappa[r], mubar[r] per right-hand side (twisted mass or Wilson) and slot
 * refills through `refill`; runs until every right-hand side, the refilled ones included, has converged.
 * Returns the lockstep iterations. */
int cg_reliable_update_block(const std::vector<const lattices::Spinorfield_eo*>& x,
                             const fermionmatrix::Fermionmatrix_eo& A, const lattices::Gaugefield& gf,
                             const std::vector<const lattices::Spinorfield_eo*>& b, const hardware::System& system,
                             hmc_float rel_tol, std::vector<hmc_float> kappa, std::vector<hmc_float> mubar,
                             const ReliableUpdateOptions& options, BlockRefill& refill,
                             std::vector<SolveStats>* stats = nullptr);

}  // namespace solvers
}  // namespace algorithms
}  // namespace physics
