// Inversion of the Wilson / twisted-mass operator on sources and the pseudoscalar correlator
// (physics/algorithms/inversion.cpp:44-242, physics/sources.cpp:35-160,
// physics/observables/wilsonTwoFlavourCorrelators.cpp:263-388).
#pragma once

#include <vector>

#include "../sources.hpp"
#include "solvers/solvers.hpp"

namespace physics {
namespace algorithms {

/* what invert_eo did */
struct InversionStats {
    int iterations          = 0;      // of the solver whose solution was kept
    int bicgstab_iterations = 0;      // BiCGStab's count, also when it threw and CG took over
    bool cg_fallback        = false;  // solver=cg: BiCGStab threw a SolverException, CG on QplusQminus_eo solved
};

/* phi = M^-1 eta for the Wilson or twisted-mass operator M = M_site + D (M_site = 1 for Wilson,
 * 1 + i mubar gamma5 for twisted mass), eo preconditioned on the odd sublattice as
 * invert_M_nf2_upperflavour (physics/algorithms/inversion.cpp:126-228):
 *   b = D_eo M_site^-1 eta_e - eta_o   (the reference's sign, so its solvers iterate identically),
 *   Aee x = b from the cold start, phi_o = -x, phi_e = M_site^-1 (eta_e - D_oe phi_o).
 * solver=bicgstab / bicgstab_save: BiCGStab on Aee.  solver=cg: BiCGStab (bicgstab_fast) on Aee first; if
 * it throws a SolverException, CG on QplusQminus_eo with source gamma5 b from BiCGStab's last iterate,
 * then x = Qminus y (inversion.cpp:175-195).  eta and phi are eo pairs (even, odd); prec is the solvers'
 * absolute stopping value |r|^2 < prec.  Unlike the reference, whose result is -M^-1 eta (its even/odd
 * swap workaround), phi is M^-1 eta.  Returns the iterations of the solver that produced x. */
int invert_eo(const lattices::Spinorfield_eo* phi_even, const lattices::Spinorfield_eo* phi_odd,
              const lattices::Gaugefield& gf, const lattices::Spinorfield_eo& eta_even,
              const lattices::Spinorfield_eo& eta_odd, const hardware::System& system,
              physics::InterfacesHandler& interfacesHandler, hmc_float prec,
              const physics::AdditionalParameters& additionalParameters, InversionStats* stats = nullptr);

/* No counterpart in the reference: the inverter's sources solved together.  The Schur sources of up to
 * `block` eta pairs go through cg_reliable_update_block on QplusQminus_eo (sources gamma5 b_k, one fused
 * multi-RHS operator per iteration, options.inner_precision inner loop), each to
 * |gamma5 b_k - QplusQminus_eo y_k| <= rel_tol |b_k| in FP64; then phi_o = -Qminus y_k and phi_e as above. */
struct BlockInversionOptions {
    solvers::ReliableUpdateOptions solver;
    hmc_float rel_tol = 1e-10;
    int block         = 12;  // right-hand sides per lockstep block (1, 2, 4, 8, 12 or 16)
};
void invert_eo_block(const std::vector<const lattices::Spinorfield_eo*>& phi_even,
                     const std::vector<const lattices::Spinorfield_eo*>& phi_odd, const lattices::Gaugefield& gf,
                     const std::vector<const lattices::Spinorfield_eo*>& eta_even,
                     const std::vector<const lattices::Spinorfield_eo*>& eta_odd, const hardware::System& system,
                     physics::InterfacesHandler& interfacesHandler,
                     const physics::AdditionalParameters& additionalParameters, const BlockInversionOptions& options,
                     std::vector<solvers::SolveStats>* stats = nullptr);

/* the pseudoscalar correlator along correlatorDirection (0: t, 3: z, the default) summed over the nSources
 * sources of sourceType / sourceContent (create_sources), each inverted with invert_eo at the parameters'
 * solverPrecision (or, with `block`, with invert_eo_block):
 *   C(t) = (2 kappa)^2 / NS^3 sum_k sum_{x,y,z} |phi_k(x, y, z, t)|^2               (NT entries)
 *   C(z) = (2 kappa)^2 / (NS^2 NT) sum_k w_k sum_{x,y,t} |phi_k(x, y, z, t)|^2     (NS entries)
 * with w_k = 1 for point sources (ocl_kernel/fermionobservables/correlator_ps_point.cl) and w_k = |eta_k|^2
 * over the lattice for the other sources (ocl_kernel/fermionobservables_correlators_stochastic.cl:78-180,
 * whose double sum over the sites of the slice and of the source factorises this way). */
std::vector<hmc_float> ps_correlator(const lattices::Gaugefield& gf, const hardware::System& system,
                                     physics::InterfacesHandler& interfacesHandler,
                                     const physics::AdditionalParameters& additionalParameters, const PRNG& prng,
                                     const BlockInversionOptions* block = nullptr);

/* ps_correlator for the 12 spin-colour point sources at (sourceX, sourceY, sourceZ, sourceT)
 * (sourceType point, nSources 12) */
std::vector<hmc_float> point_source_ps_correlator(const lattices::Gaugefield& gf, const hardware::System& system,
                                                  physics::InterfacesHandler& interfacesHandler,
                                                  const physics::AdditionalParameters& additionalParameters,
                                                  const BlockInversionOptions* block = nullptr);

}  // namespace algorithms
}  // namespace physics
