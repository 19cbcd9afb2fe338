// Point-source inversion and the pseudoscalar correlator
// (physics/algorithms/inversion.cpp:44-242, physics/sources.cpp:35-57,
// physics/observables/wilsonTwoFlavourCorrelators.cpp:263-388).
#pragma once

#include <vector>

#include "solvers/solvers.hpp"

namespace physics {
namespace algorithms {

/* phi = M^-1 eta for the Wilson or twisted-mass operator M = M_site + D (M_site = 1 for Wilson,
 * 1 + i mubar gamma5 for twisted mass), eo preconditioned on the odd sublattice:
 *   b'_o = eta_o - D_eo M_site^-1 eta_e,  Aee phi_o = b'_o,  phi_e = M_site^-1 (eta_e - D_oe phi_o)
 * with BiCGStab on Aee (the reference's default) or, for solver=cg, CG on QplusQminus_eo with source
 * gamma5 b'_o followed by phi_o = Qminus x.  eta and phi are eo pairs (even, odd); prec is the solver's
 * absolute stopping value |r|^2 < prec. */
int invert_eo(const lattices::Spinorfield_eo* phi_even, const lattices::Spinorfield_eo* phi_odd,
              const lattices::Gaugefield& gf, const lattices::Spinorfield_eo& eta_even,
              const lattices::Spinorfield_eo& eta_odd, const hardware::System& system,
              physics::InterfacesHandler& interfacesHandler, hmc_float prec,
              const physics::AdditionalParameters& additionalParameters);

/* the 12 spin-colour point sources at (sourceX, sourceY, sourceZ, sourceT) (physics/sources.cpp:35-57,
 * ocl_kernel/spinorfield_point_source.cl, content 'one'), inverted with invert_eo (prec: the parameters'
 * solverPrecision), and the pseudoscalar correlator along correlatorDirection (0: t, 3: z, the default)
 *   C(t) = (2 kappa)^2 / NS^3 sum_k sum_{x,y,z} |phi_k(x, y, z, t)|^2      (NT entries)
 *   C(z) = (2 kappa)^2 / (NS^2 NT) sum_k sum_{x,y,t} |phi_k(x, y, z, t)|^2  (NS entries)
 * (ocl_kernel/fermionobservables/correlator_ps_point.cl correlator_ps_t, correlator_ps_z) */
std::vector<hmc_float> point_source_ps_correlator(const lattices::Gaugefield& gf, const hardware::System& system,
                                                  physics::InterfacesHandler& interfacesHandler,
                                                  const physics::AdditionalParameters& additionalParameters);

}  // namespace algorithms
}  // namespace physics
