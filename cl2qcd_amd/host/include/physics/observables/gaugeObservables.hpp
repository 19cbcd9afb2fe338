// Plaquette of a gauge field (physics/observables/gaugeObservables.hpp:47-51).
#pragma once

#include "../lattices/lattices.hpp"

namespace physics {
namespace observables {

/* (1/(18 V)) sum_x sum_{mu<nu} Re tr P_mu,nu(x): 1 on a cold field */
hmc_float measurePlaquette(const physics::lattices::Gaugefield* gf);
/* sum_x sum_{mu<nu} Re tr P_mu,nu(x) */
hmc_float measurePlaquetteWithoutNormalization(const physics::lattices::Gaugefield* gf);

}  // namespace observables
}  // namespace physics
