// SU(3) heatbath / overrelaxation updates of a gauge field
// (physics/algorithms/su3heatbath.hpp:36-48).  The field must use the 18-real link format.
#pragma once

#include "../lattices/lattices.hpp"

#include <set>

namespace physics {
namespace algorithms {

/* one heatbath sweep over all links, then `overrelax` overrelaxation sweeps
 * (physics/algorithms/su3heatbath.cpp:33-59); fixed time slices are not supported and must be empty */
void su3heatbath(physics::lattices::Gaugefield& gf, physics::PRNG& prng, int overrelax = 0,
                 const std::set<int>& fixed_timeslices = std::set<int>());

/* `steps` overrelaxation sweeps (physics/algorithms/su3heatbath.cpp:61-73) */
void overrelax(physics::lattices::Gaugefield& gf, physics::PRNG& prng, int steps = 1);

}  // namespace algorithms
}  // namespace physics
