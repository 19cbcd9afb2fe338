// Heatbath, overrelaxation and plaquette on the C ABI.
#include "../include/physics/algorithms/su3heatbath.hpp"
#include "../include/physics/observables/gaugeObservables.hpp"

#include <stdexcept>

namespace {
void require_18(const physics::lattices::Gaugefield& gf)
{
    if (gf.format() != CL2_GAUGE_F64_18)
        throw std::logic_error("gauge updates and observables need the 18-real link format (useReconstruct12=false)");
}
}  // namespace

void physics::algorithms::su3heatbath(physics::lattices::Gaugefield& gf, physics::PRNG& prng, int overrelax,
                                      const std::set<int>& fixed_timeslices)
{
    if (overrelax < 0) throw std::invalid_argument("negative number of overrelaxation steps");
    if (!fixed_timeslices.empty()) throw std::logic_error("heatbath with fixed timeslices is not implemented");
    require_18(gf);
    const hardware::System& sys = gf.get_system();
    sys.check(cl2_gauge_heatbath(sys.ctx(), gf.dev_mutable(), sys.get_inputparameters().get_beta(), 1, overrelax, prng.next_seed()));
}

void physics::algorithms::overrelax(physics::lattices::Gaugefield& gf, physics::PRNG& /*prng*/, int steps)
{
    if (steps <= 0) throw std::invalid_argument("number of overrelaxation steps must be positive");
    require_18(gf);
    const hardware::System& sys = gf.get_system();
    sys.check(cl2_gauge_overrelax(sys.ctx(), gf.dev_mutable(), steps));
}

hmc_float physics::observables::measurePlaquetteWithoutNormalization(const physics::lattices::Gaugefield* gf)
{
    require_18(*gf);
    const hardware::System& sys = gf->get_system();
    double p = 0.;
    sys.check(cl2_gauge_plaquette(sys.ctx(), gf->dev(), &p));
    return p * 18. * 2. * sys.eo_sites();
}

hmc_float physics::observables::measurePlaquette(const physics::lattices::Gaugefield* gf)
{
    require_18(*gf);
    const hardware::System& sys = gf->get_system();
    double p = 0.;
    sys.check(cl2_gauge_plaquette(sys.ctx(), gf->dev(), &p));
    return p;
}
