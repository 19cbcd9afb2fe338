// physics::AdditionalParameters (physics/additionalParameters.hpp:29-37): the
// operator parameters handed to every fermion-matrix call; getters a fermion
// type does not define throw.
#pragma once

#include <stdexcept>

#include "../common_types.hpp"

namespace physics {

class AdditionalParameters {
  public:
    virtual ~AdditionalParameters() = default;
    virtual hmc_float getKappa() const { throw std::logic_error("AdditionalParameters: no kappa"); }
    virtual hmc_float getMubar() const { throw std::logic_error("AdditionalParameters: no mubar"); }
    virtual hmc_float getMass() const { throw std::logic_error("AdditionalParameters: no mass"); }
    virtual bool getConservative() const { throw std::logic_error("AdditionalParameters: no conservative flag"); }
};

/* interfaceImplementations/physicsParameters.hpp:35-63 */
class WilsonAdditionalParameters : public AdditionalParameters {
  public:
    WilsonAdditionalParameters(hmc_float kappa, hmc_float mubar) : kappa(kappa), mubar(mubar) {}
    hmc_float getKappa() const override { return kappa; }
    hmc_float getMubar() const override { return mubar; }

  private:
    hmc_float kappa, mubar;
};

class StaggeredAdditionalParameters : public AdditionalParameters {
  public:
    StaggeredAdditionalParameters(hmc_float mass, bool conservative = false) : mass(mass), conservative(conservative)
    {
    }
    hmc_float getMass() const override { return mass; }
    bool getConservative() const override { return conservative; }

  private:
    hmc_float mass;
    bool conservative;
};

}  // namespace physics
