// Parameter interfaces handed to lattices, operators and solvers
// (physics/interfacesHandler.hpp:130-233, physics/fermionmatrix/fermionmatrixInterfaces.hpp:30-35,
// physics/algorithms/algorithmsInterfaces.hpp:30-39), implemented on meta::Inputparameters.
#pragma once

#include <memory>

#include "../meta/inputparameters.hpp"
#include "additionalParameters.hpp"

namespace hardware {
class System;
}

namespace physics {

namespace fermionmatrix {
class FermionmatrixParametersInterface {
  public:
    virtual ~FermionmatrixParametersInterface() = default;
    virtual common::action getFermionicActionType() const = 0;
    virtual bool useMergedFermionicKernels() const = 0;
};
class FermionEoParametersInterface : public virtual FermionmatrixParametersInterface {};
class FermionStaggeredEoParametersInterface : public virtual FermionmatrixParametersInterface {};
}  // namespace fermionmatrix

namespace lattices {
class GaugefieldParametersInterface {
  public:
    virtual ~GaugefieldParametersInterface() = default;
    virtual int getNs() const = 0;
    virtual int getNt() const = 0;
    virtual common::startcondition getStartcondition() const = 0;
    virtual std::string getSourcefileName() const = 0;
    virtual bool useReconstruct12() const = 0;
};
class SpinorfieldParametersInterface {
  public:
    virtual ~SpinorfieldParametersInterface() = default;
    virtual size_t getNumberOfElements() const = 0;
};
class SpinorfieldEoParametersInterface {
  public:
    virtual ~SpinorfieldEoParametersInterface() = default;
    virtual size_t getEoprecSpinorfieldSize() const = 0;
};
class StaggeredfieldEoParametersInterface {
  public:
    virtual ~StaggeredfieldEoParametersInterface() = default;
    virtual size_t getEoprecSpinorfieldSize() const = 0;
};
}  // namespace lattices

namespace algorithms {
class SolversParametersInterface {
  public:
    virtual ~SolversParametersInterface() = default;
    virtual int getCgMax() const = 0;
    virtual int getIterRefresh() const = 0;
    virtual common::solver getSolver() const = 0;
    virtual int getCgIterationBlockSize() const = 0;
    virtual int getCgMinimumIterationCount() const = 0;
    virtual bool getCgUseAsyncCopy() const = 0;
    virtual bool getUseMergeKernelsSpinor() const = 0;
};
}  // namespace algorithms

namespace lattices {
class Gaugefield;
class Spinorfield;
class Spinorfield_eo;
class Staggeredfield_eo;
}  // namespace lattices
namespace fermionmatrix {
class Aee;
class Aee_AND_gamma5_eo;
class Aee_minus;
class Aee_minus_AND_gamma5_eo;
class Qplus_eo;
class Qminus_eo;
class QplusQminus_eo;
class D_KS_eo;
class MdagM_eo;
}  // namespace fermionmatrix

/* physics/sourcesInterface.hpp:26-37: where and what the inverter's sources are.  The extents return int
 * (as the gauge-field interface's, which the same object implements); the source content, a compile-time
 * kernel parameter in the reference (SOURCE_CONTENT, hardware/code/opencl_module.cpp), is read here too. */
class SourcesParametersInterface {
  public:
    virtual ~SourcesParametersInterface() = default;
    virtual bool placeSourcesOnHost() const                 = 0;
    virtual common::sourcetypes getSourceType() const       = 0;
    virtual common::sourcecontents getSourceContent() const = 0;
    virtual unsigned getSourceT() const                     = 0;
    virtual unsigned getSourceX() const                     = 0;
    virtual unsigned getSourceY() const                     = 0;
    virtual unsigned getSourceZ() const                     = 0;
    virtual int getNt() const                               = 0;
    virtual int getNs() const                               = 0;
};

/* one object that implements every interface from the run parameters */
class ParametersImplementation : public lattices::GaugefieldParametersInterface,
                                 public lattices::SpinorfieldParametersInterface,
                                 public lattices::SpinorfieldEoParametersInterface,
                                 public lattices::StaggeredfieldEoParametersInterface,
                                 public fermionmatrix::FermionEoParametersInterface,
                                 public fermionmatrix::FermionStaggeredEoParametersInterface,
                                 public algorithms::SolversParametersInterface,
                                 public SourcesParametersInterface {
  public:
    explicit ParametersImplementation(const meta::Inputparameters& p) : p(p) {}
    int getNs() const override { return p.get_nspace(); }
    int getNt() const override { return p.get_ntime(); }
    common::startcondition getStartcondition() const override { return p.get_startcondition(); }
    std::string getSourcefileName() const override { return p.get_sourcefile(); }
    bool useReconstruct12() const override { return p.get_use_rec12(); }
    size_t getNumberOfElements() const override { return meta::get_vol4d(p); }
    size_t getEoprecSpinorfieldSize() const override { return meta::get_vol4d(p) / 2; }
    common::action getFermionicActionType() const override { return p.get_fermact(); }
    bool useMergedFermionicKernels() const override { return p.get_use_merge_kernels_fermion(); }
    int getCgMax() const override { return p.get_cgmax(); }
    int getIterRefresh() const override { return p.get_iter_refresh(); }
    common::solver getSolver() const override { return p.get_solver(); }
    int getCgIterationBlockSize() const override { return p.get_cg_iteration_block_size(); }
    int getCgMinimumIterationCount() const override { return p.get_cg_minimum_iteration_count(); }
    bool getCgUseAsyncCopy() const override { return p.get_cg_use_async_copy(); }
    bool getUseMergeKernelsSpinor() const override { return p.get_use_merge_kernels_spinor(); }
    bool placeSourcesOnHost() const override { return true; }  // full-lattice fields are host arrays here
    common::sourcetypes getSourceType() const override { return p.get_sourcetype(); }
    common::sourcecontents getSourceContent() const override { return p.get_sourcecontent(); }
    unsigned getSourceT() const override { return (unsigned)p.get_source_t(); }
    unsigned getSourceX() const override { return (unsigned)p.get_source_x(); }
    unsigned getSourceY() const override { return (unsigned)p.get_source_y(); }
    unsigned getSourceZ() const override { return (unsigned)p.get_source_z(); }

  private:
    const meta::Inputparameters& p;
};

template<class T>
struct InterfaceType;
template<>
struct InterfaceType<lattices::Gaugefield> {
    using value = lattices::GaugefieldParametersInterface;
};
template<>
struct InterfaceType<lattices::Spinorfield> {
    using value = lattices::SpinorfieldParametersInterface;
};
template<>
struct InterfaceType<lattices::Spinorfield_eo> {
    using value = lattices::SpinorfieldEoParametersInterface;
};
template<>
struct InterfaceType<lattices::Staggeredfield_eo> {
    using value = lattices::StaggeredfieldEoParametersInterface;
};
#define CL2_FM_IFACE(C, I)                  \
    template<>                              \
    struct InterfaceType<fermionmatrix::C> { \
        using value = fermionmatrix::I;     \
    };
CL2_FM_IFACE(Aee, FermionEoParametersInterface)
CL2_FM_IFACE(Aee_AND_gamma5_eo, FermionEoParametersInterface)
CL2_FM_IFACE(Aee_minus, FermionEoParametersInterface)
CL2_FM_IFACE(Aee_minus_AND_gamma5_eo, FermionEoParametersInterface)
CL2_FM_IFACE(Qplus_eo, FermionEoParametersInterface)
CL2_FM_IFACE(Qminus_eo, FermionEoParametersInterface)
CL2_FM_IFACE(QplusQminus_eo, FermionEoParametersInterface)
CL2_FM_IFACE(D_KS_eo, FermionStaggeredEoParametersInterface)
CL2_FM_IFACE(MdagM_eo, FermionStaggeredEoParametersInterface)
#undef CL2_FM_IFACE

class InterfacesHandler {
  public:
    virtual ~InterfacesHandler() = default;
    virtual const algorithms::SolversParametersInterface& getSolversParametersInterface() = 0;
    virtual const SourcesParametersInterface& getSourcesParametersInterface() = 0;
    template<class T>
    const typename InterfaceType<T>::value& getInterface()
    {
        return parameters();
    }
    template<class T>
    const AdditionalParameters& getAdditionalParameters();

  protected:
    virtual const ParametersImplementation& parameters() const = 0;
    virtual const AdditionalParameters& wilsonAdditionalParameters() const = 0;
    virtual const AdditionalParameters& staggeredAdditionalParameters() const = 0;
};
template<>
inline const AdditionalParameters& InterfacesHandler::getAdditionalParameters<lattices::Spinorfield>()
{
    return wilsonAdditionalParameters();
}
template<>
inline const AdditionalParameters& InterfacesHandler::getAdditionalParameters<lattices::Spinorfield_eo>()
{
    return wilsonAdditionalParameters();
}
template<>
inline const AdditionalParameters& InterfacesHandler::getAdditionalParameters<lattices::Staggeredfield_eo>()
{
    return staggeredAdditionalParameters();
}

/* interfaceImplementations/interfacesHandler.hpp */
class InterfacesHandlerImplementation final : public InterfacesHandler {
  public:
    explicit InterfacesHandlerImplementation(const meta::Inputparameters& p)
        : impl(p), wilson(p.get_kappa(), meta::get_mubar(p)), staggered(p.get_mass())
    {
    }
    const algorithms::SolversParametersInterface& getSolversParametersInterface() override { return impl; }
    const SourcesParametersInterface& getSourcesParametersInterface() override { return impl; }

  protected:
    const ParametersImplementation& parameters() const override { return impl; }
    const AdditionalParameters& wilsonAdditionalParameters() const override { return wilson; }
    const AdditionalParameters& staggeredAdditionalParameters() const override { return staggered; }

  private:
    ParametersImplementation impl;
    WilsonAdditionalParameters wilson;
    StaggeredAdditionalParameters staggered;
};

/* physics::PRNG (physics/prng.hpp): seeds the device generator of hot starts.
 * The reference's ranluxcl streams depend on the device's thread count, so they
 * are not reproduced; the counter-based generator is keyed by this seed. */
class PrngParametersInterface {
  public:
    virtual ~PrngParametersInterface() = default;
    virtual uint32_t getHostSeed() const = 0;
};
class PrngParametersImplementation : public PrngParametersInterface {
  public:
    explicit PrngParametersImplementation(const meta::Inputparameters& p) : seed(p.get_host_seed()) {}
    uint32_t getHostSeed() const override { return seed; }

  private:
    uint32_t seed;
};
class PRNG {
  public:
    PRNG(const hardware::System& /*system*/, const PrngParametersInterface* p) : seed(p ? p->getHostSeed() : 0) {}
    uint64_t next_seed() const { return seed + 0x9e3779b97f4a7c15ull * (++counter); }

  private:
    uint64_t seed;
    mutable uint64_t counter = 0;
};

}  // namespace physics
