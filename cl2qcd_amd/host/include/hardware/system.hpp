// hardware::System for one MI355X: owns the C-ABI context (device, HIP stream,
// lattice constants) that every lattice object and operator of the drop-in
// enqueues on (replaces hardware/system.hpp:64-107 and hardware/device.hpp).
#pragma once

#include <stdexcept>
#include <string>

#include "../../../../include/cl2qcd.h"
#include "../meta/inputparameters.hpp"

namespace hardware {

/* Error raised when a C-ABI call fails (hardware::OpenclException / Opencl_Error in the reference). */
class HipException : public std::runtime_error {
  public:
    HipException(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
    int code;
};

class HardwareParametersInterface {
  public:
    virtual ~HardwareParametersInterface() = default;
    virtual int getNs() const = 0;
    virtual int getNt() const = 0;
    virtual int getDeviceId() const = 0;
};
class HardwareParametersImplementation : public HardwareParametersInterface {
  public:
    explicit HardwareParametersImplementation(const meta::Inputparameters* p) : params(*p) {}
    int getNs() const override { return params.get_nspace(); }
    int getNt() const override { return params.get_ntime(); }
    int getDeviceId() const override { return params.get_device_id(); }
    const meta::Inputparameters& params;
};

namespace code {
class OpenClKernelParametersInterface {
  public:
    virtual ~OpenClKernelParametersInterface() = default;
    virtual const meta::Inputparameters& getInputparameters() const = 0;
};
class OpenClKernelParametersImplementation : public OpenClKernelParametersInterface {
  public:
    explicit OpenClKernelParametersImplementation(const meta::Inputparameters& p) : params(p) {}
    const meta::Inputparameters& getInputparameters() const override { return params; }
    const meta::Inputparameters& params;
};
}  // namespace code

class System {
  public:
    System(const HardwareParametersInterface& hp, const code::OpenClKernelParametersInterface& kp);
    explicit System(const meta::Inputparameters& params);
    ~System();
    System(const System&) = delete;
    System& operator=(const System&) = delete;

    cl2_ctx* ctx() const noexcept { return context; }
    const meta::Inputparameters& get_inputparameters() const noexcept { return params; }
    long eo_sites() const noexcept;
    void synchronize() const;
    /* throw HipException with the context's message when rc != 0 */
    void check(int rc) const;

  private:
    void init();
    const meta::Inputparameters& params;
    cl2_ctx* context = nullptr;
};

}  // namespace hardware
