#include "../include/hardware/system.hpp"

namespace hardware {

System::System(const HardwareParametersInterface&, const code::OpenClKernelParametersInterface& kp)
    : params(kp.getInputparameters())
{
    init();
}

System::System(const meta::Inputparameters& p) : params(p)
{
    init();
}

void System::init()
{
    cl2_lattice lat{};
    lat.ns        = params.get_nspace();
    lat.nt        = params.get_ntime();
    lat.theta_s   = params.get_theta_fermion_spatial();
    lat.theta_t   = params.get_theta_fermion_temporal();
    lat.use_cp_re = params.get_use_chem_pot_re();
    lat.cp_re     = params.get_chem_pot_re();
    lat.use_cp_im = params.get_use_chem_pot_im();
    lat.cp_im     = params.get_chem_pot_im();
    const int rc  = cl2_ctx_create(&context, params.get_device_id(), &lat);
    if (rc) throw HipException(rc, std::string("cl2_ctx_create: ") + cl2_last_error(nullptr));
}

System::~System()
{
    cl2_ctx_destroy(context);
}

long System::eo_sites() const noexcept
{
    return cl2_eo_sites(context);
}

void System::synchronize() const
{
    check(cl2_sync(context));
}

void System::check(int rc) const
{
    if (rc) throw HipException(rc, cl2_last_error(context));
}

}  // namespace hardware
