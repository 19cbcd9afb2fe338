// Point-source inversion and pseudoscalar correlator on the C ABI.
#include "../include/physics/algorithms/inversion.hpp"

#include <stdexcept>

namespace physics {
namespace algorithms {

using namespace physics::lattices;
namespace fm = physics::fermionmatrix;

int invert_eo(const Spinorfield_eo* phi_e, const Spinorfield_eo* phi_o, const Gaugefield& gf,
              const Spinorfield_eo& eta_e, const Spinorfield_eo& eta_o, const hardware::System& system,
              physics::InterfacesHandler& ih, hmc_float prec, const physics::AdditionalParameters& ap)
{
    const meta::Inputparameters& params = system.get_inputparameters();
    const bool tm                       = params.get_fermact() == common::twistedmass;
    if (!tm && params.get_fermact() != common::wilson) throw std::invalid_argument("invert_eo: Wilson or twisted mass");
    const hmc_float kappa = ap.getKappa(), mubar = tm ? ap.getMubar() : 0.;
    const hmc_complex one{1., 0.};
    const Spinorfield_eo t1(system), t2(system), b(system);
    // M_site^-1 x (identity for Wilson)
    auto minv = [&](const Spinorfield_eo* out, const Spinorfield_eo& in) {
        if (tm)
            fm::M_tm_inverse_sitediagonal(out, in, mubar);
        else
            copyData(out, in);
    };
    // b'_o = eta_o - D_eo M^-1 eta_e   (dslash with evenodd=EVEN writes the odd sites)
    minv(&t1, eta_e);
    fm::dslash(&t2, gf, t1, EVEN, kappa);
    saxpy(&b, one, t2, eta_o);
    // solve on the odd sublattice
    int iters = 0;
    phi_o->cold();
    if (ih.getSolversParametersInterface().getSolver() == common::cg) {
        b.gamma5();
        const fm::QplusQminus_eo qq(system, ih.getInterface<fm::QplusQminus_eo>());
        const Spinorfield_eo x(system);
        x.cold();
        iters = solvers::cg(&x, qq, gf, b, system, ih, prec, ap);
        const fm::Qminus_eo qm(system, ih.getInterface<fm::Qminus_eo>());
        qm(phi_o, gf, x, ap);
    } else {
        const fm::Aee aee(system, ih.getInterface<fm::Aee>());
        iters = solvers::bicgstab(phi_o, aee, gf, b, system, ih, prec, ap);
    }
    // phi_e = M^-1 (eta_e - D_oe phi_o)   (dslash with evenodd=ODD writes the even sites)
    fm::dslash(&t1, gf, *phi_o, ODD, kappa);
    saxpy(&t2, one, t1, eta_e);
    minv(phi_e, t2);
    return iters;
}

std::vector<hmc_float> point_source_ps_correlator(const Gaugefield& gf, const hardware::System& system,
                                                  physics::InterfacesHandler& ih,
                                                  const physics::AdditionalParameters& ap)
{
    const meta::Inputparameters& p = system.get_inputparameters();
    const int ns = p.get_nspace(), nt = p.get_ntime();
    const int x = p.get_source_x(), y = p.get_source_y(), z = p.get_source_z(), t = p.get_source_t();
    if (x < 0 || x >= ns || y < 0 || y >= ns || z < 0 || z >= ns || t < 0 || t >= nt)
        throw std::invalid_argument("point source outside the lattice");
    if (p.get_num_sources() != 12) throw std::invalid_argument("point sources: nSources must be 12");
    const long site   = x + ns * (y + ns * (z + (long)ns * t));
    const bool even   = ((x + y + z + t) & 1) == 0;
    const long veo    = system.eo_sites();
    const int dir = p.get_corr_dir();
    const Vector<hmc_float> corr(dir == 0 ? nt : ns, system);
    const Spinorfield_eo eta_e(system), eta_o(system), phi_e(system), phi_o(system);
    for (int k = 0; k < 12; ++k) {
        std::vector<spinor> h(veo);
        std::fill(reinterpret_cast<double*>(h.data()), reinterpret_cast<double*>(h.data()) + 24 * veo, 0.);
        reinterpret_cast<double*>(h.data())[24 * (site / 2) + 2 * k] = 1.;  // component k = 3 spin + colour
        if (even) {
            eta_e.import(h.data());
            eta_o.zero();
        } else {
            eta_o.import(h.data());
            eta_e.zero();
        }
        invert_eo(&phi_e, &phi_o, gf, eta_e, eta_o, system, ih, p.get_solver_prec(), ap);
        system.check(cl2_slice_squarenorm_eo(system.ctx(), corr.dev(), phi_e.dev(), CL2_F64, dir, k > 0));
        system.check(cl2_slice_squarenorm_eo(system.ctx(), corr.dev(), phi_o.dev(), CL2_F64, dir, 1));
    }
    std::vector<hmc_float> c = corr.get();
    // one factor 2 kappa per field over the sites of a slice (correlator_ps_point.cl: NS^3 for t, NS^2 NT for z)
    const double slice_sites = dir == 0 ? (double)ns * ns * ns : (double)ns * ns * nt;
    const hmc_float fac      = 2. * ap.getKappa() * 2. * ap.getKappa() / slice_sites;
    for (auto& v : c) v *= fac;
    return c;
}

}  // namespace algorithms
}  // namespace physics
