// Inversion on sources and the pseudoscalar correlator on the C ABI.
#include "../include/physics/algorithms/inversion.hpp"

#include <memory>
#include <stdexcept>

namespace physics {
namespace algorithms {

using namespace physics::lattices;
namespace fm = physics::fermionmatrix;

namespace {

constexpr hmc_complex one{1., 0.}, mone{-1., 0.};

struct Action {
    bool tm;
    hmc_float kappa, mubar;
};
Action action_of(const hardware::System& system, const physics::AdditionalParameters& ap)
{
    const common::action a = system.get_inputparameters().get_fermact();
    if (a != common::twistedmass && a != common::wilson) throw std::invalid_argument("inversion: Wilson or twisted mass");
    const bool tm = a == common::twistedmass;
    return Action{tm, ap.getKappa(), tm ? ap.getMubar() : 0.};
}

/* M_site^-1 in (a copy for Wilson) */
void minv(const Action& a, const Spinorfield_eo* out, const Spinorfield_eo& in)
{
    if (a.tm)
        fm::M_tm_inverse_sitediagonal(out, in, a.mubar);
    else
        copyData(out, in);
}

/* b = D_eo M_site^-1 eta_e - eta_o: inversion.cpp:151-161, where saxpy(&source_even, one, source_even, tmp1) is
 * tmp1 - source_even (Wilson saxpy is y - alpha x) and the reference's "source_even" holds the odd sites */
void schur_source(const Action& a, const Spinorfield_eo* b, const Gaugefield& gf, const Spinorfield_eo& eta_e,
                  const Spinorfield_eo& eta_o, const Spinorfield_eo& t1, const Spinorfield_eo& t2)
{
    minv(a, &t1, eta_e);
    fm::dslash(&t2, gf, t1, EVEN, a.kappa);  // evenodd EVEN writes the odd sites
    saxpy(b, one, eta_o, t2);
}

/* phi_o = -x, phi_e = M_site^-1 (eta_e - D_oe phi_o) (the reference's odd solution, inversion.cpp:198-213, is
 * -(M^-1 D x + M^-1 eta) = -phi_e: the overall sign of its result) */
void reconstruct(const Action& a, const Spinorfield_eo* phi_e, const Spinorfield_eo* phi_o, const Gaugefield& gf,
                 const Spinorfield_eo& x, const Spinorfield_eo& eta_e, const Spinorfield_eo& t1,
                 const Spinorfield_eo& t2)
{
    sax(phi_o, mone, x);
    fm::dslash(&t1, gf, *phi_o, ODD, a.kappa);  // evenodd ODD writes the even sites
    saxpy(&t2, one, t1, eta_e);
    minv(a, phi_e, t2);
}

/* largest supported block size <= n */
int block_size(int n, int cap)
{
    for (int s : {16, 12, 8, 4, 2, 1})
        if (s <= n && s <= cap) return s;
    return 1;
}

}  // namespace

int invert_eo(const Spinorfield_eo* phi_e, const Spinorfield_eo* phi_o, const Gaugefield& gf,
              const Spinorfield_eo& eta_e, const Spinorfield_eo& eta_o, const hardware::System& system,
              physics::InterfacesHandler& ih, hmc_float prec, const physics::AdditionalParameters& ap,
              InversionStats* stats)
{
    const Action a = action_of(system, ap);
    const Spinorfield_eo t1(system), t2(system), b(system), x(system);
    InversionStats st;
    schur_source(a, &b, gf, eta_e, eta_o, t1, t2);
    x.cold();  // trial solution (inversion.cpp:173)
    const fm::Aee aee(system, ih.getInterface<fm::Aee>());
    if (ih.getSolversParametersInterface().getSolver() == common::cg) {
        try {
            st.iterations = st.bicgstab_iterations = solvers::bicgstab(&x, aee, gf, b, system, ih, prec, ap);
        } catch (const solvers::SolverException& e) {
            // inversion.cpp:182-194: CG on QplusQminus_eo from BiCGStab's last iterate, source gamma5 b,
            // then x = Qminus y solves Aee x = b
            st.bicgstab_iterations = e.get_iterations();
            st.cg_fallback         = true;
            b.gamma5();
            const fm::QplusQminus_eo qq(system, ih.getInterface<fm::QplusQminus_eo>());
            st.iterations = solvers::cg(&x, qq, gf, b, system, ih, prec, ap);
            const fm::Qminus_eo qm(system, ih.getInterface<fm::Qminus_eo>());
            qm(&t1, gf, x, ap);
            copyData(&x, t1);
        }
    } else {
        st.iterations = st.bicgstab_iterations = solvers::bicgstab(&x, aee, gf, b, system, ih, prec, ap);
    }
    reconstruct(a, phi_e, phi_o, gf, x, eta_e, t1, t2);
    if (stats) *stats = st;
    return st.iterations;
}

void invert_eo_block(const std::vector<const Spinorfield_eo*>& phi_e, const std::vector<const Spinorfield_eo*>& phi_o,
                     const Gaugefield& gf, const std::vector<const Spinorfield_eo*>& eta_e,
                     const std::vector<const Spinorfield_eo*>& eta_o, const hardware::System& system,
                     physics::InterfacesHandler& ih, const physics::AdditionalParameters& ap,
                     const BlockInversionOptions& o, std::vector<solvers::SolveStats>* stats)
{
    const size_t n = eta_e.size();
    if (eta_o.size() != n || phi_e.size() != n || phi_o.size() != n)
        throw std::invalid_argument("invert_eo_block: fields of different counts");
    const Action a = action_of(system, ap);
    const Spinorfield_eo t1(system), t2(system);
    const fm::QplusQminus_eo qq(system, ih.getInterface<fm::QplusQminus_eo>());
    const fm::Qminus_eo qm(system, ih.getInterface<fm::Qminus_eo>());
    if (stats) stats->clear();
    for (size_t k0 = 0; k0 < n;) {
        const int nb = block_size((int)(n - k0), o.block);
        std::vector<std::unique_ptr<Spinorfield_eo>> bs, ys;
        std::vector<const Spinorfield_eo*> bp, yp;
        for (int r = 0; r < nb; ++r) {
            bs.emplace_back(new Spinorfield_eo(system));
            ys.emplace_back(new Spinorfield_eo(system));
            schur_source(a, bs.back().get(), gf, *eta_e[k0 + r], *eta_o[k0 + r], t1, t2);
            bs.back()->gamma5();
            ys.back()->zero();
            bp.push_back(bs.back().get());
            yp.push_back(ys.back().get());
        }
        std::vector<solvers::SolveStats> st;
        solvers::cg_reliable_update_block(yp, qq, gf, bp, system, o.rel_tol, ap, o.solver, &st);
        for (int r = 0; r < nb; ++r) {
            qm(bs[r].get(), gf, *ys[r], ap);  // x = Qminus y
            reconstruct(a, phi_e[k0 + r], phi_o[k0 + r], gf, *bs[r], *eta_e[k0 + r], t1, t2);
        }
        if (stats) stats->insert(stats->end(), st.begin(), st.end());
        k0 += (size_t)nb;
    }
}

std::vector<hmc_float> ps_correlator(const Gaugefield& gf, const hardware::System& system,
                                     physics::InterfacesHandler& ih, const physics::AdditionalParameters& ap,
                                     const PRNG& prng, const BlockInversionOptions* block)
{
    const meta::Inputparameters& p = system.get_inputparameters();
    const int ns = p.get_nspace(), nt = p.get_ntime(), dir = p.get_corr_dir();
    const int nsrc = p.get_num_sources();
    if (nsrc <= 0) throw std::invalid_argument("nSources must be positive");
    const bool stochastic_z = p.get_sourcetype() != common::point && dir == 3;
    const size_t ncorr      = dir == 0 ? (size_t)nt : (size_t)ns;
    std::vector<hmc_float> corr(ncorr, 0.);
    const Vector<hmc_float> slices(ncorr, system);

    std::vector<std::unique_ptr<Spinorfield>> src;
    for (auto* f : create_sources(system, prng, (size_t)nsrc, ih)) src.emplace_back(f);
    std::vector<hmc_float> weight(src.size(), 1.);
    std::vector<std::unique_ptr<Spinorfield_eo>> ee, eo, pe, po;
    std::vector<const Spinorfield_eo*> vee, veo, vpe, vpo;
    for (size_t k = 0; k < src.size(); ++k) {
        if (stochastic_z) {
            hmc_float s = 0.;
            for (const spinor& v : src[k]->get_host()) {
                const auto* c = reinterpret_cast<const hmc_float*>(&v);
                for (int j = 0; j < 24; ++j) s += c[j] * c[j];
            }
            weight[k] = s;
        }
        for (auto* v : {&ee, &eo, &pe, &po}) v->emplace_back(new Spinorfield_eo(system));
        convert_to_eoprec(ee.back().get(), eo.back().get(), *src[k]);
        vee.push_back(ee.back().get());
        veo.push_back(eo.back().get());
        vpe.push_back(pe.back().get());
        vpo.push_back(po.back().get());
    }
    if (block)
        invert_eo_block(vpe, vpo, gf, vee, veo, system, ih, ap, *block);
    else
        for (size_t k = 0; k < src.size(); ++k) invert_eo(vpe[k], vpo[k], gf, *vee[k], *veo[k], system, ih, p.get_solver_prec(), ap);
    for (size_t k = 0; k < src.size(); ++k) {
        system.check(cl2_slice_squarenorm_eo(system.ctx(), slices.dev(), vpe[k]->dev(), CL2_F64, dir, 0));
        system.check(cl2_slice_squarenorm_eo(system.ctx(), slices.dev(), vpo[k]->dev(), CL2_F64, dir, 1));
        const std::vector<hmc_float> s = slices.get();
        for (size_t i = 0; i < ncorr; ++i) corr[i] += weight[k] * s[i];
    }
    // one factor 2 kappa per field over the sites of a slice (NS^3 for t, NS^2 NT for z)
    const double slice_sites = dir == 0 ? (double)ns * ns * ns : (double)ns * ns * nt;
    const hmc_float fac      = 2. * ap.getKappa() * 2. * ap.getKappa() / slice_sites;
    for (auto& v : corr) v *= fac;
    return corr;
}

std::vector<hmc_float> point_source_ps_correlator(const Gaugefield& gf, const hardware::System& system,
                                                  physics::InterfacesHandler& ih,
                                                  const physics::AdditionalParameters& ap,
                                                  const BlockInversionOptions* block)
{
    const meta::Inputparameters& p = system.get_inputparameters();
    if (p.get_sourcetype() != common::point) throw std::invalid_argument("point_source_ps_correlator: sourceType point");
    if (p.get_num_sources() != 12) throw std::invalid_argument("point sources: nSources must be 12");
    const PrngParametersImplementation pp(p);
    const PRNG prng(system, &pp);  // point sources draw nothing
    return ps_correlator(gf, system, ih, ap, prng, block);
}

}  // namespace algorithms
}  // namespace physics
