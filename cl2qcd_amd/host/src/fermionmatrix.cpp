// Fermion matrices of the drop-in (physics/fermionmatrix/fermionmatrix.cpp:224-676,
// explicit.cpp:26-222, fermionmatrix_stagg.cpp:50-114, explicit_stagg.cpp:26-45).
#include <stdexcept>

#include "../include/physics/fermionmatrix/fermionmatrix.hpp"

namespace physics {
namespace fermionmatrix {

using lattices::Gaugefield;
using lattices::Spinorfield_eo;
using lattices::Staggeredfield_eo;

namespace {

/* gauge pointer and format matching the spinor precision of an application */
struct Links {
    const void* ptr;
    int fmt;
};
Links links_for(const Gaugefield& gf, int prec)
{
    if (prec == CL2_F32) return Links{gf.dev_f32(), CL2_GAUGE_F32_12};
    if (prec == CL2_F16) return Links{gf.dev_f16(), CL2_GAUGE_F16_12};
    return Links{gf.dev(), gf.format()};
}

void check_parity(int evenodd)
{
    if (evenodd != EVEN && evenodd != ODD) throw std::invalid_argument("evenodd must be EVEN or ODD");
}

void sitediag(const Spinorfield_eo* out, const Spinorfield_eo& in, hmc_float mubar, int variant)
{
    if (out->precision() != in.precision()) throw std::invalid_argument("fields of different precision");
    in.get_system().check(
        cl2_M_tm_sitediagonal_eo(in.get_system().ctx(), out->dev(), in.dev(), in.precision(), mubar, variant));
}

/* per-site, per-application models of the reference, FP64 with 18-real links:
 *   dslash_eo 1632 flop / 2880 B, M_tm_sitediagonal 72 / 384, M_tm_inverse_sitediagonal 96 / 384,
 *   gamma5_eo 12 / 384, saxpy_AND_gamma5_eo 108 / 592 (hardware/code/fermions.cpp:811-975);
 *   saxpy_eoprec 96 / 592 (hardware/code/spinors.cpp:969-972,1085-1088);
 *   D_KS_eo 570 / 1584 (hardware/code/fermions_staggered.cpp:189-193,259-263);
 *   saxpby_cplx_staggered_eoprec 42 / 176 (hardware/code/spinors_staggered.cpp:1712-1715,1853-1856) */
constexpr cl_ulong DSLASH_FLOPS = 1632, DSLASH_BYTES = 2880, DIAG_FLOPS = 72, DIAG_INV_FLOPS = 96,
                   DIAG_BYTES = 384, SAXPY_FLOPS = 96, SAXPY_BYTES = 592, G5SAXPY_FLOPS = 108,
                   G5SAXPY_BYTES = 592, G5_FLOPS = 12, G5_BYTES = 384, DKS_FLOPS = 570, DKS_BYTES = 1584,
                   SAXPBY_STAGG_FLOPS = 42, SAXPBY_STAGG_BYTES = 176;

}  // namespace

/* ---------------- explicit kernels ---------------- */

void dslash(const Spinorfield_eo* out, const Gaugefield& gf, const Spinorfield_eo& in, int evenodd, hmc_float kappa)
{
    check_parity(evenodd);
    if (out->precision() != in.precision()) throw std::invalid_argument("fields of different precision");
    const Links l = links_for(gf, in.precision());
    in.get_system().check(cl2_dslash_eo(in.get_system().ctx(), out->dev(), in.dev(), l.ptr, l.fmt, evenodd, kappa));
}
void M_tm_inverse_sitediagonal(const Spinorfield_eo* out, const Spinorfield_eo& in, hmc_float mubar)
{
    sitediag(out, in, mubar, CL2_MTM_INVERSE_SITEDIAGONAL);
}
void M_tm_sitediagonal(const Spinorfield_eo* out, const Spinorfield_eo& in, hmc_float mubar)
{
    sitediag(out, in, mubar, CL2_MTM_SITEDIAGONAL);
}
void M_tm_inverse_sitediagonal_minus(const Spinorfield_eo* out, const Spinorfield_eo& in, hmc_float mubar)
{
    sitediag(out, in, mubar, CL2_MTM_INVERSE_SITEDIAGONAL_MINUS);
}
void M_tm_sitediagonal_minus(const Spinorfield_eo* out, const Spinorfield_eo& in, hmc_float mubar)
{
    sitediag(out, in, mubar, CL2_MTM_SITEDIAGONAL_MINUS);
}
void M_tm_sitediagonal_AND_gamma5_eo(const Spinorfield_eo* out, const Spinorfield_eo& in, hmc_float mubar)
{
    sitediag(out, in, mubar, CL2_MTM_SITEDIAGONAL_AND_GAMMA5);
}
void M_tm_sitediagonal_minus_AND_gamma5_eo(const Spinorfield_eo* out, const Spinorfield_eo& in, hmc_float mubar)
{
    sitediag(out, in, mubar, CL2_MTM_SITEDIAGONAL_MINUS_AND_GAMMA5);
}
static void dslash_minv(const Spinorfield_eo* out, const Gaugefield& gf, const Spinorfield_eo& in, int evenodd,
                        hmc_float kappa, hmc_float mubar, int minus)
{
    check_parity(evenodd);
    if (out->precision() != in.precision()) throw std::invalid_argument("fields of different precision");
    const Links l = links_for(gf, in.precision());
    in.get_system().check(cl2_dslash_AND_M_tm_inverse_sitediagonal_eo(in.get_system().ctx(), out->dev(), in.dev(),
                                                                      l.ptr, l.fmt, evenodd, kappa, mubar, minus));
}
void dslash_AND_M_tm_inverse_sitediagonal(const Spinorfield_eo* out, const Gaugefield& gf, const Spinorfield_eo& in,
                                          int evenodd, hmc_float kappa, hmc_float mubar)
{
    dslash_minv(out, gf, in, evenodd, kappa, mubar, 0);
}
void dslash_AND_M_tm_inverse_sitediagonal_minus(const Spinorfield_eo* out, const Gaugefield& gf,
                                                const Spinorfield_eo& in, int evenodd, hmc_float kappa,
                                                hmc_float mubar)
{
    dslash_minv(out, gf, in, evenodd, kappa, mubar, 1);
}
void DKS_eo(const Staggeredfield_eo* out, const Gaugefield& gf, const Staggeredfield_eo& in, int evenodd)
{
    check_parity(evenodd);
    if (out->precision() != in.precision()) throw std::invalid_argument("fields of different precision");
    const Links l = links_for(gf, in.precision());
    in.get_system().check(cl2_D_KS_eo(in.get_system().ctx(), out->dev(), in.dev(), l.ptr, l.fmt, evenodd));
}

/* ---------------- Wilson / tm eo operators ---------------- */

BuiltinEoOperator::BuiltinEoOperator(int op, bool herm, const hardware::System& system,
                                     const FermionEoParametersInterface& p)
    : Fermionmatrix_eo(herm, system, p), op(op),
      tmp(system, 2 * cl2_field_bytes(system.ctx(), CL2_FIELD_SPINOR_EO, CL2_F64))
{
}

namespace {
int abi_action_of(const FermionmatrixParametersInterface& p)
{
    const common::action a = p.getFermionicActionType();
    if (a == common::wilson) return CL2_ACTION_WILSON;
    if (a == common::twistedmass) return CL2_ACTION_TWISTEDMASS;
    throw std::invalid_argument("Unkown fermion action! wilson or twistedmass");
}
}  // namespace

void BuiltinEoOperator::operator()(const Spinorfield_eo* out, const Gaugefield& gf, const Spinorfield_eo& in,
                                   const physics::AdditionalParameters& ap) const
{
    const int abi_action = abi_action_of(fermionmatrixParametersInterface);
    if (out->precision() != in.precision()) throw std::invalid_argument("fields of different precision");
    const hmc_float kappa = ap.getKappa();
    const hmc_float mubar = abi_action == CL2_ACTION_TWISTEDMASS ? ap.getMubar() : 0.;
    const Links l         = links_for(gf, in.precision());
    get_system().check(cl2_fermionmatrix_eo(get_system().ctx(), op, out->dev(), in.dev(), l.ptr, l.fmt, abi_action,
                                            kappa, mubar, tmp.ptr()));
}

void BuiltinEoOperator::apply_and_dot(const Spinorfield_eo* out, const Gaugefield& gf, const Spinorfield_eo& in,
                                      const physics::AdditionalParameters& ap,
                                      const lattices::Scalar<hmc_complex>& dot) const
{
    if (op != CL2_OP_QPLUSQMINUS_EO) throw std::invalid_argument("apply_and_dot: QplusQminus_eo only");
    const int abi_action = abi_action_of(fermionmatrixParametersInterface);
    if (out->precision() != in.precision()) throw std::invalid_argument("fields of different precision");
    const hmc_float kappa = ap.getKappa();
    const hmc_float mubar = abi_action == CL2_ACTION_TWISTEDMASS ? ap.getMubar() : 0.;
    const Links l         = links_for(gf, in.precision());
    get_system().check(cl2_fermionmatrix_eo_dot(get_system().ctx(), op, out->dev(), in.dev(), l.ptr, l.fmt,
                                                abi_action, kappa, mubar, tmp.ptr(), dot.dev()));
}

void BuiltinEoOperator::cg_step(const Spinorfield_eo* r, const Gaugefield& gf, const Spinorfield_eo& p,
                                const physics::AdditionalParameters& ap, const lattices::Scalar<hmc_complex>& r2,
                                const lattices::Scalar<hmc_complex>& pAp, const lattices::Scalar<hmc_complex>& alpha,
                                const lattices::Vector<hmc_float>& norms) const
{
    if (op != CL2_OP_QPLUSQMINUS_EO) throw std::invalid_argument("cg_step: QplusQminus_eo only");
    if (norms.get_vector_size() < 4) throw std::invalid_argument("cg_step needs a result vector of 4 elements");
    const int abi_action = abi_action_of(fermionmatrixParametersInterface);
    if (r->precision() != p.precision()) throw std::invalid_argument("fields of different precision");
    const hmc_float kappa = ap.getKappa();
    const hmc_float mubar = abi_action == CL2_ACTION_TWISTEDMASS ? ap.getMubar() : 0.;
    const Links l         = links_for(gf, p.precision());
    get_system().check(cl2_fermionmatrix_eo_cg_step(get_system().ctx(), op, r->dev(), p.dev(), l.ptr, l.fmt,
                                                    abi_action, kappa, mubar, tmp.ptr(), r2.dev(), pAp.dev(),
                                                    alpha.dev(), norms.dev()));
}

cl_ulong BuiltinEoOperator::get_flops() const
{
    const cl_ulong S = get_system().eo_sites();
    const bool tm    = fermionmatrixParametersInterface.getFermionicActionType() == common::twistedmass;
    auto aee = [&](bool g5) {
        cl_ulong f = 2 * DSLASH_FLOPS + (g5 ? G5SAXPY_FLOPS : SAXPY_FLOPS);
        if (tm) f += DIAG_FLOPS + DIAG_INV_FLOPS;
        return f;
    };
    // Qplus_eo / Qminus_eo count Aee (or Aee_minus) plus gamma5_eo whether or not the merged kernel runs
    // (physics/fermionmatrix/fermionmatrix.cpp:603-613,625-635)
    auto q = [&]() { return aee(false) + G5_FLOPS; };
    switch (op) {
        case CL2_OP_AEE:
        case CL2_OP_AEE_MINUS: return S * aee(false);
        case CL2_OP_AEE_AND_GAMMA5:
        case CL2_OP_AEE_MINUS_AND_GAMMA5: return S * aee(true);
        case CL2_OP_QPLUS_EO:
        case CL2_OP_QMINUS_EO: return S * q();
        default: return 2 * S * q();
    }
}

cl_ulong BuiltinEoOperator::get_read_write_size() const
{
    const cl_ulong S = get_system().eo_sites();
    const bool tm    = fermionmatrixParametersInterface.getFermionicActionType() == common::twistedmass;
    auto aee = [&](bool g5) {
        cl_ulong b = 2 * DSLASH_BYTES + (g5 ? G5SAXPY_BYTES : SAXPY_BYTES);
        if (tm) b += 2 * DIAG_BYTES;
        return b;
    };
    auto q = [&]() { return aee(false) + G5_BYTES; };  // fermionmatrix.cpp:636-657
    switch (op) {
        case CL2_OP_AEE:
        case CL2_OP_AEE_MINUS: return S * aee(false);
        case CL2_OP_AEE_AND_GAMMA5:
        case CL2_OP_AEE_MINUS_AND_GAMMA5: return S * aee(true);
        case CL2_OP_QPLUS_EO:
        case CL2_OP_QMINUS_EO: return S * q();
        default: return 2 * S * q();
    }
}

double autotune_dslash(const Gaugefield& gf, const hardware::System& system, int prec, int nrhs)
{
    const Links l = links_for(gf, prec);
    double ms     = 0.;
    system.check(cl2_autotune_dslash(system.ctx(), l.ptr, l.fmt, nrhs, &ms, nullptr, nullptr, nullptr, nullptr, nullptr));
    return ms;
}

double autotune_dks(const Gaugefield& gf, const hardware::System& system, int prec)
{
    if (prec != CL2_F64 && prec != CL2_F32) throw std::invalid_argument("autotune_dks: FP64 or FP32 links");
    const Links l = links_for(gf, prec);
    double ms     = 0.;
    system.check(cl2_autotune_dks(system.ctx(), l.ptr, l.fmt, &ms, nullptr, nullptr, nullptr, nullptr));
    return ms;
}

void apply_block(const Fermionmatrix_eo& A, const lattices::Spinorfield_eo_block* out, const Gaugefield& gf,
                 const lattices::Spinorfield_eo_block& in, const physics::AdditionalParameters& ap,
                 const lattices::DeviceBuffer& tmp)
{
    const int op = A.abi_operator();
    if (op < 0) throw std::invalid_argument("apply_block: not a built-in eo operator");
    if (out->precision() != in.precision() || out->nrhs() != in.nrhs())
        throw std::invalid_argument("apply_block: blocks of different shape");
    if (tmp.bytes() < 2 * in.bytes()) throw std::invalid_argument("apply_block: tmp must hold two blocks");
    const common::action a = A.action();
    if (a != common::wilson && a != common::twistedmass)
        throw std::invalid_argument("Unkown fermion action! wilson or twistedmass");
    const int abi_action  = a == common::twistedmass ? CL2_ACTION_TWISTEDMASS : CL2_ACTION_WILSON;
    const hmc_float mubar = a == common::twistedmass ? ap.getMubar() : 0.;
    const Links l         = links_for(gf, in.precision());
    in.get_system().check(cl2_fermionmatrix_eo_mrhs(in.get_system().ctx(), op, out->dev(), in.dev(), l.ptr, l.fmt,
                                                    abi_action, ap.getKappa(), mubar, tmp.ptr(), in.nrhs()));
}

void apply_block(const Fermionmatrix_eo& A, const lattices::Spinorfield_eo_block* out, const Gaugefield& gf,
                 const lattices::Spinorfield_eo_block& in, const std::vector<hmc_float>& kappa,
                 const std::vector<hmc_float>& mubar, const lattices::DeviceBuffer& tmp)
{
    const int op = A.abi_operator();
    if (op < 0) throw std::invalid_argument("apply_block: not a built-in eo operator");
    if (out->precision() != in.precision() || out->nrhs() != in.nrhs())
        throw std::invalid_argument("apply_block: blocks of different shape");
    if (tmp.bytes() < 2 * in.bytes()) throw std::invalid_argument("apply_block: tmp must hold two blocks");
    if ((int)kappa.size() != in.nrhs() || (int)mubar.size() != in.nrhs())
        throw std::invalid_argument("apply_block: one kappa and one mubar per right-hand side");
    const common::action a = A.action();
    if (a != common::wilson && a != common::twistedmass)
        throw std::invalid_argument("Unkown fermion action! wilson or twistedmass");
    const int abi_action = a == common::twistedmass ? CL2_ACTION_TWISTEDMASS : CL2_ACTION_WILSON;
    const Links l        = links_for(gf, in.precision());
    in.get_system().check(cl2_fermionmatrix_eo_mrhs_params(in.get_system().ctx(), op, out->dev(), in.dev(), l.ptr,
                                                           l.fmt, abi_action, kappa.data(), mubar.data(), tmp.ptr(),
                                                           in.nrhs()));
}

bool cg_step_block(const Fermionmatrix_eo& A, const lattices::Spinorfield_eo_block* r, const Gaugefield& gf,
                   const lattices::Spinorfield_eo_block& p, const physics::AdditionalParameters* ap,
                   const std::vector<hmc_float>& kappa, const std::vector<hmc_float>& mubar,
                   const lattices::DeviceBuffer& tmp, const lattices::Vector<hmc_float>& r2,
                   const lattices::Vector<hmc_float>& active, const lattices::Vector<hmc_float>& pAp,
                   const lattices::Vector<hmc_float>& alpha, const lattices::Vector<hmc_float>& norms)
{
    const int n = p.nrhs();
    if (A.abi_operator() != CL2_OP_QPLUSQMINUS_EO || n < 2) return false;
    if (r->precision() != p.precision() || r->nrhs() != n) throw std::invalid_argument("cg_step_block: blocks of different shape");
    if (tmp.bytes() < 2 * p.bytes()) throw std::invalid_argument("cg_step_block: tmp must hold two blocks");
    const bool per_rhs = !kappa.empty();
    if (per_rhs && ((int)kappa.size() != n || (int)mubar.size() != n))
        throw std::invalid_argument("cg_step_block: one kappa and one mubar per right-hand side");
    if (!per_rhs && ap == nullptr) throw std::invalid_argument("cg_step_block: no operator parameters");
    const common::action a = A.action();
    if (a != common::wilson && a != common::twistedmass)
        throw std::invalid_argument("Unkown fermion action! wilson or twistedmass");
    const int abi_action  = a == common::twistedmass ? CL2_ACTION_TWISTEDMASS : CL2_ACTION_WILSON;
    const hmc_float kap   = per_rhs ? 0. : ap->getKappa();
    const hmc_float mubar0 = (!per_rhs && a == common::twistedmass) ? ap->getMubar() : 0.;
    const Links l         = links_for(gf, p.precision());
    p.get_system().check(cl2_fermionmatrix_eo_mrhs_cg_step(
        p.get_system().ctx(), CL2_OP_QPLUSQMINUS_EO, r->dev(), p.dev(), l.ptr, l.fmt, abi_action, kap, mubar0,
        per_rhs ? kappa.data() : nullptr, per_rhs ? mubar.data() : nullptr, tmp.ptr(), n, r2.dev(), active.dev(),
        pAp.dev(), alpha.dev(), norms.dev()));
    return true;
}

/* ---------------- staggered ---------------- */

void apply_block(const MdagM_eo& A, const lattices::Staggeredfield_eo_block* out, const Gaugefield& gf,
                 const lattices::Staggeredfield_eo_block& in, const physics::AdditionalParameters& ap,
                 const std::vector<hmc_float>& sigma, const lattices::DeviceBuffer& tmp)
{
    if (out->precision() != in.precision() || out->nrhs() != in.nrhs())
        throw std::invalid_argument("apply_block: blocks of different shape");
    if (tmp.bytes() < in.bytes()) throw std::invalid_argument("apply_block: tmp must hold one block");
    if (!sigma.empty() && (int)sigma.size() != in.nrhs())
        throw std::invalid_argument("apply_block: one shift per right-hand side");
    const Links l = links_for(gf, in.precision());
    in.get_system().check(cl2_MdagM_eo_mrhs(in.get_system().ctx(), out->dev(), in.dev(), l.ptr, l.fmt, ap.getMass(),
                                            A.get_upper_left() ? 1 : 0, sigma.empty() ? nullptr : sigma.data(),
                                            tmp.ptr(), in.nrhs()));
}

void cgm_step(const MdagM_eo& A, const lattices::Staggeredfield_eo* r, const Gaugefield& gf,
              const lattices::Staggeredfield_eo& p, const physics::AdditionalParameters& ap, hmc_float sigma,
              const lattices::Staggeredfield_eo& tmp, const lattices::Scalar<hmc_float>& rr_old,
              const lattices::Scalar<hmc_float>& pap, const lattices::Scalar<hmc_float>& beta,
              const lattices::Scalar<hmc_float>& beta_prev, const lattices::Scalar<hmc_float>& rr_new)
{
    if (r->precision() != p.precision() || tmp.precision() != p.precision())
        throw std::invalid_argument("cgm_step: fields of different precision");
    const Links l = links_for(gf, p.precision());
    p.get_system().check(cl2_MdagM_eo_cgm_step(p.get_system().ctx(), r->dev(), p.dev(), l.ptr, l.fmt, ap.getMass(),
                                               sigma, A.get_upper_left() ? 1 : 0, tmp.dev(), rr_old.dev(), pap.dev(),
                                               beta.dev(), beta_prev.dev(), rr_new.dev()));
}

bool cg_step_block(const MdagM_eo& A, const lattices::Staggeredfield_eo_block* r, const Gaugefield& gf,
                   const lattices::Staggeredfield_eo_block& p, const physics::AdditionalParameters& ap,
                   const std::vector<hmc_float>& sigma, const lattices::DeviceBuffer& tmp,
                   const lattices::Vector<hmc_float>& r2, const lattices::Vector<hmc_float>& active,
                   const lattices::Vector<hmc_float>& pAp, const lattices::Vector<hmc_float>& alpha,
                   const lattices::Vector<hmc_float>& norms)
{
    if (r->precision() != p.precision() || r->nrhs() != p.nrhs())
        throw std::invalid_argument("cg_step_block: blocks of different shape");
    if (tmp.bytes() < p.bytes()) throw std::invalid_argument("cg_step_block: tmp must hold one block");
    if (!sigma.empty() && (int)sigma.size() != p.nrhs())
        throw std::invalid_argument("cg_step_block: one shift per right-hand side");
    const Links l = links_for(gf, p.precision());
    p.get_system().check(cl2_MdagM_eo_mrhs_cg_step(
        p.get_system().ctx(), r->dev(), p.dev(), l.ptr, l.fmt, ap.getMass(), A.get_upper_left() ? 1 : 0,
        sigma.empty() ? nullptr : sigma.data(), tmp.ptr(), p.nrhs(), r2.dev(), active.dev(), pAp.dev(), alpha.dev(),
        norms.dev()));
    return true;
}

hmc_float Fermionmatrix_stagg_eo::getThresholdForMinimumEigenvalue(hmc_float) const
{
    throw std::logic_error("Threshold for minimum eigenvalue not existing or not implemented!");
}

void D_KS_eo::operator()(const Staggeredfield_eo* out, const Gaugefield& gf, const Staggeredfield_eo& in,
                         const physics::AdditionalParameters* ap) const
{
    if (ap != nullptr)
        throw std::invalid_argument(
            "D_KS_eo operator applied passing to it some additional parameters! This should not happen!");
    DKS_eo(out, gf, in, evenodd);
}
cl_ulong D_KS_eo::get_flops() const
{
    return get_system().eo_sites() * DKS_FLOPS;
}
cl_ulong D_KS_eo::get_read_write_size() const
{
    return get_system().eo_sites() * DKS_BYTES;
}

MdagM_eo::MdagM_eo(const hardware::System& system, const FermionStaggeredEoParametersInterface& p, bool upper_left)
    : Fermionmatrix_stagg_eo(true, system, p), upper_left(upper_left), tmp(system)
{
}

void MdagM_eo::operator()(const Staggeredfield_eo* out, const Gaugefield& gf, const Staggeredfield_eo& in,
                          const physics::AdditionalParameters* ap) const
{
    if (ap == nullptr)
        throw std::invalid_argument("MdagM_eo operator applied without the MANDATORY additional parameters!");
    if (out->precision() != in.precision()) throw std::invalid_argument("fields of different precision");
    const hmc_float mass = ap->getMass();
    const Links l        = links_for(gf, in.precision());
    if (in.precision() == tmp.precision()) {
        get_system().check(
            cl2_MdagM_eo(get_system().ctx(), out->dev(), in.dev(), l.ptr, l.fmt, mass, upper_left, tmp.dev()));
    } else {
        if (!tmp_lo || tmp_lo->precision() != in.precision())
            tmp_lo.reset(new Staggeredfield_eo(get_system(), in.precision()));
        get_system().check(
            cl2_MdagM_eo(get_system().ctx(), out->dev(), in.dev(), l.ptr, l.fmt, mass, upper_left, tmp_lo->dev()));
    }
}
cl_ulong MdagM_eo::get_flops() const
{
    return get_system().eo_sites() * (2 * DKS_FLOPS + SAXPBY_STAGG_FLOPS);
}
cl_ulong MdagM_eo::get_read_write_size() const
{
    return get_system().eo_sites() * (2 * DKS_BYTES + SAXPBY_STAGG_BYTES);
}

}  // namespace fermionmatrix
}  // namespace physics
