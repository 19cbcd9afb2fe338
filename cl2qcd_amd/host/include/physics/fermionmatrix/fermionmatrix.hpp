// Even-odd fermion matrices and the explicit kernels under them
// (physics/fermionmatrix/fermionmatrix.hpp:46-310, explicit.cpp:26-222,
// fermionmatrix_stagg.hpp:43-152, explicit_stagg.cpp:26-45).  Each operator owns
// its temporaries, as the reference's classes do; each call runs fused gfx950
// kernels through the C ABI.
#pragma once

#include "../lattices/lattices.hpp"

namespace physics {
namespace fermionmatrix {

/* ---- explicit kernels ---- */
void dslash(const lattices::Spinorfield_eo* out, const lattices::Gaugefield& gf, const lattices::Spinorfield_eo& in,
            int evenodd, hmc_float kappa);
void M_tm_inverse_sitediagonal(const lattices::Spinorfield_eo* out, const lattices::Spinorfield_eo& in,
                               hmc_float mubar);
void M_tm_sitediagonal(const lattices::Spinorfield_eo* out, const lattices::Spinorfield_eo& in, hmc_float mubar);
void M_tm_inverse_sitediagonal_minus(const lattices::Spinorfield_eo* out, const lattices::Spinorfield_eo& in,
                                     hmc_float mubar);
void M_tm_sitediagonal_minus(const lattices::Spinorfield_eo* out, const lattices::Spinorfield_eo& in,
                             hmc_float mubar);
/* fused dslash + inverse site diagonal (fermionmatrix_eo_dslash_AND_M_tm_inverse_sitediagonal{,_minus}.cl) */
void dslash_AND_M_tm_inverse_sitediagonal(const lattices::Spinorfield_eo* out, const lattices::Gaugefield& gf,
                                          const lattices::Spinorfield_eo& in, int evenodd, hmc_float kappa,
                                          hmc_float mubar);
void dslash_AND_M_tm_inverse_sitediagonal_minus(const lattices::Spinorfield_eo* out, const lattices::Gaugefield& gf,
                                                const lattices::Spinorfield_eo& in, int evenodd, hmc_float kappa,
                                                hmc_float mubar);
void M_tm_sitediagonal_AND_gamma5_eo(const lattices::Spinorfield_eo* out, const lattices::Spinorfield_eo& in,
                                     hmc_float mubar);
void M_tm_sitediagonal_minus_AND_gamma5_eo(const lattices::Spinorfield_eo* out, const lattices::Spinorfield_eo& in,
                                           hmc_float mubar);
void DKS_eo(const lattices::Staggeredfield_eo* out, const lattices::Gaugefield& gf,
            const lattices::Staggeredfield_eo& in, int evenodd);

class Fermionmatrix_basic {
  public:
    virtual ~Fermionmatrix_basic() = default;
    bool isHermitian() const noexcept { return _is_hermitian; }
    /* the reference's per-application flop and byte models (FP64, 18-real links) */
    virtual cl_ulong get_flops() const = 0;
    virtual cl_ulong get_read_write_size() const = 0;

  protected:
    Fermionmatrix_basic(const hardware::System& system, const FermionmatrixParametersInterface& p, bool herm)
        : fermionmatrixParametersInterface(p), _is_hermitian(herm), system(system)
    {
    }
    const hardware::System& get_system() const noexcept { return system; }
    const FermionmatrixParametersInterface& fermionmatrixParametersInterface;

  private:
    const bool _is_hermitian;
    const hardware::System& system;
};

class Fermionmatrix_eo : public Fermionmatrix_basic {
  public:
    virtual void operator()(const lattices::Spinorfield_eo* out, const lattices::Gaugefield& gf,
                            const lattices::Spinorfield_eo& in,
                            const physics::AdditionalParameters& additionalParameters) const = 0;
    /* the C-ABI operator code (CL2_OP_*) when the operator is one of the built-in classes, else -1;
     * lets solvers switch the whole operator to another precision */
    virtual int abi_operator() const noexcept { return -1; }
    common::action action() const { return fermionmatrixParametersInterface.getFermionicActionType(); }

  protected:
    Fermionmatrix_eo(bool herm, const hardware::System& system, const FermionmatrixParametersInterface& p)
        : Fermionmatrix_basic(system, p, herm)
    {
    }
};

/* one built-in Wilson/tm eo operator; the classes below only fix the C-ABI code */
class BuiltinEoOperator : public Fermionmatrix_eo {
  public:
    void operator()(const lattices::Spinorfield_eo* out, const lattices::Gaugefield& gf,
                    const lattices::Spinorfield_eo& in,
                    const physics::AdditionalParameters& additionalParameters) const override;
    cl_ulong get_flops() const override;
    cl_ulong get_read_write_size() const override;
    int abi_operator() const noexcept override { return op; }
    /* QplusQminus_eo only: out = A in and dot = <in, out> = |Qminus_eo in|^2 (a complex device scalar, imaginary
     * part 0) from the kernel that writes Qminus_eo in; throws for the other operators */
    void apply_and_dot(const lattices::Spinorfield_eo* out, const lattices::Gaugefield& gf,
                       const lattices::Spinorfield_eo& in, const physics::AdditionalParameters& additionalParameters,
                       const lattices::Scalar<hmc_complex>& dot) const;
    /* QplusQminus_eo only: the operator half of a CG step without storing A p: pAp = <p, A p>, alpha = r2 / pAp,
     * r <- r - alpha A p and norms (4 values) = (|r_new|^2, Re <r_new, r_old>, |r_new|^2, 0) */
    void cg_step(const lattices::Spinorfield_eo* r, const lattices::Gaugefield& gf, const lattices::Spinorfield_eo& p,
                 const physics::AdditionalParameters& additionalParameters, const lattices::Scalar<hmc_complex>& r2,
                 const lattices::Scalar<hmc_complex>& pAp, const lattices::Scalar<hmc_complex>& alpha,
                 const lattices::Vector<hmc_float>& norms) const;

  protected:
    BuiltinEoOperator(int op, bool herm, const hardware::System& system, const FermionEoParametersInterface& p);

  private:
    int op;
    lattices::DeviceBuffer tmp;  // two eo spinor temporaries (fermionmatrix.hpp:207-208,309)
};

#define CL2_EO_OPERATOR(NAME, OP, HERM)                                                            \
    class NAME final : public BuiltinEoOperator {                                                  \
      public:                                                                                      \
        NAME(const hardware::System& system, const FermionEoParametersInterface& p)                \
            : BuiltinEoOperator(OP, HERM, system, p)                                               \
        {                                                                                          \
        }                                                                                          \
    };
CL2_EO_OPERATOR(Aee, CL2_OP_AEE, false)
CL2_EO_OPERATOR(Aee_AND_gamma5_eo, CL2_OP_AEE_AND_GAMMA5, false)
CL2_EO_OPERATOR(Aee_minus, CL2_OP_AEE_MINUS, false)
CL2_EO_OPERATOR(Aee_minus_AND_gamma5_eo, CL2_OP_AEE_MINUS_AND_GAMMA5, false)
CL2_EO_OPERATOR(Qplus_eo, CL2_OP_QPLUS_EO, false)
CL2_EO_OPERATOR(Qminus_eo, CL2_OP_QMINUS_EO, false)
CL2_EO_OPERATOR(QplusQminus_eo, CL2_OP_QPLUSQMINUS_EO, true)
#undef CL2_EO_OPERATOR

/* autotune the hopping kernel's launch shape for links matching prec and nrhs right-hand sides on gf's
 * device (no counterpart in the reference); returns the best time per application in ms */
double autotune_dslash(const lattices::Gaugefield& gf, const hardware::System& system, int prec, int nrhs = 1);
/* the same for the staggered hopping kernel with links matching prec (FP64 or FP32) */
double autotune_dks(const lattices::Gaugefield& gf, const hardware::System& system, int prec);

/* A applied to every right-hand side of a block in one pass over the links (A must be one of the built-in
 * classes; tmp holds two block fields of in's shape).  No counterpart in the reference. */
void apply_block(const Fermionmatrix_eo& A, const lattices::Spinorfield_eo_block* out, const lattices::Gaugefield& gf,
                 const lattices::Spinorfield_eo_block& in, const physics::AdditionalParameters& ap,
                 const lattices::DeviceBuffer& tmp);
/* the same with A's parameters per right-hand side: kappa[r] and mubar[r] (mubar unused for Wilson) */
void apply_block(const Fermionmatrix_eo& A, const lattices::Spinorfield_eo_block* out, const lattices::Gaugefield& gf,
                 const lattices::Spinorfield_eo_block& in, const std::vector<hmc_float>& kappa,
                 const std::vector<hmc_float>& mubar, const lattices::DeviceBuffer& tmp);
/* The operator half of a lockstep CG step for A = QplusQminus_eo (false for any other operator, nothing
 * done): pAp[r] = <p_r, A p_r>, alpha[r] = r2[r] / pAp[r] where active[r] != 0, r_r -= alpha[r] A p_r and
 * norms[r] = (|r_new|^2, Re <r_new, r_old>), A p never stored.  kappa/mubar: one per right-hand side, or empty
 * for ap's.  No counterpart in the reference. */
bool cg_step_block(const Fermionmatrix_eo& A, const lattices::Spinorfield_eo_block* r, const lattices::Gaugefield& gf,
                   const lattices::Spinorfield_eo_block& p, const physics::AdditionalParameters* ap,
                   const std::vector<hmc_float>& kappa, const std::vector<hmc_float>& mubar,
                   const lattices::DeviceBuffer& tmp, const lattices::Vector<hmc_float>& r2,
                   const lattices::Vector<hmc_float>& active, const lattices::Vector<hmc_float>& pAp,
                   const lattices::Vector<hmc_float>& alpha, const lattices::Vector<hmc_float>& norms);

/* ---- staggered (fermionmatrix_stagg.hpp:49-152) ---- */
class Fermionmatrix_stagg_eo : public Fermionmatrix_basic {
  public:
    virtual void operator()(const lattices::Staggeredfield_eo* out, const lattices::Gaugefield& gf,
                            const lattices::Staggeredfield_eo& in,
                            const physics::AdditionalParameters* additionalParameters = nullptr) const = 0;
    virtual bool hasMinimumEigenvalueThreshold() const = 0;
    virtual hmc_float getThresholdForMinimumEigenvalue(hmc_float mass) const;

  protected:
    Fermionmatrix_stagg_eo(bool herm, const hardware::System& system, const FermionmatrixParametersInterface& p)
        : Fermionmatrix_basic(system, p, herm)
    {
    }
};

class D_KS_eo final : public Fermionmatrix_stagg_eo {
  public:
    D_KS_eo(const hardware::System& system, const FermionStaggeredEoParametersInterface& p, bool evenodd)
        : Fermionmatrix_stagg_eo(false, system, p), evenodd(evenodd)
    {
    }
    void operator()(const lattices::Staggeredfield_eo* out, const lattices::Gaugefield& gf,
                    const lattices::Staggeredfield_eo& in,
                    const physics::AdditionalParameters* additionalParameters = nullptr) const override;
    cl_ulong get_flops() const override;
    cl_ulong get_read_write_size() const override;
    bool hasMinimumEigenvalueThreshold() const override { return false; }
    bool get_parity() const noexcept { return evenodd; }

  private:
    const bool evenodd;
};

class MdagM_eo final : public Fermionmatrix_stagg_eo {
  public:
    MdagM_eo(const hardware::System& system, const FermionStaggeredEoParametersInterface& p, bool upper_left = EVEN);
    void operator()(const lattices::Staggeredfield_eo* out, const lattices::Gaugefield& gf,
                    const lattices::Staggeredfield_eo& in,
                    const physics::AdditionalParameters* additionalParameters = nullptr) const override;
    cl_ulong get_flops() const override;
    cl_ulong get_read_write_size() const override;
    bool hasMinimumEigenvalueThreshold() const override { return true; }
    hmc_float getThresholdForMinimumEigenvalue(hmc_float mass) const override { return mass * mass; }
    bool get_upper_left() const noexcept { return upper_left; }

  private:
    const bool upper_left;
    lattices::Staggeredfield_eo tmp;
    mutable std::unique_ptr<lattices::Staggeredfield_eo> tmp_lo;  // for reduced-precision applications
};

/* (A + sigma[r]) on every right-hand side of a staggered block with one pass over the links per D_KS
 * (nrhs 2, 4, 8, 12, 16; tmp holds one block of in's shape).  No counterpart in the reference. */
void apply_block(const MdagM_eo& A, const lattices::Staggeredfield_eo_block* out, const lattices::Gaugefield& gf,
                 const lattices::Staggeredfield_eo_block& in, const physics::AdditionalParameters& ap,
                 const std::vector<hmc_float>& sigma, const lattices::DeviceBuffer& tmp);
/* The operator half of a CG-M iteration on A + sigma (cl2_MdagM_eo_cgm_step): beta_prev = beta,
 * beta = -rr_old / <p, (A + sigma) p>, r += beta (A + sigma) p (never stored; tmp is a field of p's precision) and
 * rr_new = |r|^2, all device scalars.  No counterpart in the reference. */
void cgm_step(const MdagM_eo& A, const lattices::Staggeredfield_eo* r, const lattices::Gaugefield& gf,
              const lattices::Staggeredfield_eo& p, const physics::AdditionalParameters& ap, hmc_float sigma,
              const lattices::Staggeredfield_eo& tmp, const lattices::Scalar<hmc_float>& rr_old,
              const lattices::Scalar<hmc_float>& pap, const lattices::Scalar<hmc_float>& beta,
              const lattices::Scalar<hmc_float>& beta_prev, const lattices::Scalar<hmc_float>& rr_new);
/* The operator half of a lockstep CG step on (A + sigma[r]) (cl2_MdagM_eo_mrhs_cg_step): pAp, alpha,
 * r_r -= alpha_r (A + sigma_r) p_r and the residual norms, without storing A p.  No counterpart in the
 * reference; returns true (the staggered operator always has it). */
bool cg_step_block(const MdagM_eo& A, const lattices::Staggeredfield_eo_block* r, const lattices::Gaugefield& gf,
                   const lattices::Staggeredfield_eo_block& p, const physics::AdditionalParameters& ap,
                   const std::vector<hmc_float>& sigma, const lattices::DeviceBuffer& tmp,
                   const lattices::Vector<hmc_float>& r2, const lattices::Vector<hmc_float>& active,
                   const lattices::Vector<hmc_float>& pAp, const lattices::Vector<hmc_float>& alpha,
                   const lattices::Vector<hmc_float>& norms);

}  // namespace fermionmatrix
}  // namespace physics
