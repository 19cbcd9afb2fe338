/*
 * cl2qcd.h -- C ABI of the MI355X (gfx950) fermion-inversion kernels.
 *
 * This is the drop-in boundary under the C++ operator/solver layer
 * (physics::fermionmatrix, physics::algorithms::solvers).  Every entry point
 * enqueues on the context's HIP stream, takes plain device pointers, and
 * returns 0 on success or a CL2_E* code; cl2_last_error() gives the message.
 * Each declaration names the reference interface it replaces (paths relative
 * to the CL2QCD source tree).
 *
 * Device layouts (spinor fields are dense, stride = number of eo sites Veo):
 *   Wilson eo spinor  F64: double2[12][Veo]  component c = 3*spin + colour
 *                     F32: float4 [6][Veo]   components 2q, 2q+1 as (re, im, re, im), 16-byte aligned
 *                     F16: half[3][Veo][8] (components 4j..4j+3 as re,im pairs) then float norm[Veo];
 *                          component = half * norm (norm = largest |re|,|im| on the site)
 *   staggered eo vec  F64: double2[3][Veo],  F32: float2[3][Veo]
 *   gauge field      F64_18: double2[4 dir][2 parity][Vp/64][9][64]   (9 lanes = row-major 3x3)
 *                    F64_12: double2[4][2][Vp/64][6][64]   rows 0,1; row 2 = conj(row0 x row1)
 *                    F32_12: float4 [4][2][Vp/64][3][64]   rows 0,1 packed two entries per float4
 *                            (tiles of 64 eo sites, lane-major inside a tile; Vp = Veo rounded up to 64)
 *                    F16_12: per (dir, parity) tiles of 64 eo sites: half[64][8] (entries 0..3 of rows 0,1)
 *                            then half[64][4] (entries 4, 5); 24 B per link
 *   device scalars: complex = 2 doubles (re, im); real = 1 double.
 * Host arrays use the reference's AoS order: site = x + NS*(y + NS*(z + NS*t)),
 * eo index = site/2, link index = mu + 4*site with mu = T,X,Y,Z.
 */
#ifndef CL2QCD_H
#define CL2QCD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define CL2_OK 0
#define CL2_E_ARG 1      /* invalid argument (std::invalid_argument / Invalid_Parameters in the reference) */
#define CL2_E_HIP 2      /* HIP runtime error (OpenclException in the reference) */
#define CL2_E_STUCK 3    /* solver stuck: NaN or vanishing rho (SolverStuck, solvers/exceptions.hpp:35-50) */
#define CL2_E_NOT_CONVERGED 4 /* maximum iterations reached (SolverDidNotSolve, solvers/exceptions.hpp:52-64) */
#define CL2_E_UNSUPPORTED 5

/* ---- parities (common_header_files/globaldefs.hpp:38-39) ---- */
#define CL2_EVEN 0
#define CL2_ODD 1

/* ---- precisions and gauge formats ---- */
#define CL2_F64 0
#define CL2_F32 1
#define CL2_F16 2 /* Wilson eo spinors only: half storage with one float scale per site; BLAS in FP32, the
                     hopping term with half spin projections and half dot products accumulated in FP32 */
#define CL2_GAUGE_F64_18 0
#define CL2_GAUGE_F64_12 1
#define CL2_GAUGE_F32_12 2
#define CL2_GAUGE_F16_12 3 /* Wilson operators on F16 spinors */

/* ---- field kinds for cl2_field_bytes ---- */
#define CL2_FIELD_SPINOR_EO 0
#define CL2_FIELD_SU3VEC_EO 1
#define CL2_FIELD_GAUGE 2

/* ---- fermion actions (common::action) and eo operators ---- */
#define CL2_ACTION_WILSON 0
#define CL2_ACTION_TWISTEDMASS 1
#define CL2_OP_AEE 0
#define CL2_OP_AEE_AND_GAMMA5 1
#define CL2_OP_AEE_MINUS 2
#define CL2_OP_AEE_MINUS_AND_GAMMA5 3
#define CL2_OP_QPLUS_EO 4
#define CL2_OP_QMINUS_EO 5
#define CL2_OP_QPLUSQMINUS_EO 6

/* ---- twisted-mass site-diagonal variants ---- */
#define CL2_MTM_SITEDIAGONAL 0
#define CL2_MTM_INVERSE_SITEDIAGONAL 1
#define CL2_MTM_SITEDIAGONAL_MINUS 2
#define CL2_MTM_INVERSE_SITEDIAGONAL_MINUS 3
#define CL2_MTM_SITEDIAGONAL_AND_GAMMA5 4
#define CL2_MTM_SITEDIAGONAL_MINUS_AND_GAMMA5 5

/* ---- scalar arguments: a host value or a device scalar ---- */
#define CL2_SCALAR_VALUE 0      /* use re, im */
#define CL2_SCALAR_DEV_COMPLEX 1 /* dev -> (re, im)          Scalar<hmc_complex> */
#define CL2_SCALAR_DEV_REAL 2    /* dev -> re                 Scalar<hmc_float> */
#define CL2_SCALAR_DEV_VECTOR 3  /* dev[index] (real)         Vector<hmc_float> element */
typedef struct {
    int kind;
    double re, im;
    const double* dev;
    int index;
} cl2_scalar;

/* ---- lattice: the reference's compile-time kernel macros as runtime values
 *      (hardware/code/opencl_module.cpp:61-225) ---- */
typedef struct {
    int ns;            /* NSPACE, even */
    int nt;            /* NTIME, even */
    double theta_s;    /* fermion BC: every spatial link weighted by exp(i pi theta_s / ns) */
    double theta_t;    /* every temporal link weighted by exp(i pi theta_t / nt) */
    int use_cp_re;     /* _CP_REAL_: T links times exp(+-cp_re) */
    double cp_re;
    int use_cp_im;     /* _CP_IMAG_: T links times exp(i cp_im) */
    double cp_im;
} cl2_lattice;

typedef struct cl2_ctx cl2_ctx;

/* ==== context, memory, transfer ==== */

/* hardware::System / hardware::Device construction (hardware/system.hpp:64-107,
 * hardware/device.cpp:33-143) for one GPU; the context owns a HIP stream
 * (the in-order queue, device.cpp:79-80). */
int cl2_ctx_create(cl2_ctx** ctx, int device, const cl2_lattice* lat);
int cl2_ctx_destroy(cl2_ctx* ctx);
const char* cl2_last_error(const cl2_ctx* ctx);
/* Device::synchronize (hardware/device.cpp:147-150) */
int cl2_sync(cl2_ctx* ctx);
void* cl2_get_stream(cl2_ctx* ctx);
long cl2_eo_sites(const cl2_ctx* ctx);
/* bytes of one field of the given kind and precision / gauge format */
size_t cl2_field_bytes(const cl2_ctx* ctx, int kind, int prec_or_format);
/* buffer creation (hardware/buffers/buffer.cpp); memory is zeroed */
int cl2_alloc(cl2_ctx* ctx, void** dev, size_t bytes);
int cl2_free(cl2_ctx* ctx, void* dev);
/* hardware::buffers::copyData (physics/lattices/util.hpp:64-78) */
int cl2_copy(cl2_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Spinorfield_eo::import / export with AoS<->SoA conversion
 * (physics/lattices/spinorfield_eo.cpp:54-120, ocl_kernel/spinorfield_eo_convert.cl:53-67) */
int cl2_spinor_import(cl2_ctx* ctx, void* dev, const double* host_aos, int prec);
int cl2_spinor_export(cl2_ctx* ctx, double* host_aos, const void* dev, int prec);
/* Staggeredfield_eo::import / export (physics/lattices/staggeredfield_eo.cpp) */
int cl2_su3vec_import(cl2_ctx* ctx, void* dev, const double* host_aos, int prec);
int cl2_su3vec_export(cl2_ctx* ctx, double* host_aos, const void* dev, int prec);
/* Gaugefield upload and SoA conversion (hardware/lattices/gaugefield.cpp:73-102,
 * ocl_kernel/gaugefield_convert.cl:22-48) into the chosen format */
int cl2_gauge_import(cl2_ctx* ctx, void* dev, const double* host_aos, int format);
/* Gaugefield hot / cold start (physics/lattices/gaugefield.cpp:35-72, hardware::lattices::Gaugefield::set_hot /
 * set_cold): every link U = reunitarised (1 + epsilon * complex Gaussian) from a counter-based generator keyed by
 * seed; epsilon = 0 gives the cold (unit) field, epsilon >~ 10 a Haar-like hot field. */
int cl2_gauge_set_random(cl2_ctx* ctx, void* dev, int format, uint64_t seed, double epsilon);
/* su3heatbath + overrelax (physics/algorithms/su3heatbath.cpp:33-73, hardware/code/heatbath.cpp) for the
 * Wilson gauge action at coupling beta on an F64_18 field: n_sweeps times, one heatbath sweep over all links
 * followed by n_overrelax overrelaxation sweeps.  Cabibbo-Marinari SU(2) subgroups, Kennedy-Pendleton
 * sampling, random numbers from a counter hash keyed by seed (the reference's ranluxcl streams depend on
 * the device's thread count and are not reproduced). */
int cl2_gauge_heatbath(cl2_ctx* ctx, void* gauge, double beta, int n_sweeps, int n_overrelax, uint64_t seed);
/* n_steps overrelaxation sweeps (physics/algorithms/su3heatbath.cpp:61-73); deterministic */
int cl2_gauge_overrelax(cl2_ctx* ctx, void* gauge, int n_steps);
/* average plaquette (1/(18 V)) sum_x sum_{mu<nu} Re tr P_mu,nu(x) of an F64_18 field
 * (physics/observables/gaugeObservables.cpp, ocl_kernel/gaugeobservables_plaquette.cl) */
int cl2_gauge_plaquette(cl2_ctx* ctx, const void* gauge, double* plaquette);
/* Re-layout a resident F64_18 field into another format (feeds reduced-precision solvers). */
int cl2_gauge_convert(cl2_ctx* ctx, void* dst, int dst_format, const void* src_f64_18);
/* Gaugefield export to host AoS (hardware/lattices/gaugefield.cpp:104-124); 12-real formats are
 * returned with the reconstructed third row. */
int cl2_gauge_export(cl2_ctx* ctx, double* host_aos, const void* dev, int format);
/* Scalar<T>::store / get (hardware/lattices/scalar.hpp:41-110) */
int cl2_scalar_store(cl2_ctx* ctx, double* dev, const double* host, int n);
int cl2_scalar_get(cl2_ctx* ctx, double* host, const double* dev, int n);

/* ==== Wilson / twisted-mass kernels (hardware/code/fermions.hpp:47-198) ==== */

/* Fermions::dslash_eo_device (fermions.cpp:430-468): -kappa * hopping term.
 * evenodd == CL2_ODD writes the even sites from odd neighbours, CL2_EVEN the odd
 * sites (ocl_kernel/fermionmatrix_eo_dslash.cl:50-58).  The gauge format selects
 * the arithmetic: F64_* on F64 spinors, F32_12 on F32 spinors. */
int cl2_dslash_eo(cl2_ctx* ctx, void* out, const void* in, const void* gauge, int format, int evenodd,
                  double kappa);
/* Fermions::dslash_AND_M_tm_inverse_sitediagonal[_minus]_eo_device (fermions.cpp:550-624) */
int cl2_dslash_AND_M_tm_inverse_sitediagonal_eo(cl2_ctx* ctx, void* out, const void* in, const void* gauge,
                                                int format, int evenodd, double kappa, double mubar, int minus);
/* Fused second half of Aee / Aee_minus [/ _AND_gamma5] (physics/fermionmatrix/fermionmatrix.cpp:250-257,
 * 338, 432-439, 522): out = diag(f1 on spins 0,1; f2 on spins 2,3) * x - dslash(in) [then gamma5].
 * No single reference kernel; replaces M_tm_sitediagonal + dslash_eo + saxpy[_AND_gamma5]. */
int cl2_dslash_AND_sitediagonal_saxpy_eo(cl2_ctx* ctx, void* out, const void* in, const void* x, const void* gauge,
                                         int format, int evenodd, double kappa, double f1_re, double f1_im,
                                         double f2_re, double f2_im, int gamma5);
/* Fermions::M_tm_sitediagonal*, M_tm_inverse_sitediagonal*, *_AND_gamma5_eo (fermions.cpp:626-809) */
int cl2_M_tm_sitediagonal_eo(cl2_ctx* ctx, void* out, const void* in, int prec, double mubar, int variant);
/* Fermions::gamma5_eo_device (fermions.cpp:384-396) */
int cl2_gamma5_eo(cl2_ctx* ctx, void* inout, int prec);
/* Fermions::saxpy_AND_gamma5_eo_device (fermions.cpp:398-428): out = gamma5 (y - alpha x) */
int cl2_saxpy_AND_gamma5_eo(cl2_ctx* ctx, void* out, const void* x, const void* y, int prec, const cl2_scalar* alpha);
/* Whole eo operator with fused kernels: Aee, Aee_AND_gamma5_eo, Aee_minus, Aee_minus_AND_gamma5_eo,
 * Qplus_eo, Qminus_eo, QplusQminus_eo (physics/fermionmatrix/fermionmatrix.cpp:224-664).
 * tmp: caller-provided scratch of 2 spinor fields of the operator's precision. */
int cl2_fermionmatrix_eo(cl2_ctx* ctx, int op, void* out, const void* in, const void* gauge, int format, int action,
                         double kappa, double mubar, void* tmp);
/* QplusQminus_eo (op must be CL2_OP_QPLUSQMINUS_EO) as cl2_fermionmatrix_eo, and dot (2 device doubles) =
 * <in, out> = |Qminus_eo in|^2 (Qplus_eo = Qminus_eo^dagger), summed in FP64 by the kernel that writes
 * Qminus_eo in.  Replaces the operator call and scalar_product(p, v) of cg (solvers/cg.cpp:327-330), one
 * reading pass over in and out fewer. */
int cl2_fermionmatrix_eo_dot(cl2_ctx* ctx, int op, void* out, const void* in, const void* gauge, int format,
                             int action, double kappa, double mubar, void* tmp, double* dot);
/* The operator half of a CG step on QplusQminus_eo (op must be CL2_OP_QPLUSQMINUS_EO), replacing
 * solvers/cg.cpp:327-341 (v = A p, rho = <p, v>, alpha = omega / rho, r -= alpha v with its norm):
 * pAp (2 device doubles) = <p, A p> = |Qminus_eo p|^2, alpha (2) = r2 / pAp (r2: complex device scalar), and
 * r <- r - alpha A p, with norms (4) = (|r_new|^2, Re <r_new, r_old>, |r_new|^2, 0) as
 * cl2_saxpy_AND_cgnorm_eo gives them.  A p is never stored: the last hopping kernel updates r in its
 * epilogue.  tmp: 2 spinor fields; r, p and tmp distinct. */
int cl2_fermionmatrix_eo_cg_step(cl2_ctx* ctx, int op, void* r, const void* p, const void* gauge, int format,
                                 int action, double kappa, double mubar, void* tmp, const double* r2, double* pAp,
                                 double* alpha, double* norms);

/* ==== Wilson eo BLAS (hardware/code/spinors.hpp:50-224) ==== */

/* saxpy_eoprec / saxpy_arg_eoprec (spinors.cpp:462-521): out = y - alpha x */
int cl2_saxpy_eo(cl2_ctx* ctx, void* out, const void* x, const void* y, int prec, const cl2_scalar* alpha);
/* sax_eoprec (spinors.cpp:523-551): out = alpha x */
int cl2_sax_eo(cl2_ctx* ctx, void* out, const void* x, int prec, const cl2_scalar* alpha);
/* saxsbypz_eoprec (spinors.cpp:553-623): out = alpha x + beta y + z */
int cl2_saxsbypz_eo(cl2_ctx* ctx, void* out, const void* x, const void* y, const void* z, int prec,
                    const cl2_scalar* alpha, const cl2_scalar* beta);
/* scalar_product_eoprec + reduction (spinors.cpp:670-713): res = sum conj(x) y, complex, on device */
int cl2_scalar_product_eo(cl2_ctx* ctx, double* res, const void* x, const void* y, int prec);
/* No counterpart in the reference: cl2_scalar_product_eo for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same result, bit
 * for bit; FP64 or FP32 fields.  Its partial sums go to a library-owned 1 MiB device array (65536 partials,
 * the most one reduction launch writes, whatever the field size) whose address is looked up once per device
 * by cl2_ctx_create; the call itself allocates nothing.  Every call on one device shares that array, so
 * calls that overlap on one device (different streams or contexts, concurrently) race. */
int cl2_scalar_product_eo_noalloc(cl2_ctx* ctx, double* res, const void* x, const void* y, int prec);
/* global_squarenorm_eoprec + reduction (spinors.cpp:715-789): res = sum |x|^2, real, on device */
int cl2_squarenorm_eo(cl2_ctx* ctx, double* res, const void* x, int prec);
/* saxpy_AND_squarenorm_eo (spinors.cpp:858-916): out = y - alpha x; res = |out|^2 (complex (norm, 0)) */
int cl2_saxpy_AND_squarenorm_eo(cl2_ctx* ctx, void* out, const void* x, const void* y, int prec,
                                const cl2_scalar* alpha, double* res);
/* No counterpart in the reference: the fused update of the reduced-precision CG.  out = y - alpha x (out may
 * alias y); res (4 device doubles) = (|out|^2, Re <out, y>, |out|^2, 0) with y read before it is overwritten:
 * res[0] - res[1] = <r_new, r_new - r_old> gives the Polak-Ribiere beta of a CG step, and res + 2 is the new
 * |r|^2 as a complex device scalar for the next step's alpha. */
int cl2_saxpy_AND_cgnorm_eo(cl2_ctx* ctx, void* out, const void* x, const void* y, int prec, const cl2_scalar* alpha,
                            double* res);
/* The two search-space updates of a CG step in one pass (solvers/cg.cpp:334 and :149, Wilson saxpy
 * convention y - a x): x <- x - alpha p, then p <- r - beta p, both from the same p.  kind:
 * CL2_FIELD_SPINOR_EO or CL2_FIELD_SU3VEC_EO; x, p, r distinct fields of precision prec. */
int cl2_cg_update_xp(cl2_ctx* ctx, int kind, void* x, void* p, const void* r, int prec, const cl2_scalar* alpha,
                     const cl2_scalar* beta);
/* set_zero_spinorfield_eoprec / set_eoprec_spinorfield_cold (spinors.cpp:448-460,805-817):
 * cold = every complex component 1/sqrt(12 V) */
/* Spinorfield_eo::gaussian / Staggeredfield_eo::set_gaussian (physics/lattices/spinorfield_eo.hpp:93,
 * staggeredfield_eo.hpp:87; ocl_kernel/spinorfield_eo_gaussian.cl):
 * every real component N(0, 1) from a counter hash keyed by seed (the reference draws from ranluxcl with
 * variance 1/2; fields here are reproducible from the seed alone).  kind: CL2_FIELD_SPINOR_EO or
 * CL2_FIELD_SU3VEC_EO, FP64 or FP32. */
int cl2_gaussian_eo(cl2_ctx* ctx, int kind, void* out, int prec, uint64_t seed);
/* the summand of the pseudoscalar point-source correlator (ocl_kernel/fermionobservables/correlator_ps_point.cl
 * correlator_ps_t, correlator_ps_z; direction dispatch physics/observables/wilsonTwoFlavourCorrelators.cpp:357):
 * res[s] (= or +=, device doubles) of sum over the eo sites of slice s of |x|^2, slices along dir = 0 (t, nt
 * entries) or dir = 3 (z, ns entries); FP64 or FP32 Wilson eo field */
int cl2_slice_squarenorm_eo(cl2_ctx* ctx, double* res, const void* x, int prec, int dir, int accumulate);
int cl2_set_zero_eo(cl2_ctx* ctx, void* out, int prec);
int cl2_set_cold_eo(cl2_ctx* ctx, void* out, int prec);

/* Copy a spinor / staggered field into another precision (F64 <-> F32, or a plain copy);
 * no reference counterpart: the reference runs in one precision throughout
 * (CMakeLists.txt:127-132).  Feeds the reduced-precision inner solvers. */
int cl2_convert_precision_eo(cl2_ctx* ctx, int kind, void* dst, int dst_prec, const void* src, int src_prec);

/* ==== blocks of right-hand sides ====
 * No counterpart in the reference, which inverts the 12 spin-colour sources of a propagator one after
 * the other (physics/algorithms/inversion.cpp:78-242).  A block holds nrhs Wilson eo spinors as one
 * double2[12][Veo][nrhs] (FP64) or float4[6][Veo][nrhs] (FP32, component pairs as in a single field)
 * array, right-hand side fastest, or (FP16, F16_12 links) as one FP16 field of Veo*nrhs sites, right-hand
 * side r of eo site i at site i*nrhs + r, so one launch of an operator reuses every loaded link for all
 * right-hand sides.  Per-right-hand-side scalars are device arrays (complex: 2 doubles each). */
/* the eo operators of cl2_fermionmatrix_eo / cl2_dslash_eo on nrhs = 1, 2, 4, 8, 12 or 16 right-hand
 * sides; tmp = 2 block fields */
int cl2_fermionmatrix_eo_mrhs(cl2_ctx* ctx, int op, void* out, const void* in, const void* gauge, int format,
                              int action, double kappa, double mubar, void* tmp, int nrhs);
/* the same with an operator per right-hand side on one gauge field: kappa[r] and mubar[r] (host arrays of
 * nrhs values; mubar = 2 kappa mu, unused for Wilson) for right-hand side r, nrhs = 2 .. 16.  Lets requests
 * with different quark masses on one configuration share every link load (serving). */
int cl2_fermionmatrix_eo_mrhs_params(cl2_ctx* ctx, int op, void* out, const void* in, const void* gauge,
                                     int format, int action, const double* kappa, const double* mubar, void* tmp,
                                     int nrhs);
/* The operator half of a lockstep CG step on nrhs right-hand sides, QplusQminus_eo only (op must be
 * CL2_OP_QPLUSQMINUS_EO), as cl2_fermionmatrix_eo_cg_step per right-hand side: pAp[r] (complex, 2 doubles
 * each) = |Qminus_eo p_r|^2, alpha[r] = r2[r] / pAp[r] where active[r] != 0 (else 0, as cl2_block_cg_alpha),
 * r_r <- r_r - alpha[r] A p_r, norms[r] (2 doubles) = (|r_new|^2, Re <r_new, r_old>) as cl2_block_axpy_eo
 * gives them.  Replaces, for the block solver, the block operator call, cl2_block_dot_eo, cl2_block_cg_alpha
 * and the residual cl2_block_axpy_eo; A p is never stored.  kappas/mubars: host arrays of nrhs values for an
 * operator per right-hand side, or null for kappa/mubar.  Any link format (FP16 blocks with F16_12 links);
 * r2, active, pAp, alpha, norms are device arrays; tmp: 2 blocks. */
int cl2_fermionmatrix_eo_mrhs_cg_step(cl2_ctx* ctx, int op, void* r, const void* p, const void* gauge, int format,
                                      int action, double kappa, double mubar, const double* kappas,
                                      const double* mubars, void* tmp, int nrhs, const double* r2,
                                      const double* active, double* pAp, double* alpha, double* norms);
int cl2_dslash_eo_mrhs(cl2_ctx* ctx, void* out, const void* in, const void* gauge, int format, int evenodd,
                       double kappa, int nrhs);
/* copy a single field into / out of slot r of a block (nrhs in [1, 64]; FP64, FP32 or FP16 Wilson blocks, and
 * likewise for the block BLAS below) */
int cl2_block_insert_eo(cl2_ctx* ctx, void* block, const void* single, int prec, int nrhs, int r);
int cl2_block_extract_eo(cl2_ctx* ctx, void* single, const void* block, int prec, int nrhs, int r);
/* out_r = y_r + sign * coef_r x_r (coef: nrhs complex); out may alias y.  If norms is not null it receives
 * per right-hand side (|out_r|^2, Re <out_r, y_r>) with y read before it is overwritten. */
int cl2_block_axpy_eo(cl2_ctx* ctx, void* out, const void* x, const void* y, int prec, int nrhs, const double* coef,
                      double sign, double* norms);
/* No reference counterpart (the lockstep block CG's x and p updates, solvers/cg.cpp:334 and :149 for every
 * right-hand side, in one pass): x_r = x_r + alpha_r p_r, then p_r = r_r + beta_r p_r (alpha, beta: nrhs
 * complex on the device); x, p, r distinct blocks. */
int cl2_block_cg_update_xp_eo(cl2_ctx* ctx, void* x, void* p, const void* r, int prec, int nrhs, const double* alpha,
                              const double* beta);
/* res[r] = <x_r, y_r> (complex, 2 doubles each) / |x_r|^2 (one double each) */
int cl2_block_dot_eo(cl2_ctx* ctx, double* res, const void* x, const void* y, int prec, int nrhs);
int cl2_block_squarenorm_eo(cl2_ctx* ctx, double* res, const void* x, int prec, int nrhs);
int cl2_block_convert_eo(cl2_ctx* ctx, void* dst, int dst_prec, const void* src, int src_prec, int nrhs);
int cl2_block_set_zero_eo(cl2_ctx* ctx, void* out, int prec, int nrhs);
/* Blocks of nrhs staggered eo vectors [3][Veo][nrhs] (no reference counterpart; the operations of the Wilson
 * blocks above on staggered fields, and the lockstep refinement of the shifts of cg_m, solver_shifted.cpp):
 * D_KS_eo and MdagM_eo on every right-hand side with one pass over the links (nrhs 2, 4, 8, 12, 16, FP64 or
 * FP32 links), MdagM with a shift per right-hand side: out_r = (mass^2 + sigma[r]) in_r - D_KS D_KS in_r
 * (sigma: host array of nrhs values, or null for none); slot r of D_KS_eo_mrhs equals cl2_D_KS_eo bit for bit. */
int cl2_D_KS_eo_mrhs(cl2_ctx* ctx, void* out, const void* in, const void* gauge, int format, int evenodd, int nrhs);
int cl2_MdagM_eo_mrhs(cl2_ctx* ctx, void* out, const void* in, const void* gauge, int format, double mass,
                      int upper_left, const double* sigma, void* tmp, int nrhs);
/* The operator half of a lockstep CG step for MdagM_eo + sigma_r on nrhs right-hand sides, as
 * cl2_fermionmatrix_eo_mrhs_cg_step is for QplusQminus_eo; no counterpart in the reference.  The first D_KS_eo
 * kernel (w = D p into tmp) also sums |w_r|^2 and |p_r|^2, so pAp[r] (complex, 2 doubles) = (m^2 + sigma[r])
 * |p_r|^2 + |w_r|^2 (D_KS_eo is anti-hermitian) and alpha[r] = r2[r] / pAp[r] where active[r] != 0 (else 0) are
 * formed before the second kernel, whose epilogue updates r_r <- r_r - alpha[r] A p_r and sums norms[r] =
 * (|r_new|^2, Re <r_new, r_old>); A p is never stored.  Replaces, for the block solver, cl2_MdagM_eo_mrhs,
 * cl2_block_dot_stagg_eo, cl2_block_cg_alpha and the residual cl2_block_axpy_stagg_eo (values agree to rounding).
 * sigma: host array of nrhs shifts or null; r2, active, pAp, alpha, norms: device arrays; tmp: one block. */
int cl2_MdagM_eo_mrhs_cg_step(cl2_ctx* ctx, void* r, const void* p, const void* gauge, int format, double mass,
                              int upper_left, const double* sigma, void* tmp, int nrhs, const double* r2,
                              const double* active, double* pAp, double* alpha, double* norms);
int cl2_block_insert_stagg_eo(cl2_ctx* ctx, void* block, const void* single, int prec, int nrhs, int r);
int cl2_block_extract_stagg_eo(cl2_ctx* ctx, void* single, const void* block, int prec, int nrhs, int r);
int cl2_block_axpy_stagg_eo(cl2_ctx* ctx, void* out, const void* x, const void* y, int prec, int nrhs,
                            const double* coef, double sign, double* norms);
int cl2_block_dot_stagg_eo(cl2_ctx* ctx, double* res, const void* x, const void* y, int prec, int nrhs);
int cl2_block_squarenorm_stagg_eo(cl2_ctx* ctx, double* res, const void* x, int prec, int nrhs);
int cl2_block_convert_stagg_eo(cl2_ctx* ctx, void* dst, int dst_prec, const void* src, int src_prec, int nrhs);
int cl2_block_set_zero_stagg_eo(cl2_ctx* ctx, void* out, int prec, int nrhs);
int cl2_block_cg_update_xp_stagg_eo(cl2_ctx* ctx, void* x, void* p, const void* r, int prec, int nrhs,
                                    const double* alpha, const double* beta);
/* lockstep CG scalars: alpha_r = r2_r / pAp_r (complex) where active_r != 0, else 0;
 * beta_r = Polak-Ribiere (norms_r[0] - norms_r[1]) / r2_r (Fletcher-Reeves norms_r[0] / r2_r if negative)
 * where active_r != 0, else 0, and then r2_r = norms_r[0] */
int cl2_block_cg_alpha(cl2_ctx* ctx, double* alpha, const double* r2, const double* pAp, const double* active,
                       int nrhs);
int cl2_block_cg_beta(cl2_ctx* ctx, double* beta, double* r2, const double* norms, const double* active, int nrhs);

/* ==== fused multi-shift update ====
 * SolverShifted::updateVectorQuantities (physics/algorithms/solver_shifted.cpp:174-198), which launches
 * saxpy_stagg_eo and saxpby_stagg_eo once per unconverged shift, as one pass: for every k < nshift with
 * active[k] != 0,  x[k] = beta[k] p[k] + x[k]  and then  p[k] = zeta[k] r + alpha[k] p[k].  x and p are
 * host arrays of nshift device pointers to staggered eo fields; beta, zeta, alpha, active are device arrays
 * (the reference's beta_vec, zeta_foll, alpha_vec and a convergence mask).  r and p are in precision prec,
 * the x[k] in x_prec (FP64 x with FP32 r, p accumulates a reduced-precision recurrence's solutions in FP64). */
int cl2_multishift_update_stagg_eo(cl2_ctx* ctx, void* const* x, int x_prec, void* const* p, const void* r, int prec,
                                   int nshift, const double* beta, const double* zeta, const double* alpha,
                                   const double* active);

/* ==== staggered (hardware/code/fermions_staggered.hpp:46-91, spinors_staggered.hpp) ==== */

/* Fermions_staggered::D_KS_eo_device (fermions_staggered.cpp:141-168); output parity == evenodd
 * (ocl_kernel/fermionmatrix_staggered_eo_DKS.cl:50-69) */
int cl2_D_KS_eo(cl2_ctx* ctx, void* out, const void* in, const void* gauge, int format, int evenodd);
/* MdagM_eo (physics/fermionmatrix/fermionmatrix_stagg.cpp:72-91): out = m^2 in - D D in;
 * upper_left = CL2_EVEN: D_eo D_oe.  tmp: one staggered field. */
int cl2_MdagM_eo(cl2_ctx* ctx, void* out, const void* in, const void* gauge, int format, double mass,
                 int upper_left, void* tmp);
/* saxpy_{cplx,real,real_vec}_staggered_eoprec: out = alpha x + y */
int cl2_saxpy_stagg_eo(cl2_ctx* ctx, void* out, const void* x, const void* y, int prec, const cl2_scalar* alpha);
/* saxpby_{cplx,real,real_vec}_staggered_eoprec: out = alpha x + beta y */
int cl2_saxpby_stagg_eo(cl2_ctx* ctx, void* out, const void* x, const void* y, int prec, const cl2_scalar* alpha,
                        const cl2_scalar* beta);
/* sax_{cplx,real,real_vec}_staggered_eoprec: out = alpha x */
int cl2_sax_stagg_eo(cl2_ctx* ctx, void* out, const void* x, int prec, const cl2_scalar* alpha);
/* saxpbypz_cplx_staggered_eoprec: out = alpha x + beta y + z */
int cl2_saxpbypz_stagg_eo(cl2_ctx* ctx, void* out, const void* x, const void* y, const void* z, int prec,
                          const cl2_scalar* alpha, const cl2_scalar* beta);
/* scalar_product_staggered_eoprec (complex) and _real_part (real) */
int cl2_scalar_product_stagg_eo(cl2_ctx* ctx, double* res, const void* x, const void* y, int prec);
int cl2_scalar_product_real_part_stagg_eo(cl2_ctx* ctx, double* res, const void* x, const void* y, int prec);
/* global_squarenorm_staggered_eoprec */
int cl2_squarenorm_stagg_eo(cl2_ctx* ctx, double* res, const void* x, int prec);
/* sax_vectorized_and_squarenorm_eoprec (spinorfield_staggered_eo_sax_AND_squarenorm.cl:33-154):
 * res[k] = |alpha[k] x|^2 for k < n, one pass over x */
int cl2_sax_vec_AND_squarenorm_stagg_eo(cl2_ctx* ctx, double* res, const double* alpha, int n, const void* x,
                                        int prec);
/* set_zero / set_cold_spinorfield_stagg_eoprec: cold = 1/sqrt(3 V) per component */
int cl2_set_zero_stagg_eo(cl2_ctx* ctx, void* out, int prec);
int cl2_set_cold_stagg_eo(cl2_ctx* ctx, void* out, int prec);

/* ==== device scalars (hardware/code/complex.cpp:123-170, real.cpp:189-400) ==== */
#define CL2_SCALAR_OP_PRODUCT 0
#define CL2_SCALAR_OP_RATIO 1
#define CL2_SCALAR_OP_SUM 2
#define CL2_SCALAR_OP_SUBTRACTION 3
/* complex_{product,ratio,sum,subtraction}: out = a (op) b on device complex scalars */
int cl2_complex_op(cl2_ctx* ctx, int op, double* out, const double* a, const double* b);
/* real_{product,ratio,sum,subtraction} */
int cl2_real_op(cl2_ctx* ctx, int op, double* out, const double* a, const double* b);
/* access_real_vector_element both ways (real_access_vector_element.cl) */
int cl2_real_get_vector_element(cl2_ctx* ctx, double* out, const double* vec, int index);
int cl2_real_set_vector_element(cl2_ctx* ctx, double* vec, const double* in, int index);
/* update_{zeta,beta,alpha}_cgm (real.cpp:285-400, ocl_kernel/operations_real.cl:20-39) */
int cl2_update_zeta_cgm(cl2_ctx* ctx, double* out, const double* zeta_prev, const double* zeta_prev_prev,
                        const double* sbeta_prev, const double* sbeta_pres, const double* salpha_prev,
                        const double* sigma, int numeq);
int cl2_update_beta_cgm(cl2_ctx* ctx, double* out, const double* sbeta_pres, const double* zeta_pres,
                        const double* zeta_prev, int numeq);
int cl2_update_alpha_cgm(cl2_ctx* ctx, double* out, const double* salpha_pres, const double* zeta_pres,
                         const double* beta_pres, const double* zeta_prev, const double* sbeta_pres, int numeq);
/* The scalar work of one CG-M iteration (solver_shifted.cpp:136-198) fused into two launches; no counterpart in
 * the reference, which makes thirteen (copies, ratio, subtraction and the three updates above).  Every value
 * equals what that sequence gives.  All arguments are device scalars or device vectors of numeq (1..64).
 * beta step, after <p, A p>:  beta_prev = beta; beta = 0 - rr / pap.
 * alpha step, after |r|^2:   alpha_prev = alpha; alpha = rr_new / rr_old; zeta_foll, beta_vec, alpha_vec as
 *   update_{zeta,beta,alpha}_cgm; then zeta_prev = zeta, zeta = zeta_foll, rr_old = rr_new. */
int cl2_cgm_beta_step(cl2_ctx* ctx, double* beta, double* beta_prev, const double* rr, const double* pap);
int cl2_cgm_alpha_step(cl2_ctx* ctx, int numeq, double* alpha, double* alpha_prev, const double* rr_new,
                       double* rr_old, const double* beta, const double* beta_prev, const double* sigma, double* zeta,
                       double* zeta_prev, double* zeta_foll, double* beta_vec, double* alpha_vec);
/* The operator half of a CG-M iteration on A = MdagM_eo + sigma (solver_shifted.cpp:136-198, from A p to |r|^2;
 * sigma 0 there) in two D_KS_eo kernels; no counterpart in the reference.  w = D_KS_eo p into tmp (one field of
 * p's precision) with (|w|^2, |p|^2) summed; pap = (m^2 + sigma) |p|^2 + |w|^2 (= <p, A p>, D_KS_eo
 * anti-hermitian) and the beta step (beta_prev = beta, beta = 0 - rr_old / pap); then r += beta A p (never
 * stored) and rr_new = |r|^2.  Replaces the operator, the <p, A p> reduction, cl2_cgm_beta_step, the residual
 * saxpy and the |r|^2 reduction (values agree to rounding).  The reduced-precision CG on a shifted system takes
 * the same step (beta = -alpha).  rr_old, pap, beta, beta_prev, rr_new: device scalars. */
int cl2_MdagM_eo_cgm_step(cl2_ctx* ctx, void* r, const void* p, const void* gauge, int format, double mass,
                          double sigma, int upper_left, void* tmp, const double* rr_old, double* pap, double* beta,
                          double* beta_prev, double* rr_new);

/* ==== captured launch sequences ====
 * No counterpart in the reference.  Everything enqueued on the context's stream between cl2_graph_begin and
 * cl2_graph_end is recorded instead of run (HIP stream capture) and replayed by cl2_graph_launch with one
 * submission: the fixed kernel sequence of a solver iteration.  Only kernel-launching entry points may be
 * recorded (no transfers, synchronisations or allocations); device pointers and device scalars keep the
 * values they have at replay time, host-value scalars the value they had when recorded. */
typedef struct cl2_graph cl2_graph;
int cl2_graph_begin(cl2_ctx* ctx);
int cl2_graph_end(cl2_ctx* ctx, cl2_graph** graph);
int cl2_graph_launch(cl2_ctx* ctx, cl2_graph* graph, int times);
int cl2_graph_destroy(cl2_graph* graph);
/* No counterpart in the reference: node census of a captured sequence -- its kernel nodes and all its
 * nodes -- so a caller can check that a capture recorded kernel launches only (no memset, copy or host
 * nodes, which replay differently). */
int cl2_graph_node_count(const cl2_graph* graph, int* kernel_nodes, int* all_nodes);
/* copyData between device scalars / vectors (physics/lattices/scalar.hpp, vector.hpp): n doubles, done by a
 * kernel so that it can be part of a recorded sequence */
int cl2_scalar_copy(cl2_ctx* ctx, double* dst, const double* src, int n);
/* No counterpart in the reference: the beta step of the reduced-precision CG on the device, so that a whole
 * iteration can be recorded.  norms = (|r_new|^2, Re <r_new, r_old>) from cl2_saxpy_AND_cgnorm_eo; mbeta
 * (complex) = -beta with the Polak-Ribiere beta (Fletcher-Reeves when negative), r2 (complex) = (|r_new|^2, 0)
 * after reading the old value. */
int cl2_cg_beta(cl2_ctx* ctx, double* mbeta, double* r2, const double* norms);

/* ==== launch tuning (no reference counterpart: the reference picks work sizes in
 *      hardware/device_info.cpp:111-128 and hardware/code/opencl_module.cpp get_work_sizes) ==== */
/* keys: "hop.zplanes" -- z-planes per slab of the t-fastest traversal of the hopping kernels
 *       (-1 heuristic, 0 lexicographic order, else a divisor of ns);
 *       "hop.xcd" -- 1: give each XCD a contiguous range of blocks, 0: hardware round robin;
 *       "hop.block" -- threads per block of the single-field hopping kernel (128, 256, 512);
 *       "hop.occupancy" -- its register budget: 0 (<= 2 waves per SIMD, the two hops of a direction in
 *       flight), 1 (one wave per SIMD, four hops in flight), 2 (<= 2 waves per SIMD with four hops in
 *       flight; FP32 links, else as 0), 4 (one hop at a time: at least 4 waves per SIMD for FP32 links, the
 *       compiler's choice for FP16, else as 0), -1 the format's default (1 for FP64 links, 2 for FP32, 0 for
 *       FP16);
 *       "mrhs.rb" -- right-hand sides per thread of the multi-right-hand-side kernel (0: by format, 1, 2, 4);
 *       "mrhs.lds" -- with one right-hand side per thread, 1 (default): a block computes each site's neighbour
 *       indices and loads its links once into LDS for all its right-hand-side threads (16 sites per block for
 *       8 or more right-hand sides, else 32), 2: the same with twice the sites per block, 0: every thread
 *       does;
 *       "mrhs.off32" -- 1 (default): the staged kernel's FP64 neighbour gathers use a scalar base per spinor
 *       component and 32-bit lane offsets while Veo * nrhs * 16 bytes < 4 GiB (beyond that, and with 0,
 *       64-bit offsets and one-size staged blocks).
 * Every setting changes only which thread computes which site and when: results are bit identical. */
int cl2_tune_set(cl2_ctx* ctx, const char* key, long value);
long cl2_tune_get(const cl2_ctx* ctx, const char* key);

/* No counterpart in the reference (which fixes work-group sizes per device in hardware/device.cpp:286-334):
 * autotune the hopping kernel's launch shape for (format, nrhs) on this context -- z-planes per slab, XCD
 * mapping, right-hand sides per thread, and for single fields the block size (128, 256, 512) and register
 * budget -- by timing every candidate on scratch fields with the given gauge field; the fastest is kept for
 * later launches (unless cl2_tune_set fixed the shape) and returned.  occupancy: the register budget, as
 * the "hop.occupancy" key of cl2_tune_set; rhs_per_thread: 1, 2 or 4, or -1 / -2 for one right-hand side per
 * thread with the links staged in LDS ("mrhs.lds" 1 / 2).  Output pointers may be null. */
int cl2_autotune_dslash(cl2_ctx* ctx, const void* gauge, int format, int nrhs, double* best_ms, int* zplanes,
                        int* xcd, int* rhs_per_thread, int* block, int* occupancy);

/* No counterpart in the reference: the same autotuning for the staggered hopping kernel (cl2_D_KS_eo and the
 * staggered operators built on it) with links of the given format (FP64 or FP32): z-planes per slab, XCD
 * mapping, block size and register budget. */
int cl2_autotune_dks(cl2_ctx* ctx, const void* gauge, int format, double* best_ms, int* zplanes, int* xcd,
                     int* block, int* occupancy);

/* ==== timing on the context stream (Device profiling events, hardware/device.cpp:233-242) ==== */
int cl2_timer_start(cl2_ctx* ctx);
int cl2_timer_stop(cl2_ctx* ctx, double* ms);

/* ==== byte / flop models (Opencl_Module::get_read_write_size / get_flop_size,
 *      hardware/code/fermions.cpp:811-975, fermions_staggered.cpp:180-280) ==== */
/* The reference's own counts for a named kernel at this lattice (FP64, 18-real links). */
uint64_t cl2_reference_read_write_size(const cl2_ctx* ctx, const char* kernel);
uint64_t cl2_reference_flop_size(const cl2_ctx* ctx, const char* kernel);
/* Compulsory HBM bytes of one launch of a named kernel in this build's layout
 * (each link, input and output moved once); the roofline numerator. */
uint64_t cl2_algorithmic_bytes(const cl2_ctx* ctx, const char* kernel, int format);

#ifdef __cplusplus
}
#endif
#endif
