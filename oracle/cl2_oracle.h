/*
 * CPU oracle for the CL2QCD fermion-inversion path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain FP64 C++ restatement of the
 * reference algorithms, used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py as the checker.  The product library never
 * links, loads or calls it.
 *
 * Host layouts (the reference's CPU/AoS layouts):
 *   gauge   : double[4*V][18]  link index = mu + 4*site, mu = T,X,Y,Z = 0..3,
 *             3x3 row-major complex (re,im)            (types.hpp:129-139)
 *   spinor  : double[Veo][24]  component c = 3*spin + colour, (re,im)
 *   su3vec  : double[Veo][6]
 *   full spinor field: double[V][24] indexed by site.
 *   site = x + NS*y + NS^2*z + NS^3*t; eo index = site/2
 *                                   (ocl_kernel/operations_geometry.cl:122-300)
 */
#ifndef CL2_ORACLE_H
#define CL2_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Lattice and action-independent kernel parameters
 * (hardware/code/opencl_module.cpp:61-225 compile-time macros). */
typedef struct {
    int ns;            /* spatial extent, even */
    int nt;            /* temporal extent, even */
    double theta_s;    /* spatial fermion BC phase: b_S = exp(i*pi*theta_s/ns) on every link */
    double theta_t;    /* temporal fermion BC phase */
    int use_cp_re;     /* real chemical potential on T links */
    double cp_re;
    int use_cp_im;     /* imaginary chemical potential on T links */
    double cp_im;
} orc_lattice;

/* status codes shared with the solvers */
#define ORC_OK 0
#define ORC_ERR_ARG 1
#define ORC_SOLVER_STUCK 3
#define ORC_SOLVER_DID_NOT_SOLVE 4

/* operators (physics/fermionmatrix/fermionmatrix.hpp:194-310) */
#define ORC_OP_AEE 0
#define ORC_OP_AEE_AND_GAMMA5 1
#define ORC_OP_AEE_MINUS 2
#define ORC_OP_AEE_MINUS_AND_GAMMA5 3
#define ORC_OP_QPLUS_EO 4
#define ORC_OP_QMINUS_EO 5
#define ORC_OP_QPLUSQMINUS_EO 6

#define ORC_ACTION_WILSON 0
#define ORC_ACTION_TWISTEDMASS 1

/* site-diagonal twisted-mass variants (ocl_kernel/fermionmatrix_eo_m*.cl) */
#define ORC_MTM_SITEDIAG 0
#define ORC_MTM_INVERSE 1
#define ORC_MTM_SITEDIAG_MINUS 2
#define ORC_MTM_INVERSE_MINUS 3
#define ORC_MTM_SITEDIAG_AND_GAMMA5 4
#define ORC_MTM_SITEDIAG_MINUS_AND_GAMMA5 5

long orc_volume(const orc_lattice* lat);
long orc_veo(const orc_lattice* lat);
int orc_even_site(const orc_lattice* lat, long eo_idx, long* site);
int orc_odd_site(const orc_lattice* lat, long eo_idx, long* site);

/* deterministic fills (meta/type_ops.cpp:29-79, physics/lattices/util.hpp:83-91) */
void orc_fill_doubles_drand48(double* out, long n, int seed);

/* gauge fields */
void orc_gauge_cold(const orc_lattice* lat, double* gauge);
int orc_ildg_read(const char* path, const orc_lattice* lat, double* gauge, char* err, int errlen);
double orc_plaquette(const orc_lattice* lat, const double* gauge);

/* full <-> eo (ocl_kernel/spinorfield_eo_convert.cl:23-51) */
void orc_convert_to_eo(const orc_lattice* lat, const double* full, double* even, double* odd);
void orc_convert_from_eo(const orc_lattice* lat, const double* even, const double* odd, double* full);

/* Wilson / twisted-mass kernels */
void orc_dslash_eo(const orc_lattice* lat, double* out, const double* in, const double* gauge, int evenodd,
                   double kappa);
/* dslash_eo at the n output eo indices eo[0..n-1] only; out holds n spinors */
void orc_dslash_eo_sites(const orc_lattice* lat, double* out, const double* in, const double* gauge, int evenodd,
                         double kappa, const long* eo, long n);
void orc_m_tm(const orc_lattice* lat, double* out, const double* in, double mubar, int variant);
void orc_dslash_minv_eo(const orc_lattice* lat, double* out, const double* in, const double* gauge, int evenodd,
                        double kappa, double mubar, int minus);
void orc_gamma5_eo(const orc_lattice* lat, double* inout);
void orc_saxpy_and_gamma5_eo(const orc_lattice* lat, double* out, const double* x, const double* y, double are,
                             double aim);
void orc_m_full(const orc_lattice* lat, double* out, const double* in, const double* gauge, double kappa,
                double mubar, int action, int minus);

/* Wilson eo operators; merged selects the *_AND_gamma5 path of Qplus/Qminus */
int orc_apply_eo(const orc_lattice* lat, int op, double* out, const double* in, const double* gauge, int action,
                 double kappa, double mubar, int merged);

/* Wilson eo BLAS (n = number of spinors, 24 doubles each) */
void orc_saxpy(double* out, const double* x, const double* y, double are, double aim, long n, int ncomp);
void orc_sax(double* out, const double* x, double are, double aim, long n, int ncomp);
void orc_saxsbypz(double* out, const double* x, const double* y, const double* z, double are, double aim, double bre,
                  double bim, long n, int ncomp);
void orc_dot(const double* x, const double* y, long n, int ncomp, double* re, double* im);
double orc_norm2(const double* x, long n, int ncomp);

/* staggered */
void orc_dks_eo(const orc_lattice* lat, double* out, const double* in, const double* gauge, int evenodd);
void orc_mdagm_eo(const orc_lattice* lat, double* out, const double* in, const double* gauge, double mass,
                  int upper_left);
void orc_stagg_saxpby(double* out, const double* x, const double* y, double are, double aim, double bre, double bim,
                      long n);

/* solvers.  Solver parameters (meta/parametersSolver.cpp:74-94):
 *   cgmax, iter_refresh, check_every (cgIterationBlockSize), merged (useKernelMergingSpinor).
 * Return ORC_OK, ORC_SOLVER_STUCK or ORC_SOLVER_DID_NOT_SOLVE; *iters receives the iteration index. */
int orc_cg_eo(const orc_lattice* lat, int op, const double* gauge, int action, double kappa, double mubar,
              int merged_fermion, double* x, const double* b, double prec, int cgmax, int iter_refresh,
              int check_every, int merged_spinor, int* iters);
int orc_bicgstab_eo(const orc_lattice* lat, int op, const double* gauge, int action, double kappa, double mubar,
                    int merged_fermion, double* x, const double* b, double prec, int cgmax, int iter_refresh,
                    int save_variant, int* iters);
int orc_cgm(const orc_lattice* lat, const double* gauge, double mass, int upper_left, double* const* x,
            const double* b, const double* sigma, int nshift, double prec, int cgmax, int check_every,
            int merged_spinor, int* iters);

/* OpenMP thread count used by the site loops (0 = library default) */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
