// Host-side internals shared by the kernel launchers and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "cl2_types.hpp"

/* launch shape of the hopping kernels */
struct cl2_hop_tuning {
    int zplanes;  // z-planes per slab (-1: heuristic, 0: identity order)
    int xcd;      // XCD-contiguous block assignment
    int rb;       // right-hand sides per thread of the multi-RHS kernel
    int lds;      // 1: the multi-RHS kernel stages each site's links and neighbour indices in LDS once
    int block;    // threads per block of the single-field kernel
    int occ;      // register budget of the single-field kernel: 0 (<= 2 waves per SIMD, hop pairs), 1 (1 wave,
                  // four hops in flight), 4 (>= 4 waves, one hop at a time); -1: the format's default
};

struct cl2_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    cl2_lattice lat{};
    cl2::LatticeGeom g{};
    double bs[2] = {1., 0.}, bt[2] = {1., 0.};      // boundary phases exp(i pi theta / N)
    double cp_fwd[2] = {1., 0.}, cp_bwd[2] = {1., 0.};  // factors on forward / backward T links
    double* partials = nullptr;  // reduction scratch: one double2 per block
    size_t npartials = 0;
    double* dscratch = nullptr;  // a few device doubles for host-read results
    void** ptr_table = nullptr;  // device copy of the field-pointer table of the multi-shift update (2 x 64)
    std::vector<void*> ptr_table_host;  // what ptr_table holds
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    int tune_zplanes = -1;  // z-planes per slab of the hopping kernels (-1: heuristic, 0: identity order)
    int tune_xcd     = 1;   // XCD-contiguous block assignment
    int tune_mrhs_rb = 0;   // right-hand sides per thread of the multi-RHS hopping kernel (0: by format)
    int tune_block   = 256; // threads per block of the single-field hopping kernel
    int tune_occ     = -1;  // register budget of the single-field hopping kernel (-1: the format's default)
    int tune_mrhs_lds = 1;  // multi-RHS hopping kernel: links and neighbour indices staged in LDS per site
    int tune_mrhs_off32 = 1;  // staged multi-RHS kernel: FP64 gathers with 32-bit offsets where they fit
    bool tune_manual = false;  // set by cl2_tune_set: the values above override autotuned ones
    std::map<std::pair<int, int>, cl2_hop_tuning> tuned;  // (link format, right-hand sides) -> autotuned shape
    std::string err;
};

namespace cl2 {

constexpr int DSLASH_BLOCK = 256;
constexpr int BLAS_BLOCK   = 256;
constexpr int REDUCE_MAX_BLOCKS = 2048;
/* most blocks (and partials) of one elementwise BLAS or reduction launch; each thread streams a few 16-byte
 * chunks, and larger fields loop over the grid */
constexpr unsigned BLAS_MAX_BLOCKS = 65536;

inline unsigned blocks_for(size_t n, int block)
{
    return (unsigned)((n + block - 1) / block);
}

/* launch shape for (format, right-hand sides): manual settings, else the autotuned entry, else the
 * measured defaults */
cl2_hop_tuning hop_tuning(const cl2_ctx* c, int fmt, int nrhs);
/* work order of the hopping kernels for this context's lattice and a launch shape */
WorkMap work_map(const cl2_ctx* c, int block, int nrhs, const cl2_hop_tuning& t);
/* tuning-table key of the staggered hopping kernel: its link format | DKS_TUNING_KEY */
constexpr int DKS_TUNING_KEY = 0x100;
/* sites per block of the multi-RHS kernel that stages links in LDS (blocks of sites x nrhs threads) */
int mrhs_sites_per_block(int nrhs, int lds);
/* the staged multi-RHS kernel's FP64 gathers use 32-bit byte offsets within a component: veo * nrhs complex
 * numbers of complex_bytes each must stay below 4 GiB, else it takes 64-bit offsets (and lds 1 only) */
bool mrhs_offsets_narrow(uint32_t veo, int nrhs, size_t complex_bytes);
/* right-hand-side counts with a compiled multi-RHS hopping kernel */
bool supported_nrhs(int n);

/* hopping weights -kappa * b_mu * c_mu (forward) and -kappa * conj(b_mu * c_mu') (backward) */
HopWeights<double> wilson_weights(const cl2_ctx* c, double kappa);

/* launchers (return hipSuccess or the launch error) */
hipError_t launch_dslash(cl2_ctx* c, const DslashArgs& a, int format);
hipError_t launch_sitediag(cl2_ctx* c, void* out, const void* in, int prec, const double f1[2], const double f2[2],
                           double scale);
hipError_t launch_dks(cl2_ctx* c, void* out, const void* in, const void* gauge, int format, int evenodd);
/* nrhs = 2, 4, 8, 12, 16 staggered vectors [3][Veo][nrhs]; with x: out_r = coef[r] x_r - D in_r */
hipError_t launch_dks_step(cl2_ctx* c, void* out, const void* in, const void* gauge, int format, int evenodd,
                           const void* own, const void* x, double coef, void* cg_r, const double* beta,
                           unsigned* nblocks);
hipError_t launch_dks_mrhs(cl2_ctx* c, void* out, const void* in, const void* gauge, int format, int evenodd, int nrhs,
                           const void* x, const double* coef, double2* norm_partials = nullptr,
                           const void* own = nullptr, void* cg_r = nullptr, const double* cg_alpha = nullptr,
                           unsigned* nblocks = nullptr);

/* elementwise complex BLAS over n complex numbers */
enum BlasMode : int {
    B_Y_MINUS_AX = 0,  // y - a x          (Wilson saxpy)
    B_AX         = 1,  // a x
    B_AX_BY_Z    = 2,  // b y + a x + z     (Wilson saxsbypz, staggered saxpbypz)
    B_AX_PLUS_Y  = 3,  // a x + y          (staggered saxpy)
    B_AX_BY      = 4,  // a x + b y        (staggered saxpby)
    B_G5_Y_MINUS_AX = 5,  // gamma5 (y - a x) on a Wilson spinor
    B_NEG_TAIL   = 6,  // gamma5 in place: negate complex elements [tail, n)
    B_CG_XP      = 7,  // out = z - a x and out2 = y - b x, both from the same x (the x and p updates of CG)
};
struct BlasArgs {
    size_t n;     // complex elements
    size_t tail;  // first element with a negative gamma5 sign (Wilson: 6 * Veo)
    void* out;
    void* out2;  // B_CG_XP: second output
    const void* x;
    const void* y;
    const void* z;
    cl2_scalar a, b;
    int mode;
    double* norm_partials;  // if set: per-block sum |out|^2
    int cg_norm;            // with norm_partials: second partial = sum Re conj(out) y (y read before out is written)
};
hipError_t launch_blas(cl2_ctx* c, const BlasArgs& a, int prec, unsigned* nblocks_out);

enum ReduceMode : int { R_DOT = 0, R_DOT_RE = 1, R_NORM = 2 };
/* stage 1 of a two-stage reduction; returns the number of partials written (double2 each) to parts, or to
 * the context's partials when parts is null */
hipError_t launch_reduce_stage1(cl2_ctx* c, const void* x, const void* y, size_t n, int prec, int mode,
                                unsigned* nblocks, double* parts = nullptr);
/* partials launch_reduce_stage1 writes for these operands (at most BLAS_MAX_BLOCKS) */
unsigned reduce_stage1_blocks(size_t n, int prec, const void* x, const void* y);
/* stage 2: sum nparts double2 partials in fixed order.  out_kind 0: res[0] = real part;
 * 1: res[0..1] = complex sum;  2: res[k] = alpha_vec[k]^2 * real part for k < nvec;
 * 3: res[0..3] = (re, im, re, 0) */
hipError_t launch_reduce_stage2(cl2_ctx* c, unsigned nparts, double* res, int out_kind, const double* alpha_vec,
                                int nvec, const double* parts = nullptr);

hipError_t launch_gauge_update(cl2_ctx* c, void* gauge, double beta, int mode, uint64_t key);
hipError_t launch_plaquette(cl2_ctx* c, const void* gauge, double* res);
hipError_t launch_gauge_random(cl2_ctx* c, void* dev, int fmt, uint64_t seed, double eps);
hipError_t launch_gauge_convert(cl2_ctx* c, void* dst, int dfmt, const void* src);
hipError_t launch_gaussian(cl2_ctx* c, void* dev, int prec, size_t n, uint64_t seed, bool wilson);

hipError_t launch_fill(cl2_ctx* c, void* out, size_t n, int prec, double re, double im);
hipError_t launch_copy16(cl2_ctx* c, void* dst, const void* src, size_t n);
hipError_t launch_slice_norm(cl2_ctx* c, double* res, const void* x, int prec, int dir, int accumulate);

/* FP16 Wilson spinors (per-site scale) and F16_12 links */
hipError_t launch_dslash_h(cl2_ctx* c, const DslashArgs& a);
hipError_t launch_blas_h(cl2_ctx* c, const BlasArgs& a, unsigned* nblocks_out);  // a.n = sites
hipError_t launch_reduce_stage1_h(cl2_ctx* c, const void* x, const void* y, int mode, unsigned* nblocks);
hipError_t launch_sitediag_h(cl2_ctx* c, void* out, const void* in, const double f1[2], const double f2[2], double s);
/* sites: the site count of the field (0: Veo; Veo nrhs for a block of FP16 spinors) */
hipError_t launch_fill_h(cl2_ctx* c, void* out, double re, double im, uint32_t sites = 0);
hipError_t launch_convert_h(cl2_ctx* c, void* out, int out_prec, const void* in, int in_prec, uint32_t sites = 0);
/* blocks of FP16 spinors: the multi-RHS hopping kernel and the block BLAS */
hipError_t launch_dslash_mrhs_h(cl2_ctx* c, const DslashArgs& a);
hipError_t launch_block_axpy_h(cl2_ctx* c, void* out, const void* x, const void* y, int nrhs, const double* coef,
                               double sign, double* norms);
hipError_t launch_block_xp_h(cl2_ctx* c, void* x, void* p, const void* r, int nrhs, const double* alpha,
                             const double* beta);
hipError_t launch_block_dot_h(cl2_ctx* c, double* res, const void* x, const void* y, int nrhs, int mode);
hipError_t launch_block_slot_h(cl2_ctx* c, void* block, const void* single_in, void* single_out, int nrhs, int r);
hipError_t launch_convert(cl2_ctx* c, void* out, int out_prec, const void* in, int in_prec, size_t n);
hipError_t launch_convert_spinor(cl2_ctx* c, void* out, int out_prec, const void* in, int in_prec, size_t cs);

/* blocks of nrhs eo fields [ncomp][Veo][nrhs]: Wilson spinors (ncomp 12) or staggered vectors (3) */
hipError_t launch_block_axpy(cl2_ctx* c, void* out, const void* x, const void* y, int prec, int nrhs,
                             const double* coef, double sign, double* norms, int ncomp = 12);
hipError_t launch_block_dot(cl2_ctx* c, double* res, const void* x, const void* y, int prec, int nrhs, int mode,
                            int ncomp = 12);
/* stage 2 of the per-right-hand-side reductions: block r sums partials[r * nb .. r * nb + nb) in order;
 * out_kind 0: res[2r..2r+1] = complex sum, 1: res[r] = real part */
hipError_t launch_block_reduce2(cl2_ctx* c, const double2* parts, unsigned nb, double* res, int out_kind, int nrhs);
hipError_t launch_block_xp(cl2_ctx* c, void* x, void* p, const void* r, int prec, int nrhs, const double* alpha,
                           const double* beta, int ncomp = 12);
hipError_t launch_block_slot(cl2_ctx* c, void* block, const void* single_in, void* single_out, int prec, int nrhs,
                             int r, int ncomp = 12);
hipError_t launch_scalar_copy(cl2_ctx* c, double* dst, const double* src, int n);
hipError_t launch_cg_beta(cl2_ctx* c, double* mbeta, double* r2, const double* norms);
hipError_t launch_block_cg_scalars(cl2_ctx* c, int which, double* out, double* r2, const double* in,
                                   const double* active, int nrhs);
hipError_t launch_block_cg_alpha_mdagm(cl2_ctx* c, double* alpha, const double* r2, double* pAp, const double* coef,
                                       const double* active, int nrhs);

/* fused multi-shift CG vector update over n complex elements */
hipError_t launch_multishift(cl2_ctx* c, void* const* xs_dev, int x_prec, void* const* ps_dev, const void* r, int prec,
                             size_t n, int nshift, const double* beta, const double* zeta, const double* alpha,
                             const double* active);

hipError_t launch_scalar_op(cl2_ctx* c, int op, int is_complex, double* out, const double* a, const double* b);
hipError_t launch_vec_element(cl2_ctx* c, double* out, const double* in, int out_index, int in_index);
hipError_t launch_cgm_update(cl2_ctx* c, int which, double* out, const double* a, const double* b, const double* d,
                             const double* e, const double* f, const double* g, int numeq);
hipError_t launch_cgm_beta_step(cl2_ctx* c, double* beta, double* beta_prev, const double* rr, const double* pap);
hipError_t launch_cgm_beta_step_mdagm(cl2_ctx* c, double* beta, double* beta_prev, const double* rr, double* pap,
                                      const double* wp, double m2);
hipError_t launch_cgm_alpha_step(cl2_ctx* c, int numeq, double* alpha, double* alpha_prev, const double* rr_new,
                                 double* rr_old, const double* beta, const double* beta_prev, const double* sigma,
                                 double* zeta, double* zeta_prev, double* zeta_foll, double* beta_vec,
                                 double* alpha_vec);

}  // namespace cl2
