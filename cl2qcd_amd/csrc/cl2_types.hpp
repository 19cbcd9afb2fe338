// Types shared by host launch code and device kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cl2qcd.h"

namespace cl2 {

enum GaugeFmt : int {
    GAUGE_F64_18 = CL2_GAUGE_F64_18,
    GAUGE_F64_12 = CL2_GAUGE_F64_12,
    GAUGE_F32_12 = CL2_GAUGE_F32_12,
    GAUGE_F16_12 = CL2_GAUGE_F16_12,
};

struct LatticeGeom {
    uint32_t ns, nt;
    uint32_t vs;   // ns^3
    uint32_t veo;  // ns^3*nt/2, also the field stride (no padding)
};

/* Work order of the hopping kernels.  Logical work item l enumerates
 * (z-chunk, t, site-in-slab) with the site fastest, then t, then the z-chunk;
 * a slab is zplanes consecutive z-planes of one t-slice (P = zplanes*ns^2/2 eo
 * sites).  Each of the 8 XCDs gets a contiguous range of logical blocks, so one
 * XCD sweeps its z-chunk through all t and the t+-1 / z+-1 neighbours of a site
 * are still in that XCD's L2. */
struct WorkMap {
    uint32_t P;        // eo sites per slab (0: identity order)
    uint32_t col;      // P * nt
    uint32_t half_vs;  // eo sites per t-slice
    uint32_t xcd_q;    // logical blocks per XCD (0: no XCD remap)
};

/* Epilogues fused into the Wilson hopping kernel. */
enum DslashEpilogue : int {
    EPI_PLAIN = 0,  // out = D in                          (dslash_eo)
    EPI_MINV  = 1,  // out = s * diag(f1,f2) (D in)          (dslash_AND_M_tm_inverse_sitediagonal[_minus])
    EPI_AEE   = 2,  // out = diag(f1,f2) x - D in [, gamma5] (M_tm_sitediagonal + saxpy [+ gamma5])
};

template<typename R>
struct HopWeights {
    // -kappa * b_mu * (chemical-potential factor) for the forward hop and the
    // corresponding factor for the backward hop, per direction T,X,Y,Z.
    R wf[4][2];
    R wb[4][2];
};

struct DslashArgs {
    const void* in;
    void* out;
    const void* gauge;
    const void* x;  // EPI_AEE: field multiplied by the site-diagonal term
    LatticeGeom g;
    int out_parity;
    int epilogue;
    int gamma5;
    double f1[2], f2[2];  // site-diagonal factors
    double scale;
    double* norm_partials;  // if non-null: per-block sum |out|^2
    HopWeights<double> w;
    WorkMap map;
    int nrhs;  // right-hand sides per launch (fields [12][Veo][nrhs]); 1 for single fields
    int rb;    // right-hand sides per thread of the multi-RHS kernel
    int block; // threads per block of the single-field kernel (128, 256 or 512)
    int occ;   // 4: single-field kernel compiled for at least 4 waves per SIMD (<= 128 VGPRs), else 0
};

}  // namespace cl2
