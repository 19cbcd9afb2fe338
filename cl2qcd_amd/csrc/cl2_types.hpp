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
};

struct LatticeGeom {
    uint32_t ns, nt;
    uint32_t vs;   // ns^3
    uint32_t veo;  // ns^3*nt/2, also the field stride (no padding)
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
};

}  // namespace cl2
