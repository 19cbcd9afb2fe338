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

/* Layout of the FP64 and FP32 gauge-link fields.  For each (direction, parity) pair dp = 2 dir + parity
 * the LANES 16-byte lanes of a link (9 complex entries for 18-real FP64; 6 for 12-real FP64; 3 float4
 * holding 6 complex entries for 12-real FP32) are stored in tiles of 64 eo sites, lane-major inside a
 * tile: element (dp, lane k, eo site e) sits at
 *   dp * LANES * Vp + (e / 64) * LANES * 64 + k * 64 + e % 64,    Vp = Veo rounded up to 64,
 * so one wavefront reading the link of 64 consecutive sites reads one contiguous LANES x 1 KiB block
 * instead of LANES blocks a field apart. */
constexpr uint32_t LINK_TILE = 64;
__host__ __device__ inline uint32_t link_sites(uint32_t veo)
{
    return (veo + LINK_TILE - 1) / LINK_TILE * LINK_TILE;
}
__host__ __device__ inline size_t link_elem(uint32_t dp, uint32_t lanes, uint32_t k, uint32_t eo, uint32_t veo)
{
    return (size_t)dp * lanes * link_sites(veo) + (size_t)(eo / LINK_TILE) * lanes * LINK_TILE +
           (size_t)k * LINK_TILE + eo % LINK_TILE;
}

/* FP16 12-real links (24 bytes): per (dir, parity) dp, tiles of 64 eo sites, each tile 64 16-byte lanes
 * (entries 0..3 of rows 0,1) followed by 64 8-byte lanes (entries 4, 5).  Byte offsets of site e's two
 * lanes: */
__host__ __device__ inline size_t link_h16_offset(uint32_t dp, uint32_t eo, uint32_t veo)
{
    return ((size_t)dp * link_sites(veo) + (size_t)(eo / LINK_TILE) * LINK_TILE) * 24 + (size_t)(eo % LINK_TILE) * 16;
}
__host__ __device__ inline size_t link_h8_offset(uint32_t dp, uint32_t eo, uint32_t veo)
{
    return ((size_t)dp * link_sites(veo) + (size_t)(eo / LINK_TILE) * LINK_TILE) * 24 + (size_t)LINK_TILE * 16 +
           (size_t)(eo % LINK_TILE) * 8;
}

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
    double* norm_partials;  // if non-null: per-block sum |out|^2 (with cg_r: |r_new|^2, Re <r_new, r_old>)
    void* cg_r;             // single fields: if set, r <- r - (*cg_alpha) out in place of storing out
    const double* cg_alpha; // complex device scalar
    HopWeights<double> w;
    WorkMap map;
    int nrhs;  // right-hand sides per launch (fields [12][Veo][nrhs]); 1 for single fields
    int rb;    // right-hand sides per thread of the multi-RHS kernel
    int lds;   // multi-RHS kernel with one right-hand side per thread: 1 stages links and neighbour indices in LDS
    int block; // threads per block of the single-field kernel (128, 256 or 512)
    int occ;   // register budget of the single-field kernel: 0 (<= 2 waves per SIMD), 1 (1 wave), 4 (>= 4 waves)
    // Operators that differ per right-hand side (multi-RHS launches, one right-hand side per thread): the hop
    // weights w are those of kappa = 1, the hop sum of right-hand side r is scaled by rhs_kappa[r], and
    // rhs_f1/rhs_f2/rhs_scale replace f1/f2/scale.
    int per_rhs;
    double rhs_kappa[16];
    double rhs_f1[16][2], rhs_f2[16][2], rhs_scale[16];
};

}  // namespace cl2
