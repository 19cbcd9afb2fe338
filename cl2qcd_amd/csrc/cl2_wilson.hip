// Wilson / twisted-mass even-odd hopping kernels for gfx950.
//
// One thread per output eo site.  The input spinor and the gauge links are
// SoA (component-major) so every load of a wavefront is a contiguous 1 KiB
// (FP64) or 512 B (FP32) segment; the forward link of a site shares the
// site's own eo index and the backward link shares the neighbour's, so link
// loads are as coalesced as the spinor gathers.  The kernel is HBM bound
// (8 links + 2 spinors per site), so no MFMA.
//
// What bounds it on gfx950 is the reuse of the neighbour spinors in each
// XCD's 4 MB L2: every input site is read by 8 output sites, and the fewer
// sites an XCD has in flight the more of those reads hit.  The single-field
// kernel therefore keeps few sites in flight (about 4 waves per CU) and gets
// its memory parallelism from each thread instead: both hops of a direction
// load their links and full neighbour spinors before either is applied, with
// a register budget of up to 256 VGPRs, and results leave by streaming
// stores so they do not evict neighbour spinors from the L2.
#include <type_traits>

#include "cl2_device.hpp"
#include "cl2_internal.hpp"

namespace cl2 {

/* One hop, on a neighbour spinor s already in registers: the two half-spinor rows (A,B,C) of the
 * projector
 *   h = s_A + C * s_B,   r = w * (U or U^dagger) h,
 *   acc_A += r,  acc_B += conj(C) * r
 * reproduce dslash_eoprec_unified_local (ocl_kernel/fermionmatrix_eo.cl:30-253);
 * w already carries -kappa and the boundary phase. */
template<int A, int B, int CU, bool FWD, typename R>
__device__ __forceinline__ void half_hop_reg(Cx<R> (&acc)[12], const vec2<R> (&s)[12], const Su3<R>& u, Cx<R> w)
{
    Cx<R> h[3], r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        h[c] = add_unit<CU>(Cx<R>{s[3 * A + c].x, s[3 * A + c].y}, Cx<R>{s[3 * B + c].x, s[3 * B + c].y});
    if constexpr (FWD)
        su3_times(u, h, r);
    else
        su3dag_times(u, h, r);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Cx<R> t        = cmul(r[c], w);
        acc[3 * A + c] = acc[3 * A + c] + t;
        acc[3 * B + c] = add_unit<UnitConj<CU>::v>(acc[3 * B + c], t);
    }
}

/* projector rows of (1 - gamma_mu) forward and (1 + gamma_mu) backward, chiral basis */
template<int DIR, bool FWD, typename R>
__device__ __forceinline__ void apply_hop(Cx<R> (&acc)[12], const vec2<R> (&s)[12], const Su3<R>& u, Cx<R> w)
{
    if constexpr (DIR == 0) {
        half_hop_reg<0, 2, FWD ? P1 : M1, FWD>(acc, s, u, w);
        half_hop_reg<1, 3, FWD ? P1 : M1, FWD>(acc, s, u, w);
    } else if constexpr (DIR == 1) {
        half_hop_reg<0, 3, FWD ? PI : MI, FWD>(acc, s, u, w);
        half_hop_reg<1, 2, FWD ? PI : MI, FWD>(acc, s, u, w);
    } else if constexpr (DIR == 2) {
        half_hop_reg<0, 3, FWD ? P1 : M1, FWD>(acc, s, u, w);
        half_hop_reg<1, 2, FWD ? M1 : P1, FWD>(acc, s, u, w);
    } else {
        half_hop_reg<0, 2, FWD ? PI : MI, FWD>(acc, s, u, w);
        half_hop_reg<1, 3, FWD ? MI : PI, FWD>(acc, s, u, w);
    }
}

/* forward and backward hop of direction DIR: both links and both neighbour spinors (elements nf, nbk
 * of the spinor field) are loaded before either hop is applied.  The forward link is the output
 * site's own (eo index i, parity P), the backward one the neighbour's (eo index linkbk, parity P^1). */
/* forward and backward hops of directions D0 and D1: four links and four neighbour spinors in flight
 * before any hop is applied (the 1-wave-per-SIMD variants, which may also stage in AGPRs) */
template<int D0, int D1, int FMT, int NRHS, typename R>
__device__ __forceinline__ void hop_quad(Cx<R> (&acc)[12], const vec2<R>* __restrict__ in, size_t SS, const SiteNb& nb,
                                         uint32_t rhs, const void* gauge, uint32_t S, int P, uint32_t i,
                                         const HopWeights<double>& w)
{
    Su3<R> u0f, u0b, u1f, u1b;
    vec2<R> s0f[12], s0b[12], s1f[12], s1b[12];
    GaugeLoad<FMT>::load(gauge, S, D0, P, i, u0f);
    GaugeLoad<FMT>::load(gauge, S, D0, P ^ 1, nb.dn[D0], u0b);
    GaugeLoad<FMT>::load(gauge, S, D1, P, i, u1f);
    GaugeLoad<FMT>::load(gauge, S, D1, P ^ 1, nb.dn[D1], u1b);
    load_spinor<R>(s0f, in, SS, (size_t)nb.up[D0] * NRHS + rhs);
    load_spinor<R>(s0b, in, SS, (size_t)nb.dn[D0] * NRHS + rhs);
    load_spinor<R>(s1f, in, SS, (size_t)nb.up[D1] * NRHS + rhs);
    load_spinor<R>(s1b, in, SS, (size_t)nb.dn[D1] * NRHS + rhs);
    apply_hop<D0, true>(acc, s0f, u0f, Cx<R>{R(w.wf[D0][0]), R(w.wf[D0][1])});
    apply_hop<D0, false>(acc, s0b, u0b, Cx<R>{R(w.wb[D0][0]), R(w.wb[D0][1])});
    apply_hop<D1, true>(acc, s1f, u1f, Cx<R>{R(w.wf[D1][0]), R(w.wf[D1][1])});
    apply_hop<D1, false>(acc, s1b, u1b, Cx<R>{R(w.wb[D1][0]), R(w.wb[D1][1])});
}

/* the same two hops one after the other, for the 128-VGPR variants: each loads only its own link and
 * neighbour spinor */
template<int DIR, int FMT, typename R>
__device__ __forceinline__ void hop_seq(Cx<R> (&acc)[12], const vec2<R>* __restrict__ in, size_t SS, uint32_t nf,
                                        uint32_t nbk, const void* gauge, uint32_t S, int P, uint32_t i,
                                        uint32_t linkbk, Cx<R> wf, Cx<R> wb)
{
    {
        Su3<R> u;
        vec2<R> sp[12];
        GaugeLoad<FMT>::load(gauge, S, DIR, P, i, u);
        load_spinor<R>(sp, in, SS, nf);
        apply_hop<DIR, true>(acc, sp, u, wf);
    }
    {
        Su3<R> u;
        vec2<R> sp[12];
        GaugeLoad<FMT>::load(gauge, S, DIR, P ^ 1, linkbk, u);
        load_spinor<R>(sp, in, SS, nbk);
        apply_hop<DIR, false>(acc, sp, u, wb);
    }
}

template<int DIR, int FMT, typename R>
__device__ __forceinline__ void hop_pair(Cx<R> (&acc)[12], const vec2<R>* __restrict__ in, size_t SS, uint32_t nf,
                                         uint32_t nbk, const void* gauge, uint32_t S, int P, uint32_t i,
                                         uint32_t linkbk, Cx<R> wf, Cx<R> wb)
{
    Su3<R> uf, ub;
    vec2<R> sf[12], sb[12];
    GaugeLoad<FMT>::load(gauge, S, DIR, P, i, uf);
    GaugeLoad<FMT>::load(gauge, S, DIR, P ^ 1, linkbk, ub);
    load_spinor<R>(sf, in, SS, nf);
    load_spinor<R>(sb, in, SS, nbk);
    apply_hop<DIR, true>(acc, sf, uf, wf);
    apply_hop<DIR, false>(acc, sb, ub, wb);
}

/* NRHS right-hand sides in one launch: the fields are [12][Veo][NRHS] (right-hand side fastest) and
 * thread l works on site l / NRHS, right-hand side l % NRHS, so the NRHS lanes of one site issue the same
 * link addresses and each link leaves HBM once per launch.  NRHS = 1 is the single-field kernel. */
/* Epilogue (per-right-hand-side kappa when the operator differs per right-hand side, site-diagonal factors,
 * Aee saxpy, gamma5) and streaming store of output element O (right-hand side rhs) of a field of component
 * stride SS, adding |out|^2 of the site to NRM.  With a.cg_r set the result v is not stored: the CG residual
 * is updated in its place, r <- r - alpha v (alpha = a.cg_alpha, as the elementwise kernel computes it), and
 * NRM, RY receive |r_new|^2 and Re <r_new, r_old>.  A macro rather than a function: passing the
 * accumulator array to a helper, even a force-inlined one, made the compiler keep it in scratch memory. */
#define CL2_FINISH_SITE(O, NRM, RY)                                                                               \
    {                                                                                                             \
        const size_t o_ = (O);                                                                                    \
        Cx<R> f1{R(a.f1[0]), R(a.f1[1])}, f2{R(a.f2[0]), R(a.f2[1])};                                             \
        R s_ = R(a.scale);                                                                                        \
        if (a.per_rhs) {                                                                                          \
            const R kr = R(a.rhs_kappa[rhs]);                                                                     \
            _Pragma("unroll") for (int k = 0; k < 12; ++k) acc[k] = scale(acc[k], kr);                            \
            f1 = Cx<R>{R(a.rhs_f1[rhs][0]), R(a.rhs_f1[rhs][1])};                                                 \
            f2 = Cx<R>{R(a.rhs_f2[rhs][0]), R(a.rhs_f2[rhs][1])};                                                 \
            s_ = R(a.rhs_scale[rhs]);                                                                             \
        }                                                                                                         \
        if (a.epilogue == EPI_MINV) {                                                                             \
            _Pragma("unroll") for (int k = 0; k < 12; ++k) acc[k] = scale(cmul(acc[k], k < 6 ? f1 : f2), s_);    \
        } else if (a.epilogue == EPI_AEE) {                                                                       \
            V xs_[12];                                                                                            \
            load_spinor<R>(xs_, reinterpret_cast<const V*>(a.x), SS, o_);                                         \
            _Pragma("unroll") for (int k = 0; k < 12; ++k) {                                                      \
                acc[k] = cmul(Cx<R>{xs_[k].x, xs_[k].y}, k < 6 ? f1 : f2) - acc[k];                               \
                if (a.gamma5 && k >= 6) acc[k] = Cx<R>{-acc[k].re, -acc[k].im};                                   \
            }                                                                                                     \
        }                                                                                                         \
        V ov_[12];                                                                                                \
        if (a.cg_r) {                                                                                             \
            V* r_ = reinterpret_cast<V*>(a.cg_r);                                                                 \
            const Cx<R> al_{R(a.cg_alpha[2 * rhs]), R(a.cg_alpha[2 * rhs + 1])};                                  \
            V rv_[12];                                                                                            \
            load_spinor_stream<R>(rv_, r_, SS, o_);                                                               \
            _Pragma("unroll") for (int k = 0; k < 12; ++k) {                                                      \
                const Cx<R> rn = Cx<R>{rv_[k].x, rv_[k].y} - cmul(acc[k], al_);                                   \
                ov_[k]         = V{rn.re, rn.im};                                                                 \
                NRM += (double)rn.re * (double)rn.re + (double)rn.im * (double)rn.im;                             \
                RY += (double)rn.re * (double)rv_[k].x + (double)rn.im * (double)rv_[k].y;                        \
            }                                                                                                     \
            store_spinor_stream<R>(r_, SS, o_, ov_);                                                              \
        } else {                                                                                                  \
            _Pragma("unroll") for (int k = 0; k < 12; ++k) ov_[k] = V{acc[k].re, acc[k].im};                      \
            if (a.norm_partials)                                                                                  \
                _Pragma("unroll") for (int k = 0; k < 12; ++k)                                                    \
                    NRM += (double)acc[k].re * (double)acc[k].re + (double)acc[k].im * (double)acc[k].im;         \
            store_spinor_stream<R>(reinterpret_cast<V*>(a.out), SS, o_, ov_);                                     \
        }                                                                                                         \
    }

/* Register budget and hops in flight per thread:
 *   OCC = 0: at most 2 waves per SIMD (up to 256 VGPRs), the two hops of one direction in flight;
 *   OCC = 1: 1 wave per SIMD, the four hops of two directions in flight (blocks of at most 256);
 *   OCC = 2: at most 2 waves per SIMD with the four hops of two directions in flight (FP32);
 *   OCC = 4: at least 4 waves per SIMD (at most 128 VGPRs), one hop at a time.
 * PLAIN: the plain hopping term (no epilogue, no norms), compiled without the epilogue's branches and the
 * block reduction; same arithmetic (a standalone copy measured 0.8 % faster than the general kernel for FP64
 * 18-real links at 48^3x96 in one run, within the spread between boxes in the benchmark). */
template<int FMT, int BLOCK, int NRHS, int OCC = 0, bool PLAIN = false>
__global__ __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(OCC == 4 ? 4 : 1, OCC == 4 ? 8 : (OCC == 1 ? 1 : 2))))
void k_dslash_eo(DslashArgs a)
{
    using R = typename GaugeLoad<FMT>::R;
    using V = vec2<R>;
    __shared__ double lds[2 * BLOCK / 64];
    const uint32_t S = a.g.veo;
    const size_t SS  = (size_t)S * NRHS;
    uint32_t i = 0, rhs = 0;
    double nrm = 0., ry = 0.;
    if (map_site_rhs<BLOCK, NRHS>(a.map, S, blockIdx.x, threadIdx.x, i, rhs)) {
        SiteNb nb;
        site_neighbours(a.g, i, a.out_parity, nb);
        const V* in    = reinterpret_cast<const V*>(a.in);
        const int P    = a.out_parity;
        Cx<R> acc[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[k] = Cx<R>{R(0), R(0)};
        // order T, X, Y, Z as fermionmatrix_eo_dslash.cl:38-45
#define CL2_W(t, m) Cx<R>{R(a.w.t[m][0]), R(a.w.t[m][1])}
#define CL2_DIR(m)                                                                                                  \
    if constexpr (OCC == 4)                                                                                         \
        hop_seq<m, FMT>(acc, in, SS, nb.up[m] * NRHS + rhs, nb.dn[m] * NRHS + rhs, a.gauge, S, P, i, nb.dn[m],      \
                        CL2_W(wf, m), CL2_W(wb, m));                                                                \
    else                                                                                                            \
        hop_pair<m, FMT>(acc, in, SS, nb.up[m] * NRHS + rhs, nb.dn[m] * NRHS + rhs, a.gauge, S, P, i, nb.dn[m],     \
                         CL2_W(wf, m), CL2_W(wb, m));
        if constexpr (OCC == 1 || OCC == 2) {
            hop_quad<0, 1, FMT, NRHS>(acc, in, SS, nb, rhs, a.gauge, S, P, i, a.w);
            hop_quad<2, 3, FMT, NRHS>(acc, in, SS, nb, rhs, a.gauge, S, P, i, a.w);
        } else {
            CL2_DIR(0)
            CL2_DIR(1)
            CL2_DIR(2)
            CL2_DIR(3)
        }
#undef CL2_DIR
#undef CL2_W
        if constexpr (PLAIN) {
            V ov[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) ov[k] = V{acc[k].re, acc[k].im};
            store_spinor_stream<R>(reinterpret_cast<V*>(a.out), SS, (size_t)i * NRHS + rhs, ov);
        } else {
            CL2_FINISH_SITE((size_t)i * NRHS + rhs, nrm, ry)
        }
    }
    if constexpr (!PLAIN) {
        if (a.norm_partials) {
            const double2 s = block_sum2<BLOCK>(make_double2(nrm, ry), lds);
            if (threadIdx.x == 0) a.norm_partials[2 * blockIdx.x] = s.x, a.norm_partials[2 * blockIdx.x + 1] = s.y;
        }
    }
}

bool mrhs_offsets_narrow(uint32_t veo, int nrhs, size_t complex_bytes)
{
    return (uint64_t)veo * (uint64_t)nrhs * complex_bytes <= 0xffffffffull;
}

int mrhs_sites_per_block(int nrhs, int lds)
{
    return (nrhs >= 8 ? 16 : 32) * (lds == 2 ? 2 : 1);
}

/* sites per block of the staged kernel: 16 (8 or more right-hand sides) or 32, times SBX (1 or 2) */
template<int NRHS, int SBX = 1>
struct MrhsLds {
    static constexpr int SB    = (NRHS >= 8 ? 16 : 32) * SBX;  // sites per block (mrhs_sites_per_block)
    static constexpr int BLOCK = SB * NRHS;
};

/* Multi-RHS hopping kernel with per-site staging.  A block owns SB consecutive sites of the work order and
 * NRHS threads per site (right-hand side fastest).  Each site's neighbour indices are computed once and its 8
 * links loaded and reconstructed once, by SB x 8 threads, into LDS; every right-hand-side thread then takes
 * them from LDS and gathers only its own neighbour spinors, both hops of a direction in flight together.
 * This removes the per-lane index arithmetic, link loads and 12-real reconstruction that otherwise repeat
 * NRHS times per site.  Per right-hand side the arithmetic is the single-field kernel's, in the same order. */
template<int FMT, int NRHS, int SBX, bool WIDE = false>
__global__ __launch_bounds__((MrhsLds<NRHS, SBX>::BLOCK)) void k_dslash_eo_mrhs_lds(DslashArgs a)
{
    using R = typename GaugeLoad<FMT>::R;
    using V = vec2<R>;
    constexpr int SB    = MrhsLds<NRHS, SBX>::SB;
    constexpr int BLOCK = MrhsLds<NRHS, SBX>::BLOCK;
    constexpr uint32_t NONE = 0xffffffffu;
    __shared__ Su3<R> lu[SB * 8];     // [site][hop], hop = 2 dir + backward
    __shared__ uint32_t lnb[SB * 8];  // neighbour eo index of each hop
    __shared__ uint32_t lsite[SB];    // eo index of the site (NONE past the end of the lattice)
    __shared__ double2 lred[BLOCK];   // per-thread (|out|^2, Re <r_new, r_old>) for the per-RHS block sums
    const uint32_t S  = a.g.veo;
    const size_t SS   = (size_t)S * NRHS;
    const uint32_t Lb = logical_block(a.map, blockIdx.x);
    const int P       = a.out_parity;
    if (threadIdx.x < SB) {
        const uint32_t l = Lb * SB + threadIdx.x;
        uint32_t i       = NONE;
        if (l < S) {
            i = logical_site(a.map, l);
            SiteNb nb;
            site_neighbours(a.g, i, P, nb);
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                lnb[threadIdx.x * 8 + 2 * d]     = nb.up[d];
                lnb[threadIdx.x * 8 + 2 * d + 1] = nb.dn[d];
            }
        }
        lsite[threadIdx.x] = i;
    }
    __syncthreads();
    for (int w = threadIdx.x; w < SB * 8; w += BLOCK) {
        const uint32_t i = lsite[w >> 3];
        if (i != NONE) {
            const int h = w & 7;
            Su3<R> u;
            // forward link: the site's own; backward link: the neighbour's, of the other parity
            GaugeLoad<FMT>::load(a.gauge, S, h >> 1, (h & 1) ? P ^ 1 : P, (h & 1) ? lnb[w] : i, u);
            lu[w] = u;
        }
    }
    __syncthreads();
    const int s      = threadIdx.x / NRHS;
    const int rhs    = threadIdx.x - s * NRHS;
    const uint32_t i = lsite[s];
    double nrm = 0., ry = 0.;
    if (i != NONE) {
    // FP64 gathers (unless WIDE): component k of a neighbour is read at a scalar base + k * Veo * NRHS
    // complex numbers, plus a 32-bit per-lane byte offset, which the host allows only while
    // Veo * NRHS * sizeof(V) < 2^32.  That removes the per-lane 64-bit address arithmetic and brings the
    // kernel from 3 to 4 waves per SIMD (measured 25 % faster at 12 right-hand sides); for FP32 the
    // 64-bit form measured 1-3 % faster.
    constexpr bool NARROW = sizeof(R) == 8 && !WIDE;
    const V* in = reinterpret_cast<const V*>(a.in);
    gchar_ptr cb[12];
#pragma unroll
    for (int k = 0; k < 12; ++k)
        cb[k] = NARROW ? uniform_ptr(reinterpret_cast<const char*>(a.in) + (size_t)k * SS * sizeof(V)) : nullptr;
    Cx<R> acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = Cx<R>{R(0), R(0)};
    // order T, X, Y, Z, forward then backward, as fermionmatrix_eo_dslash.cl:38-45
#define CL2_DIR(m)                                                                                              \
    {                                                                                                           \
        vec2<R> sf[12], sb[12];                                                                                 \
        if constexpr (NARROW) {                                                                                 \
            const uint32_t ef = (lnb[s * 8 + 2 * (m)] * NRHS + rhs) * (uint32_t)sizeof(V);                      \
            const uint32_t eb = (lnb[s * 8 + 2 * (m) + 1] * NRHS + rhs) * (uint32_t)sizeof(V);                  \
            _Pragma("unroll") for (int k = 0; k < 12; ++k) sf[k] = load_at<R>(cb[k], ef);                       \
            _Pragma("unroll") for (int k = 0; k < 12; ++k) sb[k] = load_at<R>(cb[k], eb);                       \
        } else {                                                                                                \
            const size_t ef = (size_t)lnb[s * 8 + 2 * (m)] * NRHS + rhs;                                        \
            const size_t eb = (size_t)lnb[s * 8 + 2 * (m) + 1] * NRHS + rhs;                                    \
            load_spinor<R>(sf, in, SS, ef);                                                                     \
            load_spinor<R>(sb, in, SS, eb);                                                                     \
        }                                                                                                       \
        const Su3<R> uf = lu[s * 8 + 2 * (m)], ub = lu[s * 8 + 2 * (m) + 1];                                    \
        apply_hop<m, true>(acc, sf, uf, Cx<R>{R(a.w.wf[m][0]), R(a.w.wf[m][1])});                             \
        apply_hop<m, false>(acc, sb, ub, Cx<R>{R(a.w.wb[m][0]), R(a.w.wb[m][1])});                            \
    }
    CL2_DIR(0)
    CL2_DIR(1)
    CL2_DIR(2)
    CL2_DIR(3)
#undef CL2_DIR
    CL2_FINISH_SITE((size_t)i * NRHS + rhs, nrm, ry)
    }
    if (a.norm_partials) {
        // per right-hand side: thread r < NRHS sums the SB site threads of its column in site order, into
        // partials[r * gridDim.x + block] (double2)
        lred[threadIdx.x] = make_double2(nrm, ry);
        __syncthreads();
        if ((int)threadIdx.x < NRHS) {
            double2 t = make_double2(0., 0.);
            for (int q = 0; q < SB; ++q) {
                t.x += lred[q * NRHS + threadIdx.x].x;
                t.y += lred[q * NRHS + threadIdx.x].y;
            }
            reinterpret_cast<double2*>(a.norm_partials)[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
        }
    }
}

/* Register-blocked multi-RHS variant: one thread per (site, group of RB right-hand sides); each link is
 * loaded and reconstructed once per thread and applied to RB neighbour spinors, so link loads and SU(3)
 * reconstruction are amortised over RB right-hand sides as well as shared across the NRHS/RB lanes of a
 * site.  Per right-hand side it performs the single-field kernel's operations in the same order. */
template<int A, int B, int CU, bool FWD, int RB, typename R>
__device__ __forceinline__ void half_hop_rb(Cx<R> (&acc)[RB][12], const vec2<R>* __restrict__ in, size_t SS,
                                            uint32_t nb, const Su3<R>& u, Cx<R> w)
{
#pragma unroll
    for (int q = 0; q < RB; ++q) {
        Cx<R> h[3], r[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            vec2<R> a = in[spinor_cx<R>(3 * A + c, SS, (size_t)nb + q)];
            vec2<R> b = in[spinor_cx<R>(3 * B + c, SS, (size_t)nb + q)];
            h[c]      = add_unit<CU>(Cx<R>{a.x, a.y}, Cx<R>{b.x, b.y});
        }
        if constexpr (FWD)
            su3_times(u, h, r);
        else
            su3dag_times(u, h, r);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Cx<R> t           = cmul(r[c], w);
            acc[q][3 * A + c] = acc[q][3 * A + c] + t;
            acc[q][3 * B + c] = add_unit<UnitConj<CU>::v>(acc[q][3 * B + c], t);
        }
    }
}

template<int DIR, bool FWD, int FMT, int RB, typename R>
__device__ __forceinline__ void hop_rb(Cx<R> (&acc)[RB][12], const vec2<R>* __restrict__ in, size_t SS, uint32_t nb,
                                       const void* gauge, uint32_t S, int linkpar, uint32_t linkeo, Cx<R> w)
{
    Su3<R> u;
    GaugeLoad<FMT>::load(gauge, S, DIR, linkpar, linkeo, u);
    if constexpr (DIR == 0) {
        half_hop_rb<0, 2, FWD ? P1 : M1, FWD, RB>(acc, in, SS, nb, u, w);
        half_hop_rb<1, 3, FWD ? P1 : M1, FWD, RB>(acc, in, SS, nb, u, w);
    } else if constexpr (DIR == 1) {
        half_hop_rb<0, 3, FWD ? PI : MI, FWD, RB>(acc, in, SS, nb, u, w);
        half_hop_rb<1, 2, FWD ? PI : MI, FWD, RB>(acc, in, SS, nb, u, w);
    } else if constexpr (DIR == 2) {
        half_hop_rb<0, 3, FWD ? P1 : M1, FWD, RB>(acc, in, SS, nb, u, w);
        half_hop_rb<1, 2, FWD ? M1 : P1, FWD, RB>(acc, in, SS, nb, u, w);
    } else {
        half_hop_rb<0, 2, FWD ? PI : MI, FWD, RB>(acc, in, SS, nb, u, w);
        half_hop_rb<1, 3, FWD ? MI : PI, FWD, RB>(acc, in, SS, nb, u, w);
    }
}

template<int FMT, int BLOCK, int NRHS, int RB>
__global__ __launch_bounds__(BLOCK) void k_dslash_eo_rb(DslashArgs a)
{
    using R = typename GaugeLoad<FMT>::R;
    using V = vec2<R>;
    constexpr int G = NRHS / RB;  // thread groups per site
    const uint32_t S = a.g.veo;
    const size_t SS  = (size_t)S * NRHS;
    uint32_t i = 0, grp = 0;
    if (!map_site_rhs<BLOCK, G>(a.map, S, blockIdx.x, threadIdx.x, i, grp)) return;
    const uint32_t r0 = grp * RB;
    SiteNb nb;
    site_neighbours(a.g, i, a.out_parity, nb);
    const V* in  = reinterpret_cast<const V*>(a.in);
    const int P  = a.out_parity;
    const int Pn = P ^ 1;
    Cx<R> acc[RB][12];
#pragma unroll
    for (int q = 0; q < RB; ++q)
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[q][k] = Cx<R>{R(0), R(0)};
    hop_rb<0, true, FMT, RB>(acc, in, SS, nb.up[0] * NRHS + r0, a.gauge, S, P, i, Cx<R>{R(a.w.wf[0][0]), R(a.w.wf[0][1])});
    hop_rb<0, false, FMT, RB>(acc, in, SS, nb.dn[0] * NRHS + r0, a.gauge, S, Pn, nb.dn[0],
                              Cx<R>{R(a.w.wb[0][0]), R(a.w.wb[0][1])});
    hop_rb<1, true, FMT, RB>(acc, in, SS, nb.up[1] * NRHS + r0, a.gauge, S, P, i, Cx<R>{R(a.w.wf[1][0]), R(a.w.wf[1][1])});
    hop_rb<1, false, FMT, RB>(acc, in, SS, nb.dn[1] * NRHS + r0, a.gauge, S, Pn, nb.dn[1],
                              Cx<R>{R(a.w.wb[1][0]), R(a.w.wb[1][1])});
    hop_rb<2, true, FMT, RB>(acc, in, SS, nb.up[2] * NRHS + r0, a.gauge, S, P, i, Cx<R>{R(a.w.wf[2][0]), R(a.w.wf[2][1])});
    hop_rb<2, false, FMT, RB>(acc, in, SS, nb.dn[2] * NRHS + r0, a.gauge, S, Pn, nb.dn[2],
                              Cx<R>{R(a.w.wb[2][0]), R(a.w.wb[2][1])});
    hop_rb<3, true, FMT, RB>(acc, in, SS, nb.up[3] * NRHS + r0, a.gauge, S, P, i, Cx<R>{R(a.w.wf[3][0]), R(a.w.wf[3][1])});
    hop_rb<3, false, FMT, RB>(acc, in, SS, nb.dn[3] * NRHS + r0, a.gauge, S, Pn, nb.dn[3],
                              Cx<R>{R(a.w.wb[3][0]), R(a.w.wb[3][1])});
    const Cx<R> f1{R(a.f1[0]), R(a.f1[1])}, f2{R(a.f2[0]), R(a.f2[1])};
    const uint32_t o = i * NRHS + r0;
    V* out           = reinterpret_cast<V*>(a.out);
    const V* x       = reinterpret_cast<const V*>(a.x);
#pragma unroll
    for (int q = 0; q < RB; ++q) {
        if (a.epilogue == EPI_MINV) {
            const R s = R(a.scale);
#pragma unroll
            for (int k = 0; k < 12; ++k) acc[q][k] = scale(cmul(acc[q][k], k < 6 ? f1 : f2), s);
        } else if (a.epilogue == EPI_AEE) {
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                V xv      = x[spinor_cx<R>(k, SS, (size_t)o + q)];
                acc[q][k] = cmul(Cx<R>{xv.x, xv.y}, k < 6 ? f1 : f2) - acc[q][k];
                if (a.gamma5 && k >= 6) acc[q][k] = Cx<R>{-acc[q][k].re, -acc[q][k].im};
            }
        }
#pragma unroll
        for (int k = 0; k < 12; ++k) out[spinor_cx<R>(k, SS, (size_t)o + q)] = V{acc[q][k].re, acc[q][k].im};
    }
}

/* Site-diagonal twisted-mass factors, gamma5 variants included
 * (ocl_kernel/fermionmatrix_eo_m.cl:22-109, fermionmatrix_eo_m_merged.cl:22-64):
 * spins 0,1 times f1, spins 2,3 times f2, everything times scale. */
template<typename R, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sitediag(vec2<R>* __restrict__ out, const vec2<R>* __restrict__ in,
                                                    uint32_t S, Cx<R> f1, Cx<R> f2, R s)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= S) return;
    vec2<R> v[12];
    load_spinor<R>(v, in, S, i);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        Cx<R> r = scale(cmul(Cx<R>{v[k].x, v[k].y}, k < 6 ? f1 : f2), s);
        v[k]    = vec2<R>{r.re, r.im};
    }
    store_spinor<R>(out, S, i, v);
}

cl2_hop_tuning hop_tuning(const cl2_ctx* c, int fmt, int nrhs)
{
    if (!c->tune_manual) {
        auto it = c->tuned.find({fmt, nrhs});
        if (it != c->tuned.end()) return it->second;
    }
    // default register budget (the autotuner's choices at 48^3x96): four hops in flight for the FP64 links (one
    // wave per SIMD) and for FP32 (two waves per SIMD), hop pairs for FP16
    const int occ = c->tune_occ >= 0
                        ? c->tune_occ
                        : ((fmt == GAUGE_F64_18 || fmt == GAUGE_F64_12) ? 1 : (fmt == GAUGE_F32_12 ? 2 : 0));
    return cl2_hop_tuning{c->tune_zplanes, c->tune_xcd, c->tune_mrhs_rb, c->tune_mrhs_lds, c->tune_block, occ};
}

WorkMap work_map(const cl2_ctx* c, int block, int nrhs, const cl2_hop_tuning& t)
{
    WorkMap m{0u, 0u, c->g.vs / 2, 0u};
    const uint32_t ns = c->g.ns, plane = ns * ns / 2;
    int zp            = t.zplanes;
    if (zp < 0) {
        // one z-chunk per XCD when ns divides by 8 (measured best at 48^3x96:
        // 8 chunks of 6 planes); else the largest divisor of ns whose slab
        // stays within 8192 eo sites
        if (ns % 8 == 0) {
            zp = (int)(ns / 8);
        } else {
            zp = 1;
            for (uint32_t d = 1; d <= ns; ++d)
                if (ns % d == 0 && d * plane <= 8192u) zp = (int)d;
        }
    }
    if (zp > 0 && ns % (uint32_t)zp == 0) {
        m.P   = (uint32_t)zp * plane;
        m.col = m.P * c->g.nt;
    }
    const uint32_t nblocks = (uint32_t)(((size_t)c->g.veo * nrhs + block - 1) / block);
    if (t.xcd && nblocks >= 64) m.xcd_q = nblocks / 8;
    return m;
}

HopWeights<double> wilson_weights(const cl2_ctx* c, double kappa)
{
    HopWeights<double> w;
    for (int mu = 0; mu < 4; ++mu) {
        const double* b = (mu == 0) ? c->bt : c->bs;
        double fr = b[0], fi = b[1];
        double br = b[0], bi = b[1];
        if (mu == 0) {  // chemical potential multiplies the T links (fermionmatrix_eo.cl:133-139,224-231)
            double tr = fr * c->cp_fwd[0] - fi * c->cp_fwd[1];
            fi        = fr * c->cp_fwd[1] + fi * c->cp_fwd[0];
            fr        = tr;
            tr        = br * c->cp_bwd[0] - bi * c->cp_bwd[1];
            bi        = br * c->cp_bwd[1] + bi * c->cp_bwd[0];
            br        = tr;
        }
        w.wf[mu][0] = -kappa * fr;
        w.wf[mu][1] = -kappa * fi;
        // backward hop uses conj(b) and (c U)^dagger = conj(c) U^dagger
        w.wb[mu][0] = -kappa * br;
        w.wb[mu][1] = kappa * bi;
    }
    return w;
}

template<int FMT, int NRHS, int RB>
static hipError_t launch_rb(cl2_ctx* c, const DslashArgs& a)
{
    const unsigned nb = blocks_for((size_t)a.g.veo * (NRHS / RB), DSLASH_BLOCK);
    // one right-hand side per thread: per-site LDS staging of links and neighbour indices, or the single-field
    // kernel over the [12][Veo][NRHS] block; both keep the hops of a direction in flight together
    if constexpr (RB == 1) {
        using R = typename GaugeLoad<FMT>::R;
        if (a.lds && sizeof(R) == 8 && !(c->tune_mrhs_off32 && mrhs_offsets_narrow(a.g.veo, NRHS, sizeof(vec2<R>)))) {
            // FP64 field above 4 GiB per component (or "mrhs.off32" 0): 64-bit gathers, SB sites per block
            constexpr int SB = MrhsLds<NRHS>::SB;
            if (a.lds != 1) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_dslash_eo_mrhs_lds<FMT, NRHS, 1, true>), dim3((a.g.veo + SB - 1) / SB),
                               dim3(MrhsLds<NRHS>::BLOCK), 0, c->stream, a);
        } else if (a.lds == 2) {
            constexpr int SB = MrhsLds<NRHS, 2>::SB;
            hipLaunchKernelGGL((k_dslash_eo_mrhs_lds<FMT, NRHS, 2>), dim3((a.g.veo + SB - 1) / SB),
                               dim3(MrhsLds<NRHS, 2>::BLOCK), 0, c->stream, a);
        } else if (a.lds) {
            constexpr int SB = MrhsLds<NRHS>::SB;
            hipLaunchKernelGGL((k_dslash_eo_mrhs_lds<FMT, NRHS, 1>), dim3((a.g.veo + SB - 1) / SB),
                               dim3(MrhsLds<NRHS>::BLOCK), 0, c->stream, a);
        } else {
            hipLaunchKernelGGL((k_dslash_eo<FMT, DSLASH_BLOCK, NRHS, 0>), dim3(nb), dim3(DSLASH_BLOCK), 0, c->stream,
                               a);
        }
    } else
        hipLaunchKernelGGL((k_dslash_eo_rb<FMT, DSLASH_BLOCK, NRHS, RB>), dim3(nb), dim3(DSLASH_BLOCK), 0, c->stream,
                           a);
    return hipGetLastError();
}

/* right-hand sides per thread of the multi-RHS kernel: the tuned value if it divides NRHS, else 1 (one
 * thread per site and right-hand side, links shared across the lanes of a site), which measured fastest on
 * gfx950 for every format: RB > 1 keeps RB x 12 accumulators live and drops to one wave per SIMD */
template<int FMT, int NRHS>
static int rhs_per_thread(const DslashArgs& a)
{
    const int want = a.rb > 0 ? a.rb : 1;
    for (int rb : {want, 4, 2, 1})
        if (rb <= 4 && rb != 3 && NRHS % rb == 0) return rb;
    return 1;
}

template<int NRHS>
static hipError_t launch_dslash_n(cl2_ctx* c, const DslashArgs& a, int format)
{
    if constexpr (NRHS == 1) {
        if (format == GAUGE_F16_12) return launch_dslash_h(c, a);
        auto by_shape = [&](auto fmt_tag) -> hipError_t {
            constexpr int F = decltype(fmt_tag)::value;
            const unsigned bs = (unsigned)a.block;
            const unsigned nb = blocks_for((size_t)a.g.veo, (int)bs);
            // the 1-wave-per-SIMD shapes (blocks of 128 or 256), the 128-VGPR shapes and the four-hops-in-
            // 256-VGPRs shapes for FP32 only (the FP64 kernels spill at those budgets)
            const bool plain = a.epilogue == EPI_PLAIN && !a.norm_partials && !a.cg_r && !a.per_rhs && !a.gamma5;
            if (a.occ == 1 && bs == 128) {
                hipLaunchKernelGGL((k_dslash_eo<F, 128, 1, 1>), dim3(nb), dim3(128), 0, c->stream, a);
                return hipGetLastError();
            }
            if (a.occ == 1 && bs != 512) {
                if (plain)
                    hipLaunchKernelGGL((k_dslash_eo<F, 256, 1, 1, true>), dim3(nb), dim3(256), 0, c->stream, a);
                else
                    hipLaunchKernelGGL((k_dslash_eo<F, 256, 1, 1>), dim3(nb), dim3(256), 0, c->stream, a);
                return hipGetLastError();
            }
            if constexpr (F == GAUGE_F32_12) {
                if (a.occ == 2 && bs == 128) {
                    hipLaunchKernelGGL((k_dslash_eo<F, 128, 1, 2>), dim3(nb), dim3(128), 0, c->stream, a);
                    return hipGetLastError();
                }
                if (a.occ == 2 && bs != 512) {
                    if (plain)
                        hipLaunchKernelGGL((k_dslash_eo<F, 256, 1, 2, true>), dim3(nb), dim3(256), 0, c->stream, a);
                    else
                        hipLaunchKernelGGL((k_dslash_eo<F, 256, 1, 2>), dim3(nb), dim3(256), 0, c->stream, a);
                    return hipGetLastError();
                }
                if (a.occ == 4 && bs == 128) {
                    hipLaunchKernelGGL((k_dslash_eo<F, 128, 1, 4>), dim3(nb), dim3(128), 0, c->stream, a);
                    return hipGetLastError();
                }
                if (a.occ == 4 && bs != 512) {
                    hipLaunchKernelGGL((k_dslash_eo<F, 256, 1, 4>), dim3(nb), dim3(256), 0, c->stream, a);
                    return hipGetLastError();
                }
            }
            if (bs == 128)
                hipLaunchKernelGGL((k_dslash_eo<F, 128, 1>), dim3(nb), dim3(128), 0, c->stream, a);
            else if (bs == 512)
                hipLaunchKernelGGL((k_dslash_eo<F, 512, 1>), dim3(nb), dim3(512), 0, c->stream, a);
            else if (plain)
                hipLaunchKernelGGL((k_dslash_eo<F, 256, 1, 0, true>), dim3(nb), dim3(256), 0, c->stream, a);
            else
                hipLaunchKernelGGL((k_dslash_eo<F, 256, 1>), dim3(nb), dim3(256), 0, c->stream, a);
            return hipGetLastError();
        };
        switch (format) {
            case GAUGE_F64_18: return by_shape(std::integral_constant<int, GAUGE_F64_18>{});
            case GAUGE_F64_12: return by_shape(std::integral_constant<int, GAUGE_F64_12>{});
            case GAUGE_F32_12: return by_shape(std::integral_constant<int, GAUGE_F32_12>{});
            default: return hipErrorInvalidValue;
        }
    } else {
        auto by_rb = [&](auto fmt_tag) -> hipError_t {
            constexpr int F = decltype(fmt_tag)::value;
            switch (rhs_per_thread<F, NRHS>(a)) {
                case 4:
                    if constexpr (NRHS % 4 == 0) return launch_rb<F, NRHS, 4>(c, a);
                    return hipErrorInvalidValue;
                case 2:
                    if constexpr (NRHS % 2 == 0) return launch_rb<F, NRHS, 2>(c, a);
                    return hipErrorInvalidValue;
                default: return launch_rb<F, NRHS, 1>(c, a);
            }
        };
        switch (format) {
            case GAUGE_F64_18: return by_rb(std::integral_constant<int, GAUGE_F64_18>{});
            case GAUGE_F64_12: return by_rb(std::integral_constant<int, GAUGE_F64_12>{});
            case GAUGE_F32_12: return by_rb(std::integral_constant<int, GAUGE_F32_12>{});
            case GAUGE_F16_12: return launch_dslash_mrhs_h(c, a);
            default: return hipErrorInvalidValue;
        }
    }
}

bool supported_nrhs(int n)
{
    return n == 1 || n == 2 || n == 4 || n == 8 || n == 12 || n == 16;
}

hipError_t launch_dslash(cl2_ctx* c, const DslashArgs& a, int format)
{
    switch (a.nrhs) {
        case 1: return launch_dslash_n<1>(c, a, format);
        case 2: return launch_dslash_n<2>(c, a, format);
        case 4: return launch_dslash_n<4>(c, a, format);
        case 8: return launch_dslash_n<8>(c, a, format);
        case 12: return launch_dslash_n<12>(c, a, format);
        case 16: return launch_dslash_n<16>(c, a, format);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_sitediag(cl2_ctx* c, void* out, const void* in, int prec, const double f1[2], const double f2[2],
                           double s)
{
    const uint32_t S  = c->g.veo;
    const unsigned nb = blocks_for(S, BLAS_BLOCK);
    if (prec == CL2_F64)
        hipLaunchKernelGGL((k_sitediag<double, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream,
                           reinterpret_cast<double2*>(out), reinterpret_cast<const double2*>(in), S,
                           Cx<double>{f1[0], f1[1]}, Cx<double>{f2[0], f2[1]}, s);
    else
        hipLaunchKernelGGL((k_sitediag<float, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream,
                           reinterpret_cast<float2*>(out), reinterpret_cast<const float2*>(in), S,
                           Cx<float>{(float)f1[0], (float)f1[1]}, Cx<float>{(float)f2[0], (float)f2[1]}, (float)s);
    return hipGetLastError();
}

}  // namespace cl2
