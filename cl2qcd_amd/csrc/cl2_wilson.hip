// Wilson / twisted-mass even-odd hopping kernels for gfx950.
//
// One thread per output eo site.  The input spinor and the gauge links are
// SoA (component-major) so every load of a wavefront is a contiguous 1 KiB
// (FP64) or 512 B (FP32) segment; the forward link of a site shares the
// site's own eo index and the backward link shares the neighbour's, so link
// loads are as coalesced as the spinor gathers.  The kernel is HBM bound
// (8 links + 2 spinors per site), so no MFMA.
#include "cl2_device.hpp"
#include "cl2_internal.hpp"

namespace cl2 {

/* One hop direction: the two half-spinor rows (A,B,C) of the projector
 *   h = psi_A + C * psi_B,   r = w * (U or U^dagger) h,
 *   acc_A += r,  acc_B += conj(C) * r
 * reproduce dslash_eoprec_unified_local (ocl_kernel/fermionmatrix_eo.cl:30-253);
 * w already carries -kappa and the boundary phase. */
template<int A, int B, int CU, bool FWD, typename R>
__device__ __forceinline__ void half_hop(Cx<R> (&acc)[12], const vec2<R>* __restrict__ in, uint32_t S, uint32_t nb,
                                         const Su3<R>& u, Cx<R> w)
{
    Cx<R> h[3], r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        vec2<R> a = in[(size_t)(3 * A + c) * S + nb];
        vec2<R> b = in[(size_t)(3 * B + c) * S + nb];
        h[c]      = add_unit<CU>(Cx<R>{a.x, a.y}, Cx<R>{b.x, b.y});
    }
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

template<int DIR, bool FWD, int FMT, typename R>
__device__ __forceinline__ void hop(Cx<R> (&acc)[12], const vec2<R>* __restrict__ in, uint32_t S, uint32_t nb,
                                    const void* gauge, int linkpar, uint32_t linkeo, Cx<R> w)
{
    Su3<R> u;
    GaugeLoad<FMT>::load(gauge, S, DIR, linkpar, linkeo, u);
    // projector rows of (1 - gamma_mu) forward and (1 + gamma_mu) backward, chiral basis
    if constexpr (DIR == 0) {
        half_hop<0, 2, FWD ? P1 : M1, FWD>(acc, in, S, nb, u, w);
        half_hop<1, 3, FWD ? P1 : M1, FWD>(acc, in, S, nb, u, w);
    } else if constexpr (DIR == 1) {
        half_hop<0, 3, FWD ? PI : MI, FWD>(acc, in, S, nb, u, w);
        half_hop<1, 2, FWD ? PI : MI, FWD>(acc, in, S, nb, u, w);
    } else if constexpr (DIR == 2) {
        half_hop<0, 3, FWD ? P1 : M1, FWD>(acc, in, S, nb, u, w);
        half_hop<1, 2, FWD ? M1 : P1, FWD>(acc, in, S, nb, u, w);
    } else {
        half_hop<0, 2, FWD ? PI : MI, FWD>(acc, in, S, nb, u, w);
        half_hop<1, 3, FWD ? MI : PI, FWD>(acc, in, S, nb, u, w);
    }
}

template<int FMT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_dslash_eo(DslashArgs a)
{
    using R = typename GaugeLoad<FMT>::R;
    using V = vec2<R>;
    __shared__ double lds[2 * BLOCK / 64];
    const uint32_t S = a.g.veo;
    uint32_t i       = 0;
    double nrm       = 0.;
    if (map_site<BLOCK>(a.map, S, blockIdx.x, threadIdx.x, i)) {
        SiteNb nb;
        site_neighbours(a.g, i, a.out_parity, nb);
        const V* in    = reinterpret_cast<const V*>(a.in);
        const int P    = a.out_parity;
        const int Pn   = P ^ 1;
        Cx<R> acc[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[k] = Cx<R>{R(0), R(0)};
        // order T, X, Y, Z as fermionmatrix_eo_dslash.cl:38-45
        hop<0, true, FMT>(acc, in, S, nb.up[0], a.gauge, P, i, Cx<R>{R(a.w.wf[0][0]), R(a.w.wf[0][1])});
        hop<0, false, FMT>(acc, in, S, nb.dn[0], a.gauge, Pn, nb.dn[0], Cx<R>{R(a.w.wb[0][0]), R(a.w.wb[0][1])});
        hop<1, true, FMT>(acc, in, S, nb.up[1], a.gauge, P, i, Cx<R>{R(a.w.wf[1][0]), R(a.w.wf[1][1])});
        hop<1, false, FMT>(acc, in, S, nb.dn[1], a.gauge, Pn, nb.dn[1], Cx<R>{R(a.w.wb[1][0]), R(a.w.wb[1][1])});
        hop<2, true, FMT>(acc, in, S, nb.up[2], a.gauge, P, i, Cx<R>{R(a.w.wf[2][0]), R(a.w.wf[2][1])});
        hop<2, false, FMT>(acc, in, S, nb.dn[2], a.gauge, Pn, nb.dn[2], Cx<R>{R(a.w.wb[2][0]), R(a.w.wb[2][1])});
        hop<3, true, FMT>(acc, in, S, nb.up[3], a.gauge, P, i, Cx<R>{R(a.w.wf[3][0]), R(a.w.wf[3][1])});
        hop<3, false, FMT>(acc, in, S, nb.dn[3], a.gauge, Pn, nb.dn[3], Cx<R>{R(a.w.wb[3][0]), R(a.w.wb[3][1])});

        const Cx<R> f1{R(a.f1[0]), R(a.f1[1])}, f2{R(a.f2[0]), R(a.f2[1])};
        if (a.epilogue == EPI_MINV) {
            const R s = R(a.scale);
#pragma unroll
            for (int k = 0; k < 12; ++k) acc[k] = scale(cmul(acc[k], k < 6 ? f1 : f2), s);
        } else if (a.epilogue == EPI_AEE) {
            const V* x = reinterpret_cast<const V*>(a.x);
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                V xv   = x[(size_t)k * S + i];
                acc[k] = cmul(Cx<R>{xv.x, xv.y}, k < 6 ? f1 : f2) - acc[k];
                if (a.gamma5 && k >= 6) acc[k] = Cx<R>{-acc[k].re, -acc[k].im};
            }
        }
        V* out = reinterpret_cast<V*>(a.out);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            out[(size_t)k * S + i] = V{acc[k].re, acc[k].im};
            nrm += (double)acc[k].re * (double)acc[k].re + (double)acc[k].im * (double)acc[k].im;
        }
    }
    if (a.norm_partials) {
        double s = block_sum<BLOCK>(nrm, lds);
        if (threadIdx.x == 0) a.norm_partials[2 * blockIdx.x] = s, a.norm_partials[2 * blockIdx.x + 1] = 0.;
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
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        vec2<R> v = in[(size_t)k * S + i];
        Cx<R> r   = scale(cmul(Cx<R>{v.x, v.y}, k < 6 ? f1 : f2), s);
        out[(size_t)k * S + i] = vec2<R>{r.re, r.im};
    }
}

WorkMap work_map(const cl2_ctx* c, int block)
{
    WorkMap m{0u, 0u, c->g.vs / 2, 0u};
    const uint32_t ns = c->g.ns, plane = ns * ns / 2;
    int zp            = c->tune_zplanes;
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
    const uint32_t nblocks = (c->g.veo + block - 1) / block;
    if (c->tune_xcd && nblocks >= 64) m.xcd_q = nblocks / 8;
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

hipError_t launch_dslash(cl2_ctx* c, const DslashArgs& a, int format)
{
    const unsigned nb = blocks_for(a.g.veo, DSLASH_BLOCK);
    switch (format) {
        case GAUGE_F64_18:
            hipLaunchKernelGGL((k_dslash_eo<GAUGE_F64_18, DSLASH_BLOCK>), dim3(nb), dim3(DSLASH_BLOCK), 0, c->stream,
                               a);
            break;
        case GAUGE_F64_12:
            hipLaunchKernelGGL((k_dslash_eo<GAUGE_F64_12, DSLASH_BLOCK>), dim3(nb), dim3(DSLASH_BLOCK), 0, c->stream,
                               a);
            break;
        case GAUGE_F32_12:
            hipLaunchKernelGGL((k_dslash_eo<GAUGE_F32_12, DSLASH_BLOCK>), dim3(nb), dim3(DSLASH_BLOCK), 0, c->stream,
                               a);
            break;
        case GAUGE_F16_12: return launch_dslash_h(c, a);
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
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
