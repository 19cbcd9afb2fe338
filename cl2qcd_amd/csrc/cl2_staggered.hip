// Staggered even-odd hopping kernel D_KS_eo for gfx950.  Same site-per-thread
// SoA scheme as the Wilson kernel; links are 92 % of the traffic.
#include "cl2_device.hpp"
#include "cl2_internal.hpp"

namespace cl2 {

struct DksArgs {
    const void* in;
    void* out;
    const void* gauge;
    LatticeGeom g;
    int out_parity;
    double sf[4][2];  // 0.5 * b_mu (* imaginary chemical potential on T)
    WorkMap map;
};

/* D_KS_eo_local (ocl_kernel/fermionmatrix_staggered_eo_DKS_local.cl:65-136):
 *   sum_mu eta_mu(x) [ sf_mu U_mu(x) chi(x+mu) - conj(sf_mu) U_mu(x-mu)^dagger chi(x-mu) ]
 * with eta_X = 1, eta_Y = (-1)^x, eta_Z = (-1)^(x+y), eta_T = (-1)^(x+y+z)
 * (ocl_kernel/operations_staggered.cl:56-104). */
template<int FMT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_dks_eo(DksArgs a)
{
    using R = typename GaugeLoad<FMT>::R;
    using V = vec2<R>;
    const uint32_t S = a.g.veo;
    uint32_t i       = 0;
    if (!map_site<BLOCK>(a.map, S, blockIdx.x, threadIdx.x, i)) return;
    uint32_t x, y, z, t;
    site_coords(a.g, i, a.out_parity, x, y, z, t);
    SiteNb nb;
    site_neighbours(a.g, i, a.out_parity, nb);
    const V* in = reinterpret_cast<const V*>(a.in);
    const int P = a.out_parity;
    Cx<R> acc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[k] = Cx<R>{R(0), R(0)};
    const R eta[4] = {R(((x + y + z) & 1u) ? -1 : 1), R(1), R((x & 1u) ? -1 : 1), R(((x + y) & 1u) ? -1 : 1)};
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
        Su3<R> u;
        Cx<R> v[3], r[3];
        const Cx<R> w{R(a.sf[mu][0]) * eta[mu], R(a.sf[mu][1]) * eta[mu]};
        GaugeLoad<FMT>::load(a.gauge, S, mu, P, i, u);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            V q  = in[(size_t)c * S + nb.up[mu]];
            v[c] = Cx<R>{q.x, q.y};
        }
        su3_times(u, v, r);
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] = acc[c] + cmul(r[c], w);
        GaugeLoad<FMT>::load(a.gauge, S, mu, P ^ 1, nb.dn[mu], u);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            V q  = in[(size_t)c * S + nb.dn[mu]];
            v[c] = Cx<R>{q.x, q.y};
        }
        su3dag_times(u, v, r);
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] = acc[c] - cmulconj(r[c], w);
    }
    V* out = reinterpret_cast<V*>(a.out);
#pragma unroll
    for (int c = 0; c < 3; ++c) out[(size_t)c * S + i] = V{acc[c].re, acc[c].im};
}

hipError_t launch_dks(cl2_ctx* c, void* out, const void* in, const void* gauge, int format, int evenodd)
{
    DksArgs a;
    a.in         = in;
    a.out        = out;
    a.gauge      = gauge;
    a.g          = c->g;
    a.out_parity = evenodd;
    a.map        = work_map(c, DSLASH_BLOCK, 1, hop_tuning(c, format | 0x100, 1));
    for (int mu = 0; mu < 4; ++mu) {
        const double* b = (mu == 0) ? c->bt : c->bs;
        double re = 0.5 * b[0], im = 0.5 * b[1];
        if (mu == 0 && c->lat.use_cp_im) {  // only the imaginary chemical potential enters D_KS
            double cr = std::cos(c->lat.cp_im), ci = std::sin(c->lat.cp_im);
            double tr = re * cr - im * ci;
            im        = re * ci + im * cr;
            re        = tr;
        }
        a.sf[mu][0] = re;
        a.sf[mu][1] = im;
    }
    const unsigned nb = blocks_for(c->g.veo, DSLASH_BLOCK);
    switch (format) {
        case GAUGE_F64_18:
            hipLaunchKernelGGL((k_dks_eo<GAUGE_F64_18, DSLASH_BLOCK>), dim3(nb), dim3(DSLASH_BLOCK), 0, c->stream, a);
            break;
        case GAUGE_F64_12:
            hipLaunchKernelGGL((k_dks_eo<GAUGE_F64_12, DSLASH_BLOCK>), dim3(nb), dim3(DSLASH_BLOCK), 0, c->stream, a);
            break;
        case GAUGE_F32_12:
            hipLaunchKernelGGL((k_dks_eo<GAUGE_F32_12, DSLASH_BLOCK>), dim3(nb), dim3(DSLASH_BLOCK), 0, c->stream, a);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cl2
