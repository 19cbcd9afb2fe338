// Staggered even-odd hopping kernel D_KS_eo for gfx950.  Same site-per-thread
// SoA scheme, tiled links, work order and launch variants as the Wilson kernel;
// links are 92 % of the traffic.
#include <type_traits>

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
    // multi-RHS launches ([3][Veo][nrhs] fields): with x set, out_r = coef[r] x_r - D in_r (the last step of
    // MdagM_eo with a shift per right-hand side)
    const void* x;
    double coef[16];
    // the fused CG step of the lockstep solver (multi-RHS only): norm_partials[r * gridDim.x + block] gets per
    // right-hand side (|out_r|^2, |own_r|^2) — own is summed at the thread's eo index, so every site of its
    // field once — or, with cg_r set, r_r <- r_r - cg_alpha[r] out_r is stored instead of out (never written)
    // and the partials are (|r_new|^2, Re <r_new, r_old>)
    double2* norm_partials = nullptr;
    const void* own        = nullptr;
    void* cg_r             = nullptr;
    const double* cg_alpha = nullptr;
};

/* forward and backward hop of one direction on neighbour vectors already in registers:
 *   acc += w U_mu(x) chi(x+mu) - conj(w) U_mu(x-mu)^dagger chi(x-mu) */
template<typename R>
__device__ __forceinline__ void dks_hops(Cx<R> (&acc)[3], const Su3<R>& uf, const vec2<R> (&vf)[3], const Su3<R>& ub,
                                         const vec2<R> (&vb)[3], Cx<R> w)
{
    Cx<R> v[3], r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = Cx<R>{vf[c].x, vf[c].y};
    su3_times(uf, v, r);
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] = acc[c] + cmul(r[c], w);
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = Cx<R>{vb[c].x, vb[c].y};
    su3dag_times(ub, v, r);
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] = acc[c] - cmulconj(r[c], w);
}

/* D_KS_eo_local (ocl_kernel/fermionmatrix_staggered_eo_DKS_local.cl:65-136):
 *   sum_mu eta_mu(x) [ sf_mu U_mu(x) chi(x+mu) - conj(sf_mu) U_mu(x-mu)^dagger chi(x-mu) ]
 * with eta_X = 1, eta_Y = (-1)^x, eta_Z = (-1)^(x+y), eta_T = (-1)^(x+y+z)
 * (ocl_kernel/operations_staggered.cl:56-104).  Directions T, X, Y, Z, forward then backward.  Register
 * budget and loads in flight as the Wilson kernel: OCC = 0 both hops of a direction (<= 2 waves per SIMD),
 * 1 the four hops of two directions (1 wave), 4 one hop at a time (>= 4 waves); every variant computes the
 * same bits.  Results leave by streaming stores. */
template<int FMT, int BLOCK, int OCC>
__global__ __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(OCC == 4 ? 4 : 1, OCC == 4 ? 8 : (OCC == 1 ? 1 : 2))))
void k_dks_eo(DksArgs a)
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
    auto wt = [&](int mu) { return Cx<R>{R(a.sf[mu][0]) * eta[mu], R(a.sf[mu][1]) * eta[mu]}; };
    if constexpr (OCC == 1) {
#pragma unroll
        for (int m0 = 0; m0 < 4; m0 += 2) {
            Su3<R> u[4];
            V v[4][3];
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                GaugeLoad<FMT>::load(a.gauge, S, m0 + d, P, i, u[2 * d]);
                GaugeLoad<FMT>::load(a.gauge, S, m0 + d, P ^ 1, nb.dn[m0 + d], u[2 * d + 1]);
            }
#pragma unroll
            for (int d = 0; d < 2; ++d)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    v[2 * d][c]     = in[(size_t)c * S + nb.up[m0 + d]];
                    v[2 * d + 1][c] = in[(size_t)c * S + nb.dn[m0 + d]];
                }
            dks_hops(acc, u[0], v[0], u[1], v[1], wt(m0));
            dks_hops(acc, u[2], v[2], u[3], v[3], wt(m0 + 1));
        }
    } else {
#pragma unroll
        for (int mu = 0; mu < 4; ++mu) {
            Su3<R> uf, ub;
            V vf[3], vb[3];
            GaugeLoad<FMT>::load(a.gauge, S, mu, P, i, uf);
#pragma unroll
            for (int c = 0; c < 3; ++c) vf[c] = in[(size_t)c * S + nb.up[mu]];
            if constexpr (OCC == 4) {  // one hop at a time: the forward hop is applied before the backward loads
                Cx<R> vv[3], r[3];
                const Cx<R> w = wt(mu);
#pragma unroll
                for (int c = 0; c < 3; ++c) vv[c] = Cx<R>{vf[c].x, vf[c].y};
                su3_times(uf, vv, r);
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[c] = acc[c] + cmul(r[c], w);
                GaugeLoad<FMT>::load(a.gauge, S, mu, P ^ 1, nb.dn[mu], ub);
#pragma unroll
                for (int c = 0; c < 3; ++c) vb[c] = in[(size_t)c * S + nb.dn[mu]];
#pragma unroll
                for (int c = 0; c < 3; ++c) vv[c] = Cx<R>{vb[c].x, vb[c].y};
                su3dag_times(ub, vv, r);
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[c] = acc[c] - cmulconj(r[c], w);
            } else {
                GaugeLoad<FMT>::load(a.gauge, S, mu, P ^ 1, nb.dn[mu], ub);
#pragma unroll
                for (int c = 0; c < 3; ++c) vb[c] = in[(size_t)c * S + nb.dn[mu]];
                dks_hops(acc, uf, vf, ub, vb, wt(mu));
            }
        }
    }
    V* out = reinterpret_cast<V*>(a.out);
#pragma unroll
    for (int c = 0; c < 3; ++c) store_stream(out + (size_t)c * S + i, V{acc[c].re, acc[c].im});
}

/* Multi-RHS D_KS_eo with per-site staging, as the Wilson multi-RHS kernel: a block owns SB consecutive sites of
 * the work order and NRHS threads per site (right-hand side fastest); SB threads compute the sites' neighbour
 * indices and SB x 8 threads load and reconstruct their 8 links once into LDS, then every right-hand-side
 * thread gathers its own neighbour vectors.  Per right-hand side the arithmetic and order are the single-field
 * kernel's, so slot r equals it bit for bit. */
template<int NRHS>
struct DksMrhs {
    static constexpr int SB    = NRHS >= 8 ? 16 : 32;
    static constexpr int BLOCK = SB * NRHS;
};

template<int FMT, int NRHS>
__global__ __launch_bounds__(DksMrhs<NRHS>::BLOCK) void k_dks_eo_mrhs(DksArgs a)
{
    using R = typename GaugeLoad<FMT>::R;
    using V = vec2<R>;
    constexpr int SB        = DksMrhs<NRHS>::SB;
    constexpr int BLOCK     = DksMrhs<NRHS>::BLOCK;
    constexpr uint32_t NONE = 0xffffffffu;
    __shared__ Su3<R> lu[SB * 8];
    __shared__ uint32_t lnb[SB * 8];
    __shared__ uint32_t lsite[SB];
    __shared__ double2 lred[BLOCK];
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
        uint32_t x, y, z, t;
        site_coords(a.g, i, P, x, y, z, t);
        const R eta[4] = {R(((x + y + z) & 1u) ? -1 : 1), R(1), R((x & 1u) ? -1 : 1), R(((x + y) & 1u) ? -1 : 1)};
        const V* in    = reinterpret_cast<const V*>(a.in);
        Cx<R> acc[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] = Cx<R>{R(0), R(0)};
#pragma unroll
        for (int mu = 0; mu < 4; ++mu) {
            V vf[3], vb[3];
            const size_t ef = (size_t)lnb[s * 8 + 2 * mu] * NRHS + rhs;
            const size_t eb = (size_t)lnb[s * 8 + 2 * mu + 1] * NRHS + rhs;
#pragma unroll
            for (int c = 0; c < 3; ++c) vf[c] = in[(size_t)c * SS + ef];
#pragma unroll
            for (int c = 0; c < 3; ++c) vb[c] = in[(size_t)c * SS + eb];
            dks_hops(acc, lu[s * 8 + 2 * mu], vf, lu[s * 8 + 2 * mu + 1], vb,
                     Cx<R>{R(a.sf[mu][0]) * eta[mu], R(a.sf[mu][1]) * eta[mu]});
        }
        const size_t o = (size_t)i * NRHS + rhs;
        if (a.x) {  // the MdagM_eo epilogue as the elementwise kernel computes it: coef x - out
            const V* xv      = reinterpret_cast<const V*>(a.x);
            const Cx<R> cf   = Cx<R>{R(a.coef[rhs]), R(0)};
            const Cx<R> mone = Cx<R>{R(-1), R(0)};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const V v = xv[(size_t)c * SS + o];
                acc[c]    = cmul(Cx<R>{v.x, v.y}, cf) + cmul(acc[c], mone);
            }
        }
        if (a.cg_r) {  // r <- r - alpha A p, A p never stored
            V* rp            = reinterpret_cast<V*>(a.cg_r);
            const Cx<R> al   = Cx<R>{R(a.cg_alpha[2 * rhs]), R(a.cg_alpha[2 * rhs + 1])};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const V ro     = rp[(size_t)c * SS + o];
                const Cx<R> rn = Cx<R>{ro.x, ro.y} - cmul(acc[c], al);
                store_stream(rp + (size_t)c * SS + o, V{rn.re, rn.im});
                nrm += (double)rn.re * (double)rn.re + (double)rn.im * (double)rn.im;
                ry += (double)rn.re * (double)ro.x + (double)rn.im * (double)ro.y;
            }
        } else {
            V* out = reinterpret_cast<V*>(a.out);
#pragma unroll
            for (int c = 0; c < 3; ++c) store_stream(out + (size_t)c * SS + o, V{acc[c].re, acc[c].im});
            if (a.norm_partials) {
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    nrm += (double)acc[c].re * (double)acc[c].re + (double)acc[c].im * (double)acc[c].im;
                if (a.own) {
                    const V* ov = reinterpret_cast<const V*>(a.own);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const V v = ov[(size_t)c * SS + o];
                        ry += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
                    }
                }
            }
        }
    }
    if (a.norm_partials) {
        // per right-hand side: thread r < NRHS sums the SB site threads of its column in site order
        lred[threadIdx.x] = make_double2(nrm, ry);
        __syncthreads();
        if ((int)threadIdx.x < NRHS) {
            double2 tsum = make_double2(0., 0.);
            for (int q = 0; q < SB; ++q) {
                tsum.x += lred[q * NRHS + threadIdx.x].x;
                tsum.y += lred[q * NRHS + threadIdx.x].y;
            }
            a.norm_partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = tsum;
        }
    }
}

static void dks_weights(cl2_ctx* c, DksArgs& a)
{
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
}

hipError_t launch_dks_mrhs(cl2_ctx* c, void* out, const void* in, const void* gauge, int format, int evenodd, int nrhs,
                           const void* x, const double* coef, double2* norm_partials, const void* own, void* cg_r,
                           const double* cg_alpha, unsigned* nblocks)
{
    DksArgs a;
    a.in            = in;
    a.out           = out;
    a.gauge         = gauge;
    a.g             = c->g;
    a.out_parity    = evenodd;
    a.x             = x;
    a.norm_partials = norm_partials;
    a.own           = own;
    a.cg_r          = cg_r;
    a.cg_alpha      = cg_alpha;
    for (int r = 0; r < 16; ++r) a.coef[r] = (x && r < nrhs) ? coef[r] : 0.;
    dks_weights(c, a);
    const cl2_hop_tuning t = hop_tuning(c, format | DKS_TUNING_KEY, 1);
    auto go = [&](auto fmt_tag, auto n_tag) -> hipError_t {
        constexpr int F = decltype(fmt_tag)::value, N = decltype(n_tag)::value;
        constexpr int SB = DksMrhs<N>::SB;
        a.map = work_map(c, SB, 1, t);
        if (nblocks) *nblocks = (c->g.veo + SB - 1) / SB;
        hipLaunchKernelGGL((k_dks_eo_mrhs<F, N>), dim3((c->g.veo + SB - 1) / SB), dim3(DksMrhs<N>::BLOCK), 0, c->stream,
                           a);
        return hipGetLastError();
    };
    auto by_n = [&](auto fmt_tag) -> hipError_t {
        switch (nrhs) {
            case 2: return go(fmt_tag, std::integral_constant<int, 2>{});
            case 4: return go(fmt_tag, std::integral_constant<int, 4>{});
            case 8: return go(fmt_tag, std::integral_constant<int, 8>{});
            case 12: return go(fmt_tag, std::integral_constant<int, 12>{});
            case 16: return go(fmt_tag, std::integral_constant<int, 16>{});
            default: return hipErrorInvalidValue;
        }
    };
    switch (format) {
        case GAUGE_F64_18: return by_n(std::integral_constant<int, GAUGE_F64_18>{});
        case GAUGE_F64_12: return by_n(std::integral_constant<int, GAUGE_F64_12>{});
        case GAUGE_F32_12: return by_n(std::integral_constant<int, GAUGE_F32_12>{});
        default: return hipErrorInvalidValue;
    }
}

/* The two D_KS_eo kernels of a fused CG-M step on MdagM_eo (single field, the hop order of k_dks_eo): without
 * cg_r, out = D in is stored and each block sums (|out|^2, |own|^2), own read at the thread's eo index (every
 * site of that field once); with cg_r, v = coef[0] x - D in is formed in registers, r <- r + beta v (beta a
 * device scalar) is stored and each block sums (|r_new|^2, 0).  One double2 per block into norm_partials. */
template<int FMT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_dks_eo_step(DksArgs a, const double* beta)
{
    using R = typename GaugeLoad<FMT>::R;
    using V = vec2<R>;
    __shared__ double lds[2 * BLOCK / 64];
    const uint32_t S = a.g.veo;
    uint32_t i       = 0;
    double nrm = 0., ry = 0.;
    if (map_site<BLOCK>(a.map, S, blockIdx.x, threadIdx.x, i)) {
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
            Su3<R> uf, ub;
            V vf[3], vb[3];
            GaugeLoad<FMT>::load(a.gauge, S, mu, P, i, uf);
            GaugeLoad<FMT>::load(a.gauge, S, mu, P ^ 1, nb.dn[mu], ub);
#pragma unroll
            for (int c = 0; c < 3; ++c) vf[c] = in[(size_t)c * S + nb.up[mu]];
#pragma unroll
            for (int c = 0; c < 3; ++c) vb[c] = in[(size_t)c * S + nb.dn[mu]];
            dks_hops(acc, uf, vf, ub, vb, Cx<R>{R(a.sf[mu][0]) * eta[mu], R(a.sf[mu][1]) * eta[mu]});
        }
        if (a.cg_r) {
            const V* xv      = reinterpret_cast<const V*>(a.x);
            V* rp            = reinterpret_cast<V*>(a.cg_r);
            const Cx<R> cf   = Cx<R>{R(a.coef[0]), R(0)};
            const Cx<R> mone = Cx<R>{R(-1), R(0)};
            const Cx<R> b    = Cx<R>{R(beta[0]), R(0)};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const V xs     = xv[(size_t)c * S + i];
                const Cx<R> v  = cmul(Cx<R>{xs.x, xs.y}, cf) + cmul(acc[c], mone);
                const V ro     = rp[(size_t)c * S + i];
                const Cx<R> rn = cmul(b, v) + Cx<R>{ro.x, ro.y};
                store_stream(rp + (size_t)c * S + i, V{rn.re, rn.im});
                nrm += (double)rn.re * (double)rn.re + (double)rn.im * (double)rn.im;
            }
        } else {
            V* out = reinterpret_cast<V*>(a.out);
            const V* ov = reinterpret_cast<const V*>(a.own);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                store_stream(out + (size_t)c * S + i, V{acc[c].re, acc[c].im});
                nrm += (double)acc[c].re * (double)acc[c].re + (double)acc[c].im * (double)acc[c].im;
                const V o = ov[(size_t)c * S + i];
                ry += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
            }
        }
    }
    const double2 s2 = block_sum2<BLOCK>(make_double2(nrm, ry), lds);
    if (threadIdx.x == 0) a.norm_partials[blockIdx.x] = s2;
}

hipError_t launch_dks_step(cl2_ctx* c, void* out, const void* in, const void* gauge, int format, int evenodd,
                           const void* own, const void* x, double coef, void* cg_r, const double* beta,
                           unsigned* nblocks)
{
    constexpr int B = 256;
    DksArgs a;
    a.in            = in;
    a.out           = out;
    a.gauge         = gauge;
    a.g             = c->g;
    a.out_parity    = evenodd;
    a.x             = x;
    a.coef[0]       = coef;
    a.own           = own;
    a.cg_r          = cg_r;
    a.norm_partials = reinterpret_cast<double2*>(c->partials);
    const cl2_hop_tuning t = hop_tuning(c, format | DKS_TUNING_KEY, 1);
    a.map                  = work_map(c, B, 1, t);
    dks_weights(c, a);
    const unsigned nb = blocks_for(c->g.veo, B);
    *nblocks          = nb;
    switch (format) {
        case GAUGE_F64_18:
            hipLaunchKernelGGL((k_dks_eo_step<GAUGE_F64_18, B>), dim3(nb), dim3(B), 0, c->stream, a, beta);
            break;
        case GAUGE_F64_12:
            hipLaunchKernelGGL((k_dks_eo_step<GAUGE_F64_12, B>), dim3(nb), dim3(B), 0, c->stream, a, beta);
            break;
        case GAUGE_F32_12:
            hipLaunchKernelGGL((k_dks_eo_step<GAUGE_F32_12, B>), dim3(nb), dim3(B), 0, c->stream, a, beta);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_dks(cl2_ctx* c, void* out, const void* in, const void* gauge, int format, int evenodd)
{
    DksArgs a;
    a.in         = in;
    a.out        = out;
    a.gauge      = gauge;
    a.g          = c->g;
    a.out_parity = evenodd;
    a.x          = nullptr;
    const cl2_hop_tuning t = hop_tuning(c, format | DKS_TUNING_KEY, 1);
    const int bs           = (t.block == 128 || t.block == 512) ? t.block : 256;
    a.map                  = work_map(c, bs, 1, t);
    dks_weights(c, a);
    const unsigned nb = blocks_for(c->g.veo, bs);
    auto go = [&](auto fmt_tag, auto block_tag) {
        constexpr int F = decltype(fmt_tag)::value, B = decltype(block_tag)::value;
        if (t.occ == 4)
            hipLaunchKernelGGL((k_dks_eo<F, B, 4>), dim3(nb), dim3(B), 0, c->stream, a);
        else if (t.occ == 1 && B != 512)
            hipLaunchKernelGGL((k_dks_eo<F, B, 1>), dim3(nb), dim3(B), 0, c->stream, a);
        else
            hipLaunchKernelGGL((k_dks_eo<F, B, 0>), dim3(nb), dim3(B), 0, c->stream, a);
    };
    auto by_block = [&](auto fmt_tag) {
        if (bs == 128)
            go(fmt_tag, std::integral_constant<int, 128>{});
        else if (bs == 512)
            go(fmt_tag, std::integral_constant<int, 512>{});
        else
            go(fmt_tag, std::integral_constant<int, 256>{});
    };
    switch (format) {
        case GAUGE_F64_18: by_block(std::integral_constant<int, GAUGE_F64_18>{}); break;
        case GAUGE_F64_12: by_block(std::integral_constant<int, GAUGE_F64_12>{}); break;
        case GAUGE_F32_12: by_block(std::integral_constant<int, GAUGE_F32_12>{}); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cl2
