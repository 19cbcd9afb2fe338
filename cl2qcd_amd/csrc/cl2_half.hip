// FP16 storage for the reduced-precision inner solves: Wilson eo spinors as
// half precision with one FP32 scale per site, and 12-real links in half
// precision.  BLAS arithmetic is FP32 and reductions accumulate in FP64; the
// hopping kernels form the spin projection in half precision and multiply it by
// the link with half-precision dot products that accumulate in FP32.
//
// Spinor layout (52 bytes per eo site): 3 lanes of 8 halves (16 bytes), lane j
// holding complex components 4j..4j+3 (component = 3*spin + colour) as re, im
// pairs, stored [3][Veo]; then float norm[Veo].  The value of a component is
// half * norm, with norm the largest |re| or |im| on the site, so the stored
// halves lie in [-1, 1] and keep 11 significant bits relative to the site's
// largest entry.  16-byte lanes keep every wavefront load a 1 KiB segment.
//
// Link layout (24 bytes per link): for each (dir, parity) tiles of 64 eo sites,
// each tile 64 16-byte lanes holding entries 0..3 of rows 0,1 (row-major)
// followed by 64 8-byte lanes holding entries 4, 5; row 2 = conj(row0 x row1).
//
// The hopping kernel has the FP64/FP32 kernels' variants: H = 1, 2 or 4 hops
// with their links and neighbour sites loaded before any is applied, under a
// register budget that leaves room for them; every variant computes the same
// bits.
#include <algorithm>
#include <type_traits>

#include "cl2_device.hpp"
#include "cl2_internal.hpp"

namespace cl2 {

typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ const float* half_norms(const void* p, uint32_t S)
{
    return reinterpret_cast<const float*>(static_cast<const char*>(p) + (size_t)48 * S);
}
__device__ __forceinline__ float* half_norms(void* p, uint32_t S)
{
    return reinterpret_cast<float*>(static_cast<char*>(p) + (size_t)48 * S);
}

/* STREAM: non-temporal loads (the BLAS kernels, which read each site once) */
template<bool STREAM = false>
__device__ __forceinline__ void load_site_h(const void* p, uint32_t S, uint32_t i, Cx<float> (&v)[12])
{
    const h8* l   = reinterpret_cast<const h8*>(p);
    const float n = STREAM ? __builtin_nontemporal_load(half_norms(p, S) + i) : half_norms(p, S)[i];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const f8v f = __builtin_convertvector(STREAM ? __builtin_nontemporal_load(l + (size_t)j * S + i)
                                                     : l[(size_t)j * S + i],
                                              f8v);
#pragma unroll
        for (int q = 0; q < 4; ++q) v[4 * j + q] = Cx<float>{f[2 * q] * n, f[2 * q + 1] * n};
    }
}

/* the same without the scale: the site's 12 stored (re, im) half pairs and its norm, for the hopping kernels,
 * which fold the norm into the hop weight (the hop is linear in the spinor) and so work on the stored halves */
__device__ __forceinline__ float load_site_h_raw(const void* p, uint32_t S, uint32_t i, h2 (&v)[12])
{
    const h8* l = reinterpret_cast<const h8*>(p);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const h8 f = l[(size_t)j * S + i];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[4 * j + q] = h2{f[2 * q], f[2 * q + 1]};
    }
    return half_norms(p, S)[i];
}

/* the two operand forms of one component of h = a + C b (C in {+1, -1, +i, -i}, a and b stored halves) for
 * the dot products of the link product: (h_re, -h_im) and (h_im, h_re) for U h (FWD), (h_re, h_im) and
 * (h_im, -h_re) for U^dagger h.  Each form is one packed half add (signs and the swap are operand modifiers),
 * so h is rounded to half once and both forms carry that value. */
template<int C, bool FWD>
__device__ __forceinline__ void proj_h(h2 a, h2 b, h2& f1, h2& f2)
{
    h2 cb;  // C b
    if constexpr (C == P1) cb = b;
    else if constexpr (C == M1) cb = -b;
    else if constexpr (C == PI) cb = h2{-b.y, b.x};
    else cb = h2{b.y, -b.x};
    if constexpr (FWD) {
        f1 = h2{a.x, -a.y} + h2{cb.x, -cb.y};
        f2 = h2{a.y, a.x} + h2{cb.y, cb.x};
    } else {
        f1 = a + cb;
        f2 = h2{a.y, -a.x} + h2{cb.y, -cb.x};
    }
}

/* hop weight times a site's norm */
__device__ __forceinline__ Cx<float> weight_n(const double (&w)[2], float n)
{
    return Cx<float>{float(w[0]) * n, float(w[1]) * n};
}

/* Re <a, b> over one site's 12 components, in float by fused multiply-adds in component order: the per-site
 * terms of the hopping kernels' norm partials, which are summed over sites in FP64 (the float sum of 24 terms
 * is exact to a few float ulps, far below the FP16 values' own rounding) */
__device__ __forceinline__ float site_redot(const Cx<float> (&a)[12], const Cx<float> (&b)[12])
{
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 12; ++k) s = __builtin_fmaf(a[k].im, b[k].im, __builtin_fmaf(a[k].re, b[k].re, s));
    return s;
}

/* quantise one site; NaN and Inf stay visible as NaN after the round trip.  STREAM: non-temporal stores
 * (results of a hopping kernel, which no later work of the launch reads) */
template<bool STREAM = false>
__device__ __forceinline__ void store_site_h(void* p, uint32_t S, uint32_t i, const Cx<float> (&v)[12])
{
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < 12; ++k) m = fmaxf(m, fmaxf(fabsf(v[k].re), fabsf(v[k].im)));
    const float inv = m > 0.f ? 1.f / m : 0.f;
    h8* l           = reinterpret_cast<h8*>(p);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        f8v f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            f[2 * q]     = v[4 * j + q].re * inv;
            f[2 * q + 1] = v[4 * j + q].im * inv;
        }
        if constexpr (STREAM)
            __builtin_nontemporal_store(__builtin_convertvector(f, h8), l + (size_t)j * S + i);
        else
            l[(size_t)j * S + i] = __builtin_convertvector(f, h8);
    }
    if constexpr (STREAM)
        __builtin_nontemporal_store(m, half_norms(p, S) + i);
    else
        half_norms(p, S)[i] = m;
}

/* an FP16 link as the hopping kernels use it: rows 0 and 1 as stored, row 2 = conj(row0 x row1) formed in
 * float from them and rounded to half; entries are (re, im) half pairs */
struct Su3h {
    h2 e[3][3];
};

__device__ __forceinline__ void load_link_h(const void* g, uint32_t S, int dir, int par, uint32_t eo, Su3h& u)
{
    const char* base = static_cast<const char*>(g);
    const uint32_t dp = (uint32_t)(dir * 2 + par);
    const h8 a = __builtin_nontemporal_load(reinterpret_cast<const h8*>(base + link_h16_offset(dp, eo, S)));
    const h4 b = __builtin_nontemporal_load(reinterpret_cast<const h4*>(base + link_h8_offset(dp, eo, S)));
    u.e[0][0] = h2{a[0], a[1]};
    u.e[0][1] = h2{a[2], a[3]};
    u.e[0][2] = h2{a[4], a[5]};
    u.e[1][0] = h2{a[6], a[7]};
    u.e[1][1] = h2{b[0], b[1]};
    u.e[1][2] = h2{b[2], b[3]};
    Su3<float> f;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) f.e[r][c] = Cx<float>{(float)u.e[r][c].x, (float)u.e[r][c].y};
    su3_reconstruct_row2(f);
#pragma unroll
    for (int c = 0; c < 3; ++c) u.e[2][c] = h2{(_Float16)f.e[2][c].re, (_Float16)f.e[2][c].im};
}

/* x.lo y.lo + x.hi y.hi in float: the first dot product of a chain.  ASM: the same instruction with the zero
 * accumulator as an inline constant (the builtin's accumulating form first copies a zero into the result
 * register; the multi-RHS kernel saves those copies, while in the single-field kernels the asm form makes the
 * scheduler hold far more registers) */
template<bool ASM>
__device__ __forceinline__ float dot2_first(h2 x, h2 y)
{
    if constexpr (ASM) {
        float r;
        asm("v_dot2_f32_f16 %0, %1, %2, 0" : "=v"(r) : "v"(x), "v"(y));
        return r;
    } else {
        return __builtin_amdgcn_fdot2(x, y, 0.f, false);
    }
}

/* r = U h (FWD) or U^dagger h from the two forms of each component of h: Re r_a = sum_b dot(U_ab, f1_b),
 * Im r_a = sum_b dot(U_ab, f2_b) (U_ba for U^dagger).  Each dot takes two exact half products and accumulates
 * in float, b = 0, 1, 2 in turn. */
template<bool FWD, bool ASM>
__device__ __forceinline__ void su3_dot_h(const Su3h& u, const h2 (&f1)[3], const h2 (&f2)[3], Cx<float> (&r)[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const h2 e0 = FWD ? u.e[a][0] : u.e[0][a];
        float re    = dot2_first<ASM>(e0, f1[0]), im = dot2_first<ASM>(e0, f2[0]);
#pragma unroll
        for (int b = 1; b < 3; ++b) {
            const h2 e = FWD ? u.e[a][b] : u.e[b][a];
            re         = __builtin_amdgcn_fdot2(e, f1[b], re, false);
            im         = __builtin_amdgcn_fdot2(e, f2[b], im, false);
        }
        r[a] = Cx<float>{re, im};
    }
}

/* one projector row pair of a hop on a neighbour spinor already in registers
 * (same algebra as the FP64/FP32 kernel: h = psi_A + C psi_B, r = w U h, with h in half precision and U h
 * accumulated in float).  RW: the weight is real (a spatial hop without a boundary phase), so w r enters both
 * accumulators by fused multiply-adds instead of a complex product and two additions. */
template<int A, int B, int CU, bool FWD, bool RW, bool ASM>
__device__ __forceinline__ void half_hop_r(Cx<float> (&acc)[12], const h2 (&psi)[12], const Su3h& u, Cx<float> w)
{
    h2 f1[3], f2[3];
    Cx<float> r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) proj_h<CU, FWD>(psi[3 * A + c], psi[3 * B + c], f1[c], f2[c]);
    su3_dot_h<FWD, ASM>(u, f1, f2, r);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if constexpr (RW) {  // acc_A += w r, acc_B += conj(C) w r as fused multiply-adds
            constexpr int CB = UnitConj<CU>::v;
            Cx<float>& x     = acc[3 * A + c];
            Cx<float>& y     = acc[3 * B + c];
            const float wr   = w.re;
            x = Cx<float>{__builtin_fmaf(r[c].re, wr, x.re), __builtin_fmaf(r[c].im, wr, x.im)};
            if constexpr (CB == P1) y = Cx<float>{__builtin_fmaf(r[c].re, wr, y.re), __builtin_fmaf(r[c].im, wr, y.im)};
            if constexpr (CB == M1) y = Cx<float>{__builtin_fmaf(r[c].re, -wr, y.re), __builtin_fmaf(r[c].im, -wr, y.im)};
            if constexpr (CB == PI) y = Cx<float>{__builtin_fmaf(r[c].im, -wr, y.re), __builtin_fmaf(r[c].re, wr, y.im)};
            if constexpr (CB == MI) y = Cx<float>{__builtin_fmaf(r[c].im, wr, y.re), __builtin_fmaf(r[c].re, -wr, y.im)};
        } else {
            const Cx<float> t = cmul(r[c], w);
            acc[3 * A + c]    = acc[3 * A + c] + t;
            acc[3 * B + c]    = add_unit<UnitConj<CU>::v>(acc[3 * B + c], t);
        }
    }
}

template<int DIR, bool FWD, bool RW = false, bool ASM = false>
__device__ __forceinline__ void apply_hop_h(Cx<float> (&acc)[12], const h2 (&psi)[12], const Su3h& u, Cx<float> w)
{
    if constexpr (DIR == 0) {
        half_hop_r<0, 2, FWD ? P1 : M1, FWD, RW, ASM>(acc, psi, u, w);
        half_hop_r<1, 3, FWD ? P1 : M1, FWD, RW, ASM>(acc, psi, u, w);
    } else if constexpr (DIR == 1) {
        half_hop_r<0, 3, FWD ? PI : MI, FWD, RW, ASM>(acc, psi, u, w);
        half_hop_r<1, 2, FWD ? PI : MI, FWD, RW, ASM>(acc, psi, u, w);
    } else if constexpr (DIR == 2) {
        half_hop_r<0, 3, FWD ? P1 : M1, FWD, RW, ASM>(acc, psi, u, w);
        half_hop_r<1, 2, FWD ? M1 : P1, FWD, RW, ASM>(acc, psi, u, w);
    } else {
        half_hop_r<0, 2, FWD ? PI : MI, FWD, RW, ASM>(acc, psi, u, w);
        half_hop_r<1, 3, FWD ? MI : PI, FWD, RW, ASM>(acc, psi, u, w);
    }
}

/* hops of directions D0 .. D0 + ND - 1 (forward then backward each), with the links and neighbour sites of
 * H of them loaded before they are applied (H = 1, 2 or 4, ND * 2 a multiple of H) */
template<int D0, int ND, int H, bool RS = false>
__device__ __forceinline__ void hops_h(Cx<float> (&acc)[12], const void* in, uint32_t S, const SiteNb& nb,
                                       const void* gauge, int P, uint32_t i, const HopWeights<double>& w)
{
    if constexpr (ND > 0) {
        constexpr int G = H < 2 ? 1 : H / 2;  // directions per group
        Su3h u[2 * G];
        h2 psi[2 * G][12];
        if constexpr (H == 1) {
            load_link_h(gauge, S, D0, P, i, u[0]);
            float n = load_site_h_raw(in, S, nb.up[D0], psi[0]);
            apply_hop_h<D0, true, RS && D0 != 0>(acc, psi[0], u[0], weight_n(w.wf[D0], n));
            load_link_h(gauge, S, D0, P ^ 1, nb.dn[D0], u[0]);
            n = load_site_h_raw(in, S, nb.dn[D0], psi[0]);
            apply_hop_h<D0, false, RS && D0 != 0>(acc, psi[0], u[0], weight_n(w.wb[D0], n));
            hops_h<D0 + 1, ND - 1, H, RS>(acc, in, S, nb, gauge, P, i, w);
        } else {
            float n[2 * G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                load_link_h(gauge, S, D0 + g, P, i, u[2 * g]);
                load_link_h(gauge, S, D0 + g, P ^ 1, nb.dn[D0 + g], u[2 * g + 1]);
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                n[2 * g]     = load_site_h_raw(in, S, nb.up[D0 + g], psi[2 * g]);
                n[2 * g + 1] = load_site_h_raw(in, S, nb.dn[D0 + g], psi[2 * g + 1]);
            }
            apply_hop_h<D0, true, RS && D0 != 0>(acc, psi[0], u[0], weight_n(w.wf[D0], n[0]));
            apply_hop_h<D0, false, RS && D0 != 0>(acc, psi[1], u[1], weight_n(w.wb[D0], n[1]));
            if constexpr (G == 2) {
                apply_hop_h<D0 + 1, true, RS && D0 + 1 != 0>(acc, psi[2], u[2], weight_n(w.wf[D0 + 1], n[2]));
                apply_hop_h<D0 + 1, false, RS && D0 + 1 != 0>(acc, psi[3], u[3], weight_n(w.wb[D0 + 1], n[3]));
            }
            hops_h<D0 + G, ND - G, H, RS>(acc, in, S, nb, gauge, P, i, w);
        }
    }
}

/* H hops in flight per thread; the register budget grows with H (H = 4: up to 2 waves per SIMD, H = 2: 4,
 * H = 1: the compiler's choice) */
template<int BLOCK, int H, bool RS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(1, H == 4 ? 2 : (H == 2 ? 4 : 8))))
void k_dslash_eo_h(DslashArgs a)
{
    __shared__ double lds[2 * BLOCK / 64];
    const uint32_t S = a.g.veo;
    uint32_t i       = 0;
    double nrm = 0., ry = 0.;
    if (map_site<BLOCK>(a.map, S, blockIdx.x, threadIdx.x, i)) {
        SiteNb nb;
        site_neighbours(a.g, i, a.out_parity, nb);
        Cx<float> acc[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[k] = Cx<float>{0.f, 0.f};
        // order T, X, Y, Z, forward then backward, as fermionmatrix_eo_dslash.cl:38-45
        hops_h<0, 4, H, RS>(acc, a.in, S, nb, a.gauge, a.out_parity, i, a.w);

        const Cx<float> f1{float(a.f1[0]), float(a.f1[1])}, f2{float(a.f2[0]), float(a.f2[1])};
        if (a.epilogue == EPI_MINV) {
            const float s = float(a.scale);
#pragma unroll
            for (int k = 0; k < 12; ++k) acc[k] = scale(cmul(acc[k], k < 6 ? f1 : f2), s);
        } else if (a.epilogue == EPI_AEE) {
            Cx<float> xv[12];
            load_site_h(a.x, S, i, xv);
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                acc[k] = cmul(xv[k], k < 6 ? f1 : f2) - acc[k];
                if (a.gamma5 && k >= 6) acc[k] = Cx<float>{-acc[k].re, -acc[k].im};
            }
        }
        if (a.cg_r) {  // CG residual update in place of storing the result (see the FP64/FP32 kernel)
            Cx<float> ro[12], rn[12];
            load_site_h<true>(a.cg_r, S, i, ro);
            const Cx<float> al{float(a.cg_alpha[0]), float(a.cg_alpha[1])};
#pragma unroll
            for (int k = 0; k < 12; ++k) rn[k] = ro[k] - cmul(acc[k], al);
            store_site_h<true>(a.cg_r, S, i, rn);
            ry  = (double)site_redot(rn, ro);
            nrm = (double)site_redot(rn, rn);
        } else {
            store_site_h<true>(a.out, S, i, acc);
            if (a.norm_partials) nrm = (double)site_redot(acc, acc);
        }
    }
    if (a.norm_partials) {
        const double2 s = block_sum2<BLOCK>(make_double2(nrm, ry), lds);
        if (threadIdx.x == 0) a.norm_partials[2 * blockIdx.x] = s.x, a.norm_partials[2 * blockIdx.x + 1] = s.y;
    }
}

/* whether the X, Y, Z hop weights have no imaginary part (no spatial boundary phase): the kernels then
 * multiply by them as reals */
static bool spatial_weights_real(const HopWeights<double>& w)
{
    for (int mu = 1; mu < 4; ++mu)
        if (w.wf[mu][1] != 0. || w.wb[mu][1] != 0.) return false;
    return true;
}

/* launch shape: block size a.block (128, 256, 512), hops in flight by the register budget a.occ (4: one
 * hop at a time, 0: two, 1: four) */
hipError_t launch_dslash_h(cl2_ctx* c, const DslashArgs& a)
{
    const int bs      = (a.block == 128 || a.block == 512) ? a.block : 256;
    const unsigned nb = blocks_for(a.g.veo, bs);
    const bool rs = spatial_weights_real(a.w);
    auto go = [&](auto block_tag) {
        constexpr int B = decltype(block_tag)::value;
        auto launch = [&](auto kernel) { hipLaunchKernelGGL(kernel, dim3(nb), dim3(B), 0, c->stream, a); };
        if (a.occ == 4)
            rs ? launch(k_dslash_eo_h<B, 1, true>) : launch(k_dslash_eo_h<B, 1, false>);
        else if (a.occ == 1 && B != 512)
            rs ? launch(k_dslash_eo_h<B, 4, true>) : launch(k_dslash_eo_h<B, 4, false>);
        else
            rs ? launch(k_dslash_eo_h<B, 2, true>) : launch(k_dslash_eo_h<B, 2, false>);
    };
    if (bs == 128)
        go(std::integral_constant<int, 128>{});
    else if (bs == 512)
        go(std::integral_constant<int, 512>{});
    else
        go(std::integral_constant<int, 256>{});
    return hipGetLastError();
}

/* ---------------- per-site BLAS on FP16 spinors ---------------- */

__device__ __forceinline__ Cx<float> resolve_f(const cl2_scalar& s)
{
    switch (s.kind) {
        case CL2_SCALAR_DEV_COMPLEX: return Cx<float>{float(s.dev[0]), float(s.dev[1])};
        case CL2_SCALAR_DEV_REAL: return Cx<float>{float(s.dev[0]), 0.f};
        case CL2_SCALAR_DEV_VECTOR: return Cx<float>{float(s.dev[s.index]), 0.f};
        default: return Cx<float>{float(s.re), float(s.im)};
    }
}

/* the Wilson modes of k_blas, one thread per site (a.n is the site count); the grid covers the field up to
 * BLAS_MAX_BLOCKS blocks and loops beyond */
template<int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_blas_h(BlasArgs a)
{
    __shared__ double lds[2 * BLOCK / 64];
    const Cx<float> al = resolve_f(a.a), be = resolve_f(a.b);
    const uint32_t S   = (uint32_t)a.n;
    double nrm = 0., ry = 0.;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < S; i += gridDim.x * BLOCK) {
        Cx<float> x[12], r[12];
        load_site_h<true>(a.x, S, i, x);
        if (a.mode == B_AX) {
#pragma unroll
            for (int k = 0; k < 12; ++k) r[k] = cmul(x[k], al);
        } else if (a.mode == B_NEG_TAIL) {
#pragma unroll
            for (int k = 0; k < 12; ++k) r[k] = k < 6 ? x[k] : Cx<float>{-x[k].re, -x[k].im};
        } else {
            Cx<float> y[12];
            load_site_h<true>(a.y, S, i, y);
            if (a.mode == B_AX_BY_Z || a.mode == B_CG_XP) {
                Cx<float> z[12];
                load_site_h<true>(a.z, S, i, z);
                if (a.mode == B_CG_XP) {
                    Cx<float> r2[12];
#pragma unroll
                    for (int k = 0; k < 12; ++k) {
                        r[k]  = z[k] - cmul(x[k], al);
                        r2[k] = y[k] - cmul(x[k], be);
                    }
                    store_site_h<true>(a.out2, S, i, r2);
                } else {
#pragma unroll
                    for (int k = 0; k < 12; ++k) r[k] = cmul(y[k], be) + cmul(x[k], al) + z[k];
                }
            } else {  // B_Y_MINUS_AX, B_G5_Y_MINUS_AX
                const bool g5 = a.mode == B_G5_Y_MINUS_AX;
#pragma unroll
                for (int k = 0; k < 12; ++k) {
                    r[k] = y[k] - cmul(x[k], al);
                    if (g5 && k >= 6) r[k] = Cx<float>{-r[k].re, -r[k].im};
                    ry += (double)r[k].re * y[k].re + (double)r[k].im * y[k].im;
                }
            }
        }
        store_site_h<true>(a.out, S, i, r);
#pragma unroll
        for (int k = 0; k < 12; ++k) nrm += (double)r[k].re * r[k].re + (double)r[k].im * r[k].im;
    }
    if (a.norm_partials) {
        double2 s = block_sum2<BLOCK>(make_double2(nrm, a.cg_norm ? ry : 0.), lds);
        if (threadIdx.x == 0) a.norm_partials[2 * blockIdx.x] = s.x, a.norm_partials[2 * blockIdx.x + 1] = s.y;
    }
}

static unsigned site_blocks(uint32_t S)
{
    const unsigned nb = blocks_for(S, BLAS_BLOCK);
    return nb > BLAS_MAX_BLOCKS ? BLAS_MAX_BLOCKS : (nb == 0 ? 1u : nb);
}

hipError_t launch_blas_h(cl2_ctx* c, const BlasArgs& a, unsigned* nblocks_out)
{
    const unsigned nb = site_blocks((uint32_t)a.n);
    if (nblocks_out) *nblocks_out = nb;
    hipLaunchKernelGGL((k_blas_h<BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream, a);
    return hipGetLastError();
}

template<int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_reduce1_h(const void* x, const void* y, uint32_t S, int mode,
                                                     double* __restrict__ partials)
{
    __shared__ double lds[2 * BLOCK / 64];
    double2 acc = make_double2(0., 0.);
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < S; i += gridDim.x * BLOCK) {
        Cx<float> u[12];
        load_site_h<true>(x, S, i, u);
        if (mode == R_NORM) {
#pragma unroll
            for (int k = 0; k < 12; ++k) acc.x += (double)u[k].re * u[k].re + (double)u[k].im * u[k].im;
        } else {
            Cx<float> v[12];
            load_site_h<true>(y, S, i, v);
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                acc.x += (double)u[k].re * v[k].re + (double)u[k].im * v[k].im;
                acc.y += (double)u[k].re * v[k].im - (double)v[k].re * u[k].im;
            }
        }
    }
    double2 s = block_sum2<BLOCK>(acc, lds);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x]     = s.x;
        partials[2 * blockIdx.x + 1] = (mode == R_DOT) ? s.y : 0.;
    }
}

hipError_t launch_reduce_stage1_h(cl2_ctx* c, const void* x, const void* y, int mode, unsigned* nblocks)
{
    const unsigned nb = site_blocks(c->g.veo);
    *nblocks          = nb;
    hipLaunchKernelGGL((k_reduce1_h<BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream, x, y, c->g.veo, mode,
                       c->partials);
    return hipGetLastError();
}

/* site-diagonal twisted-mass factors: spins 0,1 times f1, spins 2,3 times f2, times s */
__global__ __launch_bounds__(BLAS_BLOCK) void k_sitediag_h(void* out, const void* in, uint32_t S, Cx<float> f1,
                                                           Cx<float> f2, float s)
{
    const uint32_t i = blockIdx.x * BLAS_BLOCK + threadIdx.x;
    if (i >= S) return;
    Cx<float> v[12];
    load_site_h(in, S, i, v);
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = scale(cmul(v[k], k < 6 ? f1 : f2), s);
    store_site_h(out, S, i, v);
}

hipError_t launch_sitediag_h(cl2_ctx* c, void* out, const void* in, const double f1[2], const double f2[2], double s)
{
    hipLaunchKernelGGL(k_sitediag_h, dim3(blocks_for(c->g.veo, BLAS_BLOCK)), dim3(BLAS_BLOCK), 0, c->stream, out, in,
                       c->g.veo, Cx<float>{float(f1[0]), float(f1[1])}, Cx<float>{float(f2[0]), float(f2[1])}, float(s));
    return hipGetLastError();
}

__global__ __launch_bounds__(BLAS_BLOCK) void k_fill_h(void* out, uint32_t S, float re, float im)
{
    const uint32_t i = blockIdx.x * BLAS_BLOCK + threadIdx.x;
    if (i >= S) return;
    Cx<float> v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = Cx<float>{re, im};
    store_site_h(out, S, i, v);
}

hipError_t launch_fill_h(cl2_ctx* c, void* out, double re, double im, uint32_t sites)
{
    const uint32_t S = sites ? sites : c->g.veo;
    hipLaunchKernelGGL(k_fill_h, dim3(blocks_for(S, BLAS_BLOCK)), dim3(BLAS_BLOCK), 0, c->stream, out, S, float(re),
                       float(im));
    return hipGetLastError();
}

/* FP64 / FP32 spinor (the layouts of load_spinor) <-> FP16 with per-site norm; with S = Veo nrhs the same
 * kernels convert blocks, whose site offset i nrhs + r is the same in every precision */
template<typename R>
__global__ __launch_bounds__(BLAS_BLOCK) void k_to_half(void* out, const vec2<R>* __restrict__ in, uint32_t S)
{
    const uint32_t i = blockIdx.x * BLAS_BLOCK + threadIdx.x;
    if (i >= S) return;
    vec2<R> q[12];
    load_spinor<R>(q, in, S, i);
    Cx<float> v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = Cx<float>{float(q[k].x), float(q[k].y)};
    store_site_h(out, S, i, v);
}
template<typename R>
__global__ __launch_bounds__(BLAS_BLOCK) void k_from_half(vec2<R>* __restrict__ out, const void* in, uint32_t S)
{
    const uint32_t i = blockIdx.x * BLAS_BLOCK + threadIdx.x;
    if (i >= S) return;
    Cx<float> v[12];
    load_site_h(in, S, i, v);
    vec2<R> q[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) q[k] = vec2<R>{R(v[k].re), R(v[k].im)};
    store_spinor<R>(out, S, i, q);
}

hipError_t launch_convert_h(cl2_ctx* c, void* out, int out_prec, const void* in, int in_prec, uint32_t sites)
{
    const uint32_t S  = sites ? sites : c->g.veo;
    const unsigned nb = blocks_for(S, BLAS_BLOCK);
    if (out_prec == CL2_F16 && in_prec == CL2_F64)
        hipLaunchKernelGGL(k_to_half<double>, dim3(nb), dim3(BLAS_BLOCK), 0, c->stream, out,
                           reinterpret_cast<const double2*>(in), S);
    else if (out_prec == CL2_F16 && in_prec == CL2_F32)
        hipLaunchKernelGGL(k_to_half<float>, dim3(nb), dim3(BLAS_BLOCK), 0, c->stream, out,
                           reinterpret_cast<const float2*>(in), S);
    else if (out_prec == CL2_F64 && in_prec == CL2_F16)
        hipLaunchKernelGGL(k_from_half<double>, dim3(nb), dim3(BLAS_BLOCK), 0, c->stream,
                           reinterpret_cast<double2*>(out), in, S);
    else if (out_prec == CL2_F32 && in_prec == CL2_F16)
        hipLaunchKernelGGL(k_from_half<float>, dim3(nb), dim3(BLAS_BLOCK), 0, c->stream,
                           reinterpret_cast<float2*>(out), in, S);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}


/* ---------------- blocks of FP16 spinors ----------------
 * A block of nrhs FP16 spinors is one FP16 field of Veo nrhs sites: right-hand side r of eo site i is site
 * offset i nrhs + r (three 16-byte lanes [3][Veo nrhs], then float norm[Veo nrhs]), so the nrhs lanes of a
 * site are adjacent as in the FP64 and FP32 blocks, and every single-field routine that takes the site
 * count applies to a block unchanged. */

/* Multi-RHS hopping kernel on FP16 blocks, the staging of k_dslash_eo_mrhs_lds: a block owns SB consecutive
 * sites of the work order and NRHS threads per site; the sites' neighbour indices and their 8 links (row 2
 * reconstructed, as half pairs) are put in LDS once, then every right-hand-side thread gathers its own FP16
 * neighbour sites.  Per right-hand side the arithmetic is k_dslash_eo_h's (same hop order, norms folded into
 * the hop weights), so slot r equals the single-field kernel bit for bit. */
template<int NRHS>
struct MrhsH {
    static constexpr int SB    = NRHS >= 8 ? 16 : 32;
    static constexpr int BLOCK = SB * NRHS;
};

template<int NRHS, bool RS>
__global__ __launch_bounds__(MrhsH<NRHS>::BLOCK) void k_dslash_eo_mrhs_h(DslashArgs a)
{
    constexpr int SB        = MrhsH<NRHS>::SB;
    constexpr int BLOCK     = MrhsH<NRHS>::BLOCK;
    constexpr uint32_t NONE = 0xffffffffu;
    __shared__ Su3h lu[SB * 8];
    __shared__ uint32_t lnb[SB * 8];
    __shared__ uint32_t lsite[SB];
    __shared__ double2 lred[BLOCK];
    const uint32_t S  = a.g.veo;
    const uint32_t SN = S * NRHS;
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
            Su3h u;
            // forward link: the site's own; backward link: the neighbour's, of the other parity
            load_link_h(a.gauge, S, h >> 1, (h & 1) ? P ^ 1 : P, (h & 1) ? lnb[w] : i, u);
            lu[w] = u;
        }
    }
    __syncthreads();
    const int s      = threadIdx.x / NRHS;
    const int rhs    = threadIdx.x - s * NRHS;
    const uint32_t i = lsite[s];
    double nrm = 0., ry = 0.;
    if (i != NONE) {
        Cx<float> acc[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[k] = Cx<float>{0.f, 0.f};
        // order T, X, Y, Z, forward then backward, as fermionmatrix_eo_dslash.cl:38-45
#define CL2_DIR_H(m)                                                                                   \
    {                                                                                                  \
        h2 pf[12], pb[12];                                                                             \
        const float nf = load_site_h_raw(a.in, SN, lnb[s * 8 + 2 * (m)] * NRHS + rhs, pf);             \
        const float nk = load_site_h_raw(a.in, SN, lnb[s * 8 + 2 * (m) + 1] * NRHS + rhs, pb);         \
        const Su3h uf = lu[s * 8 + 2 * (m)], ub = lu[s * 8 + 2 * (m) + 1];                             \
        apply_hop_h<m, true, RS && (m) != 0, true>(acc, pf, uf, weight_n(a.w.wf[m], nf));            \
        apply_hop_h<m, false, RS && (m) != 0, true>(acc, pb, ub, weight_n(a.w.wb[m], nk));           \
    }
        CL2_DIR_H(0)
        CL2_DIR_H(1)
        CL2_DIR_H(2)
        CL2_DIR_H(3)
#undef CL2_DIR_H
        const uint32_t o = i * NRHS + rhs;
        Cx<float> f1{float(a.f1[0]), float(a.f1[1])}, f2{float(a.f2[0]), float(a.f2[1])};
        float sc = float(a.scale);
        if (a.per_rhs) {  // an operator per right-hand side: the hop ran with unit weights
            const float kr = float(a.rhs_kappa[rhs]);
#pragma unroll
            for (int k = 0; k < 12; ++k) acc[k] = scale(acc[k], kr);
            f1 = Cx<float>{float(a.rhs_f1[rhs][0]), float(a.rhs_f1[rhs][1])};
            f2 = Cx<float>{float(a.rhs_f2[rhs][0]), float(a.rhs_f2[rhs][1])};
            sc = float(a.rhs_scale[rhs]);
        }
        if (a.epilogue == EPI_MINV) {
#pragma unroll
            for (int k = 0; k < 12; ++k) acc[k] = scale(cmul(acc[k], k < 6 ? f1 : f2), sc);
        } else if (a.epilogue == EPI_AEE) {
            Cx<float> xv[12];
            load_site_h(a.x, SN, o, xv);
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                acc[k] = cmul(xv[k], k < 6 ? f1 : f2) - acc[k];
                if (a.gamma5 && k >= 6) acc[k] = Cx<float>{-acc[k].re, -acc[k].im};
            }
        }
        if (a.cg_r) {  // CG residual update in place of storing the result
            Cx<float> ro[12], rn[12];
            load_site_h<true>(a.cg_r, SN, o, ro);
            const Cx<float> al{float(a.cg_alpha[2 * rhs]), float(a.cg_alpha[2 * rhs + 1])};
#pragma unroll
            for (int k = 0; k < 12; ++k) rn[k] = ro[k] - cmul(acc[k], al);
            store_site_h<true>(a.cg_r, SN, o, rn);
            ry  = (double)site_redot(rn, ro);
            nrm = (double)site_redot(rn, rn);
        } else {
            store_site_h<true>(a.out, SN, o, acc);
            if (a.norm_partials) nrm = (double)site_redot(acc, acc);
        }
    }
    if (a.norm_partials) {
        // per right-hand side: thread r < NRHS sums the SB site threads of its column in site order
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

hipError_t launch_dslash_mrhs_h(cl2_ctx* c, const DslashArgs& a)
{
    const bool rs = spatial_weights_real(a.w);
    auto go = [&](auto n_tag) -> hipError_t {
        constexpr int N = decltype(n_tag)::value;
        const dim3 grid((a.g.veo + MrhsH<N>::SB - 1) / MrhsH<N>::SB), block(MrhsH<N>::BLOCK);
        if (rs)
            hipLaunchKernelGGL((k_dslash_eo_mrhs_h<N, true>), grid, block, 0, c->stream, a);
        else
            hipLaunchKernelGGL((k_dslash_eo_mrhs_h<N, false>), grid, block, 0, c->stream, a);
        return hipGetLastError();
    };
    switch (a.nrhs) {
        case 2: return go(std::integral_constant<int, 2>{});
        case 4: return go(std::integral_constant<int, 4>{});
        case 8: return go(std::integral_constant<int, 8>{});
        case 12: return go(std::integral_constant<int, 12>{});
        case 16: return go(std::integral_constant<int, 16>{});
        default: return hipErrorInvalidValue;
    }
}

/* per-right-hand-side sums of a block's thread values, in threads 0..nrhs-1: thread t's column is right-hand
 * side t mod nrhs (block size and grid stride are multiples of nrhs), summed in thread order */
__device__ __forceinline__ double2 rhs_column_sum(double2 v, double2* lds, int nrhs)
{
    lds[threadIdx.x] = v;
    __syncthreads();
    double2 s = make_double2(0., 0.);
    if ((int)threadIdx.x < nrhs)
        for (unsigned k = threadIdx.x; k < blockDim.x; k += (unsigned)nrhs) {
            s.x += lds[k].x;
            s.y += lds[k].y;
        }
    __syncthreads();
    return s;
}

/* out_r = y_r + sign c_r x_r; with partials: per right-hand side (|out_r|^2, Re <out_r, y_r>) into
 * partials[r * nb + block].  out may alias y (each site is read before it is written). */
__global__ __launch_bounds__(256) void k_block_axpy_h(void* out, const void* x, const void* y, uint32_t SN, int nrhs,
                                                      const double* coef, double sign, double2* partials)
{
    __shared__ double2 lds[256];
    const int r = (int)(threadIdx.x % (unsigned)nrhs);
    const Cx<float> cr{float(sign * coef[2 * r]), float(sign * coef[2 * r + 1])};
    double2 acc = make_double2(0., 0.);
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < SN; e += gridDim.x * blockDim.x) {
        Cx<float> xv[12], yv[12], o[12];
        load_site_h<true>(x, SN, e, xv);
        load_site_h<true>(y, SN, e, yv);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            o[k] = yv[k] + cmul(cr, xv[k]);
            acc.x += (double)o[k].re * o[k].re + (double)o[k].im * o[k].im;
            acc.y += (double)o[k].re * yv[k].re + (double)o[k].im * yv[k].im;
        }
        store_site_h<true>(out, SN, e, o);
    }
    if (partials) {
        const double2 s = rhs_column_sum(acc, lds, nrhs);
        if ((int)threadIdx.x < nrhs) partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

/* the x and p updates of a lockstep CG step: x_r = x_r + alpha_r p_r, then p_r = r_r + beta_r p_r; U (site,
 * right-hand side) elements per thread, a block size apart, with all their loads issued first (two per thread
 * measured 7 % faster than one) */
template<int U>
__global__ __launch_bounds__(256) void k_block_xp_h(void* xo, void* p, const void* rv, uint32_t SN, int nrhs,
                                                    const double* alpha, const double* beta)
{
    const int r = (int)(threadIdx.x % (unsigned)nrhs);  // the block size is a multiple of nrhs
    const Cx<float> ca{float(alpha[2 * r]), float(alpha[2 * r + 1])}, cb{float(beta[2 * r]), float(beta[2 * r + 1])};
    const uint32_t base = blockIdx.x * blockDim.x * U + threadIdx.x;
    Cx<float> xv[U][12], pv[U][12], rr[U][12];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t e = base + u * blockDim.x;
        if (e < SN) {
            load_site_h<true>(xo, SN, e, xv[u]);
            load_site_h<true>(p, SN, e, pv[u]);
            load_site_h<true>(rv, SN, e, rr[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t e = base + u * blockDim.x;
        if (e >= SN) continue;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            xv[u][k] = xv[u][k] + cmul(ca, pv[u][k]);
            pv[u][k] = rr[u][k] + cmul(cb, pv[u][k]);
        }
        store_site_h<true>(xo, SN, e, xv[u]);
        store_site_h<true>(p, SN, e, pv[u]);
    }
}

/* per right-hand side: sum conj(x) y (mode 0) or sum |x|^2 (mode 1) */
__global__ __launch_bounds__(256) void k_block_dot_h(const void* x, const void* y, uint32_t SN, int nrhs, int mode,
                                                     double2* partials)
{
    __shared__ double2 lds[256];
    double2 acc = make_double2(0., 0.);
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < SN; e += gridDim.x * blockDim.x) {
        Cx<float> u[12];
        load_site_h<true>(x, SN, e, u);
        if (mode == 1) {
#pragma unroll
            for (int k = 0; k < 12; ++k) acc.x += (double)u[k].re * u[k].re + (double)u[k].im * u[k].im;
        } else {
            Cx<float> v[12];
            load_site_h<true>(y, SN, e, v);
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                acc.x += (double)u[k].re * v[k].re + (double)u[k].im * v[k].im;
                acc.y += (double)u[k].re * v[k].im - (double)v[k].re * u[k].im;
            }
        }
    }
    const double2 s = rhs_column_sum(acc, lds, nrhs);
    if ((int)threadIdx.x < nrhs) partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
}

/* single FP16 field <-> slot r of an FP16 block: the stored halves and norm are copied as they are */
__global__ __launch_bounds__(256) void k_block_slot_h(void* block, const void* single, void* single_out, uint32_t S,
                                                      int nrhs, int r)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= S) return;
    const uint32_t SN = S * (uint32_t)nrhs, e = i * (uint32_t)nrhs + (uint32_t)r;
    h8* bl            = reinterpret_cast<h8*>(block);
    if (single) {
        const h8* sl = reinterpret_cast<const h8*>(single);
#pragma unroll
        for (int j = 0; j < 3; ++j) bl[(size_t)j * SN + e] = sl[(size_t)j * S + i];
        half_norms(block, SN)[e] = half_norms(single, S)[i];
    } else {
        h8* sl = reinterpret_cast<h8*>(single_out);
#pragma unroll
        for (int j = 0; j < 3; ++j) sl[(size_t)j * S + i] = bl[(size_t)j * SN + e];
        half_norms(single_out, S)[i] = half_norms(static_cast<const void*>(block), SN)[e];
    }
}

/* grid of an FP16 block kernel: blocks of the largest multiple of nrhs not above 256 threads, at most what the
 * per-right-hand-side partials buffer holds (the grid loops beyond) */
static void block_shape_h(cl2_ctx* c, uint32_t SN, int nrhs, unsigned& bt, unsigned& nb)
{
    bt               = 256u - (256u % (unsigned)nrhs);
    size_t n         = (SN + bt - 1) / bt;
    const size_t cap = c->npartials / (size_t)nrhs;
    nb               = (unsigned)std::max<size_t>(1, std::min(n, std::min(cap, (size_t)BLAS_MAX_BLOCKS)));
}

hipError_t launch_block_axpy_h(cl2_ctx* c, void* out, const void* x, const void* y, int nrhs, const double* coef,
                               double sign, double* norms)
{
    const uint32_t SN = c->g.veo * (uint32_t)nrhs;
    unsigned bt, nb;
    block_shape_h(c, SN, nrhs, bt, nb);
    double2* parts = norms ? reinterpret_cast<double2*>(c->partials) : nullptr;
    hipLaunchKernelGGL(k_block_axpy_h, dim3(nb), dim3(bt), 0, c->stream, out, x, y, SN, nrhs, coef, sign, parts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !norms) return e;
    return launch_block_reduce2(c, parts, nb, norms, 0, nrhs);
}

hipError_t launch_block_xp_h(cl2_ctx* c, void* x, void* p, const void* r, int nrhs, const double* alpha,
                             const double* beta)
{
    // no partials: two (site, right-hand side) elements per thread over the full grid
    const uint32_t SN = c->g.veo * (uint32_t)nrhs;
    const unsigned bt = 256u - (256u % (unsigned)nrhs), nb = (SN + 2 * bt - 1) / (2 * bt);
    hipLaunchKernelGGL(k_block_xp_h<2>, dim3(nb), dim3(bt), 0, c->stream, x, p, r, SN, nrhs, alpha, beta);
    return hipGetLastError();
}

hipError_t launch_block_dot_h(cl2_ctx* c, double* res, const void* x, const void* y, int nrhs, int mode)
{
    const uint32_t SN = c->g.veo * (uint32_t)nrhs;
    unsigned bt, nb;
    block_shape_h(c, SN, nrhs, bt, nb);
    double2* parts = reinterpret_cast<double2*>(c->partials);
    hipLaunchKernelGGL(k_block_dot_h, dim3(nb), dim3(bt), 0, c->stream, x, y, SN, nrhs, mode, parts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_block_reduce2(c, parts, nb, res, mode == 1 ? 1 : 0, nrhs);
}

hipError_t launch_block_slot_h(cl2_ctx* c, void* block, const void* single_in, void* single_out, int nrhs, int r)
{
    hipLaunchKernelGGL(k_block_slot_h, dim3(blocks_for(c->g.veo, 256)), dim3(256), 0, c->stream, block, single_in,
                       single_out, c->g.veo, nrhs, r);
    return hipGetLastError();
}

}  // namespace cl2
