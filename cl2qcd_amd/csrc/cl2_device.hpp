// Device-side building blocks for the gfx950 kernels: complex and SU(3)
// arithmetic, eo lattice geometry, gauge-link formats and wavefront/LDS
// reductions.  All kernels are bandwidth bound; nothing here uses MFMA.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cl2_types.hpp"

namespace cl2 {

template<typename R>
struct V2;
template<>
struct V2<double> {
    using T = double2;
};
template<>
struct V2<float> {
    using T = float2;
};
template<typename R>
using vec2 = typename V2<R>::T;

template<typename R>
struct Cx {
    R re, im;
};

template<typename R>
__device__ __forceinline__ Cx<R> cx(R a, R b)
{
    return Cx<R>{a, b};
}
/* The arithmetic helpers below fix their own rounding: no implicit contraction (the pragma) and fused
 * multiply-adds only where written, so every kernel built from them computes the same operations in the
 * same order whatever the surrounding code -- results do not depend on launch shape or kernel variant. */
template<typename R>
__device__ __forceinline__ Cx<R> operator+(Cx<R> a, Cx<R> b)
{
#pragma clang fp contract(off)
    return Cx<R>{a.re + b.re, a.im + b.im};
}
template<typename R>
__device__ __forceinline__ Cx<R> operator-(Cx<R> a, Cx<R> b)
{
#pragma clang fp contract(off)
    return Cx<R>{a.re - b.re, a.im - b.im};
}
template<typename R>
__device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> b)
{
#pragma clang fp contract(off)
    return Cx<R>{fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}
template<typename R>
__device__ __forceinline__ Cx<R> cmulconj(Cx<R> a, Cx<R> b)
{  // a * conj(b)
#pragma clang fp contract(off)
    return Cx<R>{fma(a.re, b.re, a.im * b.im), fma(a.im, b.re, -(a.re * b.im))};
}
template<typename R>
__device__ __forceinline__ Cx<R> conj(Cx<R> a)
{
    return Cx<R>{a.re, -a.im};
}
template<typename R>
__device__ __forceinline__ Cx<R> scale(Cx<R> a, R s)
{
#pragma clang fp contract(off)
    return Cx<R>{a.re * s, a.im * s};
}

/* Multiplication by one of {+1, -1, +i, -i}: the only factors the Wilson spin
 * projectors (1 -/+ gamma_mu) of the chiral basis need. */
enum Unit : int { P1 = 0, M1 = 1, PI = 2, MI = 3 };
template<int U>
struct UnitConj;
template<>
struct UnitConj<P1> {
    static constexpr int v = P1;
};
template<>
struct UnitConj<M1> {
    static constexpr int v = M1;
};
template<>
struct UnitConj<PI> {
    static constexpr int v = MI;
};
template<>
struct UnitConj<MI> {
    static constexpr int v = PI;
};
template<int U, typename R>
__device__ __forceinline__ Cx<R> add_unit(Cx<R> a, Cx<R> b)
{  // a + U*b
#pragma clang fp contract(off)
    if constexpr (U == P1) return Cx<R>{a.re + b.re, a.im + b.im};
    if constexpr (U == M1) return Cx<R>{a.re - b.re, a.im - b.im};
    if constexpr (U == PI) return Cx<R>{a.re - b.im, a.im + b.re};
    return Cx<R>{a.re + b.im, a.im - b.re};
}

template<typename R>
struct Su3 {
    Cx<R> e[3][3];
};

/* r = U v and r = U^dagger v, each row an FMA chain over the three columns */
template<typename R>
__device__ __forceinline__ void su3_times(const Su3<R>& u, const Cx<R> (&v)[3], Cx<R> (&r)[3])
{
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        R re = u.e[i][0].re * v[0].re;
        R im = u.e[i][0].re * v[0].im;
        re   = fma(-u.e[i][0].im, v[0].im, re);
        im   = fma(u.e[i][0].im, v[0].re, im);
#pragma unroll
        for (int j = 1; j < 3; ++j) {
            re = fma(u.e[i][j].re, v[j].re, re);
            im = fma(u.e[i][j].re, v[j].im, im);
            re = fma(-u.e[i][j].im, v[j].im, re);
            im = fma(u.e[i][j].im, v[j].re, im);
        }
        r[i] = Cx<R>{re, im};
    }
}
template<typename R>
__device__ __forceinline__ void su3dag_times(const Su3<R>& u, const Cx<R> (&v)[3], Cx<R> (&r)[3])
{
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        R re = u.e[0][i].re * v[0].re;
        R im = u.e[0][i].re * v[0].im;
        re   = fma(u.e[0][i].im, v[0].im, re);
        im   = fma(-u.e[0][i].im, v[0].re, im);
#pragma unroll
        for (int j = 1; j < 3; ++j) {
            re = fma(u.e[j][i].re, v[j].re, re);
            im = fma(u.e[j][i].re, v[j].im, im);
            re = fma(u.e[j][i].im, v[j].im, re);
            im = fma(-u.e[j][i].im, v[j].re, im);
        }
        r[i] = Cx<R>{re, im};
    }
}

/* Third row of an SU(3) matrix from the first two: conj(row0 x row1).  The
 * reference's 12-real path (operations_gaugefield.cl:41-46) omits the
 * conjugate; this is the correct form for det U = 1. */
template<typename R>
__device__ __forceinline__ void su3_reconstruct_row2(Su3<R>& u)
{
    u.e[2][0] = conj(cmul(u.e[0][1], u.e[1][2]) - cmul(u.e[0][2], u.e[1][1]));
    u.e[2][1] = conj(cmul(u.e[0][2], u.e[1][0]) - cmul(u.e[0][0], u.e[1][2]));
    u.e[2][2] = conj(cmul(u.e[0][0], u.e[1][1]) - cmul(u.e[0][1], u.e[1][0]));
}

/* ---------------- lattice geometry ----------------
 * Reference site order: site = x + NS*(y + NS*(z + NS*t)), eo index = site/2,
 * directions T=0,X=1,Y=2,Z=3 (ocl_kernel/operations_geometry.cl:72-300).  A
 * thread that owns eo index i of parity P sits at the site of parity P among
 * {2i, 2i+1}; both share y, z, t because NS is even. */
struct SiteNb {
    uint32_t up[4];  // eo index of x + mu
    uint32_t dn[4];  // eo index of x - mu
};

__device__ __forceinline__ void site_neighbours(const LatticeGeom& g, uint32_t i, int parity, SiteNb& nb)
{
    const uint32_t s0 = 2u * i;
    const uint32_t ns = g.ns;
    uint32_t r        = s0 / ns;
    const uint32_t x0 = s0 - r * ns;
    const uint32_t y  = r % ns;
    r /= ns;
    const uint32_t z = r % ns;
    const uint32_t t = r / ns;
    const uint32_t x = x0 + (((y + z + t) & 1u) ^ (uint32_t)parity);

    const uint32_t ns2 = ns * ns;
    const uint32_t vs  = g.vs;
    const uint32_t base = ns * y + ns2 * z + vs * t;  // site of (0,y,z,t)
    const uint32_t xu = (x + 1 == ns) ? 0u : x + 1;
    const uint32_t xd = (x == 0) ? ns - 1 : x - 1;
    nb.up[1] = (base + xu) >> 1;
    nb.dn[1] = (base + xd) >> 1;
    const uint32_t site = base + x;
    nb.up[2] = (y + 1 == ns ? site - (ns - 1) * ns : site + ns) >> 1;
    nb.dn[2] = (y == 0 ? site + (ns - 1) * ns : site - ns) >> 1;
    nb.up[3] = (z + 1 == ns ? site - (ns - 1) * ns2 : site + ns2) >> 1;
    nb.dn[3] = (z == 0 ? site + (ns - 1) * ns2 : site - ns2) >> 1;
    nb.up[0] = (t + 1 == g.nt ? site - (g.nt - 1) * vs : site + vs) >> 1;
    nb.dn[0] = (t == 0 ? site + (g.nt - 1) * vs : site - vs) >> 1;
}

/* eo index handled by thread tid of hardware block h under the WorkMap;
 * returns false for threads past the end of the lattice. */
template<int BLOCK>
__device__ __forceinline__ bool map_site(const WorkMap& m, uint32_t veo, uint32_t h, uint32_t tid, uint32_t& i)
{
    uint32_t L = h;
    if (m.xcd_q && h < 8u * m.xcd_q) L = (h & 7u) * m.xcd_q + (h >> 3);
    const uint32_t l = L * BLOCK + tid;
    if (l >= veo) return false;
    if (m.P == 0) {
        i = l;
        return true;
    }
    const uint32_t zc = l / m.col;
    const uint32_t r  = l - zc * m.col;
    const uint32_t t  = r / m.P;
    i                 = t * m.half_vs + zc * m.P + (r - t * m.P);
    return true;
}

/* a wave-uniform pointer pinned to scalar registers, so that base + 32-bit lane offset addressing is
 * selected for loads from it instead of per-lane 64-bit address arithmetic */
typedef const __attribute__((address_space(1))) char* gchar_ptr;
__device__ __forceinline__ gchar_ptr uniform_ptr(const char* p)
{
    const uint64_t v  = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (gchar_ptr)(((uint64_t)hi << 32) | lo);
}

/* the complex number at byte offset off of a uniform base */
template<typename R>
__device__ __forceinline__ vec2<R> load_at(gchar_ptr base, uint32_t off)
{
    typedef R N __attribute__((ext_vector_type(2)));
    const N v = *(const __attribute__((address_space(1))) N*)(base + off);
    return vec2<R>{v.x, v.y};
}

/* eo index of logical site l of the work order (l < veo) */
__device__ __forceinline__ uint32_t logical_site(const WorkMap& m, uint32_t l)
{
    if (m.P == 0) return l;
    const uint32_t zc = l / m.col;
    const uint32_t r  = l - zc * m.col;
    const uint32_t t  = r / m.P;
    return t * m.half_vs + zc * m.P + (r - t * m.P);
}

/* logical block of hardware block h under the XCD remap */
__device__ __forceinline__ uint32_t logical_block(const WorkMap& m, uint32_t h)
{
    return (m.xcd_q && h < 8u * m.xcd_q) ? (h & 7u) * m.xcd_q + (h >> 3) : h;
}

/* the same for NRHS right-hand sides per site: logical item l covers site l / NRHS, right-hand side
 * l % NRHS, and the slab order applies to the site part */
template<int BLOCK, int NRHS>
__device__ __forceinline__ bool map_site_rhs(const WorkMap& m, uint32_t veo, uint32_t h, uint32_t tid, uint32_t& i,
                                             uint32_t& rhs)
{
    if constexpr (NRHS == 1) {
        rhs = 0;
        return map_site<BLOCK>(m, veo, h, tid, i);
    } else {
        uint32_t L = h;
        if (m.xcd_q && h < 8u * m.xcd_q) L = (h & 7u) * m.xcd_q + (h >> 3);
        const uint32_t l = L * BLOCK + tid;
        if (l >= veo * NRHS) return false;
        const uint32_t ls = l / NRHS;
        rhs               = l - ls * NRHS;
        if (m.P == 0) {
            i = ls;
            return true;
        }
        const uint32_t zc = ls / m.col;
        const uint32_t r  = ls - zc * m.col;
        const uint32_t t  = r / m.P;
        i                 = t * m.half_vs + zc * m.P + (r - t * m.P);
        return true;
    }
}

__device__ __forceinline__ void site_coords(const LatticeGeom& g, uint32_t i, int parity, uint32_t& x, uint32_t& y,
                                            uint32_t& z, uint32_t& t)
{
    const uint32_t s0 = 2u * i;
    uint32_t r        = s0 / g.ns;
    const uint32_t x0 = s0 - r * g.ns;
    y                 = r % g.ns;
    r /= g.ns;
    z = r % g.ns;
    t = r / g.ns;
    x = x0 + (((y + z + t) & 1u) ^ (uint32_t)parity);
}

/* Streaming (non-temporal) loads for data read exactly once per launch, and streaming stores for
 * results no later work of the same launch reads: both keep the L2 for data that is reused (the
 * neighbour spinors of the hopping kernels). */
typedef double nt_d2 __attribute__((ext_vector_type(2)));
typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef float nt_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ double2 load_stream(const double2* p)
{
    nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2*>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 load_stream(const float2* p)
{
    nt_f2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f2*>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 load_stream(const float4* p)
{
    nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void store_stream(double2* p, double2 v)
{
    nt_d2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<nt_d2*>(p));
}
__device__ __forceinline__ void store_stream(float2* p, float2 v)
{
    nt_f2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<nt_f2*>(p));
}
__device__ __forceinline__ void store_stream(float4* p, float4 v)
{
    nt_f4 w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<nt_f4*>(p));
}

/* ---------------- Wilson spinor storage ----------------
 * An eo Wilson spinor field, or a block of N of them, is component-major.  FP64 fields hold component k
 * (k = 3 spin + colour) of site offset o at complex index k CS + o: [12][V] (CS = V, o = site) or
 * [12][V][N] (CS = V N, o = site N + right-hand side).  FP32 fields hold the components in pairs:
 * components 2q and 2q+1 of site offset o share the 16-byte lane q CS + o, [6][V] or [6][V][N] of float4,
 * so a neighbour gather moves 16 bytes per lane in either precision (for FP32 twice the bytes per load
 * instruction of one complex number per lane, half the instructions).  In both layouts the upper spin
 * components 0-5 fill the first half of the array, which is all the elementwise kernels (gamma5, the
 * site-diagonal factors) rely on. */
template<typename R>
__device__ __forceinline__ size_t spinor_cx(int k, size_t CS, size_t o)
{
    if constexpr (sizeof(R) == 8)
        return (size_t)k * CS + o;
    else
        return (size_t)(k >> 1) * 2 * CS + 2 * o + (size_t)(k & 1);
}
/* the 12 components of site offset o (cached loads: neighbour spinors are reused through the L2) */
template<typename R>
__device__ __forceinline__ void load_spinor(vec2<R> (&s)[12], const vec2<R>* __restrict__ in, size_t CS, size_t o)
{
    if constexpr (sizeof(R) == 8) {
#pragma unroll
        for (int k = 0; k < 12; ++k) s[k] = in[(size_t)k * CS + o];
    } else {
        const float4* p = reinterpret_cast<const float4*>(in);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const float4 v = p[(size_t)q * CS + o];
            s[2 * q]       = vec2<R>{v.x, v.y};
            s[2 * q + 1]   = vec2<R>{v.z, v.w};
        }
    }
}
/* the same with streaming loads (data read once) */
template<typename R>
__device__ __forceinline__ void load_spinor_stream(vec2<R> (&s)[12], const vec2<R>* in, size_t CS, size_t o)
{
    if constexpr (sizeof(R) == 8) {
#pragma unroll
        for (int k = 0; k < 12; ++k) s[k] = load_stream(in + (size_t)k * CS + o);
    } else {
        const float4* p = reinterpret_cast<const float4*>(in);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const float4 v = load_stream(p + (size_t)q * CS + o);
            s[2 * q]       = vec2<R>{v.x, v.y};
            s[2 * q + 1]   = vec2<R>{v.z, v.w};
        }
    }
}
/* streaming store of 12 components to site offset o */
template<typename R>
__device__ __forceinline__ void store_spinor_stream(vec2<R>* out, size_t CS, size_t o, const vec2<R> (&s)[12])
{
    if constexpr (sizeof(R) == 8) {
#pragma unroll
        for (int k = 0; k < 12; ++k) store_stream(out + (size_t)k * CS + o, s[k]);
    } else {
        float4* p = reinterpret_cast<float4*>(out);
#pragma unroll
        for (int q = 0; q < 6; ++q)
            store_stream(p + (size_t)q * CS + o, make_float4(s[2 * q].x, s[2 * q].y, s[2 * q + 1].x, s[2 * q + 1].y));
    }
}
/* plain store of 12 components to site offset o */
template<typename R>
__device__ __forceinline__ void store_spinor(vec2<R>* out, size_t CS, size_t o, const vec2<R> (&s)[12])
{
    if constexpr (sizeof(R) == 8) {
#pragma unroll
        for (int k = 0; k < 12; ++k) out[(size_t)k * CS + o] = s[k];
    } else {
        float4* p = reinterpret_cast<float4*>(out);
#pragma unroll
        for (int q = 0; q < 6; ++q)
            p[(size_t)q * CS + o] = make_float4(s[2 * q].x, s[2 * q].y, s[2 * q + 1].x, s[2 * q + 1].y);
    }
}

/* ---------------- gauge-link formats ----------------
 * Device layout: link_elem (tiles of 64 eo sites per (dir, parity), lane-major inside a tile), where a
 * lane is one complex entry (FP64) or a pair of complex entries packed into a 16-byte vector (FP32);
 * gstride is Veo. */
template<int FMT>
struct GaugeLoad;

template<>
struct GaugeLoad<GAUGE_F64_18> {
    using R = double;
    static constexpr int LANES = 9;
    __device__ static __forceinline__ void load(const void* g, uint32_t gstride, int dir, int par, uint32_t eo,
                                                Su3<double>& u)
    {
        const double2* p = reinterpret_cast<const double2*>(g) + link_elem(dir * 2 + par, 9, 0, eo, gstride);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            double2 v           = load_stream(p + k * LINK_TILE);
            u.e[k / 3][k % 3] = Cx<double>{v.x, v.y};
        }
    }
};
template<>
struct GaugeLoad<GAUGE_F64_12> {
    using R = double;
    static constexpr int LANES = 6;
    __device__ static __forceinline__ void load(const void* g, uint32_t gstride, int dir, int par, uint32_t eo,
                                                Su3<double>& u)
    {
        const double2* p = reinterpret_cast<const double2*>(g) + link_elem(dir * 2 + par, 6, 0, eo, gstride);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double2 v           = load_stream(p + k * LINK_TILE);
            u.e[k / 3][k % 3] = Cx<double>{v.x, v.y};
        }
        su3_reconstruct_row2(u);
    }
};
template<>
struct GaugeLoad<GAUGE_F32_12> {
    using R = float;
    static constexpr int LANES = 3;  // 3 float4 = 6 complex
    __device__ static __forceinline__ void load(const void* g, uint32_t gstride, int dir, int par, uint32_t eo,
                                                Su3<float>& u)
    {
        const float4* p = reinterpret_cast<const float4*>(g) + link_elem(dir * 2 + par, 3, 0, eo, gstride);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float4 v                      = load_stream(p + k * LINK_TILE);
            u.e[(2 * k) / 3][(2 * k) % 3] = Cx<float>{v.x, v.y};
            u.e[(2 * k + 1) / 3][(2 * k + 1) % 3] = Cx<float>{v.z, v.w};
        }
        su3_reconstruct_row2(u);
    }
};

/* ---------------- reductions ----------------
 * Block partial sums with a fixed combination order (wavefront butterfly, then
 * LDS across the block's waves), so results are bit-reproducible for a given
 * launch shape. */
template<typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template<int BLOCK>
__device__ __forceinline__ double block_sum(double v, double* lds)
{
    v              = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = 0.;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) r += lds[w];
    }
    __syncthreads();
    return r;  // valid in thread 0
}

template<int BLOCK>
__device__ __forceinline__ double2 block_sum2(double2 v, double* lds)
{
    v.x            = wave_sum(v.x);
    v.y            = wave_sum(v.y);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        lds[2 * wid]     = v.x;
        lds[2 * wid + 1] = v.y;
    }
    __syncthreads();
    double2 r = make_double2(0., 0.);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) {
            r.x += lds[2 * w];
            r.y += lds[2 * w + 1];
        }
    }
    __syncthreads();
    return r;
}

}  // namespace cl2
