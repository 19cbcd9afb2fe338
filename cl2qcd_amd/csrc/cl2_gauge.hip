// Device-side gauge-field starts: random SU(3) links from a counter-based
// generator (the Gaugefield hot/cold starts of physics/lattices/gaugefield.cpp:35-72),
// written straight into any of the device link formats.
#include "cl2_device.hpp"
#include "cl2_internal.hpp"

namespace cl2 {

typedef _Float16 gh4 __attribute__((ext_vector_type(4)));
typedef _Float16 gh8 __attribute__((ext_vector_type(8)));

/* rows 0,1 of u as F16_12 (layout: link_h16_offset / link_h8_offset) */
template<typename R>
__device__ __forceinline__ void store_link_h(void* dev, uint32_t dp, uint32_t veo, uint32_t eo, const Su3<R>& u)
{
    char* base     = static_cast<char*>(dev);
    const Cx<R> e0 = u.e[0][0], e1 = u.e[0][1], e2 = u.e[0][2], e3 = u.e[1][0], e4 = u.e[1][1], e5 = u.e[1][2];
    *reinterpret_cast<gh8*>(base + link_h16_offset(dp, eo, veo)) =
        gh8{(_Float16)e0.re, (_Float16)e0.im, (_Float16)e1.re, (_Float16)e1.im,
            (_Float16)e2.re, (_Float16)e2.im, (_Float16)e3.re, (_Float16)e3.im};
    *reinterpret_cast<gh4*>(base + link_h8_offset(dp, eo, veo)) =
        gh4{(_Float16)e4.re, (_Float16)e4.im, (_Float16)e5.re, (_Float16)e5.im};
}

/* splitmix64 finaliser: a stateless hash of (seed, counter) */
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t h)
{
    return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

/* One link per thread: A = 1 + eps * G with G complex Gaussian, then rows 0,1
 * Gram-Schmidt orthonormalised and row 2 = conj(row0 x row1), so det U = 1. */
__global__ void k_gauge_random(void* dev, int fmt, uint32_t veo, uint64_t seed, double eps)
{
    const uint64_t link = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (dir, parity, eo)
    if (link >= 8ull * veo) return;
    Su3<double> u;
    uint64_t ctr = (seed * 0x100000001b3ull) ^ (link * 32ull);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double r1 = u01(mix64(ctr++)), r2 = u01(mix64(ctr++));
        double rad = sqrt(-2.0 * log(r1)), ph = 6.283185307179586 * r2;
        const int row = k / 3, col = k % 3;
        u.e[row][col] = Cx<double>{(row == col ? 1.0 : 0.0) + eps * rad * cos(ph), eps * rad * sin(ph)};
    }
    double n0 = 0.;
    for (int c = 0; c < 3; ++c) n0 += u.e[0][c].re * u.e[0][c].re + u.e[0][c].im * u.e[0][c].im;
    n0 = 1.0 / sqrt(n0);
    for (int c = 0; c < 3; ++c) u.e[0][c] = scale(u.e[0][c], n0);
    Cx<double> p{0., 0.};  // <row0, row1>
    for (int c = 0; c < 3; ++c) p = p + cmulconj(u.e[1][c], u.e[0][c]);
    for (int c = 0; c < 3; ++c) u.e[1][c] = u.e[1][c] - cmul(p, u.e[0][c]);
    double n1 = 0.;
    for (int c = 0; c < 3; ++c) n1 += u.e[1][c].re * u.e[1][c].re + u.e[1][c].im * u.e[1][c].im;
    n1 = 1.0 / sqrt(n1);
    for (int c = 0; c < 3; ++c) u.e[1][c] = scale(u.e[1][c], n1);
    su3_reconstruct_row2(u);

    const uint32_t dp = (uint32_t)(link / veo), eo = (uint32_t)(link % veo);
    if (fmt == GAUGE_F16_12) {
        store_link_h(dev, dp, veo, eo, u);
    } else if (fmt == GAUGE_F32_12) {
        float4* q = reinterpret_cast<float4*>(dev) + link_elem(dp, 3, 0, eo, veo);
        for (int k = 0; k < 3; ++k) {
            Cx<double> a = u.e[(2 * k) / 3][(2 * k) % 3], b = u.e[(2 * k + 1) / 3][(2 * k + 1) % 3];
            q[k * LINK_TILE] = make_float4((float)a.re, (float)a.im, (float)b.re, (float)b.im);
        }
    } else {
        const int lanes = (fmt == GAUGE_F64_18) ? 9 : 6;
        double2* q      = reinterpret_cast<double2*>(dev) + link_elem(dp, lanes, 0, eo, veo);
        for (int k = 0; k < lanes; ++k) q[k * LINK_TILE] = make_double2(u.e[k / 3][k % 3].re, u.e[k / 3][k % 3].im);
    }
}

hipError_t launch_gauge_random(cl2_ctx* c, void* dev, int fmt, uint64_t seed, double eps)
{
    const size_t n    = 8ull * c->g.veo;
    const unsigned nb = blocks_for(n, 256);
    hipLaunchKernelGGL(k_gauge_random, dim3(nb), dim3(256), 0, c->stream, dev, fmt, c->g.veo, seed, eps);
    return hipGetLastError();
}

/* Device-layout conversion between formats (F64_18 -> F64_12 / F32_12 / F16_12), so one
 * resident FP64 field can feed the reduced-precision solvers. */
__global__ void k_gauge_convert(void* dst, int dfmt, const double2* __restrict__ src, uint32_t veo)
{
    const uint64_t link = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (link >= 8ull * veo) return;
    const uint32_t dp = (uint32_t)(link / veo), eo = (uint32_t)(link % veo);
    const double2* s  = src + link_elem(dp, 9, 0, eo, veo);
    if (dfmt == GAUGE_F16_12) {
        Su3<double> u;
        for (int k = 0; k < 6; ++k) u.e[k / 3][k % 3] = Cx<double>{s[k * LINK_TILE].x, s[k * LINK_TILE].y};
        store_link_h(dst, dp, veo, eo, u);
    } else if (dfmt == GAUGE_F32_12) {
        float4* q = reinterpret_cast<float4*>(dst) + link_elem(dp, 3, 0, eo, veo);
        for (int k = 0; k < 3; ++k) {
            double2 a = s[(2 * k) * LINK_TILE], b = s[(2 * k + 1) * LINK_TILE];
            q[k * LINK_TILE] = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
        }
    } else {
        const int lanes = (dfmt == GAUGE_F64_18) ? 9 : 6;
        double2* q      = reinterpret_cast<double2*>(dst) + link_elem(dp, lanes, 0, eo, veo);
        for (int k = 0; k < lanes; ++k) q[k * LINK_TILE] = s[k * LINK_TILE];
    }
}

/* complex Gaussian field, N(0, 1) per real component, from the counter hash of (seed, element) (Box-Muller).
 * e is the element's index in the component-major order k cs + o; with pair_cs != 0 (FP32 Wilson spinors,
 * cs = pair_cs) it is stored where the component-pair layout keeps it, so a seed gives the same field in
 * either precision. */
template<typename R>
__global__ void k_gaussian(vec2<R>* out, size_t n, uint64_t seed, size_t pair_cs)
{
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += step) {
        const uint64_t ctr = (seed * 0x100000001b3ull) ^ (e * 2ull);
        const double r1 = u01(mix64(ctr)), r2 = u01(mix64(ctr + 1));
        const double rad = sqrt(-2.0 * log(r1)), ph = 6.283185307179586 * r2;
        const size_t at  = pair_cs ? spinor_cx<float>((int)(e / pair_cs), pair_cs, e % pair_cs) : e;
        out[at]          = vec2<R>{R(rad * cos(ph)), R(rad * sin(ph))};
    }
}

hipError_t launch_gaussian(cl2_ctx* c, void* dev, int prec, size_t n, uint64_t seed, bool wilson)
{
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
    if (prec == CL2_F64)
        hipLaunchKernelGGL(k_gaussian<double>, dim3(nb), dim3(256), 0, c->stream, reinterpret_cast<double2*>(dev), n, seed,
                           (size_t)0);
    else
        hipLaunchKernelGGL(k_gaussian<float>, dim3(nb), dim3(256), 0, c->stream, reinterpret_cast<float2*>(dev), n, seed,
                           wilson ? n / 12 : (size_t)0);
    return hipGetLastError();
}

hipError_t launch_gauge_convert(cl2_ctx* c, void* dst, int dfmt, const void* src)
{
    const size_t n    = 8ull * c->g.veo;
    const unsigned nb = blocks_for(n, 256);
    hipLaunchKernelGGL(k_gauge_convert, dim3(nb), dim3(256), 0, c->stream, dst, dfmt,
                       reinterpret_cast<const double2*>(src), c->g.veo);
    return hipGetLastError();
}

}  // namespace cl2
