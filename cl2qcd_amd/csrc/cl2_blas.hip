// Spinor / staggered-vector BLAS, two-stage deterministic reductions, device
// scalar arithmetic and the multi-shift CG recurrences.
//
// Fields are flat arrays of complex numbers (Wilson: 12*Veo, staggered: 3*Veo),
// so one elementwise kernel serves both; the Wilson and staggered sign
// conventions are separate modes.  Reductions accumulate in FP64 and combine in
// a fixed order (per-block butterfly + LDS, then one block over the partials),
// so results are reproducible run to run.
#include <algorithm>

#include "cl2_device.hpp"
#include "cl2_internal.hpp"

namespace cl2 {

__device__ __forceinline__ Cx<double> resolve(const cl2_scalar& s)
{
    switch (s.kind) {
        case CL2_SCALAR_DEV_COMPLEX: return Cx<double>{s.dev[0], s.dev[1]};
        case CL2_SCALAR_DEV_REAL: return Cx<double>{s.dev[0], 0.};
        case CL2_SCALAR_DEV_VECTOR: return Cx<double>{s.dev[s.index], 0.};
        default: return Cx<double>{s.re, s.im};
    }
}

/* 16-byte chunks of a complex field: one FP64 complex number, or two FP32 numbers when every operand is
 * 16-byte aligned and the length even (else one FP32 number per 8-byte chunk).  Loads and stores are
 * streaming: a BLAS pass touches each element once. */
template<typename R, int CPC>
struct Chunk;
template<>
struct Chunk<double, 1> {
    __device__ static void load(const void* p, size_t c, Cx<double> (&v)[1])
    {
        const double2 w = load_stream(reinterpret_cast<const double2*>(p) + c);
        v[0]            = Cx<double>{w.x, w.y};
    }
    __device__ static void store(void* p, size_t c, const Cx<double> (&v)[1])
    {
        store_stream(reinterpret_cast<double2*>(p) + c, make_double2(v[0].re, v[0].im));
    }
};
template<>
struct Chunk<float, 2> {
    __device__ static void load(const void* p, size_t c, Cx<float> (&v)[2])
    {
        const float4 w = load_stream(reinterpret_cast<const float4*>(p) + c);
        v[0]           = Cx<float>{w.x, w.y};
        v[1]           = Cx<float>{w.z, w.w};
    }
    __device__ static void store(void* p, size_t c, const Cx<float> (&v)[2])
    {
        store_stream(reinterpret_cast<float4*>(p) + c, make_float4(v[0].re, v[0].im, v[1].re, v[1].im));
    }
};
template<>
struct Chunk<float, 1> {
    __device__ static void load(const void* p, size_t c, Cx<float> (&v)[1])
    {
        const float2 w = load_stream(reinterpret_cast<const float2*>(p) + c);
        v[0]           = Cx<float>{w.x, w.y};
    }
    __device__ static void store(void* p, size_t c, const Cx<float> (&v)[1])
    {
        store_stream(reinterpret_cast<float2*>(p) + c, make_float2(v[0].re, v[0].im));
    }
};

/* Elementwise BLAS: each thread takes U chunks BLOCK apart, issues all their loads, then computes and
 * stores; the grid covers the field (up to BLAS_MAX_BLOCKS blocks, beyond that it loops).  Operands of
 * one element are read before the element is written, so out may alias any input (in-place updates). */
template<typename R, int CPC, int U, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_blas(BlasArgs a)
{
    using IO = Chunk<R, CPC>;
    __shared__ double lds[2 * BLOCK / 64];
    const Cx<double> A = resolve(a.a), B = resolve(a.b);
    const Cx<R> al{R(A.re), R(A.im)}, be{R(B.re), R(B.im)};
    const int mode      = a.mode;
    const bool has_y    = !(mode == B_AX || mode == B_NEG_TAIL);
    const bool has_z    = mode == B_AX_BY_Z || mode == B_CG_XP;
    const size_t nchunk = a.n / CPC;
    double nrm = 0., ry = 0.;
    const size_t step = (size_t)gridDim.x * BLOCK * U;
    for (size_t base = (size_t)blockIdx.x * BLOCK * U + threadIdx.x; base < nchunk; base += step) {
        Cx<R> xv[U][CPC], yv[U][CPC], zv[U][CPC];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t c = base + (size_t)u * BLOCK;
            if (c < nchunk) {
                IO::load(a.x, c, xv[u]);
                if (has_y) IO::load(a.y, c, yv[u]);
                if (has_z) IO::load(a.z, c, zv[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t c = base + (size_t)u * BLOCK;
            if (c >= nchunk) continue;
            Cx<R> r[CPC], r2[CPC];
#pragma unroll
            for (int j = 0; j < CPC; ++j) {
                const size_t k = c * CPC + j;
                const Cx<R> x = xv[u][j], y = yv[u][j], z = zv[u][j];
                switch (mode) {
                    case B_Y_MINUS_AX:
                        r[j] = y - cmul(x, al);
                        ry += (double)r[j].re * (double)y.re + (double)r[j].im * (double)y.im;
                        break;
                    case B_AX: r[j] = cmul(x, al); break;
                    case B_AX_BY_Z: r[j] = cmul(y, be) + cmul(x, al) + z; break;
                    case B_AX_PLUS_Y: r[j] = cmul(x, al) + y; break;
                    case B_AX_BY: r[j] = cmul(x, al) + cmul(y, be); break;
                    case B_G5_Y_MINUS_AX:
                        r[j] = y - cmul(x, al);
                        if (k >= a.tail) r[j] = Cx<R>{-r[j].re, -r[j].im};
                        break;
                    case B_CG_XP:
                        r[j]  = z - cmul(x, al);
                        r2[j] = y - cmul(x, be);
                        break;
                    default:  // B_NEG_TAIL: in-place gamma5
                        r[j] = (k >= a.tail) ? Cx<R>{-x.re, -x.im} : x;
                        break;
                }
                nrm += (double)r[j].re * (double)r[j].re + (double)r[j].im * (double)r[j].im;
            }
            IO::store(a.out, c, r);
            if (mode == B_CG_XP) IO::store(a.out2, c, r2);
        }
    }
    if (a.norm_partials) {
        double2 s = block_sum2<BLOCK>(make_double2(nrm, a.cg_norm ? ry : 0.), lds);
        if (threadIdx.x == 0) a.norm_partials[2 * blockIdx.x] = s.x, a.norm_partials[2 * blockIdx.x + 1] = s.y;
    }
}

/* chunks per thread of the elementwise kernels: one when nothing is reduced (the shape that streams
 * fastest), four with per-block partials (a quarter of the partials for stage 2 to sum) */
constexpr int BLAS_U_PLAIN = 1, BLAS_U_REDUCE = 4;

static unsigned chunk_blocks(size_t nchunk, int u)
{
    const size_t nb = (nchunk + (size_t)BLAS_BLOCK * u - 1) / ((size_t)BLAS_BLOCK * u);
    return nb == 0 ? 1u : (unsigned)std::min<size_t>(nb, BLAS_MAX_BLOCKS);
}

static bool aligned16(const void* p)
{
    return p == nullptr || (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

template<typename R, int CPC>
static hipError_t launch_blas_t(cl2_ctx* c, const BlasArgs& a, unsigned* nblocks_out)
{
    const size_t nchunk = a.n / CPC;
    if (a.norm_partials) {
        const unsigned nb = chunk_blocks(nchunk, BLAS_U_REDUCE);
        if (nblocks_out) *nblocks_out = nb;
        hipLaunchKernelGGL((k_blas<R, CPC, BLAS_U_REDUCE, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream, a);
    } else {
        const unsigned nb = chunk_blocks(nchunk, BLAS_U_PLAIN);
        if (nblocks_out) *nblocks_out = nb;
        hipLaunchKernelGGL((k_blas<R, CPC, BLAS_U_PLAIN, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream, a);
    }
    return hipGetLastError();
}

hipError_t launch_blas(cl2_ctx* c, const BlasArgs& a, int prec, unsigned* nblocks_out)
{
    if (prec == CL2_F64) return launch_blas_t<double, 1>(c, a, nblocks_out);
    const bool pairs = (a.n % 2) == 0 && aligned16(a.out) && aligned16(a.x) && aligned16(a.y) &&
                       aligned16(a.z) && (a.mode != B_CG_XP || aligned16(a.out2));
    return pairs ? launch_blas_t<float, 2>(c, a, nblocks_out) : launch_blas_t<float, 1>(c, a, nblocks_out);
}

/* stage 1: per-block sums of conj(x) y (R_DOT), its real part (R_DOT_RE) or |x|^2 (R_NORM), chunked as the
 * elementwise kernel */
template<typename R, int CPC, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_reduce1(const void* x, const void* y, size_t n, int mode,
                                                   double* __restrict__ partials)
{
    using IO = Chunk<R, CPC>;
    constexpr int U = BLAS_U_REDUCE;
    __shared__ double lds[2 * BLOCK / 64];
    double2 acc         = make_double2(0., 0.);
    const size_t nchunk = n / CPC;
    const size_t step   = (size_t)gridDim.x * BLOCK * U;
    const bool norm     = mode == R_NORM;
    for (size_t base = (size_t)blockIdx.x * BLOCK * U + threadIdx.x; base < nchunk; base += step) {
        Cx<R> u[U][CPC], v[U][CPC];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const size_t c = base + (size_t)q * BLOCK;
            if (c < nchunk) {
                IO::load(x, c, u[q]);
                if (!norm) IO::load(y, c, v[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            if (base + (size_t)q * BLOCK >= nchunk) continue;
#pragma unroll
            for (int j = 0; j < CPC; ++j) {
                if (norm) {
                    acc.x += (double)u[q][j].re * (double)u[q][j].re + (double)u[q][j].im * (double)u[q][j].im;
                } else {
                    acc.x += (double)u[q][j].re * (double)v[q][j].re + (double)u[q][j].im * (double)v[q][j].im;
                    acc.y += (double)u[q][j].re * (double)v[q][j].im - (double)v[q][j].re * (double)u[q][j].im;
                }
            }
        }
    }
    double2 s = block_sum2<BLOCK>(acc, lds);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x]     = s.x;
        partials[2 * blockIdx.x + 1] = (mode == R_DOT) ? s.y : 0.;
    }
}

static bool reduce_pairs(size_t n, int prec, const void* x, const void* y)
{
    return prec != CL2_F64 && n % 2 == 0 && aligned16(x) && aligned16(y);
}

unsigned reduce_stage1_blocks(size_t n, int prec, const void* x, const void* y)
{
    return chunk_blocks(reduce_pairs(n, prec, x, y) ? n / 2 : n, BLAS_U_REDUCE);
}

hipError_t launch_reduce_stage1(cl2_ctx* c, const void* x, const void* y, size_t n, int prec, int mode,
                                unsigned* nblocks, double* parts)
{
    if (!parts) parts = c->partials;
    const unsigned nb = reduce_stage1_blocks(n, prec, x, y);
    *nblocks          = nb;
    if (prec == CL2_F64)
        hipLaunchKernelGGL((k_reduce1<double, 1, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream, x, y, n,
                           mode, parts);
    else if (reduce_pairs(n, prec, x, y))
        hipLaunchKernelGGL((k_reduce1<float, 2, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream, x, y, n,
                           mode, parts);
    else
        hipLaunchKernelGGL((k_reduce1<float, 1, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream, x, y, n,
                           mode, parts);
    return hipGetLastError();
}

/* stage 2: one block sums the partials in a fixed order.
 * out_kind 0: res[0] = re;  1: res = (re, im);  2: res[k] = alpha[k]^2 * re for k < nvec;
 * 3: res = (re, im, re, 0) */
template<int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_reduce2(const double* __restrict__ partials, unsigned nparts,
                                                   double* __restrict__ res, int out_kind,
                                                   const double* __restrict__ alpha, int nvec)
{
    __shared__ double lds[2 * BLOCK / 64];
    double2 acc = make_double2(0., 0.);
    for (unsigned k = threadIdx.x; k < nparts; k += BLOCK) {
        acc.x += partials[2 * k];
        acc.y += partials[2 * k + 1];
    }
    double2 s = block_sum2<BLOCK>(acc, lds);
    if (threadIdx.x == 0) {
        if (out_kind == 0) {
            res[0] = s.x;
        } else if (out_kind == 1) {
            res[0] = s.x;
            res[1] = s.y;
        } else if (out_kind == 3) {
            res[0] = s.x;
            res[1] = s.y;
            res[2] = s.x;
            res[3] = 0.;
        } else {
            for (int k = 0; k < nvec; ++k) res[k] = alpha[k] * alpha[k] * s.x;
        }
    }
}

hipError_t launch_reduce_stage2(cl2_ctx* c, unsigned nparts, double* res, int out_kind, const double* alpha_vec,
                                int nvec, const double* parts)
{
    // 1024 threads: the elementwise kernels leave up to BLAS_MAX_BLOCKS partials
    hipLaunchKernelGGL((k_reduce2<1024>), dim3(1), dim3(1024), 0, c->stream, parts ? parts : c->partials, nparts, res,
                       out_kind, alpha_vec, nvec);
    return hipGetLastError();
}

template<typename R, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_fill(vec2<R>* __restrict__ out, size_t n, R re, R im)
{
    const size_t step = (size_t)gridDim.x * BLOCK;
    for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < n; k += step) out[k] = vec2<R>{re, im};
}

/* per-slice |x|^2 of an eo spinor field.  The eo sites of slice s are nchunk contiguous runs of
 * chunk sites, run j starting at s * slice_off + j * chunk_stride: the timeslice t is the one run
 * [t * vs/2, (t+1) * vs/2); the z-slice z is, for every t, the run of the NS^2/2 eo sites of the
 * (x, y) planes with that z (NS is even, so both sites of an eo pair share z).  Block s reduces
 * its runs in a fixed order. */
template<typename R>
__global__ __launch_bounds__(256) void k_slice_norm(double* res, const vec2<R>* __restrict__ x, uint32_t S,
                                                    uint32_t slice_off, uint32_t chunk, uint32_t chunk_stride,
                                                    uint32_t nchunk, int accumulate)
{
    __shared__ double lds[8];
    const uint32_t s = blockIdx.x;
    double acc       = 0.;
    for (uint32_t i = threadIdx.x; i < nchunk * chunk; i += 256) {
        const uint32_t e = s * slice_off + (i / chunk) * chunk_stride + i % chunk;
        vec2<R> v[12];
        load_spinor<R>(v, x, S, e);
#pragma unroll
        for (int k = 0; k < 12; ++k) acc += (double)v[k].x * (double)v[k].x + (double)v[k].y * (double)v[k].y;
    }
    const double sum = block_sum<256>(acc, lds);
    if (threadIdx.x == 0) res[s] = accumulate ? res[s] + sum : sum;
}

hipError_t launch_slice_norm(cl2_ctx* c, double* res, const void* x, int prec, int dir, int accumulate)
{
    const uint32_t half_vs = c->g.vs / 2, half_plane = c->g.ns * c->g.ns / 2;
    const uint32_t nslice       = dir == 0 ? c->g.nt : c->g.ns;
    const uint32_t slice_off    = dir == 0 ? half_vs : half_plane;
    const uint32_t chunk        = dir == 0 ? half_vs : half_plane;
    const uint32_t nchunk       = dir == 0 ? 1 : c->g.nt;
    if (prec == CL2_F64)
        hipLaunchKernelGGL(k_slice_norm<double>, dim3(nslice), dim3(256), 0, c->stream, res,
                           reinterpret_cast<const double2*>(x), c->g.veo, slice_off, chunk, half_vs, nchunk,
                           accumulate);
    else
        hipLaunchKernelGGL(k_slice_norm<float>, dim3(nslice), dim3(256), 0, c->stream, res,
                           reinterpret_cast<const float2*>(x), c->g.veo, slice_off, chunk, half_vs, nchunk,
                           accumulate);
    return hipGetLastError();
}

/* device-to-device copy of n 16-byte elements with streaming loads and stores, one element per
 * thread (the launch shape that measured fastest for plain copies on gfx950) */
__global__ __launch_bounds__(256) void k_copy16(double2* __restrict__ dst, const double2* __restrict__ src, size_t n)
{
    const size_t step = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += step) store_stream(dst + i, load_stream(src + i));
}

hipError_t launch_copy16(cl2_ctx* c, void* dst, const void* src, size_t n)
{
    const size_t nb = std::min<size_t>((n + 255) / 256, (size_t)1 << 22);
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)nb), dim3(256), 0, c->stream, reinterpret_cast<double2*>(dst),
                       reinterpret_cast<const double2*>(src), n);
    return hipGetLastError();
}

/* one element per thread for the streaming kernels without partials (fill, conversion) */
static unsigned stream_blocks(size_t n)
{
    const size_t nb = (n + BLAS_BLOCK - 1) / BLAS_BLOCK;
    return nb == 0 ? 1u : (unsigned)std::min<size_t>(nb, (size_t)1 << 22);
}

hipError_t launch_fill(cl2_ctx* c, void* out, size_t n, int prec, double re, double im)
{
    const unsigned nb = stream_blocks(n);
    if (prec == CL2_F64)
        hipLaunchKernelGGL((k_fill<double, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream,
                           reinterpret_cast<double2*>(out), n, re, im);
    else
        hipLaunchKernelGGL((k_fill<float, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream,
                           reinterpret_cast<float2*>(out), n, (float)re, (float)im);
    return hipGetLastError();
}

/* precision conversion between complex fields of the same layout (FP64 <-> FP32 staggered fields) */
template<typename RD, typename RS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_convert(vec2<RD>* __restrict__ out, const vec2<RS>* __restrict__ in,
                                                   size_t n)
{
    const size_t step = (size_t)gridDim.x * BLOCK;
    for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < n; k += step) {
        vec2<RS> v = in[k];
        out[k]     = vec2<RD>{RD(v.x), RD(v.y)};
    }
}

/* FP64 <-> FP32 conversion of Wilson spinors (cs = Veo) or blocks of them (cs = Veo nrhs): FP64 [12][cs]
 * complex numbers <-> FP32 [6][cs] component pairs; thread (o, q) converts components 2q and 2q+1 of
 * site offset o */
template<typename RD, typename RS>
__global__ __launch_bounds__(256) void k_convert_spinor(vec2<RD>* __restrict__ out, const vec2<RS>* __restrict__ in,
                                                        size_t cs)
{
    const size_t o = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t q = blockIdx.y;
    if (o >= cs) return;
    vec2<RS> a, b;
    if constexpr (sizeof(RS) == 8) {
        a = load_stream(in + 2 * q * cs + o);
        b = load_stream(in + (2 * q + 1) * cs + o);
    } else {
        const float4 v = load_stream(reinterpret_cast<const float4*>(in) + q * cs + o);
        a              = vec2<RS>{v.x, v.y};
        b              = vec2<RS>{v.z, v.w};
    }
    if constexpr (sizeof(RD) == 8) {
        store_stream(out + 2 * q * cs + o, vec2<RD>{RD(a.x), RD(a.y)});
        store_stream(out + (2 * q + 1) * cs + o, vec2<RD>{RD(b.x), RD(b.y)});
    } else {
        store_stream(reinterpret_cast<float4*>(out) + q * cs + o, make_float4(float(a.x), float(a.y), float(b.x), float(b.y)));
    }
}

hipError_t launch_convert_spinor(cl2_ctx* c, void* out, int out_prec, const void* in, int in_prec, size_t cs)
{
    const dim3 grid((unsigned)((cs + 255) / 256), 6);
    if (out_prec == CL2_F32 && in_prec == CL2_F64)
        hipLaunchKernelGGL((k_convert_spinor<float, double>), grid, dim3(256), 0, c->stream,
                           reinterpret_cast<float2*>(out), reinterpret_cast<const double2*>(in), cs);
    else if (out_prec == CL2_F64 && in_prec == CL2_F32)
        hipLaunchKernelGGL((k_convert_spinor<double, float>), grid, dim3(256), 0, c->stream,
                           reinterpret_cast<double2*>(out), reinterpret_cast<const float2*>(in), cs);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_convert(cl2_ctx* c, void* out, int out_prec, const void* in, int in_prec, size_t n)
{
    const unsigned nb = stream_blocks(n);
    if (out_prec == CL2_F32 && in_prec == CL2_F64)
        hipLaunchKernelGGL((k_convert<float, double, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream,
                           reinterpret_cast<float2*>(out), reinterpret_cast<const double2*>(in), n);
    else if (out_prec == CL2_F64 && in_prec == CL2_F32)
        hipLaunchKernelGGL((k_convert<double, float, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream,
                           reinterpret_cast<double2*>(out), reinterpret_cast<const float2*>(in), n);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

/* Device-resident scalar arithmetic (complex_{product,ratio,sum,subtraction}.cl,
 * real_{product,ratio,sum,subtraction}.cl); ratio follows complexdivide
 * (common_header_files/operations_complex.hpp:59-66). */
__global__ void k_scalar_op(int op, int is_complex, double* out, const double* a, const double* b)
{
    if (threadIdx.x != 0) return;
    if (!is_complex) {
        double r;
        switch (op) {
            case CL2_SCALAR_OP_PRODUCT: r = a[0] * b[0]; break;
            case CL2_SCALAR_OP_RATIO: r = a[0] / b[0]; break;
            case CL2_SCALAR_OP_SUM: r = a[0] + b[0]; break;
            default: r = a[0] - b[0]; break;
        }
        out[0] = r;
        return;
    }
    double ar = a[0], ai = a[1], br = b[0], bi = b[1], rr, ri;
    switch (op) {
        case CL2_SCALAR_OP_PRODUCT:
            rr = ar * br - ai * bi;
            ri = ar * bi + ai * br;
            break;
        case CL2_SCALAR_OP_RATIO: {
            double nrm = br * br + bi * bi;
            rr         = (ar * br + ai * bi) / nrm;
            ri         = (ai * br - ar * bi) / nrm;
            break;
        }
        case CL2_SCALAR_OP_SUM:
            rr = ar + br;
            ri = ai + bi;
            break;
        default:
            rr = ar - br;
            ri = ai - bi;
            break;
    }
    out[0] = rr;
    out[1] = ri;
}

hipError_t launch_scalar_op(cl2_ctx* c, int op, int is_complex, double* out, const double* a, const double* b)
{
    hipLaunchKernelGGL(k_scalar_op, dim3(1), dim3(64), 0, c->stream, op, is_complex, out, a, b);
    return hipGetLastError();
}

__global__ void k_vec_element(double* out, const double* in, int out_index, int in_index)
{
    if (threadIdx.x == 0) out[out_index] = in[in_index];
}

hipError_t launch_vec_element(cl2_ctx* c, double* out, const double* in, int out_index, int in_index)
{
    hipLaunchKernelGGL(k_vec_element, dim3(1), dim3(64), 0, c->stream, out, in, out_index, in_index);
    return hipGetLastError();
}

/* Multi-shift CG recurrences (ocl_kernel/operations_real.cl:20-39), without contraction so that the separate
 * launches below and the fused CG-M step compute the same bits:
 *   zeta:  (a b c) / (d e (b - a) + b c (1 - f d))
 *   beta:  -a b / c
 *   alpha: -a b c / (d e) */
__device__ __forceinline__ double cgm_zeta(double A, double B, double C, double D, double E, double F)
{
#pragma clang fp contract(off)
    return (A * B * C) / (D * E * (B - A) + B * C * (1 - F * D));
}
__device__ __forceinline__ double cgm_beta(double a, double b, double c)
{
#pragma clang fp contract(off)
    return -a * b / c;
}
__device__ __forceinline__ double cgm_alpha(double a, double b, double c, double d, double e)
{
#pragma clang fp contract(off)
    return -a * b * c / (d * e);
}

/* One thread per shift; pointer roles follow the launcher argument lists of real.cpp:285-400. */
__global__ void k_cgm(int which, double* out, const double* a, const double* b, const double* d, const double* e,
                      const double* f, const double* g, int numeq)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= numeq) return;
    if (which == 0)  // zeta(zeta_prev, zeta_prev_prev, *sbeta_prev, *sbeta_pres, *salpha_prev, sigma)
        out[k] = cgm_zeta(a[k], b[k], d[0], e[0], f[0], g[k]);
    else if (which == 1)  // beta(*sbeta_pres, zeta_pres, zeta_prev)
        out[k] = cgm_beta(a[0], b[k], d[k]);
    else  // alpha(*salpha_pres, zeta_pres, beta_pres, zeta_prev, *sbeta_pres)
        out[k] = cgm_alpha(a[0], b[k], d[k], e[k], f[0]);
}

/* The scalar work of one CG-M iteration (solver_shifted.cpp:136-198) in two launches instead of thirteen:
 * beta step:  beta_prev = beta; beta = 0 - rr / pap;
 * alpha step: alpha_prev = alpha; alpha = rr_new / rr_old; per shift zeta_foll, beta_vec, alpha_vec from
 *             (zeta, zeta_prev, beta_prev, beta, alpha_prev, sigma); then zeta_prev = zeta, zeta = zeta_foll and
 *             rr_old = rr_new.  Every value is the one the separate launches give (same operations, same
 *             order).  One block of 64 threads (numeq <= 64): all reads happen before the barrier, all
 *             writes of shared scalars after it. */
__global__ void k_cgm_beta_step(double* beta, double* beta_prev, const double* rr, const double* pap)
{
    if (threadIdx.x != 0) return;
    const double b = beta[0];
    beta_prev[0]   = b;
    beta[0]        = 0.0 - rr[0] / pap[0];
}

__global__ __launch_bounds__(64) void k_cgm_alpha_step(int numeq, double* alpha, double* alpha_prev,
                                                       const double* rr_new, double* rr_old, const double* beta,
                                                       const double* beta_prev, const double* sigma, double* zeta,
                                                       double* zeta_prev, double* zeta_foll, double* beta_vec,
                                                       double* alpha_vec)
{
    const int k        = threadIdx.x;
    const double a_old = alpha[0], rn = rr_new[0];
    const double a_new = rn / rr_old[0];
    const double b = beta[0], bp = beta_prev[0];
    double z = 0., zp = 0., s = 0.;
    if (k < numeq) z = zeta[k], zp = zeta_prev[k], s = sigma[k];
    __syncthreads();
    if (k < numeq) {
        const double zf = cgm_zeta(z, zp, bp, b, a_old, s);
        const double bv = cgm_beta(b, zf, z);
        zeta_foll[k]    = zf;
        beta_vec[k]     = bv;
        alpha_vec[k]    = cgm_alpha(a_new, zf, bv, z, b);
        zeta_prev[k]    = z;
        zeta[k]         = zf;
    }
    if (k == 0) {
        alpha_prev[0] = a_old;
        alpha[0]      = a_new;
        rr_old[0]     = rn;
    }
}

/* the beta step from the fused MdagM step's sums: pap = m2 |p|^2 + |D p|^2 from wp = (|D p|^2, |p|^2) */
__global__ void k_cgm_beta_step_mdagm(double* beta, double* beta_prev, const double* rr, double* pap, const double* wp,
                                      double m2)
{
    if (threadIdx.x != 0) return;
    const double a = m2 * wp[1] + wp[0];
    pap[0]         = a;
    const double b = beta[0];
    beta_prev[0]   = b;
    beta[0]        = 0.0 - rr[0] / a;
}

hipError_t launch_cgm_beta_step_mdagm(cl2_ctx* c, double* beta, double* beta_prev, const double* rr, double* pap,
                                      const double* wp, double m2)
{
    hipLaunchKernelGGL(k_cgm_beta_step_mdagm, dim3(1), dim3(64), 0, c->stream, beta, beta_prev, rr, pap, wp, m2);
    return hipGetLastError();
}

hipError_t launch_cgm_beta_step(cl2_ctx* c, double* beta, double* beta_prev, const double* rr, const double* pap)
{
    hipLaunchKernelGGL(k_cgm_beta_step, dim3(1), dim3(64), 0, c->stream, beta, beta_prev, rr, pap);
    return hipGetLastError();
}

hipError_t launch_cgm_alpha_step(cl2_ctx* c, int numeq, double* alpha, double* alpha_prev, const double* rr_new,
                                 double* rr_old, const double* beta, const double* beta_prev, const double* sigma,
                                 double* zeta, double* zeta_prev, double* zeta_foll, double* beta_vec,
                                 double* alpha_vec)
{
    hipLaunchKernelGGL(k_cgm_alpha_step, dim3(1), dim3(64), 0, c->stream, numeq, alpha, alpha_prev, rr_new, rr_old,
                       beta, beta_prev, sigma, zeta, zeta_prev, zeta_foll, beta_vec, alpha_vec);
    return hipGetLastError();
}

hipError_t launch_cgm_update(cl2_ctx* c, int which, double* out, const double* a, const double* b, const double* d,
                             const double* e, const double* f, const double* g, int numeq)
{
    const unsigned nb = blocks_for((size_t)numeq, 64);
    hipLaunchKernelGGL(k_cgm, dim3(nb ? nb : 1), dim3(64), 0, c->stream, which, out, a, b, d, e, f, g, numeq);
    return hipGetLastError();
}

}  // namespace cl2
