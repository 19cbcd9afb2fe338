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

template<typename R, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_blas(BlasArgs a)
{
    using V = vec2<R>;
    __shared__ double lds[2 * BLOCK / 64];
    const Cx<double> A = resolve(a.a), B = resolve(a.b);
    const Cx<R> al{R(A.re), R(A.im)}, be{R(B.re), R(B.im)};
    const V* x = reinterpret_cast<const V*>(a.x);
    const V* y = reinterpret_cast<const V*>(a.y);
    const V* z = reinterpret_cast<const V*>(a.z);
    V* out     = reinterpret_cast<V*>(a.out);
    double nrm = 0., ry = 0.;
    const size_t step = (size_t)gridDim.x * BLOCK;
    for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < a.n; k += step) {
        Cx<R> r;
        switch (a.mode) {
            case B_Y_MINUS_AX: {
                V xv = x[k], yv = y[k];
                r    = Cx<R>{yv.x, yv.y} - cmul(Cx<R>{xv.x, xv.y}, al);
                ry += (double)r.re * (double)yv.x + (double)r.im * (double)yv.y;
                break;
            }
            case B_AX: {
                V xv = x[k];
                r    = cmul(Cx<R>{xv.x, xv.y}, al);
                break;
            }
            case B_AX_BY_Z: {
                V xv = x[k], yv = y[k], zv = z[k];
                r    = cmul(Cx<R>{yv.x, yv.y}, be) + cmul(Cx<R>{xv.x, xv.y}, al) + Cx<R>{zv.x, zv.y};
                break;
            }
            case B_AX_PLUS_Y: {
                V xv = x[k], yv = y[k];
                r    = cmul(Cx<R>{xv.x, xv.y}, al) + Cx<R>{yv.x, yv.y};
                break;
            }
            case B_AX_BY: {
                V xv = x[k], yv = y[k];
                r    = cmul(Cx<R>{xv.x, xv.y}, al) + cmul(Cx<R>{yv.x, yv.y}, be);
                break;
            }
            case B_G5_Y_MINUS_AX: {
                V xv = x[k], yv = y[k];
                r    = Cx<R>{yv.x, yv.y} - cmul(Cx<R>{xv.x, xv.y}, al);
                if (k >= a.tail) r = Cx<R>{-r.re, -r.im};
                break;
            }
            default: {  // B_NEG_TAIL: in-place gamma5
                V xv = x[k];
                r    = (k >= a.tail) ? Cx<R>{-xv.x, -xv.y} : Cx<R>{xv.x, xv.y};
                break;
            }
        }
        out[k] = V{r.re, r.im};
        nrm += (double)r.re * (double)r.re + (double)r.im * (double)r.im;
    }
    if (a.norm_partials) {
        double2 s = block_sum2<BLOCK>(make_double2(nrm, a.cg_norm ? ry : 0.), lds);
        if (threadIdx.x == 0) a.norm_partials[2 * blockIdx.x] = s.x, a.norm_partials[2 * blockIdx.x + 1] = s.y;
    }
}

static unsigned reduce_blocks(size_t n)
{
    unsigned nb = blocks_for(n, BLAS_BLOCK);
    return nb > (unsigned)REDUCE_MAX_BLOCKS ? (unsigned)REDUCE_MAX_BLOCKS : (nb == 0 ? 1u : nb);
}

hipError_t launch_blas(cl2_ctx* c, const BlasArgs& a, int prec, unsigned* nblocks_out)
{
    const unsigned nb = reduce_blocks(a.n);
    if (nblocks_out) *nblocks_out = nb;
    if (prec == CL2_F64)
        hipLaunchKernelGGL((k_blas<double, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream, a);
    else
        hipLaunchKernelGGL((k_blas<float, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream, a);
    return hipGetLastError();
}

/* stage 1: per-block sums of conj(x) y (R_DOT), its real part (R_DOT_RE) or |x|^2 (R_NORM) */
template<typename R, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_reduce1(const vec2<R>* __restrict__ x, const vec2<R>* __restrict__ y,
                                                   size_t n, int mode, double* __restrict__ partials)
{
    __shared__ double lds[2 * BLOCK / 64];
    double2 acc       = make_double2(0., 0.);
    const size_t step = (size_t)gridDim.x * BLOCK;
    if (mode == R_NORM) {
        for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < n; k += step) {
            vec2<R> v = x[k];
            acc.x += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        }
    } else {
        for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < n; k += step) {
            vec2<R> u = x[k], v = y[k];
            acc.x += (double)u.x * (double)v.x + (double)u.y * (double)v.y;
            acc.y += (double)u.x * (double)v.y - (double)v.x * (double)u.y;
        }
    }
    double2 s = block_sum2<BLOCK>(acc, lds);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x]     = s.x;
        partials[2 * blockIdx.x + 1] = (mode == R_DOT) ? s.y : 0.;
    }
}

hipError_t launch_reduce_stage1(cl2_ctx* c, const void* x, const void* y, size_t n, int prec, int mode,
                                unsigned* nblocks)
{
    const unsigned nb = reduce_blocks(n);
    *nblocks          = nb;
    if (prec == CL2_F64)
        hipLaunchKernelGGL((k_reduce1<double, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream,
                           reinterpret_cast<const double2*>(x), reinterpret_cast<const double2*>(y), n, mode,
                           c->partials);
    else
        hipLaunchKernelGGL((k_reduce1<float, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream,
                           reinterpret_cast<const float2*>(x), reinterpret_cast<const float2*>(y), n, mode,
                           c->partials);
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
                                int nvec)
{
    hipLaunchKernelGGL((k_reduce2<BLAS_BLOCK>), dim3(1), dim3(BLAS_BLOCK), 0, c->stream, c->partials, nparts, res,
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
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const vec2<R> v = x[(size_t)k * S + e];
            acc += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        }
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

hipError_t launch_fill(cl2_ctx* c, void* out, size_t n, int prec, double re, double im)
{
    const unsigned nb = reduce_blocks(n);
    if (prec == CL2_F64)
        hipLaunchKernelGGL((k_fill<double, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream,
                           reinterpret_cast<double2*>(out), n, re, im);
    else
        hipLaunchKernelGGL((k_fill<float, BLAS_BLOCK>), dim3(nb), dim3(BLAS_BLOCK), 0, c->stream,
                           reinterpret_cast<float2*>(out), n, (float)re, (float)im);
    return hipGetLastError();
}

/* precision conversion between complex fields (FP64 <-> FP32) */
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

hipError_t launch_convert(cl2_ctx* c, void* out, int out_prec, const void* in, int in_prec, size_t n)
{
    const unsigned nb = reduce_blocks(n);
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

/* Multi-shift CG recurrences, one thread per shift (ocl_kernel/operations_real.cl:20-39):
 *   zeta:  (a b c) / (d e (b - a) + b c (1 - f d))
 *   beta:  -a b / c
 *   alpha: -a b c / (d e)
 * Pointer roles follow the launcher argument lists of real.cpp:285-400. */
__global__ void k_cgm(int which, double* out, const double* a, const double* b, const double* d, const double* e,
                      const double* f, const double* g, int numeq)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= numeq) return;
    if (which == 0) {  // zeta(zeta_prev, zeta_prev_prev, *sbeta_prev, *sbeta_pres, *salpha_prev, sigma)
        double A = a[k], B = b[k], C = d[0], D = e[0], E = f[0], F = g[k];
        out[k]   = (A * B * C) / (D * E * (B - A) + B * C * (1 - F * D));
    } else if (which == 1) {  // beta(*sbeta_pres, zeta_pres, zeta_prev)
        out[k] = -a[0] * b[k] / d[k];
    } else {  // alpha(*salpha_pres, zeta_pres, beta_pres, zeta_prev, *sbeta_pres)
        out[k] = -a[0] * b[k] * d[k] / (e[k] * f[0]);
    }
}

hipError_t launch_cgm_update(cl2_ctx* c, int which, double* out, const double* a, const double* b, const double* d,
                             const double* e, const double* f, const double* g, int numeq)
{
    const unsigned nb = blocks_for((size_t)numeq, 64);
    hipLaunchKernelGGL(k_cgm, dim3(nb ? nb : 1), dim3(64), 0, c->stream, which, out, a, b, d, e, f, g, numeq);
    return hipGetLastError();
}

}  // namespace cl2
