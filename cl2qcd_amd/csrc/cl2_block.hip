// Level-1 operations on blocks of nrhs Wilson eo spinors stored [12][Veo][nrhs]
// (right-hand side fastest, the layout of the multi-RHS hopping kernel), with
// one scalar per right-hand side, and the per-right-hand-side scalar
// recurrences of a lockstep CG over the block.
//
// Reductions are per right-hand side and deterministic: a thread always works on
// one right-hand side (block sizes and grid strides are multiples of nrhs), a
// block sums each right-hand side's column of its threads in a fixed order, and
// stage 2 reduces each right-hand side's block partials with one block each.
#include "cl2_device.hpp"
#include "cl2_internal.hpp"

namespace cl2 {

namespace {

/* threads per block: the largest multiple of nrhs not above 256 */
inline unsigned block_threads(int nrhs)
{
    return 256u - (256u % (unsigned)nrhs);
}

/* per-right-hand-side column sums of a block's thread values, valid in threads 0..nrhs-1 */
__device__ __forceinline__ double2 column_sum(double2 v, double2* lds, int nrhs)
{
    lds[threadIdx.x] = v;
    __syncthreads();
    double2 s = make_double2(0., 0.);
    if ((int)threadIdx.x < nrhs)
        for (unsigned k = threadIdx.x; k < blockDim.x; k += nrhs) {
            s.x += lds[k].x;
            s.y += lds[k].y;
        }
    __syncthreads();
    return s;
}

}  // namespace

/* out = y + sign * c_r x  (c: nrhs complex on the device); with norms: per right-hand side
 * (sum |out|^2, sum Re conj(out) y) into partials[r * nb + block] */
template<typename R>
__global__ __launch_bounds__(256) void k_block_axpy(vec2<R>* out, const vec2<R>* x, const vec2<R>* y, size_t n,
                                                    int nrhs, const double* coef, double sign, double2* partials)
{
    __shared__ double2 lds[256];
    const int r   = (int)(threadIdx.x % (unsigned)nrhs);
    const Cx<R> c = Cx<R>{R(sign * coef[2 * r]), R(sign * coef[2 * r + 1])};
    double2 acc   = make_double2(0., 0.);
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += step) {
        const vec2<R> xv = x[e], yv = y[e];
        const Cx<R> o    = Cx<R>{yv.x, yv.y} + cmul(c, Cx<R>{xv.x, xv.y});
        out[e]           = vec2<R>{o.re, o.im};
        acc.x += (double)o.re * (double)o.re + (double)o.im * (double)o.im;
        acc.y += (double)o.re * (double)yv.x + (double)o.im * (double)yv.y;
    }
    if (partials) {
        const double2 s = column_sum(acc, lds, nrhs);
        if ((int)threadIdx.x < nrhs) partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

/* per right-hand side: sum conj(x) y (mode 0) or sum |x|^2 (mode 1) */
template<typename R>
__global__ __launch_bounds__(256) void k_block_dot(const vec2<R>* x, const vec2<R>* y, size_t n, int nrhs, int mode,
                                                   double2* partials)
{
    __shared__ double2 lds[256];
    double2 acc       = make_double2(0., 0.);
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += step) {
        const vec2<R> u = x[e];
        if (mode == 1) {
            acc.x += (double)u.x * (double)u.x + (double)u.y * (double)u.y;
        } else {
            const vec2<R> v = y[e];
            acc.x += (double)u.x * (double)v.x + (double)u.y * (double)v.y;
            acc.y += (double)u.x * (double)v.y - (double)v.x * (double)u.y;
        }
    }
    const double2 s = column_sum(acc, lds, nrhs);
    if ((int)threadIdx.x < nrhs) partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
}

/* stage 2: block r sums right-hand side r's nb partials; out_kind 0: res[2r..2r+1] = complex sum,
 * 1: res[r] = real part */
__global__ __launch_bounds__(256) void k_block_reduce2(const double2* partials, unsigned nb, double* res, int out_kind)
{
    __shared__ double lds[8];
    const double2* p = partials + (size_t)blockIdx.x * nb;
    double2 acc      = make_double2(0., 0.);
    for (unsigned k = threadIdx.x; k < nb; k += 256) {
        acc.x += p[k].x;
        acc.y += p[k].y;
    }
    const double2 s = block_sum2<256>(acc, lds);
    if (threadIdx.x == 0) {
        if (out_kind == 0) {
            res[2 * blockIdx.x]     = s.x;
            res[2 * blockIdx.x + 1] = s.y;
        } else {
            res[blockIdx.x] = s.x;
        }
    }
}

/* copy single field <-> slot r of a block */
template<typename R>
__global__ __launch_bounds__(256) void k_block_slot(vec2<R>* block, const vec2<R>* single, vec2<R>* single_out,
                                                    size_t n_single, int nrhs, int r)
{
    const size_t step = (size_t)gridDim.x * 256;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n_single; e += step) {
        if (single)
            block[e * nrhs + r] = single[e];
        else
            single_out[e] = block[e * nrhs + r];
    }
}

/* lockstep CG scalars: alpha_r = r2_r / <p, A p>_r for active right-hand sides, else 0 */
__global__ void k_block_cg_alpha(double* alpha, const double* r2, const double* pAp, const double* active, int nrhs)
{
    const int r = threadIdx.x;
    if (r >= nrhs) return;
    double ar = 0., ai = 0.;
    if (active[r] != 0.) {
        const double a = pAp[2 * r], b = pAp[2 * r + 1], d = a * a + b * b;
        ar = r2[r] * a / d;
        ai = -r2[r] * b / d;
    }
    alpha[2 * r]     = ar;
    alpha[2 * r + 1] = ai;
}

/* beta_r = Polak-Ribiere <r_new, r_new - r_old> / |r_old|^2 (Fletcher-Reeves when that is negative) for
 * active right-hand sides, else 0; then r2_r = |r_new|^2.  norms = (|r_new|^2, Re <r_new, r_old>) per r. */
__global__ void k_block_cg_beta(double* beta, double* r2, const double* norms, const double* active, int nrhs)
{
    const int r = threadIdx.x;
    if (r >= nrhs) return;
    const double rn = norms[2 * r], ry = norms[2 * r + 1];
    double b        = (rn - ry) / r2[r];
    if (!(b >= 0.)) b = rn / r2[r];
    beta[2 * r]     = active[r] != 0. ? b : 0.;
    beta[2 * r + 1] = 0.;
    if (active[r] != 0.) r2[r] = rn;
}

/* grid of a block kernel: with per-right-hand-side partials at most REDUCE_MAX_BLOCKS blocks (and what the
 * partials buffer holds), else enough blocks to stream without a grid-stride tail */
/* single-RHS CG: mbeta = -(Polak-Ribiere beta) as a complex scalar, then r2 = (|r_new|^2, 0);
 * norms = (|r_new|^2, Re <r_new, r_old>) */
__global__ void k_cg_beta(double* mbeta, double* r2, const double* norms)
{
    if (threadIdx.x != 0) return;
    const double rn = norms[0], ry = norms[1];
    double b        = (rn - ry) / r2[0];
    if (!(b >= 0.)) b = rn / r2[0];
    mbeta[0] = -b;
    mbeta[1] = 0.;
    r2[0]    = rn;
    r2[1]    = 0.;
}

hipError_t launch_cg_beta(cl2_ctx* c, double* mbeta, double* r2, const double* norms)
{
    hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(64), 0, c->stream, mbeta, r2, norms);
    return hipGetLastError();
}

/* copy of a few device doubles as a kernel (so it can sit in a recorded iteration) */
__global__ void k_scalar_copy(double* dst, const double* src, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

hipError_t launch_scalar_copy(cl2_ctx* c, double* dst, const double* src, int n)
{
    hipLaunchKernelGGL(k_scalar_copy, dim3((n + 63) / 64), dim3(64), 0, c->stream, dst, src, n);
    return hipGetLastError();
}

static unsigned grid_for(cl2_ctx* c, size_t n, int nrhs, bool partials)
{
    const unsigned bt = block_threads(nrhs);
    size_t nb         = (n + bt - 1) / bt;
    const size_t cap  = partials ? std::min<size_t>((size_t)REDUCE_MAX_BLOCKS, c->npartials / (size_t)nrhs)
                                 : (size_t)65536;
    if (nb > cap) nb = cap;
    return nb == 0 ? 1u : (unsigned)nb;
}

hipError_t launch_block_axpy(cl2_ctx* c, void* out, const void* x, const void* y, int prec, int nrhs,
                             const double* coef, double sign, double* norms)
{
    const size_t n    = (size_t)12 * c->g.veo * nrhs;
    const unsigned nb = grid_for(c, n, nrhs, norms != nullptr), bt = block_threads(nrhs);
    double2* parts    = norms ? reinterpret_cast<double2*>(c->partials) : nullptr;
    if (prec == CL2_F64)
        hipLaunchKernelGGL(k_block_axpy<double>, dim3(nb), dim3(bt), 0, c->stream, reinterpret_cast<double2*>(out),
                           reinterpret_cast<const double2*>(x), reinterpret_cast<const double2*>(y), n, nrhs, coef,
                           sign, parts);
    else
        hipLaunchKernelGGL(k_block_axpy<float>, dim3(nb), dim3(bt), 0, c->stream, reinterpret_cast<float2*>(out),
                           reinterpret_cast<const float2*>(x), reinterpret_cast<const float2*>(y), n, nrhs, coef, sign,
                           parts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !norms) return e;
    hipLaunchKernelGGL(k_block_reduce2, dim3(nrhs), dim3(256), 0, c->stream, parts, nb, norms, 0);
    return hipGetLastError();
}

hipError_t launch_block_dot(cl2_ctx* c, double* res, const void* x, const void* y, int prec, int nrhs, int mode)
{
    const size_t n    = (size_t)12 * c->g.veo * nrhs;
    const unsigned nb = grid_for(c, n, nrhs, true), bt = block_threads(nrhs);
    double2* parts    = reinterpret_cast<double2*>(c->partials);
    if (prec == CL2_F64)
        hipLaunchKernelGGL(k_block_dot<double>, dim3(nb), dim3(bt), 0, c->stream, reinterpret_cast<const double2*>(x),
                           reinterpret_cast<const double2*>(y), n, nrhs, mode, parts);
    else
        hipLaunchKernelGGL(k_block_dot<float>, dim3(nb), dim3(bt), 0, c->stream, reinterpret_cast<const float2*>(x),
                           reinterpret_cast<const float2*>(y), n, nrhs, mode, parts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_block_reduce2, dim3(nrhs), dim3(256), 0, c->stream, parts, nb, res, mode == 1 ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_block_slot(cl2_ctx* c, void* block, const void* single_in, void* single_out, int prec, int nrhs,
                             int r)
{
    const size_t n    = (size_t)12 * c->g.veo;
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    if (prec == CL2_F64)
        hipLaunchKernelGGL(k_block_slot<double>, dim3(nb), dim3(256), 0, c->stream, reinterpret_cast<double2*>(block),
                           reinterpret_cast<const double2*>(single_in), reinterpret_cast<double2*>(single_out), n, nrhs,
                           r);
    else
        hipLaunchKernelGGL(k_block_slot<float>, dim3(nb), dim3(256), 0, c->stream, reinterpret_cast<float2*>(block),
                           reinterpret_cast<const float2*>(single_in), reinterpret_cast<float2*>(single_out), n, nrhs,
                           r);
    return hipGetLastError();
}

hipError_t launch_block_cg_scalars(cl2_ctx* c, int which, double* out, double* r2, const double* in,
                                   const double* active, int nrhs)
{
    if (which == 0)
        hipLaunchKernelGGL(k_block_cg_alpha, dim3(1), dim3(64), 0, c->stream, out, r2, in, active, nrhs);
    else
        hipLaunchKernelGGL(k_block_cg_beta, dim3(1), dim3(64), 0, c->stream, out, r2, in, active, nrhs);
    return hipGetLastError();
}

}  // namespace cl2
