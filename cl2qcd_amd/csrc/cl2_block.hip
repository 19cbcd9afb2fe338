// Level-1 operations on blocks of nrhs Wilson eo spinors stored [12][Veo][nrhs]
// (FP64) or [6][Veo][nrhs] of component pairs (FP32), right-hand side fastest,
// the layout of the multi-RHS hopping kernel, and on blocks of staggered
// fields [3][Veo][nrhs], with one scalar per right-hand side, and the
// per-right-hand-side scalar recurrences of a lockstep CG over the block.
//
// Reductions are per right-hand side and deterministic: a thread always works on
// one right-hand side (block sizes and grid strides are multiples of nrhs), a
// block sums each right-hand side's column of its threads in a fixed order, and
// stage 2 reduces each right-hand side's block partials with one block each.
#include <initializer_list>

#include "cl2_device.hpp"
#include "cl2_internal.hpp"

namespace cl2 {

/* Elements are streamed in 16-byte chunks: one FP64 complex number, or two FP32 numbers.  For FP32 Wilson
 * blocks a chunk is one component pair of one right-hand side ("pairs": both slots belong to right-hand
 * side q, the chunk index modulo ncol = nrhs); for FP32 staggered blocks it holds two adjacent right-hand
 * sides when nrhs is even or 1 and the operands are 16-byte aligned (slot j of chunk column q belongs to
 * right-hand side (q CPC + j) mod nrhs, ncol = max(1, nrhs / CPC)).  A thread keeps one column. */
template<typename R, int CPC>
struct BChunk;
template<>
struct BChunk<double, 1> {
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
struct BChunk<float, 2> {
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
struct BChunk<float, 1> {
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

namespace {

/* right-hand side of slot j of chunk column q */
template<int CPC>
__device__ __forceinline__ int slot_rhs(int q, int j, int nrhs, bool pairs)
{
    return pairs ? q : (q * CPC + j) % nrhs;
}

/* per-right-hand-side column sums of a block's thread values (CPC slots per thread), valid in threads
 * 0..nrhs-1 */
template<int CPC>
__device__ __forceinline__ double2 column_sum(const double2 (&v)[CPC], double2* lds, int nrhs, int ncol, bool pairs)
{
#pragma unroll
    for (int j = 0; j < CPC; ++j) lds[threadIdx.x * CPC + j] = v[j];
    __syncthreads();
    double2 s   = make_double2(0., 0.);
    const int t = (int)threadIdx.x;
    if (t < nrhs) {
        if (pairs) {  // column t is right-hand side t, both slots
            for (unsigned k = (unsigned)t; k < blockDim.x; k += (unsigned)ncol)
                for (int j = 0; j < CPC; ++j) {
                    s.x += lds[k * CPC + j].x;
                    s.y += lds[k * CPC + j].y;
                }
        } else if (nrhs % CPC == 0) {
            const int j = t % CPC;
            for (unsigned k = (unsigned)(t / CPC); k < blockDim.x; k += (unsigned)ncol) {
                s.x += lds[k * CPC + j].x;
                s.y += lds[k * CPC + j].y;
            }
        } else {  // nrhs == 1 < CPC: every slot is right-hand side 0
            for (unsigned k = 0; k < blockDim.x * CPC; ++k) {
                s.x += lds[k].x;
                s.y += lds[k].y;
            }
        }
    }
    __syncthreads();
    return s;
}

}  // namespace

/* out = y + sign * c_r x  (c: nrhs complex on the device); with norms: per right-hand side
 * (sum |out|^2, sum Re conj(out) y) into partials[r * nb + block].  U chunks per thread per step, their
 * loads issued first. */
template<typename R, int CPC, int U>
__global__ __launch_bounds__(256) void k_block_axpy(void* out, const void* x, const void* y, size_t nchunk, int nrhs,
                                                    int ncol, bool pairs, const double* coef, double sign,
                                                    double2* partials)
{
    using IO = BChunk<R, CPC>;
    __shared__ double2 lds[256 * CPC];
    const int q = (int)(threadIdx.x % (unsigned)ncol);
    Cx<R> c[CPC];
    double2 acc[CPC];
#pragma unroll
    for (int j = 0; j < CPC; ++j) {
        const int r = slot_rhs<CPC>(q, j, nrhs, pairs);
        c[j]        = Cx<R>{R(sign * coef[2 * r]), R(sign * coef[2 * r + 1])};
        acc[j]      = make_double2(0., 0.);
    }
    const size_t bt   = blockDim.x;
    const size_t step = (size_t)gridDim.x * bt * U;
    for (size_t base = (size_t)blockIdx.x * bt * U + threadIdx.x; base < nchunk; base += step) {
        Cx<R> xv[U][CPC], yv[U][CPC];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t e = base + u * bt;
            if (e < nchunk) {
                IO::load(x, e, xv[u]);
                IO::load(y, e, yv[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t e = base + u * bt;
            if (e >= nchunk) continue;
            Cx<R> o[CPC];
#pragma unroll
            for (int j = 0; j < CPC; ++j) {
                o[j] = yv[u][j] + cmul(c[j], xv[u][j]);
                acc[j].x += (double)o[j].re * (double)o[j].re + (double)o[j].im * (double)o[j].im;
                acc[j].y += (double)o[j].re * (double)yv[u][j].re + (double)o[j].im * (double)yv[u][j].im;
            }
            IO::store(out, e, o);
        }
    }
    if (partials) {
        const double2 s = column_sum<CPC>(acc, lds, nrhs, ncol, pairs);
        if ((int)threadIdx.x < nrhs) partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

/* the x and p updates of a lockstep CG step in one pass: x = x + alpha_r p, then p = r + beta_r p */
template<typename R, int CPC>
__global__ __launch_bounds__(256) void k_block_xp(void* xo, void* p, const void* rv, size_t nchunk, int nrhs, int ncol,
                                                  bool pairs, const double* alpha, const double* beta)
{
    using IO = BChunk<R, CPC>;
    const int q = (int)(threadIdx.x % (unsigned)ncol);
    Cx<R> ca[CPC], cb[CPC];
#pragma unroll
    for (int j = 0; j < CPC; ++j) {
        const int r = slot_rhs<CPC>(q, j, nrhs, pairs);
        ca[j]       = Cx<R>{R(alpha[2 * r]), R(alpha[2 * r + 1])};
        cb[j]       = Cx<R>{R(beta[2 * r]), R(beta[2 * r + 1])};
    }
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nchunk; e += step) {
        Cx<R> xv[CPC], pv[CPC], rr[CPC], xn[CPC], pn[CPC];
        IO::load(xo, e, xv);
        IO::load(p, e, pv);
        IO::load(rv, e, rr);
#pragma unroll
        for (int j = 0; j < CPC; ++j) {
            xn[j] = xv[j] + cmul(ca[j], pv[j]);
            pn[j] = rr[j] + cmul(cb[j], pv[j]);
        }
        IO::store(xo, e, xn);
        IO::store(p, e, pn);
    }
}

/* per right-hand side: sum conj(x) y (mode 0) or sum |x|^2 (mode 1) */
template<typename R, int CPC, int U>
__global__ __launch_bounds__(256) void k_block_dot(const void* x, const void* y, size_t nchunk, int nrhs, int ncol,
                                                   bool pairs, int mode, double2* partials)
{
    using IO = BChunk<R, CPC>;
    __shared__ double2 lds[256 * CPC];
    double2 acc[CPC];
#pragma unroll
    for (int j = 0; j < CPC; ++j) acc[j] = make_double2(0., 0.);
    const size_t bt   = blockDim.x;
    const size_t step = (size_t)gridDim.x * bt * U;
    for (size_t base = (size_t)blockIdx.x * bt * U + threadIdx.x; base < nchunk; base += step) {
        Cx<R> u[U][CPC], v[U][CPC];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const size_t e = base + k * bt;
            if (e < nchunk) {
                IO::load(x, e, u[k]);
                if (mode != 1) IO::load(y, e, v[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            if (base + k * bt >= nchunk) continue;
#pragma unroll
            for (int j = 0; j < CPC; ++j) {
                if (mode == 1) {
                    acc[j].x += (double)u[k][j].re * (double)u[k][j].re + (double)u[k][j].im * (double)u[k][j].im;
                } else {
                    acc[j].x += (double)u[k][j].re * (double)v[k][j].re + (double)u[k][j].im * (double)v[k][j].im;
                    acc[j].y += (double)u[k][j].re * (double)v[k][j].im - (double)v[k][j].re * (double)u[k][j].im;
                }
            }
        }
    }
    const double2 s = column_sum<CPC>(acc, lds, nrhs, ncol, pairs);
    if ((int)threadIdx.x < nrhs) partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
}

/* stage 2: block r sums right-hand side r's nb partials; out_kind 0: res[2r..2r+1] = complex sum,
 * 1: res[r] = real part */
template<int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_block_reduce2(const double2* partials, unsigned nb, double* res,
                                                         int out_kind)
{
    __shared__ double lds[2 * BLOCK / 64];
    const double2* p = partials + (size_t)blockIdx.x * nb;
    double2 acc      = make_double2(0., 0.);
    for (unsigned k = threadIdx.x; k < nb; k += BLOCK) {
        acc.x += p[k].x;
        acc.y += p[k].y;
    }
    const double2 s = block_sum2<BLOCK>(acc, lds);
    if (threadIdx.x == 0) {
        if (out_kind == 0) {
            res[2 * blockIdx.x]     = s.x;
            res[2 * blockIdx.x + 1] = s.y;
        } else {
            res[blockIdx.x] = s.x;
        }
    }
}

/* copy single field <-> slot r of a block, in units of T (one complex number, or one FP32 component pair) */
template<typename T>
__global__ __launch_bounds__(256) void k_block_slot(T* block, const T* single, T* single_out, size_t n_single, int nrhs,
                                                    int r)
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

/* the staggered fused CG step: pAp holds (|D p_r|^2, |p_r|^2) per right-hand side on entry and the complex
 * <p_r, (m^2 + sigma_r - D D) p_r> = coef_r |p_r|^2 + |D p_r|^2 (D_KS anti-hermitian) on exit; then alpha as
 * k_block_cg_alpha */
struct Coef16 {
    double v[16];
};
__global__ void k_block_cg_alpha_mdagm(double* alpha, const double* r2, double* pAp, Coef16 coef, const double* active,
                                       int nrhs)
{
    const int r = threadIdx.x;
    if (r >= nrhs) return;
    const double a = coef.v[r] * pAp[2 * r + 1] + pAp[2 * r];
    pAp[2 * r]     = a;
    pAp[2 * r + 1] = 0.;
    double ar = 0.;
    if (active[r] != 0.) ar = r2[r] * a / (a * a);
    alpha[2 * r]     = ar;
    alpha[2 * r + 1] = 0.;
}

hipError_t launch_block_cg_alpha_mdagm(cl2_ctx* c, double* alpha, const double* r2, double* pAp, const double* coef,
                                       const double* active, int nrhs)
{
    Coef16 k{};
    for (int r = 0; r < nrhs && r < 16; ++r) k.v[r] = coef[r];
    hipLaunchKernelGGL(k_block_cg_alpha_mdagm, dim3(1), dim3(64), 0, c->stream, alpha, r2, pAp, k, active, nrhs);
    return hipGetLastError();
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

/* launch shape of a block kernel over n complex elements: CPC, chunk columns, threads per block (the largest
 * multiple of the column count not above 256) and blocks (U chunks per thread; with per-right-hand-side
 * partials at most what the partials buffer holds, beyond that the grid loops) */
struct BlockShape {
    int cpc, ncol;
    bool pairs;
    unsigned bt, nb;
    size_t nchunk;
};

static bool aligned16(const void* p)
{
    return p == nullptr || (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

/* FP32 Wilson blocks (ncomp 12) are arrays of component pairs: always 16-byte chunks, one right-hand side each
 * (the caller has checked the 16-byte alignment) */
static BlockShape block_shape(cl2_ctx* c, size_t n, int prec, int nrhs, int u, bool partials,
                              std::initializer_list<const void*> ptrs, int ncomp)
{
    BlockShape b;
    b.pairs   = prec != CL2_F64 && ncomp == 12;
    bool wide = prec != CL2_F64 && (b.pairs || ((nrhs % 2 == 0 || nrhs == 1) && n % 2 == 0));
    for (const void* p : ptrs) wide = wide && (b.pairs || aligned16(p));
    b.cpc    = wide ? 2 : 1;
    b.ncol   = b.pairs ? nrhs : (nrhs % b.cpc == 0 ? nrhs / b.cpc : 1);
    b.bt     = 256u - (256u % (unsigned)b.ncol);
    b.nchunk = n / (size_t)b.cpc;
    size_t nb = (b.nchunk + (size_t)b.bt * u - 1) / ((size_t)b.bt * u);
    const size_t cap = partials ? c->npartials / (size_t)nrhs : ((size_t)1 << 22);
    if (nb > cap) nb = cap;
    b.nb = nb == 0 ? 1u : (unsigned)nb;
    return b;
}

constexpr int BLOCK_U_REDUCE = 4;

hipError_t launch_block_reduce2(cl2_ctx* c, const double2* parts, unsigned nb, double* res, int out_kind, int nrhs)
{
    hipLaunchKernelGGL((k_block_reduce2<1024>), dim3(nrhs), dim3(1024), 0, c->stream, parts, nb, res, out_kind);
    return hipGetLastError();
}

hipError_t launch_block_axpy(cl2_ctx* c, void* out, const void* x, const void* y, int prec, int nrhs,
                             const double* coef, double sign, double* norms, int ncomp)
{
    const size_t n = (size_t)ncomp * c->g.veo * nrhs;
    const int u    = norms ? BLOCK_U_REDUCE : 1;
    const BlockShape b = block_shape(c, n, prec, nrhs, u, norms != nullptr, {out, x, y}, ncomp);
    double2* parts     = norms ? reinterpret_cast<double2*>(c->partials) : nullptr;
    if (prec == CL2_F64) {
        if (norms)
            hipLaunchKernelGGL((k_block_axpy<double, 1, BLOCK_U_REDUCE>), dim3(b.nb), dim3(b.bt), 0, c->stream, out, x,
                               y, b.nchunk, nrhs, b.ncol, b.pairs, coef, sign, parts);
        else
            hipLaunchKernelGGL((k_block_axpy<double, 1, 1>), dim3(b.nb), dim3(b.bt), 0, c->stream, out, x, y, b.nchunk,
                               nrhs, b.ncol, b.pairs, coef, sign, parts);
    } else if (b.cpc == 2) {
        if (norms)
            hipLaunchKernelGGL((k_block_axpy<float, 2, BLOCK_U_REDUCE>), dim3(b.nb), dim3(b.bt), 0, c->stream, out, x,
                               y, b.nchunk, nrhs, b.ncol, b.pairs, coef, sign, parts);
        else
            hipLaunchKernelGGL((k_block_axpy<float, 2, 1>), dim3(b.nb), dim3(b.bt), 0, c->stream, out, x, y, b.nchunk,
                               nrhs, b.ncol, b.pairs, coef, sign, parts);
    } else {
        if (norms)
            hipLaunchKernelGGL((k_block_axpy<float, 1, BLOCK_U_REDUCE>), dim3(b.nb), dim3(b.bt), 0, c->stream, out, x,
                               y, b.nchunk, nrhs, b.ncol, b.pairs, coef, sign, parts);
        else
            hipLaunchKernelGGL((k_block_axpy<float, 1, 1>), dim3(b.nb), dim3(b.bt), 0, c->stream, out, x, y, b.nchunk,
                               nrhs, b.ncol, b.pairs, coef, sign, parts);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !norms) return e;
    return launch_block_reduce2(c, parts, b.nb, norms, 0, nrhs);
}

hipError_t launch_block_xp(cl2_ctx* c, void* x, void* p, const void* r, int prec, int nrhs, const double* alpha,
                           const double* beta, int ncomp)
{
    const size_t n     = (size_t)ncomp * c->g.veo * nrhs;
    const BlockShape b = block_shape(c, n, prec, nrhs, 1, false, {x, p, r}, ncomp);
    if (prec == CL2_F64)
        hipLaunchKernelGGL((k_block_xp<double, 1>), dim3(b.nb), dim3(b.bt), 0, c->stream, x, p, r, b.nchunk, nrhs,
                           b.ncol, b.pairs, alpha, beta);
    else if (b.cpc == 2)
        hipLaunchKernelGGL((k_block_xp<float, 2>), dim3(b.nb), dim3(b.bt), 0, c->stream, x, p, r, b.nchunk, nrhs,
                           b.ncol, b.pairs, alpha, beta);
    else
        hipLaunchKernelGGL((k_block_xp<float, 1>), dim3(b.nb), dim3(b.bt), 0, c->stream, x, p, r, b.nchunk, nrhs,
                           b.ncol, b.pairs, alpha, beta);
    return hipGetLastError();
}

hipError_t launch_block_dot(cl2_ctx* c, double* res, const void* x, const void* y, int prec, int nrhs, int mode,
                            int ncomp)
{
    const size_t n     = (size_t)ncomp * c->g.veo * nrhs;
    const BlockShape b = block_shape(c, n, prec, nrhs, BLOCK_U_REDUCE, true, {x, y}, ncomp);
    double2* parts     = reinterpret_cast<double2*>(c->partials);
    if (prec == CL2_F64)
        hipLaunchKernelGGL((k_block_dot<double, 1, BLOCK_U_REDUCE>), dim3(b.nb), dim3(b.bt), 0, c->stream, x, y,
                           b.nchunk, nrhs, b.ncol, b.pairs, mode, parts);
    else if (b.cpc == 2)
        hipLaunchKernelGGL((k_block_dot<float, 2, BLOCK_U_REDUCE>), dim3(b.nb), dim3(b.bt), 0, c->stream, x, y,
                           b.nchunk, nrhs, b.ncol, b.pairs, mode, parts);
    else
        hipLaunchKernelGGL((k_block_dot<float, 1, BLOCK_U_REDUCE>), dim3(b.nb), dim3(b.bt), 0, c->stream, x, y,
                           b.nchunk, nrhs, b.ncol, b.pairs, mode, parts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_block_reduce2(c, parts, b.nb, res, mode == 1 ? 1 : 0, nrhs);
}

hipError_t launch_block_slot(cl2_ctx* c, void* block, const void* single_in, void* single_out, int prec, int nrhs,
                             int r, int ncomp)
{
    const size_t n    = (size_t)ncomp * c->g.veo;
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    if (prec == CL2_F64)
        hipLaunchKernelGGL(k_block_slot<double2>, dim3(nb), dim3(256), 0, c->stream, reinterpret_cast<double2*>(block),
                           reinterpret_cast<const double2*>(single_in), reinterpret_cast<double2*>(single_out), n, nrhs,
                           r);
    else if (ncomp == 12)  // FP32 Wilson: component pairs
        hipLaunchKernelGGL(k_block_slot<float4>, dim3(nb), dim3(256), 0, c->stream, reinterpret_cast<float4*>(block),
                           reinterpret_cast<const float4*>(single_in), reinterpret_cast<float4*>(single_out), n / 2,
                           nrhs, r);
    else
        hipLaunchKernelGGL(k_block_slot<float2>, dim3(nb), dim3(256), 0, c->stream, reinterpret_cast<float2*>(block),
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
