// Fused vector update of multi-shift CG (SolverShifted::updateVectorQuantities,
// physics/algorithms/solver_shifted.cpp:174-198): for every active shift k
//   x_k <- beta_k p_k + x_k,   p_k <- zeta_k r + alpha_k p_k
// in one pass over the lattice, so r is read once and each x_k, p_k once, instead
// of two saxpy launches per shift that re-read r and p_k.
#include "cl2_device.hpp"
#include "cl2_internal.hpp"

namespace cl2 {

/* RX: precision of the solutions x_k (may be higher than R, the precision of r and p_k, so that a
 * reduced-precision recurrence still accumulates its solutions in FP64) */
template<typename RX, typename R>
__global__ __launch_bounds__(256) void k_multishift(void* const* __restrict__ xs, void* const* __restrict__ ps,
                                                    const vec2<R>* __restrict__ r, size_t n, int nshift,
                                                    const double* __restrict__ beta, const double* __restrict__ zeta,
                                                    const double* __restrict__ alpha, const double* __restrict__ active)
{
    const size_t step = (size_t)gridDim.x * 256;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += step) {
        const vec2<R> rv = r[e];
        for (int k = 0; k < nshift; ++k) {
            if (active[k] == 0.) continue;
            vec2<RX>* x      = reinterpret_cast<vec2<RX>*>(xs[k]);
            vec2<R>* p       = reinterpret_cast<vec2<R>*>(ps[k]);
            const RX b       = RX(beta[k]);
            const R z = R(zeta[k]), a = R(alpha[k]);
            const vec2<R> pv = p[e];
            const vec2<RX> xv = x[e];
            x[e]             = vec2<RX>{b * RX(pv.x) + xv.x, b * RX(pv.y) + xv.y};
            p[e]             = vec2<R>{z * rv.x + a * pv.x, z * rv.y + a * pv.y};
        }
    }
}

hipError_t launch_multishift(cl2_ctx* c, void* const* xs_dev, int x_prec, void* const* ps_dev, const void* r, int prec,
                             size_t n, int nshift, const double* beta, const double* zeta, const double* alpha,
                             const double* active)
{
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
    if (prec == CL2_F64)
        hipLaunchKernelGGL((k_multishift<double, double>), dim3(nb), dim3(256), 0, c->stream, xs_dev, ps_dev,
                           reinterpret_cast<const double2*>(r), n, nshift, beta, zeta, alpha, active);
    else if (x_prec == CL2_F64)
        hipLaunchKernelGGL((k_multishift<double, float>), dim3(nb), dim3(256), 0, c->stream, xs_dev, ps_dev,
                           reinterpret_cast<const float2*>(r), n, nshift, beta, zeta, alpha, active);
    else
        hipLaunchKernelGGL((k_multishift<float, float>), dim3(nb), dim3(256), 0, c->stream, xs_dev, ps_dev,
                           reinterpret_cast<const float2*>(r), n, nshift, beta, zeta, alpha, active);
    return hipGetLastError();
}

}  // namespace cl2
