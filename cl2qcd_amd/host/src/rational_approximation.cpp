// Minimax rational approximations for the multi-shift solver's shifts: Remez exchange in quad precision.
//
// Reference interface: physics/algorithms/rational_approximation.{hpp,cpp} (Rational_Coefficients,
// Rational_Approximation, Rescale_Coefficients, the text file format); the reference's generator is the
// Remez code of alg_remez.cpp on GMP floats of `precision` bits (113 by default).  This restatement is its own:
//
//   * form: r(x) = N(x) / D(x) in barycentric form, N = sum_j alpha_j / (x - z_j), D = sum_j beta_j / (x - z_j),
//     a rational function of type (d, d) for d + 1 support points z_j inside [low, high] (one between each
//     pair of neighbouring reference points, the geometric mean); the Cauchy-like columns keep the linear
//     systems well conditioned where a monomial or Chebyshev basis in x is not (reference points spread
//     over decades near low);
//   * levelled relative error on N = 2d + 2 reference points: N(x_i) - f(x_i) D(x_i) = s_i delta f(x_i) D(x_i)
//     with s_i = (-1)^i and sum_j beta_j = 1, made linear by taking the D on the right from the previous
//     solve and iterating to a fixed point;
//   * exchange: the zeros of e(x) = r(x) / f(x) - 1 between neighbouring points split [low, high] into N
//     segments; the new points are the extrema of s_i e on each segment (golden section in log x), until
//     the extrema agree to 1e-15 relative (the minimax error then lies between the smallest and the
//     largest of them), or stop improving below 1e-11;
//   * continuation in the degree: degree k starts from the alternation points of degree k - 1, resampled
//     in log x;
//   * partial fractions: the d real zeros -b_k of D (all below low) by sign changes on a logarithmic grid
//     and bisection, a0 = sum alpha / sum beta, a_k = N(-b_k) / D'(-b_k).
// All arithmetic is __float128 (113-bit significand, the reference's default precision).
#include "../include/physics/algorithms/rational_approximation.hpp"

#include <quadmath.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <ostream>
#include <stdexcept>
#include <string>

namespace physics {
namespace algorithms {

namespace {

using Q = __float128;

struct Interval {
    Q lo, hi;
    Q t(Q x) const { return (2 * x - hi - lo) / (hi - lo); }
};

/* A x = rhs by Gaussian elimination with partial pivoting (A is n x n, row-major, destroyed) */
std::vector<Q> solve(std::vector<Q> A, std::vector<Q> rhs, int n)
{
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (fabsq(A[i * n + k]) > fabsq(A[piv * n + k])) piv = i;
        if (A[piv * n + k] == 0) throw std::runtime_error("Rational_Approximation: singular Remez system");
        if (piv != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[piv * n + j]);
            std::swap(rhs[k], rhs[piv]);
        }
        for (int i = k + 1; i < n; ++i) {
            const Q m = A[i * n + k] / A[k * n + k];
            if (m == 0) continue;
            for (int j = k; j < n; ++j) A[i * n + j] -= m * A[k * n + j];
            rhs[i] -= m * rhs[k];
        }
    }
    std::vector<Q> x(n);
    for (int i = n - 1; i >= 0; --i) {
        Q s = rhs[i];
        for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * x[j];
        x[i] = s / A[i * n + i];
    }
    return x;
}

struct Remez {
    int d;
    Q e;  // exponent of f(x) = x^e
    Interval I;
    std::vector<Q> z, alpha, beta;  // support points and barycentric weights (d + 1 each)
    Q f(Q x) const { return powq(x, e); }
    /* N(x), D(x); at a support point the limits alpha_j / beta_j of the ratio are taken by the caller */
    void nd(Q x, Q& n, Q& dd) const
    {
        n = dd = 0;
        for (size_t j = 0; j < z.size(); ++j) {
            const Q c = 1 / (x - z[j]);
            n += alpha[j] * c;
            dd += beta[j] * c;
        }
    }
    Q r(Q x) const
    {
        for (size_t j = 0; j < z.size(); ++j)
            if (x == z[j]) return alpha[j] / beta[j];
        Q n, dd;
        nd(x, n, dd);
        return n / dd;
    }
    Q err(Q x) const { return r(x) / f(x) - 1; }
};

/* the levelled solve on the reference points x (support points between neighbours): alpha, beta and delta */
Q level(Remez& R, const std::vector<Q>& x)
{
    const int d = R.d, N = 2 * d + 2, M = N + 1;  // unknowns alpha (d + 1), beta (d + 1), delta
    R.z.assign(d + 1, 0);
    for (int j = 0; j <= d; ++j) R.z[j] = sqrtq(x[2 * j] * x[2 * j + 1]);
    std::vector<Q> dprev(N, 1);
    Q delta = 0;
    for (int it = 0; it < 200; ++it) {
        std::vector<Q> A((size_t)M * M, 0), rhs(M, 0);
        for (int i = 0; i < N; ++i) {
            const Q fi = R.f(x[i]), si = (i & 1) ? -1 : 1;
            for (int j = 0; j <= d; ++j) {
                const Q c                         = 1 / (x[i] - R.z[j]);
                A[(size_t)i * M + j]              = c;
                A[(size_t)i * M + d + 1 + j]      = -fi * c;
            }
            A[(size_t)i * M + 2 * d + 2] = -si * fi * dprev[i];
        }
        for (int j = 0; j <= d; ++j) A[(size_t)N * M + d + 1 + j] = 1;  // sum beta = 1
        rhs[N] = 1;
        const std::vector<Q> s = solve(A, rhs, M);
        R.alpha.assign(s.begin(), s.begin() + d + 1);
        R.beta.assign(s.begin() + d + 1, s.begin() + 2 * d + 2);
        const Q dn = s[2 * d + 2];
        Q change   = 0;
        for (int i = 0; i < N; ++i) {
            Q n, dd;
            R.nd(x[i], n, dd);
            change   = fmaxq(change, fabsq(dd / dprev[i] - 1));
            dprev[i] = dd;
        }
        const bool done = change < (Q)1e-28 && fabsq(dn - delta) <= (Q)1e-26 * fabsq(dn);
        delta           = dn;
        if (done) break;
    }
    return delta;
}

/* zero of g on (a, b) where g(a), g(b) differ in sign, by bisection in log x */
template<class G>
Q zero_log(const G& g, Q a, Q b)
{
    Q ga = g(a);
    Q la = logq(a), lb = logq(b);
    for (int it = 0; it < 160; ++it) {
        const Q lm = (la + lb) / 2, m = expq(lm), gm = g(m);
        if ((gm > 0) == (ga > 0)) {
            la = lm;
            ga = gm;
        } else {
            lb = lm;
        }
    }
    return expq((la + lb) / 2);
}

/* maximiser of g on [a, b] by golden section in log x, compared with the ends */
template<class G>
Q max_log(const G& g, Q a, Q b)
{
    const Q r = (sqrtq((Q)5.0) - 1) / 2;
    Q la = logq(a), lb = logq(b);
    Q l1 = lb - r * (lb - la), l2 = la + r * (lb - la);
    Q g1 = g(expq(l1)), g2 = g(expq(l2));
    for (int it = 0; it < 200 && lb - la > (Q)1e-30 * (fabsq(la) + 1); ++it) {
        if (g1 < g2) {
            la = l1;
            l1 = l2;
            g1 = g2;
            l2 = la + r * (lb - la);
            g2 = g(expq(l2));
        } else {
            lb = l2;
            l2 = l1;
            g2 = g1;
            l1 = lb - r * (lb - la);
            g1 = g(expq(l1));
        }
    }
    Q best = expq((la + lb) / 2), gb = g(best);
    if (g(a) > gb) best = a, gb = g(a);
    if (g(b) > gb) best = b;
    return best;
}

}  // namespace

Rational_Coefficients::Rational_Coefficients(int d) : _d(d), _a0(0.), _a(d, 0.), _b(d, 0.)
{
    if (d < 0) throw std::invalid_argument("Rational_Coefficients: negative order");
}

Rational_Coefficients::Rational_Coefficients(int d, hmc_float a0, const std::vector<hmc_float>& a,
                                             const std::vector<hmc_float>& b)
    : _d(d), _a0(a0), _a(a), _b(b)
{
    if (d < 0 || (int)a.size() != d || (int)b.size() != d)
        throw std::invalid_argument("Rational_Coefficients: order and coefficient vectors differ");
}

void Rational_Coefficients::Set_coeff(hmc_float a0, const std::vector<hmc_float>& a, const std::vector<hmc_float>& b)
{
    if ((int)a.size() != _d || (int)b.size() != _d) throw std::invalid_argument("Rational_Coefficients: wrong sizes");
    _a0 = a0;
    _a  = a;
    _b  = b;
}

hmc_float Rational_Coefficients::evaluate(hmc_float x) const
{
    hmc_float r = _a0;
    for (int k = 0; k < _d; ++k) r += _a[k] / (x + _b[k]);
    return r;
}

Rational_Approximation::Rational_Approximation(int d, int y_, int z_, hmc_float low_, hmc_float high_, bool inv_,
                                               int precision_)
    : Rational_Coefficients(d), y(y_), z(z_), precision(precision_), low(low_), high(high_), inv(inv_)
{
    if (d < 1) throw std::invalid_argument("Rational_Approximation: degree must be at least 1");
    if (y <= 0 || z <= 0 || y % z == 0) throw std::invalid_argument("Rational_Approximation: y, z > 0 and y % z != 0");
    if (!(low > 0 && high > low)) throw std::invalid_argument("Rational_Approximation: need 0 < low < high");
    Remez R;
    R.e = (inv ? -1 : 1) * (Q)y / (Q)z;
    R.I = Interval{(Q)low, (Q)high};
    const Q L0 = logq(R.I.lo), L1 = logq(R.I.hi), PI = acosq((Q)-1);
    // continuation in the degree: degree k starts from the alternation points of degree k - 1, resampled
    // (piecewise linearly in log x) from 2k to 2k + 2 points; degree 1 from Chebyshev nodes in log x
    std::vector<Q> x(4);
    for (int i = 0; i < 4; ++i) x[i] = expq(L0 + (L1 - L0) * (1 - cosq(PI * i / 3)) / 2);
    Q emax = 0;
    iters  = 0;
    for (int k = 1; k <= d; ++k) {
        const int N = 2 * k + 2;
        if (k > 1) {
            const int n = (int)x.size();
            std::vector<Q> nx(N);
            for (int j = 0; j < N; ++j) {
                const Q s  = (Q)j * (n - 1) / (N - 1);
                const int i = std::min((int)s, n - 2);
                const Q w  = s - i;
                nx[j]      = expq((1 - w) * logq(x[i]) + w * logq(x[i + 1]));
            }
            x = nx;
        }
        x[0]     = R.I.lo;
        x[N - 1] = R.I.hi;
        R.d      = k;
        Q spread_prev = 1;
        int stalled = 0;
        bool converged = false;
        for (int it = 1; it <= 200; ++it) {
            ++iters;
            const Q delta = level(R, x);
            const Q sgn   = delta >= 0 ? 1 : -1;
            // zeros of the error between neighbouring points, then the extrema of s_i e on the segments
            std::vector<Q> zs(N - 1);
            for (int i = 0; i + 1 < N; ++i) zs[i] = zero_log([&](Q v) { return R.err(v); }, x[i], x[i + 1]);
            std::vector<Q> nx(N);
            Q emin = (Q)1e300;
            emax   = 0;
            for (int i = 0; i < N; ++i) {
                const Q a = i == 0 ? R.I.lo : zs[i - 1], b = i == N - 1 ? R.I.hi : zs[i];
                const Q si = ((i & 1) ? -1 : 1) * sgn;
                nx[i]      = max_log([&](Q v) { return si * R.err(v); }, a, b);
                const Q ev = fabsq(R.err(nx[i]));
                emin       = fminq(emin, ev);
                emax       = fmaxq(emax, ev);
            }
            if (!(emax < (Q)1)) break;  // a pole inside the interval: no recovery from here
            x              = nx;
            const Q spread = (emax - emin) / emax;
            if (spread <= (Q)1e-15) {
                converged = true;
                break;
            }
            stalled     = spread >= (Q)0.5 * spread_prev ? stalled + 1 : 0;
            spread_prev = spread;
            if (stalled >= 4 && spread <= (Q)1e-11) {
                converged = true;
                break;
            }
        }
        if (!converged)
            throw std::runtime_error("Rational_Approximation: the Remez exchange did not converge at degree " +
                                     std::to_string(k) + " (largest error " + std::to_string((double)emax) +
                                     ", spread " + std::to_string((double)spread_prev) + ")");
    }
    error = (hmc_float)emax;
    // partial fractions: the zeros -b_k of D lie below low; bracket them on a log grid of |x|
    auto Dx = [&](Q v) {
        Q n, dd;
        R.nd(v, n, dd);
        return dd;
    };
    std::vector<Q> roots;
    for (int per = 400; per <= 51200 && (int)roots.size() != d; per *= 4) {
        roots.clear();
        const Q m0 = logq(R.I.lo) - 12 * logq((Q)10.0), m1 = logq(R.I.hi) + 12 * logq((Q)10.0);
        const int n = (int)((m1 - m0) / logq((Q)10.0) * per);
        Q prev_u = m0, prev = Dx(-expq(m0));
        for (int k = 1; k <= n; ++k) {
            const Q u = m0 + (m1 - m0) * k / n, v = Dx(-expq(u));
            if ((v > 0) != (prev > 0)) {
                Q la = prev_u, lb = u, ga = prev;
                for (int it = 0; it < 160; ++it) {
                    const Q lm = (la + lb) / 2, gm = Dx(-expq(lm));
                    if ((gm > 0) == (ga > 0)) {
                        la = lm;
                        ga = gm;
                    } else {
                        lb = lm;
                    }
                }
                roots.push_back(expq((la + lb) / 2));
            }
            prev_u = u;
            prev   = v;
        }
    }
    if ((int)roots.size() != d) throw std::runtime_error("Rational_Approximation: denominator zeros not all real");
    std::sort(roots.begin(), roots.end());
    std::vector<hmc_float> a(d), b(d);
    Q sa = 0, sb = 0;
    for (int j = 0; j <= d; ++j) sa += R.alpha[j], sb += R.beta[j];
    for (int k = 0; k < d; ++k) {
        const Q p = -roots[k];
        Q n = 0, dd = 0;  // N(p) and D'(p) = -sum beta_j / (p - z_j)^2
        for (int j = 0; j <= d; ++j) {
            const Q c = 1 / (p - R.z[j]);
            n += R.alpha[j] * c;
            dd -= R.beta[j] * c * c;
        }
        a[k] = (hmc_float)(n / dd);
        b[k] = (hmc_float)roots[k];
    }
    Set_coeff((hmc_float)(sa / sb), a, b);
}

Rational_Approximation::Rational_Approximation(const std::string& filename) : Rational_Coefficients(0)
{
    std::ifstream in(filename);
    if (!in) throw std::invalid_argument("Unable to open the file with Rational_Approximation data!");
    int d = 0, inv_i = 1;
    hmc_float a0 = 0.;
    in >> d >> y >> z >> low >> high >> inv_i >> precision >> error >> a0;
    if (!in || d < 1) throw std::invalid_argument("Rational_Approximation: malformed file " + filename);
    inv = inv_i != 0;
    std::vector<hmc_float> a(d), b(d);
    for (int k = 0; k < d; ++k) in >> a[k] >> b[k];
    if (!in) throw std::invalid_argument("Rational_Approximation: malformed file " + filename);
    static_cast<Rational_Coefficients&>(*this) = Rational_Coefficients(d, a0, a, b);
}

void Rational_Approximation::Save_rational_approximation(const std::string& filename) const
{
    std::ofstream out(filename);
    if (!out) throw std::invalid_argument("Unable to open " + filename);
    out << std::setprecision(16) << Get_order() << "     " << y << " " << z << "     " << low << " " << high << "     "
        << (inv ? 1 : 0) << " " << precision << "     " << error << "\n\n"
        << Get_a0() << "\n";
    const std::vector<hmc_float> a = Get_a(), b = Get_b();
    for (size_t k = 0; k < a.size(); ++k) out << a[k] << "\t\t" << b[k] << "\n";
}

Rational_Coefficients Rational_Approximation::Rescale_Coefficients(hmc_float minEigenvalue,
                                                                   hmc_float maxEigenvalue) const
{
    if (high != 1) throw std::invalid_argument("Upper bound different from 1 in rescale_coefficients!");
    if (low > minEigenvalue / maxEigenvalue)
        throw std::invalid_argument("Rational_Approximation does not respect lower_bound <= min/max");
    // r(x / max) approximates (x / max)^e, so max^e r(x / max) approximates x^e:
    // a0 -> max^e a0, a_k / (x/max + b_k) -> max^(e+1) a_k / (x + max b_k)
    const hmc_float s = std::pow(maxEigenvalue, Get_exponent());
    std::vector<hmc_float> a = Get_a(), b = Get_b();
    for (size_t k = 0; k < a.size(); ++k) {
        a[k] *= s * maxEigenvalue;
        b[k] *= maxEigenvalue;
    }
    return Rational_Coefficients((int)a.size(), Get_a0() * s, a, b);
}

std::ostream& operator<<(std::ostream& os, const Rational_Approximation& r)
{
    os << "Rational approximation of x^(" << (r.inv ? "-" : "") << r.y << "/" << r.z << ") on [" << r.low << ", "
       << r.high << "], degree " << r.Get_order() << ", relative error " << r.error << "\n  a0 = " << r.Get_a0();
    const std::vector<hmc_float> a = r.Get_a(), b = r.Get_b();
    for (size_t k = 0; k < a.size(); ++k) os << "\n  a[" << k << "] = " << a[k] << "   b[" << k << "] = " << b[k];
    return os;
}

}  // namespace algorithms
}  // namespace physics
