/* Rational approximations r(x) = a0 + sum_k a_k / (x + b_k) of x^(-y/z) or x^(y/z) for the shifts of the
 * multi-shift solver (physics/algorithms/rational_approximation.hpp:38-221 in the reference, whose
 * generator is the Remez algorithm of alg_remez.cpp on 113-bit GMP floats).
 *
 * Here the minimax approximation (relative error) is found by a Remez exchange carried out in IEEE quad
 * precision (__float128, the same 113-bit significand), with numerator and denominator represented in the
 * Chebyshev basis of the interval for conditioning; the partial fractions come from the real roots of the
 * denominator, located outside the interval.  Host-side only: it produces the shifts b_k that the GPU
 * solvers take. */
#pragma once

#include <iosfwd>
#include <string>
#include <vector>

#include "../../common_types.hpp"

namespace physics {
namespace algorithms {

/* a0 + sum_{k<d} a[k] / (x + b[k]) (rational_approximation.hpp:38-111) */
class Rational_Coefficients {
  public:
    explicit Rational_Coefficients(int d);
    Rational_Coefficients(int d, hmc_float a0, const std::vector<hmc_float>& a, const std::vector<hmc_float>& b);
    Rational_Coefficients(Rational_Coefficients&&)            = default;
    Rational_Coefficients& operator=(Rational_Coefficients&&) = default;
    Rational_Coefficients(const Rational_Coefficients&)       = delete;
    Rational_Coefficients& operator=(const Rational_Coefficients&) = delete;
    virtual ~Rational_Coefficients() = default;

    unsigned int Get_order() const { return (unsigned int)_d; }
    hmc_float Get_a0() const { return _a0; }
    std::vector<hmc_float> Get_a() const { return _a; }
    std::vector<hmc_float> Get_b() const { return _b; }
    /* r(x) in double precision */
    hmc_float evaluate(hmc_float x) const;

  protected:
    void Set_coeff(hmc_float a0, const std::vector<hmc_float>& a, const std::vector<hmc_float>& b);

  private:
    int _d;
    hmc_float _a0;
    std::vector<hmc_float> _a, _b;
};

/* The minimax rational approximation of degree (d, d) to x^(-y/z) (inv) or x^(y/z) on [low, high]
 * (rational_approximation.hpp:112-221).  y, z > 0 and y % z != 0. */
class Rational_Approximation : public Rational_Coefficients {
  public:
    Rational_Approximation(int d, int y, int z, hmc_float low, hmc_float high, bool inv = true, int precision = 113);
    /* read back what Save_rational_approximation wrote: "d y z low high inv precision error", a blank line,
     * a0, then one "a_k b_k" line per term (the reference's text format, rational_approximation.cpp:139-167) */
    explicit Rational_Approximation(const std::string& filename);

    hmc_float Get_error() const { return error; }
    hmc_float Get_lower_bound() const { return low; }
    hmc_float Get_upper_bound() const { return high; }
    hmc_float Get_exponent() const { return inv ? -(hmc_float)y / z : (hmc_float)y / z; }
    void Save_rational_approximation(const std::string& filename) const;
    /* the coefficients for an operator with spectrum in [minEigenvalue, maxEigenvalue]: the approximation must
     * live on [low, 1] with low <= min/max; x -> x / max (rational_approximation.cpp:195-219) */
    Rational_Coefficients Rescale_Coefficients(hmc_float minEigenvalue, hmc_float maxEigenvalue) const;
    /* Remez iterations used */
    int iterations() const { return iters; }

    friend std::ostream& operator<<(std::ostream&, const Rational_Approximation&);

  private:
    int y = 1, z = 2, precision = 113, iters = 0;
    hmc_float low = 0., high = 1.;
    bool inv = true;
    hmc_float error = 0.;
};

}  // namespace algorithms
}  // namespace physics
