// Device-resident lattice objects and their BLAS (physics/lattices/: gaugefield.hpp,
// spinorfield.hpp, spinorfield_eo.hpp:48-313, staggeredfield_eo.hpp:49-258,
// scalar.hpp:35-105, vector.hpp, util.hpp:64-91).  Every object owns one buffer on
// the System's GPU in the layout of include/cl2qcd.h; every call enqueues on the
// System's stream.  Output arguments are const pointers with shallow const, as in
// the reference.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../hardware/system.hpp"
#include "../interfacesHandler.hpp"

namespace physics {
namespace lattices {

/* device allocation owned by a lattice object */
class DeviceBuffer {
  public:
    DeviceBuffer(const hardware::System& system, size_t bytes);
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void* ptr() const noexcept { return p; }
    size_t bytes() const noexcept { return n; }

  private:
    const hardware::System& system;
    void* p = nullptr;
    size_t n = 0;
};

/* Scalar<T> / Vector<T>: device scalars (physics/lattices/scalar.hpp:35-105) */
template<typename T>
class Scalar {
  public:
    explicit Scalar(const hardware::System& system);
    void store(const T& val) const;
    T get() const;
    double* dev() const noexcept { return static_cast<double*>(buf.ptr()); }
    const hardware::System& get_system() const noexcept { return system; }

  private:
    const hardware::System& system;
    DeviceBuffer buf;
};

template<typename T>
class Vector {
  public:
    Vector(size_t n, const hardware::System& system);
    void store(const std::vector<T>& vals) const;
    std::vector<T> get() const;
    size_t get_vector_size() const noexcept { return n; }
    double* dev() const noexcept { return static_cast<double*>(buf.ptr()); }
    const hardware::System& get_system() const noexcept { return system; }

  private:
    const hardware::System& system;
    size_t n;
    DeviceBuffer buf;
};

class Gaugefield {
  public:
    Gaugefield(const hardware::System& system, const GaugefieldParametersInterface* parameters, const PRNG& prng);
    Gaugefield(const hardware::System& system, const GaugefieldParametersInterface* parameters, const PRNG& prng,
               bool hot);
    Gaugefield(const hardware::System& system, const GaugefieldParametersInterface* parameters, const PRNG& prng,
               std::string ildgfile);
    /* links from a host AoS array (link index mu + 4*site, row-major 3x3) */
    void import(const Matrixsu3* host) const;
    /* ILDG/LIME file of the links with trajectory number `number` in its xlf-info record
     * (physics/lattices/gaugefield.cpp:113-127; precision 64, or 32 if the run parameters say so) */
    void save(const std::string& outputfile, int number) const;
    std::vector<Matrixsu3> get_host() const;
    /* device pointer and link format of the resident FP64 field */
    const void* dev() const noexcept { return buf->ptr(); }
    /* device pointer for in-place updates; drops the reduced-precision copies */
    void* dev_mutable()
    {
        buf_f32.reset();
        buf_f16.reset();
        return buf->ptr();
    }
    int format() const noexcept { return fmt; }
    /* FP32 / FP16 12-real copies for reduced-precision operators, made on first use */
    const void* dev_f32() const;
    const void* dev_f16() const;
    const hardware::System& get_system() const noexcept { return system; }

  private:
    void initializeHotOrCold(bool hot, const PRNG& prng);
    void initializeFromILDGSourcefile(const std::string& file);
    const hardware::System& system;
    int fmt;
    std::unique_ptr<DeviceBuffer> buf;
    mutable std::unique_ptr<DeviceBuffer> buf_f32, buf_f16;
    const void* reduced_copy(std::unique_ptr<DeviceBuffer>& cache, int fmt) const;
};

/* full-lattice spinor field, host resident: source of convert_to_eoprec */
class Spinorfield {
  public:
    Spinorfield(const hardware::System& system, const SpinorfieldParametersInterface& p);
    unsigned get_elements() const noexcept { return (unsigned)host.size(); }
    void import(const spinor* h) const;
    std::vector<spinor> get_host() const { return host; }
    const hardware::System& get_system() const noexcept { return system; }

  private:
    friend void convert_to_eoprec(const class Spinorfield_eo*, const class Spinorfield_eo*, const Spinorfield&);
    friend void convert_from_eoprec(const Spinorfield*, const class Spinorfield_eo&, const class Spinorfield_eo&);
    const hardware::System& system;
    mutable std::vector<spinor> host;
};

class Spinorfield_eo {
  public:
    Spinorfield_eo(const hardware::System& system, const SpinorfieldEoParametersInterface& p, int prec = CL2_F64);
    Spinorfield_eo(const hardware::System& system, int prec = CL2_F64);
    void gamma5() const;
    void zero() const;
    void cold() const;
    /* Gaussian field, N(0, 1) per real component, seeded from the PRNG (spinorfield_eo.hpp:93) */
    void gaussian(const PRNG& prng) const;
    void gaussian_from_seed(uint64_t seed) const;
    unsigned get_elements() const noexcept;
    void import(const spinor* host) const;
    std::vector<spinor> get_host() const;
    void* dev() const noexcept { return buf.ptr(); }
    int precision() const noexcept { return prec; }
    const hardware::System& get_system() const noexcept { return system; }

  private:
    const hardware::System& system;
    int prec;
    DeviceBuffer buf;
};

class Staggeredfield_eo {
  public:
    Staggeredfield_eo(const hardware::System& system, const StaggeredfieldEoParametersInterface& p,
                      int prec = CL2_F64);
    Staggeredfield_eo(const hardware::System& system, int prec = CL2_F64);
    void set_zero() const;
    void set_cold() const;
    /* Gaussian field, N(0, 1) per real component, seeded from the PRNG (staggeredfield_eo.hpp:87) */
    void set_gaussian(const PRNG& prng) const;
    void gaussian_from_seed(uint64_t seed) const;
    unsigned get_elements() const noexcept;
    void import(const su3vec* host) const;
    std::vector<su3vec> get_host() const;
    void* dev() const noexcept { return buf.ptr(); }
    int precision() const noexcept { return prec; }
    const hardware::System& get_system() const noexcept { return system; }

  private:
    const hardware::System& system;
    int prec;
    DeviceBuffer buf;
};

/* nrhs Wilson eo spinors in one device array [12][Veo][nrhs] (right-hand side fastest), the operand of
 * the multi-right-hand-side operators; no counterpart in the reference, which holds one field per source */
class Spinorfield_eo_block {
  public:
    Spinorfield_eo_block(const hardware::System& system, int nrhs, int prec = CL2_F64);
    void insert(int r, const Spinorfield_eo& f) const;
    void extract(int r, const Spinorfield_eo* f) const;
    void zero() const;
    void* dev() const noexcept { return buf.ptr(); }
    int precision() const noexcept { return prec; }
    int nrhs() const noexcept { return n; }
    size_t bytes() const noexcept { return buf.bytes(); }
    const hardware::System& get_system() const noexcept { return system; }

  private:
    const hardware::System& system;
    int n, prec;
    DeviceBuffer buf;
};

/* nrhs staggered eo vectors in the multi-RHS layout [3][Veo][nrhs] (right-hand side fastest) */
class Staggeredfield_eo_block {
  public:
    Staggeredfield_eo_block(const hardware::System& system, int nrhs, int prec = CL2_F64);
    void insert(int r, const Staggeredfield_eo& f) const;
    void extract(int r, const Staggeredfield_eo* f) const;
    void zero() const;
    void* dev() const noexcept { return buf.ptr(); }
    int precision() const noexcept { return prec; }
    int nrhs() const noexcept { return n; }
    size_t bytes() const noexcept { return buf.bytes(); }
    const hardware::System& get_system() const noexcept { return system; }

  private:
    const hardware::System& system;
    int n, prec;
    DeviceBuffer buf;
};

/* ---- Wilson eo BLAS (physics/lattices/spinorfield_eo.hpp:161-313) ---- */
void copyData(const Spinorfield_eo* to, const Spinorfield_eo& from);
void copyData(const Spinorfield_eo* to, const Spinorfield_eo* from);
hmc_complex scalar_product(const Spinorfield_eo& left, const Spinorfield_eo& right);
void scalar_product(const Scalar<hmc_complex>* res, const Spinorfield_eo& left, const Spinorfield_eo& right);
hmc_float squarenorm(const Spinorfield_eo& field);
void squarenorm(const Scalar<hmc_float>* res, const Spinorfield_eo& field);
/* out = y - alpha x */
void saxpy(const Spinorfield_eo* out, const hmc_complex alpha, const Spinorfield_eo& x, const Spinorfield_eo& y);
void saxpy(const Spinorfield_eo* out, const Scalar<hmc_complex>& alpha, const Spinorfield_eo& x,
           const Spinorfield_eo& y);
void sax(const Spinorfield_eo* out, const hmc_complex alpha, const Spinorfield_eo& x);
void sax(const Spinorfield_eo* out, const Scalar<hmc_complex>& alpha, const Spinorfield_eo& x);
/* out = alpha x + beta y + z */
void saxsbypz(const Spinorfield_eo* out, const hmc_complex alpha, const Spinorfield_eo& x, const hmc_complex beta,
              const Spinorfield_eo& y, const Spinorfield_eo& z);
void saxsbypz(const Spinorfield_eo* out, const Scalar<hmc_complex>& alpha, const Spinorfield_eo& x,
              const Scalar<hmc_complex>& beta, const Spinorfield_eo& y, const Spinorfield_eo& z);
void saxpy_AND_squarenorm(const Spinorfield_eo* out, const Scalar<hmc_complex>& alpha, const Spinorfield_eo& x,
                          const Spinorfield_eo& y, const Scalar<hmc_complex>& squarenorm);
/* out = y - alpha x; res[0..3] = (|out|^2, Re <out, y_old>, |out|^2, 0): the fused update of the
 * reduced-precision CG */
void saxpy_AND_cgnorm(const Spinorfield_eo* out, const Scalar<hmc_complex>& alpha, const Spinorfield_eo& x,
                      const Spinorfield_eo& y, const Vector<hmc_float>& res);
/* the x and p updates of a CG step in one pass (Wilson saxpy convention): x <- x - malpha p, then
 * p <- r - mbeta p, both from the same p (x, p, r distinct) */
void cg_update_xp(const Spinorfield_eo* x, const Spinorfield_eo* p, const Spinorfield_eo& r,
                  const Scalar<hmc_complex>& malpha, const Scalar<hmc_complex>& mbeta);
void saxpy_AND_gamma5_eo(const Spinorfield_eo* out, const hmc_complex alpha, const Spinorfield_eo& x,
                         const Spinorfield_eo& y);
void convert_to_eoprec(const Spinorfield_eo* even, const Spinorfield_eo* odd, const Spinorfield& in);
void convert_from_eoprec(const Spinorfield* merged, const Spinorfield_eo& even, const Spinorfield_eo& odd);
void log_squarenorm(const std::string& msg, const Spinorfield_eo& x);
/* copy between precisions (or a plain copy when they agree) */
void convert_precision(const Spinorfield_eo* to, const Spinorfield_eo& from);
void convert_precision(const Staggeredfield_eo* to, const Staggeredfield_eo& from);
/* staggered fields, same convention and updates as the Wilson overload */
void cg_update_xp(const Staggeredfield_eo* x, const Staggeredfield_eo* p, const Staggeredfield_eo& r,
                  const Scalar<hmc_float>& malpha, const Scalar<hmc_float>& mbeta);

/* ---- staggered eo BLAS (physics/lattices/staggeredfield_eo.hpp:120-258) ---- */
void copyData(const Staggeredfield_eo* to, const Staggeredfield_eo& from);
void copyData(const Staggeredfield_eo* to, const Staggeredfield_eo* from);
hmc_complex scalar_product(const Staggeredfield_eo& left, const Staggeredfield_eo& right);
void scalar_product(const Scalar<hmc_complex>* res, const Staggeredfield_eo& left, const Staggeredfield_eo& right);
hmc_float scalar_product_real_part(const Staggeredfield_eo& left, const Staggeredfield_eo& right);
void scalar_product_real_part(const Scalar<hmc_float>* res, const Staggeredfield_eo& left,
                              const Staggeredfield_eo& right);
hmc_float squarenorm(const Staggeredfield_eo& field);
void squarenorm(const Scalar<hmc_float>* res, const Staggeredfield_eo& field);
void sax(const Staggeredfield_eo* out, const hmc_complex alpha, const Staggeredfield_eo& x);
void sax(const Staggeredfield_eo* out, const Scalar<hmc_complex>& alpha, const Staggeredfield_eo& x);
void sax(const Staggeredfield_eo* out, const hmc_float alpha, const Staggeredfield_eo& x);
void sax(const Staggeredfield_eo* out, const Scalar<hmc_float>& alpha, const Staggeredfield_eo& x);
void sax(const Staggeredfield_eo* out, const Vector<hmc_float>& alpha, const int index_alpha,
         const Staggeredfield_eo& x);
/* out = alpha x + y */
void saxpy(const Staggeredfield_eo* out, const hmc_complex alpha, const Staggeredfield_eo& x,
           const Staggeredfield_eo& y);
void saxpy(const Staggeredfield_eo* out, const Scalar<hmc_complex>& alpha, const Staggeredfield_eo& x,
           const Staggeredfield_eo& y);
void saxpy(const Staggeredfield_eo* out, const Scalar<hmc_float>& alpha, const Staggeredfield_eo& x,
           const Staggeredfield_eo& y);
void saxpy(const Staggeredfield_eo* out, const hmc_float alpha, const Staggeredfield_eo& x,
           const Staggeredfield_eo& y);
void saxpy(const Staggeredfield_eo* out, const Vector<hmc_float>& alpha, const int index_alpha,
           const Staggeredfield_eo& x, const Staggeredfield_eo& y);
/* out = alpha x + beta y */
void saxpby(const Staggeredfield_eo* out, const hmc_complex alpha, const Staggeredfield_eo& x, const hmc_complex beta,
            const Staggeredfield_eo& y);
void saxpby(const Staggeredfield_eo* out, const Scalar<hmc_complex>& alpha, const Staggeredfield_eo& x,
            const Scalar<hmc_complex>& beta, const Staggeredfield_eo& y);
void saxpby(const Staggeredfield_eo* out, const hmc_float alpha, const Staggeredfield_eo& x, const hmc_float beta,
            const Staggeredfield_eo& y);
void saxpby(const Staggeredfield_eo* out, const Scalar<hmc_float>& alpha, const Staggeredfield_eo& x,
            const Scalar<hmc_float>& beta, const Staggeredfield_eo& y);
void saxpby(const Staggeredfield_eo* out, const Vector<hmc_float>& alpha, const int index_alpha,
            const Staggeredfield_eo& x, const Vector<hmc_float>& beta, const int index_beta,
            const Staggeredfield_eo& y);
void saxpbypz(const Staggeredfield_eo* out, const hmc_complex alpha, const Staggeredfield_eo& x, const hmc_complex beta,
              const Staggeredfield_eo& y, const Staggeredfield_eo& z);
void saxpbypz(const Staggeredfield_eo* out, const Scalar<hmc_complex>& alpha, const Staggeredfield_eo& x,
              const Scalar<hmc_complex>& beta, const Staggeredfield_eo& y, const Staggeredfield_eo& z);
/* res[k] = |alpha[k] x|^2 */
void sax_vec_and_squarenorm(const Vector<hmc_float>* res, const Vector<hmc_float>& alpha, const Staggeredfield_eo& x);

/* ---- device scalar arithmetic (physics/lattices/algebra_real.hpp, scalar.hpp) ---- */
void add(const Scalar<hmc_complex>* dest, const Scalar<hmc_complex>& left, const Scalar<hmc_complex>& right);
void subtract(const Scalar<hmc_complex>* dest, const Scalar<hmc_complex>& minuend,
              const Scalar<hmc_complex>& subtrahend);
void multiply(const Scalar<hmc_complex>* dest, const Scalar<hmc_complex>& left, const Scalar<hmc_complex>& right);
void divide(const Scalar<hmc_complex>* dest, const Scalar<hmc_complex>& numerator,
            const Scalar<hmc_complex>& denominator);
void copyData(const Scalar<hmc_complex>* to, const Scalar<hmc_complex>& from);
void add(const Scalar<hmc_float>* dest, const Scalar<hmc_float>& left, const Scalar<hmc_float>& right);
void subtract(const Scalar<hmc_float>* dest, const Scalar<hmc_float>& minuend, const Scalar<hmc_float>& subtrahend);
void multiply(const Scalar<hmc_float>* dest, const Scalar<hmc_float>& left, const Scalar<hmc_float>& right);
void divide(const Scalar<hmc_float>* dest, const Scalar<hmc_float>& numerator, const Scalar<hmc_float>& denominator);
void copyData(const Scalar<hmc_float>* to, const Scalar<hmc_float>& from);
void copyData(const Vector<hmc_float>* to, const Vector<hmc_float>& from);
void access_real_vector_element(const Scalar<hmc_float>* out, const Vector<hmc_float>& in, const int index);
void access_real_vector_element(const Vector<hmc_float>* out, const Scalar<hmc_float>& in, const int index);
/* CG-M recurrences (physics/lattices/algebra_real.hpp) */
void update_zeta_cgm(const Vector<hmc_float>* out, const Vector<hmc_float>& zeta_prev,
                     const Vector<hmc_float>& zeta_prev_prev, const Scalar<hmc_float>& sbeta_prev,
                     const Scalar<hmc_float>& sbeta_pres, const Scalar<hmc_float>& salpha_prev,
                     const Vector<hmc_float>& sigma, const int numeq);
void update_beta_cgm(const Vector<hmc_float>* out, const Scalar<hmc_float>& sbeta_pres,
                     const Vector<hmc_float>& zeta_pres, const Vector<hmc_float>& zeta_prev, const int numeq);
void update_alpha_cgm(const Vector<hmc_float>* out, const Scalar<hmc_float>& salpha_pres,
                      const Vector<hmc_float>& zeta_pres, const Vector<hmc_float>& beta_pres,
                      const Vector<hmc_float>& zeta_prev, const Scalar<hmc_float>& sbeta_pres, const int numeq);

/* deterministic fills (physics/lattices/util.hpp:83-91, meta/type_ops.cpp:29-79) */
void pseudo_randomize(const Spinorfield* to, int seed);
void pseudo_randomize(const Spinorfield_eo* to, int seed);
void pseudo_randomize(const Staggeredfield_eo* to, int seed);
template<class Lattice, typename Basetype>
void pseudo_randomize(const Lattice* to, int seed)
{
    pseudo_randomize(to, seed);
}

}  // namespace lattices
}  // namespace physics
