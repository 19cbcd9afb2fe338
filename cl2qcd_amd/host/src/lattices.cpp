// Lattice objects, BLAS and device scalars of the drop-in, on the C ABI.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>

#include "../include/ildg_io/ildgIo.hpp"
#include "../include/physics/lattices/lattices.hpp"

namespace physics {
namespace lattices {

namespace {

cl2_ctx* ctx(const hardware::System& s)
{
    return s.ctx();
}

cl2_scalar val(hmc_complex a)
{
    cl2_scalar s{};
    s.kind = CL2_SCALAR_VALUE;
    s.re   = a.re;
    s.im   = a.im;
    return s;
}
cl2_scalar val(hmc_float a)
{
    return val(hmc_complex{a, 0.});
}
cl2_scalar dev(const Scalar<hmc_complex>& a)
{
    cl2_scalar s{};
    s.kind = CL2_SCALAR_DEV_COMPLEX;
    s.dev  = a.dev();
    return s;
}
cl2_scalar dev(const Scalar<hmc_float>& a)
{
    cl2_scalar s{};
    s.kind = CL2_SCALAR_DEV_REAL;
    s.dev  = a.dev();
    return s;
}
cl2_scalar dev(const Vector<hmc_float>& a, int index)
{
    if (index < 0 || (size_t)index >= a.get_vector_size()) throw std::out_of_range("vector index out of range");
    cl2_scalar s{};
    s.kind  = CL2_SCALAR_DEV_VECTOR;
    s.dev   = a.dev();
    s.index = index;
    return s;
}

template<class F>
void same_prec(const F& a, const F& b)
{
    if (a.precision() != b.precision()) throw std::invalid_argument("fields of different precision");
}

/* site of parity par with site/2 == eo (ocl_kernel/operations_geometry.cl:305-395) */
size_t eo_site(int ns, size_t eo, int par)
{
    const size_t s = 2 * eo;
    size_t r       = s / ns;
    const size_t x = s % ns, y = r % ns;
    r /= ns;
    const size_t z = r % ns, t = r / ns;
    const int p    = (int)((x + y + z + t) & 1);
    return p == par ? s : s + 1;
}

}  // namespace

/* ---------------- buffers and scalars ---------------- */

DeviceBuffer::DeviceBuffer(const hardware::System& system, size_t bytes) : system(system), n(bytes)
{
    system.check(cl2_alloc(system.ctx(), &p, bytes));
}
DeviceBuffer::~DeviceBuffer()
{
    if (p) cl2_free(system.ctx(), p);
}

template<typename T>
Scalar<T>::Scalar(const hardware::System& system) : system(system), buf(system, sizeof(T))
{
}
template<typename T>
void Scalar<T>::store(const T& v) const
{
    system.check(cl2_scalar_store(ctx(system), dev(), reinterpret_cast<const double*>(&v), sizeof(T) / 8));
}
template<typename T>
T Scalar<T>::get() const
{
    T v;
    system.check(cl2_scalar_get(ctx(system), reinterpret_cast<double*>(&v), dev(), sizeof(T) / 8));
    return v;
}
template class Scalar<hmc_complex>;
template class Scalar<hmc_float>;

template<typename T>
Vector<T>::Vector(size_t n, const hardware::System& system) : system(system), n(n), buf(system, sizeof(T) * n)
{
}
template<typename T>
void Vector<T>::store(const std::vector<T>& v) const
{
    if (v.size() != n) throw std::invalid_argument("Vector::store: size mismatch");
    system.check(cl2_scalar_store(ctx(system), dev(), reinterpret_cast<const double*>(v.data()),
                                  (int)(n * sizeof(T) / 8)));
}
template<typename T>
std::vector<T> Vector<T>::get() const
{
    std::vector<T> v(n);
    system.check(cl2_scalar_get(ctx(system), reinterpret_cast<double*>(v.data()), dev(), (int)(n * sizeof(T) / 8)));
    return v;
}
template class Vector<hmc_float>;

/* ---------------- gauge field ---------------- */

Gaugefield::Gaugefield(const hardware::System& system, const GaugefieldParametersInterface* p, const PRNG& prng)
    : system(system), fmt(p->useReconstruct12() ? CL2_GAUGE_F64_12 : CL2_GAUGE_F64_18)
{
    buf.reset(new DeviceBuffer(system, cl2_field_bytes(system.ctx(), CL2_FIELD_GAUGE, fmt)));
    switch (p->getStartcondition()) {
        case common::start_from_source: initializeFromILDGSourcefile(p->getSourcefileName()); break;
        case common::hot_start: initializeHotOrCold(true, prng); break;
        default: initializeHotOrCold(false, prng); break;
    }
}
Gaugefield::Gaugefield(const hardware::System& system, const GaugefieldParametersInterface* p, const PRNG& prng,
                       bool hot)
    : system(system), fmt(p->useReconstruct12() ? CL2_GAUGE_F64_12 : CL2_GAUGE_F64_18)
{
    buf.reset(new DeviceBuffer(system, cl2_field_bytes(system.ctx(), CL2_FIELD_GAUGE, fmt)));
    initializeHotOrCold(hot, prng);
}
Gaugefield::Gaugefield(const hardware::System& system, const GaugefieldParametersInterface* p, const PRNG&,
                       std::string ildgfile)
    : system(system), fmt(p->useReconstruct12() ? CL2_GAUGE_F64_12 : CL2_GAUGE_F64_18)
{
    buf.reset(new DeviceBuffer(system, cl2_field_bytes(system.ctx(), CL2_FIELD_GAUGE, fmt)));
    initializeFromILDGSourcefile(ildgfile);
}

void Gaugefield::initializeHotOrCold(bool hot, const PRNG& prng)
{
    system.check(cl2_gauge_set_random(system.ctx(), buf->ptr(), fmt, hot ? prng.next_seed() : 0, hot ? 10.0 : 0.0));
}

void Gaugefield::import(const Matrixsu3* host) const
{
    system.check(cl2_gauge_import(system.ctx(), buf->ptr(), reinterpret_cast<const double*>(host), fmt));
    buf_f32.reset();
    buf_f16.reset();
}

std::vector<Matrixsu3> Gaugefield::get_host() const
{
    std::vector<Matrixsu3> h(4 * 2 * system.eo_sites());
    system.check(cl2_gauge_export(system.ctx(), reinterpret_cast<double*>(h.data()), buf->ptr(), fmt));
    return h;
}

const void* Gaugefield::reduced_copy(std::unique_ptr<DeviceBuffer>& cache, int rfmt) const
{
    if (!cache) {
        std::unique_ptr<DeviceBuffer> b(new DeviceBuffer(system, cl2_field_bytes(system.ctx(), CL2_FIELD_GAUGE, rfmt)));
        if (fmt == CL2_GAUGE_F64_18) {
            system.check(cl2_gauge_convert(system.ctx(), b->ptr(), rfmt, buf->ptr()));
        } else {
            std::vector<Matrixsu3> h = get_host();
            system.check(cl2_gauge_import(system.ctx(), b->ptr(), reinterpret_cast<const double*>(h.data()), rfmt));
        }
        cache = std::move(b);
    }
    return cache->ptr();
}
const void* Gaugefield::dev_f32() const
{
    return reduced_copy(buf_f32, CL2_GAUGE_F32_12);
}
const void* Gaugefield::dev_f16() const
{
    return reduced_copy(buf_f16, CL2_GAUGE_F16_12);
}

/* links of an ILDG/LIME file (ildg_io/ildgIo_gaugefield.cpp:44-67, 142-221), checksum verified */
void Gaugefield::initializeFromILDGSourcefile(const std::string& file)
{
    const meta::Inputparameters& p = system.get_inputparameters();
    const std::vector<Matrixsu3> host = ildgIo::readGaugefieldFromSourcefile(file, p.get_nspace(), p.get_ntime());
    system.check(cl2_gauge_import(system.ctx(), buf->ptr(), reinterpret_cast<const double*>(host.data()), fmt));
}

void Gaugefield::save(const std::string& outputfile, int number) const
{
    const meta::Inputparameters& p = system.get_inputparameters();
    ildgIo::XlfInfo info;
    info.trajectory = number;
    info.beta       = p.get_beta();
    info.kappa      = p.get_kappa();
    info.mu         = p.get_mu();
    info.mubar      = meta::get_mubar(p);
    std::vector<Matrixsu3> host = get_host();
    if (fmt == CL2_GAUGE_F64_18) {
        double plaq = 0.;
        system.check(cl2_gauge_plaquette(system.ctx(), buf->ptr(), &plaq));
        info.plaquette = plaq;
    }
    ildgIo::writeGaugefieldToFile(outputfile, host, p.get_nspace(), p.get_ntime(), info, p.get_precision());
}

/* ---------------- spinor fields ---------------- */

Spinorfield::Spinorfield(const hardware::System& system, const SpinorfieldParametersInterface& p)
    : system(system), host(p.getNumberOfElements())
{
}
void Spinorfield::import(const spinor* h) const
{
    std::memcpy(host.data(), h, host.size() * sizeof(spinor));
}

Spinorfield_eo::Spinorfield_eo(const hardware::System& system, const SpinorfieldEoParametersInterface&, int prec)
    : Spinorfield_eo(system, prec)
{
}
Spinorfield_eo::Spinorfield_eo(const hardware::System& system, int prec)
    : system(system), prec(prec), buf(system, cl2_field_bytes(system.ctx(), CL2_FIELD_SPINOR_EO, prec))
{
}
void Spinorfield_eo::gaussian(const PRNG& prng) const
{
    gaussian_from_seed(prng.next_seed());
}
void Spinorfield_eo::gaussian_from_seed(uint64_t seed) const
{
    if (prec == CL2_F16) throw std::invalid_argument("gaussian: FP64 or FP32 field");
    system.check(cl2_gaussian_eo(system.ctx(), CL2_FIELD_SPINOR_EO, buf.ptr(), prec, seed));
}
unsigned Spinorfield_eo::get_elements() const noexcept
{
    return (unsigned)system.eo_sites();
}
void Spinorfield_eo::gamma5() const
{
    system.check(cl2_gamma5_eo(ctx(system), dev(), prec));
}
void Spinorfield_eo::zero() const
{
    system.check(cl2_set_zero_eo(ctx(system), dev(), prec));
}
void Spinorfield_eo::cold() const
{
    system.check(cl2_set_cold_eo(ctx(system), dev(), prec));
}
void Spinorfield_eo::import(const spinor* host) const
{
    system.check(cl2_spinor_import(ctx(system), dev(), reinterpret_cast<const double*>(host), prec));
}
std::vector<spinor> Spinorfield_eo::get_host() const
{
    std::vector<spinor> h(get_elements());
    system.check(cl2_spinor_export(ctx(system), reinterpret_cast<double*>(h.data()), dev(), prec));
    return h;
}

Spinorfield_eo_block::Spinorfield_eo_block(const hardware::System& system, int nrhs, int prec)
    : system(system), n(nrhs), prec(prec),
      buf(system, (nrhs >= 1 && nrhs <= 64 && (prec == CL2_F64 || prec == CL2_F32 || prec == CL2_F16))
                      ? (size_t)nrhs * cl2_field_bytes(system.ctx(), CL2_FIELD_SPINOR_EO, prec)
                      : 0)
{
}
void Spinorfield_eo_block::insert(int r, const Spinorfield_eo& f) const
{
    if (f.precision() != prec) throw std::invalid_argument("block and field of different precision");
    system.check(cl2_block_insert_eo(system.ctx(), buf.ptr(), f.dev(), prec, n, r));
}
void Spinorfield_eo_block::extract(int r, const Spinorfield_eo* f) const
{
    if (f->precision() != prec) throw std::invalid_argument("block and field of different precision");
    system.check(cl2_block_extract_eo(system.ctx(), f->dev(), buf.ptr(), prec, n, r));
}
void Spinorfield_eo_block::zero() const
{
    system.check(cl2_block_set_zero_eo(system.ctx(), buf.ptr(), prec, n));
}

Staggeredfield_eo_block::Staggeredfield_eo_block(const hardware::System& system, int nrhs, int prec)
    : system(system), n(nrhs), prec(prec),
      buf(system, (nrhs >= 1 && nrhs <= 64 && (prec == CL2_F64 || prec == CL2_F32))
                      ? (size_t)nrhs * cl2_field_bytes(system.ctx(), CL2_FIELD_SU3VEC_EO, prec)
                      : 0)
{
}
void Staggeredfield_eo_block::insert(int r, const Staggeredfield_eo& f) const
{
    if (f.precision() != prec) throw std::invalid_argument("block and field of different precision");
    system.check(cl2_block_insert_stagg_eo(system.ctx(), buf.ptr(), f.dev(), prec, n, r));
}
void Staggeredfield_eo_block::extract(int r, const Staggeredfield_eo* f) const
{
    if (f->precision() != prec) throw std::invalid_argument("block and field of different precision");
    system.check(cl2_block_extract_stagg_eo(system.ctx(), f->dev(), buf.ptr(), prec, n, r));
}
void Staggeredfield_eo_block::zero() const
{
    system.check(cl2_block_set_zero_stagg_eo(system.ctx(), buf.ptr(), prec, n));
}

Staggeredfield_eo::Staggeredfield_eo(const hardware::System& system, const StaggeredfieldEoParametersInterface&,
                                     int prec)
    : Staggeredfield_eo(system, prec)
{
}
Staggeredfield_eo::Staggeredfield_eo(const hardware::System& system, int prec)
    : system(system), prec(prec), buf(system, cl2_field_bytes(system.ctx(), CL2_FIELD_SU3VEC_EO, prec))
{
}
void Staggeredfield_eo::set_gaussian(const PRNG& prng) const
{
    gaussian_from_seed(prng.next_seed());
}
void Staggeredfield_eo::gaussian_from_seed(uint64_t seed) const
{
    system.check(cl2_gaussian_eo(system.ctx(), CL2_FIELD_SU3VEC_EO, buf.ptr(), prec, seed));
}
unsigned Staggeredfield_eo::get_elements() const noexcept
{
    return (unsigned)system.eo_sites();
}
void Staggeredfield_eo::set_zero() const
{
    system.check(cl2_set_zero_stagg_eo(ctx(system), dev(), prec));
}
void Staggeredfield_eo::set_cold() const
{
    system.check(cl2_set_cold_stagg_eo(ctx(system), dev(), prec));
}
void Staggeredfield_eo::import(const su3vec* host) const
{
    system.check(cl2_su3vec_import(ctx(system), dev(), reinterpret_cast<const double*>(host), prec));
}
std::vector<su3vec> Staggeredfield_eo::get_host() const
{
    std::vector<su3vec> h(get_elements());
    system.check(cl2_su3vec_export(ctx(system), reinterpret_cast<double*>(h.data()), dev(), prec));
    return h;
}

/* ---------------- Wilson eo BLAS ---------------- */

void copyData(const Spinorfield_eo* to, const Spinorfield_eo& from)
{
    same_prec(*to, from);
    const auto& s = from.get_system();
    s.check(cl2_copy(ctx(s), to->dev(), from.dev(), cl2_field_bytes(ctx(s), CL2_FIELD_SPINOR_EO, from.precision())));
}
void copyData(const Spinorfield_eo* to, const Spinorfield_eo* from)
{
    copyData(to, *from);
}
hmc_complex scalar_product(const Spinorfield_eo& left, const Spinorfield_eo& right)
{
    Scalar<hmc_complex> r(left.get_system());
    scalar_product(&r, left, right);
    return r.get();
}
void scalar_product(const Scalar<hmc_complex>* res, const Spinorfield_eo& left, const Spinorfield_eo& right)
{
    same_prec(left, right);
    const auto& s = left.get_system();
    s.check(cl2_scalar_product_eo(ctx(s), res->dev(), left.dev(), right.dev(), left.precision()));
}
hmc_float squarenorm(const Spinorfield_eo& field)
{
    Scalar<hmc_float> r(field.get_system());
    squarenorm(&r, field);
    return r.get();
}
void squarenorm(const Scalar<hmc_float>* res, const Spinorfield_eo& field)
{
    const auto& s = field.get_system();
    s.check(cl2_squarenorm_eo(ctx(s), res->dev(), field.dev(), field.precision()));
}
static void wsaxpy(const Spinorfield_eo* out, const cl2_scalar& a, const Spinorfield_eo& x, const Spinorfield_eo& y)
{
    same_prec(x, y);
    const auto& s = x.get_system();
    s.check(cl2_saxpy_eo(ctx(s), out->dev(), x.dev(), y.dev(), x.precision(), &a));
}
void saxpy(const Spinorfield_eo* out, const hmc_complex alpha, const Spinorfield_eo& x, const Spinorfield_eo& y)
{
    wsaxpy(out, val(alpha), x, y);
}
void saxpy(const Spinorfield_eo* out, const Scalar<hmc_complex>& alpha, const Spinorfield_eo& x,
           const Spinorfield_eo& y)
{
    wsaxpy(out, dev(alpha), x, y);
}
void sax(const Spinorfield_eo* out, const hmc_complex alpha, const Spinorfield_eo& x)
{
    cl2_scalar a = val(alpha);
    x.get_system().check(cl2_sax_eo(ctx(x.get_system()), out->dev(), x.dev(), x.precision(), &a));
}
void sax(const Spinorfield_eo* out, const Scalar<hmc_complex>& alpha, const Spinorfield_eo& x)
{
    cl2_scalar a = dev(alpha);
    x.get_system().check(cl2_sax_eo(ctx(x.get_system()), out->dev(), x.dev(), x.precision(), &a));
}
void saxsbypz(const Spinorfield_eo* out, const hmc_complex alpha, const Spinorfield_eo& x, const hmc_complex beta,
              const Spinorfield_eo& y, const Spinorfield_eo& z)
{
    cl2_scalar a = val(alpha), b = val(beta);
    x.get_system().check(
        cl2_saxsbypz_eo(ctx(x.get_system()), out->dev(), x.dev(), y.dev(), z.dev(), x.precision(), &a, &b));
}
void saxsbypz(const Spinorfield_eo* out, const Scalar<hmc_complex>& alpha, const Spinorfield_eo& x,
              const Scalar<hmc_complex>& beta, const Spinorfield_eo& y, const Spinorfield_eo& z)
{
    cl2_scalar a = dev(alpha), b = dev(beta);
    x.get_system().check(
        cl2_saxsbypz_eo(ctx(x.get_system()), out->dev(), x.dev(), y.dev(), z.dev(), x.precision(), &a, &b));
}
void saxpy_AND_squarenorm(const Spinorfield_eo* out, const Scalar<hmc_complex>& alpha, const Spinorfield_eo& x,
                          const Spinorfield_eo& y, const Scalar<hmc_complex>& sq)
{
    cl2_scalar a = dev(alpha);
    x.get_system().check(
        cl2_saxpy_AND_squarenorm_eo(ctx(x.get_system()), out->dev(), x.dev(), y.dev(), x.precision(), &a, sq.dev()));
}
void saxpy_AND_cgnorm(const Spinorfield_eo* out, const Scalar<hmc_complex>& alpha, const Spinorfield_eo& x,
                      const Spinorfield_eo& y, const Vector<hmc_float>& res)
{
    if (res.get_vector_size() < 4) throw std::invalid_argument("saxpy_AND_cgnorm needs a result vector of 4 elements");
    cl2_scalar a = dev(alpha);
    x.get_system().check(
        cl2_saxpy_AND_cgnorm_eo(ctx(x.get_system()), out->dev(), x.dev(), y.dev(), x.precision(), &a, res.dev()));
}
void cg_update_xp(const Spinorfield_eo* x, const Spinorfield_eo* p, const Spinorfield_eo& r,
                  const Scalar<hmc_complex>& malpha, const Scalar<hmc_complex>& mbeta)
{
    if (x->precision() != p->precision() || p->precision() != r.precision())
        throw std::invalid_argument("cg_update_xp: fields of different precision");
    cl2_scalar a = dev(malpha), b = dev(mbeta);
    r.get_system().check(cl2_cg_update_xp(ctx(r.get_system()), CL2_FIELD_SPINOR_EO, x->dev(), p->dev(), r.dev(),
                                          r.precision(), &a, &b));
}
void cg_update_xp(const Staggeredfield_eo* x, const Staggeredfield_eo* p, const Staggeredfield_eo& r,
                  const Scalar<hmc_float>& malpha, const Scalar<hmc_float>& mbeta)
{
    if (x->precision() != p->precision() || p->precision() != r.precision())
        throw std::invalid_argument("cg_update_xp: fields of different precision");
    cl2_scalar a = dev(malpha), b = dev(mbeta);
    r.get_system().check(cl2_cg_update_xp(ctx(r.get_system()), CL2_FIELD_SU3VEC_EO, x->dev(), p->dev(), r.dev(),
                                          r.precision(), &a, &b));
}
void saxpy_AND_gamma5_eo(const Spinorfield_eo* out, const hmc_complex alpha, const Spinorfield_eo& x,
                         const Spinorfield_eo& y)
{
    cl2_scalar a = val(alpha);
    x.get_system().check(
        cl2_saxpy_AND_gamma5_eo(ctx(x.get_system()), out->dev(), x.dev(), y.dev(), x.precision(), &a));
}

void convert_to_eoprec(const Spinorfield_eo* even, const Spinorfield_eo* odd, const Spinorfield& in)
{
    const int ns  = even->get_system().get_inputparameters().get_nspace();
    const size_t S = even->get_elements();
    std::vector<spinor> e(S), o(S);
    for (size_t i = 0; i < S; ++i) {
        e[i] = in.host[eo_site(ns, i, 0)];
        o[i] = in.host[eo_site(ns, i, 1)];
    }
    even->import(e.data());
    odd->import(o.data());
}
void convert_from_eoprec(const Spinorfield* merged, const Spinorfield_eo& even, const Spinorfield_eo& odd)
{
    const int ns             = even.get_system().get_inputparameters().get_nspace();
    const std::vector<spinor> e = even.get_host(), o = odd.get_host();
    for (size_t i = 0; i < e.size(); ++i) {
        merged->host[eo_site(ns, i, 0)] = e[i];
        merged->host[eo_site(ns, i, 1)] = o[i];
    }
}
void convert_precision(const Spinorfield_eo* to, const Spinorfield_eo& from)
{
    const auto& s = from.get_system();
    s.check(cl2_convert_precision_eo(ctx(s), CL2_FIELD_SPINOR_EO, to->dev(), to->precision(), from.dev(),
                                     from.precision()));
}
void convert_precision(const Staggeredfield_eo* to, const Staggeredfield_eo& from)
{
    const auto& s = from.get_system();
    s.check(cl2_convert_precision_eo(ctx(s), CL2_FIELD_SU3VEC_EO, to->dev(), to->precision(), from.dev(),
                                     from.precision()));
}
void log_squarenorm(const std::string& msg, const Spinorfield_eo& x)
{
    if (std::getenv("CL2QCD_DEBUG")) std::fprintf(stderr, "%s%.16g\n", msg.c_str(), squarenorm(x));
}

/* ---------------- staggered BLAS ---------------- */

void copyData(const Staggeredfield_eo* to, const Staggeredfield_eo& from)
{
    same_prec(*to, from);
    const auto& s = from.get_system();
    s.check(cl2_copy(ctx(s), to->dev(), from.dev(), cl2_field_bytes(ctx(s), CL2_FIELD_SU3VEC_EO, from.precision())));
}
void copyData(const Staggeredfield_eo* to, const Staggeredfield_eo* from)
{
    copyData(to, *from);
}
hmc_complex scalar_product(const Staggeredfield_eo& l, const Staggeredfield_eo& r)
{
    Scalar<hmc_complex> res(l.get_system());
    scalar_product(&res, l, r);
    return res.get();
}
void scalar_product(const Scalar<hmc_complex>* res, const Staggeredfield_eo& l, const Staggeredfield_eo& r)
{
    same_prec(l, r);
    l.get_system().check(cl2_scalar_product_stagg_eo(ctx(l.get_system()), res->dev(), l.dev(), r.dev(), l.precision()));
}
hmc_float scalar_product_real_part(const Staggeredfield_eo& l, const Staggeredfield_eo& r)
{
    Scalar<hmc_float> res(l.get_system());
    scalar_product_real_part(&res, l, r);
    return res.get();
}
void scalar_product_real_part(const Scalar<hmc_float>* res, const Staggeredfield_eo& l, const Staggeredfield_eo& r)
{
    same_prec(l, r);
    l.get_system().check(
        cl2_scalar_product_real_part_stagg_eo(ctx(l.get_system()), res->dev(), l.dev(), r.dev(), l.precision()));
}
hmc_float squarenorm(const Staggeredfield_eo& f)
{
    Scalar<hmc_float> res(f.get_system());
    squarenorm(&res, f);
    return res.get();
}
void squarenorm(const Scalar<hmc_float>* res, const Staggeredfield_eo& f)
{
    f.get_system().check(cl2_squarenorm_stagg_eo(ctx(f.get_system()), res->dev(), f.dev(), f.precision()));
}
static void ssax(const Staggeredfield_eo* out, const cl2_scalar& a, const Staggeredfield_eo& x)
{
    x.get_system().check(cl2_sax_stagg_eo(ctx(x.get_system()), out->dev(), x.dev(), x.precision(), &a));
}
void sax(const Staggeredfield_eo* out, const hmc_complex a, const Staggeredfield_eo& x)
{
    ssax(out, val(a), x);
}
void sax(const Staggeredfield_eo* out, const Scalar<hmc_complex>& a, const Staggeredfield_eo& x)
{
    ssax(out, dev(a), x);
}
void sax(const Staggeredfield_eo* out, const hmc_float a, const Staggeredfield_eo& x)
{
    ssax(out, val(a), x);
}
void sax(const Staggeredfield_eo* out, const Scalar<hmc_float>& a, const Staggeredfield_eo& x)
{
    ssax(out, dev(a), x);
}
void sax(const Staggeredfield_eo* out, const Vector<hmc_float>& a, const int idx, const Staggeredfield_eo& x)
{
    ssax(out, dev(a, idx), x);
}
static void ssaxpy(const Staggeredfield_eo* out, const cl2_scalar& a, const Staggeredfield_eo& x,
                   const Staggeredfield_eo& y)
{
    same_prec(x, y);
    x.get_system().check(cl2_saxpy_stagg_eo(ctx(x.get_system()), out->dev(), x.dev(), y.dev(), x.precision(), &a));
}
void saxpy(const Staggeredfield_eo* out, const hmc_complex a, const Staggeredfield_eo& x, const Staggeredfield_eo& y)
{
    ssaxpy(out, val(a), x, y);
}
void saxpy(const Staggeredfield_eo* out, const Scalar<hmc_complex>& a, const Staggeredfield_eo& x,
           const Staggeredfield_eo& y)
{
    ssaxpy(out, dev(a), x, y);
}
void saxpy(const Staggeredfield_eo* out, const Scalar<hmc_float>& a, const Staggeredfield_eo& x,
           const Staggeredfield_eo& y)
{
    ssaxpy(out, dev(a), x, y);
}
void saxpy(const Staggeredfield_eo* out, const hmc_float a, const Staggeredfield_eo& x, const Staggeredfield_eo& y)
{
    ssaxpy(out, val(a), x, y);
}
void saxpy(const Staggeredfield_eo* out, const Vector<hmc_float>& a, const int idx, const Staggeredfield_eo& x,
           const Staggeredfield_eo& y)
{
    ssaxpy(out, dev(a, idx), x, y);
}
static void ssaxpby(const Staggeredfield_eo* out, const cl2_scalar& a, const Staggeredfield_eo& x,
                    const cl2_scalar& b, const Staggeredfield_eo& y)
{
    same_prec(x, y);
    x.get_system().check(
        cl2_saxpby_stagg_eo(ctx(x.get_system()), out->dev(), x.dev(), y.dev(), x.precision(), &a, &b));
}
void saxpby(const Staggeredfield_eo* out, const hmc_complex a, const Staggeredfield_eo& x, const hmc_complex b,
            const Staggeredfield_eo& y)
{
    ssaxpby(out, val(a), x, val(b), y);
}
void saxpby(const Staggeredfield_eo* out, const Scalar<hmc_complex>& a, const Staggeredfield_eo& x,
            const Scalar<hmc_complex>& b, const Staggeredfield_eo& y)
{
    ssaxpby(out, dev(a), x, dev(b), y);
}
void saxpby(const Staggeredfield_eo* out, const hmc_float a, const Staggeredfield_eo& x, const hmc_float b,
            const Staggeredfield_eo& y)
{
    ssaxpby(out, val(a), x, val(b), y);
}
void saxpby(const Staggeredfield_eo* out, const Scalar<hmc_float>& a, const Staggeredfield_eo& x,
            const Scalar<hmc_float>& b, const Staggeredfield_eo& y)
{
    ssaxpby(out, dev(a), x, dev(b), y);
}
void saxpby(const Staggeredfield_eo* out, const Vector<hmc_float>& a, const int ia, const Staggeredfield_eo& x,
            const Vector<hmc_float>& b, const int ib, const Staggeredfield_eo& y)
{
    ssaxpby(out, dev(a, ia), x, dev(b, ib), y);
}
void saxpbypz(const Staggeredfield_eo* out, const hmc_complex a, const Staggeredfield_eo& x, const hmc_complex b,
              const Staggeredfield_eo& y, const Staggeredfield_eo& z)
{
    cl2_scalar sa = val(a), sb = val(b);
    x.get_system().check(
        cl2_saxpbypz_stagg_eo(ctx(x.get_system()), out->dev(), x.dev(), y.dev(), z.dev(), x.precision(), &sa, &sb));
}
void saxpbypz(const Staggeredfield_eo* out, const Scalar<hmc_complex>& a, const Staggeredfield_eo& x,
              const Scalar<hmc_complex>& b, const Staggeredfield_eo& y, const Staggeredfield_eo& z)
{
    cl2_scalar sa = dev(a), sb = dev(b);
    x.get_system().check(
        cl2_saxpbypz_stagg_eo(ctx(x.get_system()), out->dev(), x.dev(), y.dev(), z.dev(), x.precision(), &sa, &sb));
}
void sax_vec_and_squarenorm(const Vector<hmc_float>* res, const Vector<hmc_float>& alpha, const Staggeredfield_eo& x)
{
    if (res->get_vector_size() != alpha.get_vector_size()) throw std::invalid_argument("vector sizes differ");
    x.get_system().check(cl2_sax_vec_AND_squarenorm_stagg_eo(ctx(x.get_system()), res->dev(), alpha.dev(),
                                                              (int)alpha.get_vector_size(), x.dev(), x.precision()));
}

/* ---------------- device scalar arithmetic ---------------- */

#define CL2_SCALAR_OP(FN, OP, T, CALL)                                                               \
    void FN(const Scalar<T>* dest, const Scalar<T>& a, const Scalar<T>& b)                           \
    {                                                                                                \
        dest->get_system().check(CALL(ctx(dest->get_system()), OP, dest->dev(), a.dev(), b.dev())); \
    }
CL2_SCALAR_OP(add, CL2_SCALAR_OP_SUM, hmc_complex, cl2_complex_op)
CL2_SCALAR_OP(subtract, CL2_SCALAR_OP_SUBTRACTION, hmc_complex, cl2_complex_op)
CL2_SCALAR_OP(multiply, CL2_SCALAR_OP_PRODUCT, hmc_complex, cl2_complex_op)
CL2_SCALAR_OP(divide, CL2_SCALAR_OP_RATIO, hmc_complex, cl2_complex_op)
CL2_SCALAR_OP(add, CL2_SCALAR_OP_SUM, hmc_float, cl2_real_op)
CL2_SCALAR_OP(subtract, CL2_SCALAR_OP_SUBTRACTION, hmc_float, cl2_real_op)
CL2_SCALAR_OP(multiply, CL2_SCALAR_OP_PRODUCT, hmc_float, cl2_real_op)
CL2_SCALAR_OP(divide, CL2_SCALAR_OP_RATIO, hmc_float, cl2_real_op)
#undef CL2_SCALAR_OP

void copyData(const Scalar<hmc_complex>* to, const Scalar<hmc_complex>& from)
{
    to->get_system().check(cl2_scalar_copy(ctx(to->get_system()), to->dev(), from.dev(), 2));
}
void copyData(const Scalar<hmc_float>* to, const Scalar<hmc_float>& from)
{
    to->get_system().check(cl2_scalar_copy(ctx(to->get_system()), to->dev(), from.dev(), 1));
}
void copyData(const Vector<hmc_float>* to, const Vector<hmc_float>& from)
{
    if (to->get_vector_size() != from.get_vector_size()) throw std::invalid_argument("vector sizes differ");
    to->get_system().check(
        cl2_scalar_copy(ctx(to->get_system()), to->dev(), from.dev(), (int)from.get_vector_size()));
}
void access_real_vector_element(const Scalar<hmc_float>* out, const Vector<hmc_float>& in, const int index)
{
    if (index < 0 || (size_t)index >= in.get_vector_size()) throw std::out_of_range("vector index");
    out->get_system().check(cl2_real_get_vector_element(ctx(out->get_system()), out->dev(), in.dev(), index));
}
void access_real_vector_element(const Vector<hmc_float>* out, const Scalar<hmc_float>& in, const int index)
{
    if (index < 0 || (size_t)index >= out->get_vector_size()) throw std::out_of_range("vector index");
    in.get_system().check(cl2_real_set_vector_element(ctx(in.get_system()), out->dev(), in.dev(), index));
}
void update_zeta_cgm(const Vector<hmc_float>* out, const Vector<hmc_float>& zeta_prev,
                     const Vector<hmc_float>& zeta_prev_prev, const Scalar<hmc_float>& sbeta_prev,
                     const Scalar<hmc_float>& sbeta_pres, const Scalar<hmc_float>& salpha_prev,
                     const Vector<hmc_float>& sigma, const int numeq)
{
    sbeta_prev.get_system().check(cl2_update_zeta_cgm(ctx(sbeta_prev.get_system()), out->dev(), zeta_prev.dev(),
                                                      zeta_prev_prev.dev(), sbeta_prev.dev(), sbeta_pres.dev(),
                                                      salpha_prev.dev(), sigma.dev(), numeq));
}
void update_beta_cgm(const Vector<hmc_float>* out, const Scalar<hmc_float>& sbeta_pres,
                     const Vector<hmc_float>& zeta_pres, const Vector<hmc_float>& zeta_prev, const int numeq)
{
    sbeta_pres.get_system().check(cl2_update_beta_cgm(ctx(sbeta_pres.get_system()), out->dev(), sbeta_pres.dev(),
                                                      zeta_pres.dev(), zeta_prev.dev(), numeq));
}
void update_alpha_cgm(const Vector<hmc_float>* out, const Scalar<hmc_float>& salpha_pres,
                      const Vector<hmc_float>& zeta_pres, const Vector<hmc_float>& beta_pres,
                      const Vector<hmc_float>& zeta_prev, const Scalar<hmc_float>& sbeta_pres, const int numeq)
{
    salpha_pres.get_system().check(cl2_update_alpha_cgm(ctx(salpha_pres.get_system()), out->dev(), salpha_pres.dev(),
                                                        zeta_pres.dev(), beta_pres.dev(), zeta_prev.dev(),
                                                        sbeta_pres.dev(), numeq));
}

/* ---------------- deterministic fills ---------------- */

template<typename T>
static std::vector<T> drand48_fill(size_t n, int seed)
{
    std::vector<T> v(n);
    double* d = reinterpret_cast<double*>(v.data());
    srand48(seed);
    for (size_t k = 0; k < n * sizeof(T) / sizeof(double); ++k) d[k] = drand48();
    return v;
}
void pseudo_randomize(const Spinorfield* to, int seed)
{
    auto v = drand48_fill<spinor>(to->get_elements(), seed);
    to->import(v.data());
}
void pseudo_randomize(const Spinorfield_eo* to, int seed)
{
    auto v = drand48_fill<spinor>(to->get_elements(), seed);
    to->import(v.data());
}
void pseudo_randomize(const Staggeredfield_eo* to, int seed)
{
    auto v = drand48_fill<su3vec>(to->get_elements(), seed);
    to->import(v.data());
}

}  // namespace lattices
}  // namespace physics
