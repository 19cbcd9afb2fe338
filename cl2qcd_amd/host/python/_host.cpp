// Python binding of the C++ drop-in (physics::lattices, physics::fermionmatrix,
// physics::algorithms::solvers) so Python tests and tools drive the same
// classes a C++ caller of the reference would.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>

#include "../include/ildg_io/ildgIo.hpp"
#include "../include/physics/algorithms/inversion.hpp"
#include "../include/physics/algorithms/rational_approximation.hpp"
#include "../include/physics/algorithms/solvers/solvers.hpp"
#include "../include/physics/algorithms/su3heatbath.hpp"
#include "../include/physics/observables/gaugeObservables.hpp"
#include "../include/serving/engine.hpp"

namespace py = pybind11;
using namespace physics;
using namespace physics::lattices;
namespace fm = physics::fermionmatrix;
namespace sv = physics::algorithms::solvers;

namespace {

template<typename T>
py::array_t<double> to_numpy(const std::vector<T>& v)
{
    const size_t w = sizeof(T) / sizeof(double);
    py::array_t<double> a({v.size(), w});
    std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(T));
    return a;
}
template<typename T>
std::vector<T> from_numpy(py::array_t<double, py::array::c_style | py::array::forcecast> a, size_t n)
{
    if ((size_t)a.size() != n * sizeof(T) / sizeof(double)) throw std::invalid_argument("array has the wrong size");
    std::vector<T> v(n);
    std::memcpy(v.data(), a.data(), n * sizeof(T));
    return v;
}

/* owns the parameters and the System so Python keeps them alive together */
struct Lattice {
    explicit Lattice(const std::vector<std::string>& args)
        : params(args), system(params), handler(params), prng_params(params), prng(system, &prng_params)
    {
    }
    meta::Inputparameters params;
    hardware::System system;
    InterfacesHandlerImplementation handler;
    PrngParametersImplementation prng_params;
    PRNG prng;
};

}  // namespace

PYBIND11_MODULE(_host, m)
{
    m.doc() = "C++ drop-in of the CL2QCD fermion-inversion interface on the gfx950 C ABI";
    static py::exception<sv::SolverException> exc_solver(m, "SolverException");
    static py::exception<sv::SolverStuck> exc_stuck(m, "SolverStuck", exc_solver.ptr());
    static py::exception<sv::SolverDidNotSolve> exc_dns(m, "SolverDidNotSolve", exc_solver.ptr());
    static py::exception<hardware::HipException> exc_hip(m, "HipException");
    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const sv::SolverStuck& e) {
            PyErr_SetObject(exc_stuck.ptr(), py::make_tuple(e.what(), e.get_iterations()).ptr());
        } catch (const sv::SolverDidNotSolve& e) {
            PyErr_SetObject(exc_dns.ptr(), py::make_tuple(e.what(), e.get_iterations()).ptr());
        } catch (const hardware::HipException& e) {
            PyErr_SetObject(exc_hip.ptr(), py::make_tuple(e.what(), e.code).ptr());
        }
    });
    m.attr("EVEN") = EVEN;
    m.attr("ODD")  = ODD;
    m.attr("F64")  = CL2_F64;
    m.attr("F32")  = CL2_F32;
    m.attr("F16")  = CL2_F16;

    py::class_<Lattice>(m, "Lattice", "run parameters + hardware::System + InterfacesHandler + PRNG")
        .def(py::init<const std::vector<std::string>&>(), py::arg("args") = std::vector<std::string>{})
        .def_property_readonly("eo_sites", [](const Lattice& l) { return l.system.eo_sites(); })
        .def_property_readonly("nspace", [](const Lattice& l) { return l.params.get_nspace(); })
        .def_property_readonly("ntime", [](const Lattice& l) { return l.params.get_ntime(); })
        .def_property_readonly("mubar", [](const Lattice& l) { return meta::get_mubar(l.params); })
        .def_property_readonly("kappa", [](const Lattice& l) { return l.params.get_kappa(); })
        .def_property_readonly("mass", [](const Lattice& l) { return l.params.get_mass(); })
        .def_property_readonly("solver_prec", [](const Lattice& l) { return l.params.get_solver_prec(); })
        .def("synchronize", [](const Lattice& l) { l.system.synchronize(); })
        .def("wilson_parameters",
             [](Lattice& l) -> const AdditionalParameters& {
                 return l.handler.getAdditionalParameters<Spinorfield_eo>();
             },
             py::return_value_policy::reference_internal)
        .def("staggered_parameters",
             [](Lattice& l) -> const AdditionalParameters& {
                 return l.handler.getAdditionalParameters<Staggeredfield_eo>();
             },
             py::return_value_policy::reference_internal)
        .def("tune", [](Lattice& l, const std::string& key, long v) {
            l.system.check(cl2_tune_set(l.system.ctx(), key.c_str(), v));
        });

    py::class_<AdditionalParameters>(m, "AdditionalParameters")
        .def("getKappa", &AdditionalParameters::getKappa)
        .def("getMubar", &AdditionalParameters::getMubar)
        .def("getMass", &AdditionalParameters::getMass);
    py::class_<WilsonAdditionalParameters, AdditionalParameters>(m, "WilsonAdditionalParameters")
        .def(py::init<hmc_float, hmc_float>());
    py::class_<StaggeredAdditionalParameters, AdditionalParameters>(m, "StaggeredAdditionalParameters")
        .def(py::init<hmc_float>());

    py::class_<Gaugefield>(m, "Gaugefield")
        .def(py::init([](Lattice& l, bool hot) {
                 return new Gaugefield(l.system, &l.handler.getInterface<Gaugefield>(), l.prng, hot);
             }),
             py::keep_alive<1, 2>(), py::arg("lattice"), py::arg("hot") = false)
        .def(py::init([](Lattice& l, const std::string& file) {
                 return new Gaugefield(l.system, &l.handler.getInterface<Gaugefield>(), l.prng, file);
             }),
             py::keep_alive<1, 2>())
        .def("get_host", [](const Gaugefield& g) { return to_numpy(g.get_host()); })
        .def("import_", [](const Gaugefield& g, py::array_t<double> a) {
            auto v = from_numpy<Matrixsu3>(a, 8 * g.get_system().eo_sites());
            g.import(v.data());
        })
        .def("save", &Gaugefield::save, py::arg("outputfile"), py::arg("number"))
        // plumbing for benchmarks: the resident field's device pointer and link format
        .def_property_readonly("device_ptr", [](const Gaugefield& g) { return reinterpret_cast<uintptr_t>(g.dev()); })
        .def_property_readonly("format", &Gaugefield::format)
        // this field's links := src's (an 18-real field of the same lattice size on the same device), converted to
        // this field's format on the device
        .def("convert_from", [](Gaugefield& g, const Gaugefield& src) {
            if (src.format() != CL2_GAUGE_F64_18) throw std::invalid_argument("convert_from: src must be 18-real");
            if (src.get_system().eo_sites() != g.get_system().eo_sites())
                throw std::invalid_argument("convert_from: lattices differ in size");
            src.get_system().synchronize();
            const hardware::System& sys = g.get_system();
            sys.check(cl2_gauge_convert(sys.ctx(), g.dev_mutable(), g.format(), src.dev()));
            sys.synchronize();
        });

    // ILDG/LIME files on the host (no device needed)
    m.def("ildg_checksum", [](py::bytes data, size_t nt, size_t ns) {
        const std::string d = data;
        const ildgIo::Checksum c = ildgIo::calculate_ildg_checksum(d.data(), d.size(), nt, ns);
        return py::make_tuple(c.suma, c.sumb);
    });
    m.def(
        "ildg_read",
        [](const std::string& path, int ns, int nt, bool ignore_checksum) {
            ildgIo::Sourcefileparameters info;
            std::vector<Matrixsu3> links = ildgIo::readGaugefieldFromSourcefile(path, ns, nt, &info, ignore_checksum);
            py::dict d;
            d["precision"]    = info.precision;
            d["has_checksum"] = info.has_checksum;
            d["suma"]         = info.checksum.suma;
            d["sumb"]         = info.checksum.sumb;
            d["xlf_info"]     = info.xlf_info;
            return py::make_tuple(to_numpy(links), d);
        },
        py::arg("path"), py::arg("ns"), py::arg("nt"), py::arg("ignore_checksum") = false);
    m.def(
        "ildg_write",
        [](const std::string& path, py::array_t<double> links, int ns, int nt, int trajectory, double plaquette,
           double beta, double kappa, double mu, double mubar, int precision) {
            const size_t vol = (size_t)ns * ns * ns * nt;
            ildgIo::XlfInfo info;
            info.trajectory = trajectory;
            info.plaquette  = plaquette;
            info.beta       = beta;
            info.kappa      = kappa;
            info.mu         = mu;
            info.mubar      = mubar;
            ildgIo::writeGaugefieldToFile(path, from_numpy<Matrixsu3>(links, 4 * vol), ns, nt, info, precision);
        },
        py::arg("path"), py::arg("links"), py::arg("ns"), py::arg("nt"), py::arg("trajectory") = 0,
        py::arg("plaquette") = 0., py::arg("beta") = 0., py::arg("kappa") = 0., py::arg("mu") = 0.,
        py::arg("mubar") = 0., py::arg("precision") = 64);

    m.def(
        "su3heatbath", [](Gaugefield& gf, Lattice& l, int overrelax) { physics::algorithms::su3heatbath(gf, l.prng, overrelax); },
        py::arg("gf"), py::arg("lattice"), py::arg("overrelax") = 0);
    m.def(
        "overrelax", [](Gaugefield& gf, Lattice& l, int steps) { physics::algorithms::overrelax(gf, l.prng, steps); },
        py::arg("gf"), py::arg("lattice"), py::arg("steps") = 1);
    m.def("measurePlaquette", [](const Gaugefield& gf) { return physics::observables::measurePlaquette(&gf); });
    m.def("measurePlaquetteWithoutNormalization",
          [](const Gaugefield& gf) { return physics::observables::measurePlaquetteWithoutNormalization(&gf); });

    py::class_<Spinorfield>(m, "Spinorfield")
        .def(py::init([](Lattice& l) { return new Spinorfield(l.system, l.handler.getInterface<Spinorfield>()); }),
             py::keep_alive<1, 2>())
        .def("get_host", [](const Spinorfield& f) { return to_numpy(f.get_host()); })
        .def("import_", [](const Spinorfield& f, py::array_t<double> a) {
            auto v = from_numpy<spinor>(a, f.get_elements());
            f.import(v.data());
        });

    py::class_<Spinorfield_eo>(m, "Spinorfield_eo")
        .def(py::init([](Lattice& l, int prec) { return new Spinorfield_eo(l.system, prec); }), py::keep_alive<1, 2>(),
             py::arg("lattice"), py::arg("prec") = CL2_F64)
        .def("gamma5", &Spinorfield_eo::gamma5)
        .def("gaussian_from_seed", &Spinorfield_eo::gaussian_from_seed)
        .def("zero", &Spinorfield_eo::zero)
        .def("cold", &Spinorfield_eo::cold)
        .def("get_host", [](const Spinorfield_eo& f) { return to_numpy(f.get_host()); })
        .def("import_", [](const Spinorfield_eo& f, py::array_t<double> a) {
            auto v = from_numpy<spinor>(a, f.get_elements());
            f.import(v.data());
        });

    py::class_<Staggeredfield_eo>(m, "Staggeredfield_eo")
        .def(py::init([](Lattice& l, int prec) { return new Staggeredfield_eo(l.system, prec); }),
             py::keep_alive<1, 2>(), py::arg("lattice"), py::arg("prec") = CL2_F64)
        .def("set_zero", &Staggeredfield_eo::set_zero)
        .def("set_cold", &Staggeredfield_eo::set_cold)
        .def("gaussian_from_seed", &Staggeredfield_eo::gaussian_from_seed)
        .def("get_host", [](const Staggeredfield_eo& f) { return to_numpy(f.get_host()); })
        .def("import_", [](const Staggeredfield_eo& f, py::array_t<double> a) {
            auto v = from_numpy<su3vec>(a, f.get_elements());
            f.import(v.data());
        });

    py::class_<Scalar<hmc_complex>>(m, "ComplexScalar")
        .def(py::init([](Lattice& l) { return new Scalar<hmc_complex>(l.system); }), py::keep_alive<1, 2>())
        .def("store", [](const Scalar<hmc_complex>& s, std::complex<double> v) { s.store({v.real(), v.imag()}); })
        .def("get", [](const Scalar<hmc_complex>& s) {
            hmc_complex c = s.get();
            return std::complex<double>(c.re, c.im);
        });
    py::class_<Scalar<hmc_float>>(m, "RealScalar")
        .def(py::init([](Lattice& l) { return new Scalar<hmc_float>(l.system); }), py::keep_alive<1, 2>())
        .def("store", &Scalar<hmc_float>::store)
        .def("get", &Scalar<hmc_float>::get);
    py::class_<Vector<hmc_float>>(m, "RealVector")
        .def(py::init([](Lattice& l, size_t n) { return new Vector<hmc_float>(n, l.system); }), py::keep_alive<1, 2>())
        .def("store", &Vector<hmc_float>::store)
        .def("get", &Vector<hmc_float>::get);

    auto cplx = [](std::complex<double> c) { return hmc_complex{c.real(), c.imag()}; };
    // Wilson BLAS
    m.def("copyData", [](const Spinorfield_eo& to, const Spinorfield_eo& from) { copyData(&to, from); });
    m.def("scalar_product", [](const Spinorfield_eo& a, const Spinorfield_eo& b) {
        hmc_complex c = scalar_product(a, b);
        return std::complex<double>(c.re, c.im);
    });
    m.def("squarenorm", [](const Spinorfield_eo& a) { return squarenorm(a); });
    m.def("saxpy", [cplx](const Spinorfield_eo& out, std::complex<double> a, const Spinorfield_eo& x,
                          const Spinorfield_eo& y) { saxpy(&out, cplx(a), x, y); });
    m.def("sax", [cplx](const Spinorfield_eo& out, std::complex<double> a, const Spinorfield_eo& x) {
        sax(&out, cplx(a), x);
    });
    m.def("saxsbypz", [cplx](const Spinorfield_eo& out, std::complex<double> a, const Spinorfield_eo& x,
                             std::complex<double> b, const Spinorfield_eo& y,
                             const Spinorfield_eo& z) { saxsbypz(&out, cplx(a), x, cplx(b), y, z); });
    m.def("saxpy_AND_gamma5_eo", [cplx](const Spinorfield_eo& out, std::complex<double> a, const Spinorfield_eo& x,
                                        const Spinorfield_eo& y) { saxpy_AND_gamma5_eo(&out, cplx(a), x, y); });
    m.def("saxpy_AND_squarenorm", [](const Spinorfield_eo& out, const Scalar<hmc_complex>& a, const Spinorfield_eo& x,
                                     const Spinorfield_eo& y, const Scalar<hmc_complex>& r) {
        saxpy_AND_squarenorm(&out, a, x, y, r);
    });
    m.def("convert_from_eoprec", [](const Spinorfield& out, const Spinorfield_eo& e, const Spinorfield_eo& o) {
        convert_from_eoprec(&out, e, o);
    });
    m.def("convert_to_eoprec", [](const Spinorfield_eo& e, const Spinorfield_eo& o, const Spinorfield& in) {
        convert_to_eoprec(&e, &o, in);
    });
    m.def("pseudo_randomize", [](const Spinorfield& f, int seed) { pseudo_randomize(&f, seed); });
    m.def("pseudo_randomize", [](const Spinorfield_eo& f, int seed) { pseudo_randomize(&f, seed); });
    m.def("pseudo_randomize", [](const Staggeredfield_eo& f, int seed) { pseudo_randomize(&f, seed); });
    // staggered BLAS
    m.def("copyData", [](const Staggeredfield_eo& to, const Staggeredfield_eo& from) { copyData(&to, from); });
    m.def("squarenorm", [](const Staggeredfield_eo& a) { return squarenorm(a); });
    m.def("scalar_product", [](const Staggeredfield_eo& a, const Staggeredfield_eo& b) {
        hmc_complex c = scalar_product(a, b);
        return std::complex<double>(c.re, c.im);
    });
    m.def("scalar_product_real_part",
          [](const Staggeredfield_eo& a, const Staggeredfield_eo& b) { return scalar_product_real_part(a, b); });
    m.def("saxpy", [cplx](const Staggeredfield_eo& out, std::complex<double> a, const Staggeredfield_eo& x,
                          const Staggeredfield_eo& y) { saxpy(&out, cplx(a), x, y); });
    m.def("saxpby", [cplx](const Staggeredfield_eo& out, std::complex<double> a, const Staggeredfield_eo& x,
                           std::complex<double> b, const Staggeredfield_eo& y) { saxpby(&out, cplx(a), x, cplx(b), y); });
    m.def("sax", [cplx](const Staggeredfield_eo& out, std::complex<double> a, const Staggeredfield_eo& x) {
        sax(&out, cplx(a), x);
    });

    // explicit kernels
    m.def("dslash", [](const Spinorfield_eo& out, const Gaugefield& gf, const Spinorfield_eo& in, int eo,
                       hmc_float kappa) { fm::dslash(&out, gf, in, eo, kappa); });
    m.def("M_tm_inverse_sitediagonal", [](const Spinorfield_eo& o, const Spinorfield_eo& i, hmc_float mb) {
        fm::M_tm_inverse_sitediagonal(&o, i, mb);
    });
    m.def("M_tm_sitediagonal",
          [](const Spinorfield_eo& o, const Spinorfield_eo& i, hmc_float mb) { fm::M_tm_sitediagonal(&o, i, mb); });
    m.def("M_tm_inverse_sitediagonal_minus", [](const Spinorfield_eo& o, const Spinorfield_eo& i, hmc_float mb) {
        fm::M_tm_inverse_sitediagonal_minus(&o, i, mb);
    });
    m.def("M_tm_sitediagonal_minus", [](const Spinorfield_eo& o, const Spinorfield_eo& i, hmc_float mb) {
        fm::M_tm_sitediagonal_minus(&o, i, mb);
    });
    m.def("dslash_AND_M_tm_inverse_sitediagonal",
          [](const Spinorfield_eo& o, const Gaugefield& gf, const Spinorfield_eo& i, int eo, hmc_float k, hmc_float mb) {
              fm::dslash_AND_M_tm_inverse_sitediagonal(&o, gf, i, eo, k, mb);
          });
    m.def("dslash_AND_M_tm_inverse_sitediagonal_minus",
          [](const Spinorfield_eo& o, const Gaugefield& gf, const Spinorfield_eo& i, int eo, hmc_float k, hmc_float mb) {
              fm::dslash_AND_M_tm_inverse_sitediagonal_minus(&o, gf, i, eo, k, mb);
          });
    m.def("DKS_eo", [](const Staggeredfield_eo& out, const Gaugefield& gf, const Staggeredfield_eo& in, int eo) {
        fm::DKS_eo(&out, gf, in, eo);
    });

    py::class_<fm::Fermionmatrix_eo>(m, "Fermionmatrix_eo")
        .def("__call__", [](const fm::Fermionmatrix_eo& A, const Spinorfield_eo& out, const Gaugefield& gf,
                            const Spinorfield_eo& in, const AdditionalParameters& ap) { A(&out, gf, in, ap); })
        .def("isHermitian", &fm::Fermionmatrix_eo::isHermitian)
        .def("get_flops", &fm::Fermionmatrix_eo::get_flops)
        .def("get_read_write_size", &fm::Fermionmatrix_eo::get_read_write_size);
#define CL2_BIND_EO(NAME)                                                                                  \
    py::class_<fm::NAME, fm::Fermionmatrix_eo>(m, #NAME)                                                   \
        .def(py::init([](Lattice& l) { return new fm::NAME(l.system, l.handler.getInterface<fm::NAME>()); }), \
             py::keep_alive<1, 2>());
    CL2_BIND_EO(Aee)
    CL2_BIND_EO(Aee_AND_gamma5_eo)
    CL2_BIND_EO(Aee_minus)
    CL2_BIND_EO(Aee_minus_AND_gamma5_eo)
    CL2_BIND_EO(Qplus_eo)
    CL2_BIND_EO(Qminus_eo)
    CL2_BIND_EO(QplusQminus_eo)
#undef CL2_BIND_EO

    py::class_<fm::Fermionmatrix_stagg_eo>(m, "Fermionmatrix_stagg_eo")
        .def("__call__",
             [](const fm::Fermionmatrix_stagg_eo& A, const Staggeredfield_eo& out, const Gaugefield& gf,
                const Staggeredfield_eo& in, const AdditionalParameters* ap) { A(&out, gf, in, ap); },
             py::arg("out"), py::arg("gf"), py::arg("in"), py::arg("ap") = nullptr)
        .def("get_flops", &fm::Fermionmatrix_stagg_eo::get_flops)
        .def("get_read_write_size", &fm::Fermionmatrix_stagg_eo::get_read_write_size);
    py::class_<fm::D_KS_eo, fm::Fermionmatrix_stagg_eo>(m, "D_KS_eo")
        .def(py::init([](Lattice& l, int eo) {
                 return new fm::D_KS_eo(l.system, l.handler.getInterface<fm::D_KS_eo>(), eo != 0);
             }),
             py::keep_alive<1, 2>());
    py::class_<fm::MdagM_eo, fm::Fermionmatrix_stagg_eo>(m, "MdagM_eo")
        .def(py::init([](Lattice& l, int ul) {
                 return new fm::MdagM_eo(l.system, l.handler.getInterface<fm::MdagM_eo>(), ul != 0);
             }),
             py::keep_alive<1, 2>(), py::arg("lattice"), py::arg("upper_left") = EVEN);

    m.def("cg", [](const Spinorfield_eo& x, const fm::Fermionmatrix_eo& A, const Gaugefield& gf,
                   const Spinorfield_eo& b, Lattice& l, hmc_float prec, const AdditionalParameters& ap) {
        py::gil_scoped_release rel;
        return sv::cg(&x, A, gf, b, l.system, l.handler, prec, ap);
    });
    m.def("bicgstab", [](const Spinorfield_eo& x, const fm::Fermionmatrix_eo& A, const Gaugefield& gf,
                         const Spinorfield_eo& b, Lattice& l, hmc_float prec, const AdditionalParameters& ap) {
        py::gil_scoped_release rel;
        return sv::bicgstab(&x, A, gf, b, l.system, l.handler, prec, ap);
    });
    m.def(
        "cg_reliable_update",
        [](const Spinorfield_eo& x, const fm::Fermionmatrix_eo& A, const Gaugefield& gf, const Spinorfield_eo& b,
           Lattice& l, hmc_float rel_tol, const AdditionalParameters& ap, int inner_prec, double delta, int maxit,
           bool use_graph) {
            sv::ReliableUpdateOptions o;
            o.inner_precision = inner_prec;
            o.delta           = delta;
            o.max_iterations  = maxit;
            o.use_graph       = use_graph;
            sv::SolveStats st;
            {
                py::gil_scoped_release rel;
                sv::cg_reliable_update(&x, A, gf, b, l.system, rel_tol, ap, o, &st);
            }
            py::dict d;
            d["iterations"]             = st.iterations;
            d["reliable_updates"]       = st.reliable_updates;
            d["graph_nodes"]            = st.graph_nodes;
            d["graph_kernel_nodes"]     = st.graph_kernel_nodes;
            d["true_relative_residual"] = st.true_relative_residual;
            return d;
        },
        py::arg("x"), py::arg("A"), py::arg("gf"), py::arg("b"), py::arg("lattice"), py::arg("rel_tol"), py::arg("ap"),
        py::arg("inner_prec") = CL2_F32, py::arg("delta") = 0.1, py::arg("max_iterations") = 100000,
        py::arg("use_graph") = true);
    m.def(
        "cg_reliable_update_block",
        [](std::vector<Spinorfield_eo*> xs, const fm::Fermionmatrix_eo& A, const Gaugefield& gf,
           std::vector<Spinorfield_eo*> bs, Lattice& l, hmc_float rel_tol, const AdditionalParameters& ap,
           int inner_prec, double delta, int maxit, bool use_graph) {
            sv::ReliableUpdateOptions o;
            o.inner_precision = inner_prec;
            o.delta           = delta;
            o.max_iterations  = maxit;
            o.use_graph       = use_graph;
            std::vector<const Spinorfield_eo*> x(xs.begin(), xs.end()), b(bs.begin(), bs.end());
            std::vector<sv::SolveStats> st;
            int it;
            {
                py::gil_scoped_release rel;
                it = sv::cg_reliable_update_block(x, A, gf, b, l.system, rel_tol, ap, o, &st);
            }
            py::list per;
            for (const auto& s : st) {
                py::dict d;
                d["iterations"]             = s.iterations;
                d["reliable_updates"]       = s.reliable_updates;
                d["graph_nodes"]            = s.graph_nodes;
                d["graph_kernel_nodes"]     = s.graph_kernel_nodes;
                d["true_relative_residual"] = s.true_relative_residual;
                per.append(d);
            }
            return py::make_tuple(it, per);
        },
        py::arg("x"), py::arg("A"), py::arg("gf"), py::arg("b"), py::arg("lattice"), py::arg("rel_tol"), py::arg("ap"),
        py::arg("inner_prec") = CL2_F32, py::arg("delta") = 0.1, py::arg("max_iterations") = 100000,
        py::arg("use_graph") = true);
    m.def(
        "cg_reliable_update_block_params",
        [](std::vector<Spinorfield_eo*> xs, const fm::Fermionmatrix_eo& A, const Gaugefield& gf,
           std::vector<Spinorfield_eo*> bs, Lattice& l, hmc_float rel_tol, std::vector<hmc_float> kappa,
           std::vector<hmc_float> mubar, int inner_prec, double delta, int maxit, bool use_graph) {
            if (kappa.size() != bs.size() || mubar.size() != bs.size())
                throw std::invalid_argument("one kappa and one mubar per right-hand side");
            sv::ReliableUpdateOptions o;
            o.inner_precision = inner_prec;
            o.delta           = delta;
            o.max_iterations  = maxit;
            o.use_graph       = use_graph;
            std::vector<const Spinorfield_eo*> x(xs.begin(), xs.end()), b(bs.begin(), bs.end());
            std::vector<std::unique_ptr<physics::WilsonAdditionalParameters>> own;
            std::vector<const AdditionalParameters*> aps;
            for (size_t r = 0; r < kappa.size(); ++r) {
                own.emplace_back(new physics::WilsonAdditionalParameters(kappa[r], mubar[r]));
                aps.push_back(own.back().get());
            }
            std::vector<sv::SolveStats> st;
            int it;
            {
                py::gil_scoped_release rel;
                it = sv::cg_reliable_update_block(x, A, gf, b, l.system, rel_tol, aps, o, &st);
            }
            py::list per;
            for (const auto& s : st) {
                py::dict d;
                d["iterations"]             = s.iterations;
                d["reliable_updates"]       = s.reliable_updates;
                d["graph_nodes"]            = s.graph_nodes;
                d["graph_kernel_nodes"]     = s.graph_kernel_nodes;
                d["true_relative_residual"] = s.true_relative_residual;
                per.append(d);
            }
            return py::make_tuple(it, per);
        },
        py::arg("x"), py::arg("A"), py::arg("gf"), py::arg("b"), py::arg("lattice"), py::arg("rel_tol"),
        py::arg("kappa"), py::arg("mubar"), py::arg("inner_prec") = CL2_F32, py::arg("delta") = 0.1,
        py::arg("max_iterations") = 100000, py::arg("use_graph") = true);
    auto block_opts = [](int inner_prec, double rel_tol, int block) {
        physics::algorithms::BlockInversionOptions o;
        o.solver.inner_precision = inner_prec;
        o.rel_tol                = rel_tol;
        o.block                  = block;
        return o;
    };
    m.def(
        "point_source_ps_correlator",
        [block_opts](const Gaugefield& gf, Lattice& l, bool block, int inner_prec, double rel_tol, int block_size) {
            const auto o = block_opts(inner_prec, rel_tol, block_size);
            py::gil_scoped_release rel;
            return physics::algorithms::point_source_ps_correlator(
                gf, l.system, l.handler, l.handler.getAdditionalParameters<Spinorfield_eo>(), block ? &o : nullptr);
        },
        py::arg("gf"), py::arg("lattice"), py::arg("block") = false, py::arg("inner_prec") = CL2_F32,
        py::arg("rel_tol") = 1e-10, py::arg("block_size") = 12);
    m.def(
        "ps_correlator",
        [block_opts](const Gaugefield& gf, Lattice& l, bool block, int inner_prec, double rel_tol, int block_size) {
            const auto o = block_opts(inner_prec, rel_tol, block_size);
            py::gil_scoped_release rel;
            return physics::algorithms::ps_correlator(gf, l.system, l.handler,
                                                      l.handler.getAdditionalParameters<Spinorfield_eo>(), l.prng,
                                                      block ? &o : nullptr);
        },
        py::arg("gf"), py::arg("lattice"), py::arg("block") = false, py::arg("inner_prec") = CL2_F32,
        py::arg("rel_tol") = 1e-10, py::arg("block_size") = 12);
    m.def(
        "invert_eo",
        [](const Spinorfield_eo& pe, const Spinorfield_eo& po, const Gaugefield& gf, const Spinorfield_eo& ee,
           const Spinorfield_eo& eo, Lattice& l, hmc_float prec) {
            physics::algorithms::InversionStats st;
            {
                py::gil_scoped_release rel;
                physics::algorithms::invert_eo(&pe, &po, gf, ee, eo, l.system, l.handler, prec,
                                               l.handler.getAdditionalParameters<Spinorfield_eo>(), &st);
            }
            py::dict d;
            d["iterations"]          = st.iterations;
            d["bicgstab_iterations"] = st.bicgstab_iterations;
            d["cg_fallback"]         = st.cg_fallback;
            return d;
        });
    m.def("create_sources", [](Lattice& l, size_t n) {
        py::list out;
        for (auto* f : physics::create_sources(l.system, l.prng, n, l.handler)) {
            out.append(to_numpy(f->get_host()));
            delete f;
        }
        return out;
    });
    m.def(
        "autotune_dslash",
        [](const Gaugefield& gf, Lattice& l, int prec, int nrhs) {
            py::gil_scoped_release rel;
            return fm::autotune_dslash(gf, l.system, prec, nrhs);
        },
        py::arg("gf"), py::arg("lattice"), py::arg("prec") = CL2_F64, py::arg("nrhs") = 1);
    m.def(
        "autotune_dks",
        [](const Gaugefield& gf, Lattice& l, int prec) {
            py::gil_scoped_release rel;
            return fm::autotune_dks(gf, l.system, prec);
        },
        py::arg("gf"), py::arg("lattice"), py::arg("prec") = CL2_F64);
    m.def("convert_precision",
          [](const Spinorfield_eo& to, const Spinorfield_eo& from) { convert_precision(&to, from); });
    m.def(
        "cg_reliable_update_stagg",
        [](const Staggeredfield_eo& x, const fm::Fermionmatrix_stagg_eo& A, const Gaugefield& gf, hmc_float sigma,
           const Staggeredfield_eo& b, Lattice& l, hmc_float rel_tol, const AdditionalParameters& ap, int inner_prec,
           double delta, int maxit, bool use_graph) {
            sv::ReliableUpdateOptions o;
            o.inner_precision = inner_prec;
            o.delta           = delta;
            o.max_iterations  = maxit;
            o.use_graph       = use_graph;
            sv::SolveStats st;
            {
                py::gil_scoped_release rel;
                sv::cg_reliable_update_stagg(&x, A, gf, sigma, b, l.system, rel_tol, ap, o, &st);
            }
            py::dict d;
            d["iterations"]             = st.iterations;
            d["reliable_updates"]       = st.reliable_updates;
            d["graph_nodes"]            = st.graph_nodes;
            d["graph_kernel_nodes"]     = st.graph_kernel_nodes;
            d["true_relative_residual"] = st.true_relative_residual;
            return d;
        },
        py::arg("x"), py::arg("A"), py::arg("gf"), py::arg("sigma"), py::arg("b"), py::arg("lattice"),
        py::arg("rel_tol"), py::arg("ap"), py::arg("inner_prec") = CL2_F32, py::arg("delta") = 0.1,
        py::arg("max_iterations") = 100000, py::arg("use_graph") = true);
    {
        namespace pa = physics::algorithms;
        auto coeffs = [](const pa::Rational_Coefficients& c) {
            py::dict d;
            d["order"] = c.Get_order();
            d["a0"]    = c.Get_a0();
            d["a"]     = c.Get_a();
            d["b"]     = c.Get_b();
            return d;
        };
        py::class_<pa::Rational_Approximation>(m, "Rational_Approximation")
            .def(py::init([](int d, int y, int z, hmc_float low, hmc_float high, bool inv) {
                     py::gil_scoped_release rel;
                     return new pa::Rational_Approximation(d, y, z, low, high, inv);
                 }),
                 py::arg("d"), py::arg("y"), py::arg("z"), py::arg("low"), py::arg("high"), py::arg("inv") = true)
            .def(py::init<const std::string&>(), py::arg("filename"))
            .def("order", &pa::Rational_Approximation::Get_order)
            .def("a0", &pa::Rational_Approximation::Get_a0)
            .def("a", &pa::Rational_Approximation::Get_a)
            .def("b", &pa::Rational_Approximation::Get_b)
            .def("error", &pa::Rational_Approximation::Get_error)
            .def("exponent", &pa::Rational_Approximation::Get_exponent)
            .def("iterations", &pa::Rational_Approximation::iterations)
            .def("evaluate", &pa::Rational_Approximation::evaluate)
            .def("save", &pa::Rational_Approximation::Save_rational_approximation)
            .def("rescale", [coeffs](const pa::Rational_Approximation& r, hmc_float mn, hmc_float mx) {
                return coeffs(r.Rescale_Coefficients(mn, mx));
            });
    }
    m.def(
        "cg_reliable_update_stagg_block",
        [](std::vector<Staggeredfield_eo*> xs, const fm::MdagM_eo& A, const Gaugefield& gf, std::vector<hmc_float> sigma,
           const Staggeredfield_eo& b, Lattice& l, hmc_float rel_tol, const AdditionalParameters& ap, int inner_prec,
           double delta, bool use_graph) {
            sv::ReliableUpdateOptions o;
            o.inner_precision = inner_prec;
            o.delta           = delta;
            o.use_graph       = use_graph;
            std::vector<const Staggeredfield_eo*> x(xs.begin(), xs.end());
            std::vector<sv::SolveStats> st;
            int it;
            {
                py::gil_scoped_release rel;
                it = sv::cg_reliable_update_stagg_block(x, A, gf, sigma, b, l.system, rel_tol, ap, o, &st);
            }
            py::list per;
            for (const auto& s : st) {
                py::dict d;
                d["iterations"]             = s.iterations;
                d["reliable_updates"]       = s.reliable_updates;
                d["graph_nodes"]            = s.graph_nodes;
                d["graph_kernel_nodes"]     = s.graph_kernel_nodes;
                d["true_relative_residual"] = s.true_relative_residual;
                per.append(d);
            }
            return py::make_tuple(it, per);
        },
        py::arg("x"), py::arg("A"), py::arg("gf"), py::arg("sigma"), py::arg("b"), py::arg("lattice"),
        py::arg("rel_tol"), py::arg("ap"), py::arg("inner_prec") = CL2_F32, py::arg("delta") = 0.1,
        py::arg("use_graph") = true);
    m.def(
        "cg_m_mixed",
        [](std::vector<Staggeredfield_eo*> xs, const fm::Fermionmatrix_stagg_eo& A, const Gaugefield& gf,
           std::vector<hmc_float> sigma, const Staggeredfield_eo& b, Lattice& l, hmc_float rel_tol,
           const AdditionalParameters& ap, int inner_prec, double sloppy_tol, double delta, int maxit,
           bool use_graph) {
            std::vector<std::shared_ptr<Staggeredfield_eo>> x;
            for (auto* p : xs) x.emplace_back(p, [](Staggeredfield_eo*) {});  // Python owns the fields
            sv::MultishiftOptions o;
            o.inner_precision = inner_prec;
            o.sloppy_tol      = sloppy_tol;
            o.delta           = delta;
            o.max_iterations  = maxit;
            o.use_graph       = use_graph;
            std::vector<sv::SolveStats> st;
            int total = 0, ms = 0;
            {
                py::gil_scoped_release rel;
                total = sv::cg_m_mixed(x, A, gf, sigma, b, l.system, rel_tol, ap, o, &st, &ms);
            }
            py::list per;
            for (const auto& s : st) {
                py::dict d;
                d["iterations"]             = s.iterations;
                d["reliable_updates"]       = s.reliable_updates;
                d["graph_nodes"]            = s.graph_nodes;
                d["graph_kernel_nodes"]     = s.graph_kernel_nodes;
                d["true_relative_residual"] = s.true_relative_residual;
                per.append(d);
            }
            return py::make_tuple(total, ms, per);
        },
        py::arg("x"), py::arg("A"), py::arg("gf"), py::arg("sigma"), py::arg("b"), py::arg("lattice"),
        py::arg("rel_tol"), py::arg("ap"), py::arg("inner_prec") = CL2_F32, py::arg("sloppy_tol") = 1e-7,
        py::arg("delta") = 0.1, py::arg("max_iterations") = 100000, py::arg("use_graph") = true);
    m.def("cg_m", [](std::vector<Staggeredfield_eo*> xs, const fm::Fermionmatrix_stagg_eo& A, const Gaugefield& gf,
                     std::vector<hmc_float> sigma, const Staggeredfield_eo& b, Lattice& l, hmc_float prec,
                     const AdditionalParameters& ap) {
        std::vector<std::shared_ptr<Staggeredfield_eo>> x;
        for (auto* p : xs) x.emplace_back(p, [](Staggeredfield_eo*) {});  // Python owns the fields
        py::gil_scoped_release rel;
        return sv::cg_m(x, A, gf, sigma, b, l.system, l.handler, prec, ap);
    });

    py::class_<serving::Engine>(m, "Engine")
        .def(py::init([](size_t hbm_budget_bytes, int workers, int inner_precision, int max_batch_rhs, bool mixed,
                         bool join_running) {
                 serving::EngineOptions o;
                 o.hbm_budget_bytes = hbm_budget_bytes;
                 o.workers          = workers;
                 o.inner_precision  = inner_precision;
                 o.max_batch_rhs    = max_batch_rhs;
                 o.mixed            = mixed;
                 o.join_running     = join_running;
                 return new serving::Engine(o);
             }),
             py::arg("hbm_budget_bytes") = (size_t)64 << 30, py::arg("workers") = 4,
             py::arg("inner_precision") = CL2_F32, py::arg("max_batch_rhs") = 12, py::arg("mixed") = true,
             py::arg("join_running") = true)
        .def(
            "load_config",
            [](serving::Engine& e, int ns, int nt, double beta, int sweeps, uint64_t seed) {
                py::gil_scoped_release rel;
                return e.load_config(ns, nt, beta, sweeps, seed);
            },
            py::arg("ns"), py::arg("nt"), py::arg("beta"), py::arg("sweeps"), py::arg("seed"))
        .def(
            "load_config_file",
            [](serving::Engine& e, int ns, int nt, const std::string& path) {
                py::gil_scoped_release rel;
                return e.load_config_file(ns, nt, path);
            },
            py::arg("ns"), py::arg("nt"), py::arg("path"))
        .def(
            "save_config",
            [](const serving::Engine& e, int config, const std::string& path) {
                py::gil_scoped_release rel;
                e.save_config(config, path);
            },
            py::arg("config"), py::arg("path"))
        .def("plaquette", &serving::Engine::plaquette)
        .def("eo_sites", &serving::Engine::eo_sites)
        .def("resident_bytes", &serving::Engine::resident_bytes)
        .def("peak_work_bytes", &serving::Engine::peak_work_bytes)
        .def("reset_peaks", &serving::Engine::reset_peaks)
        .def("stats",
             [](const serving::Engine& e) {
                 const serving::EngineStats s = e.stats();
                 py::dict d;
                 d["solves_started"]  = s.solves_started;
                 d["requests_merged"] = s.requests_merged;
                 d["requests_joined"] = s.requests_joined;
                 d["max_queue_depth"] = s.max_queue_depth;
                 d["queue_depth"]     = s.queue_depth;
                 return d;
             })
        .def("now", &serving::Engine::now)
        .def(
            "submit",
            [](serving::Engine& e, int config, const std::string& action, double kappa, double mu, double mass,
               std::vector<py::array_t<double, py::array::c_style | py::array::forcecast>> sources,
               std::vector<uint64_t> gaussian_seeds, double rel_tol, bool return_solutions) {
                serving::Request r;
                r.config = config;
                if (action == "twistedmass")
                    r.action = serving::Action::twistedmass;
                else if (action == "staggered")
                    r.action = serving::Action::staggered;
                else
                    throw std::invalid_argument("action must be twistedmass or staggered");
                r.kappa   = kappa;
                r.mu      = mu;
                r.mass    = mass;
                r.rel_tol = rel_tol;
                for (auto& a : sources) r.sources.emplace_back(a.data(), a.data() + a.size());
                r.gaussian_seeds   = gaussian_seeds;
                r.return_solutions = return_solutions;
                py::gil_scoped_release rel;
                return e.submit(std::move(r));
            },
            py::arg("config"), py::arg("action"), py::arg("kappa") = 0.16, py::arg("mu") = 0.005,
            py::arg("mass") = 0.01, py::arg("sources") = std::vector<py::array_t<double>>(),
            py::arg("gaussian_seeds") = std::vector<uint64_t>(), py::arg("rel_tol") = 1e-10,
            py::arg("return_solutions") = true)
        .def("warmup", &serving::Engine::warmup)
        .def("wait",
             [](serving::Engine& e, long id) {
                 serving::Result r;
                 {
                     py::gil_scoped_release rel;
                     r = e.wait(id);
                 }
                 py::dict d;
                 d["id"]      = r.id;
                 d["ok"]      = r.ok;
                 d["error"]   = r.error;
                 py::list sols;
                 for (const auto& s : r.solutions) {
                     const size_t w = (size_t)r.width;
                     py::array_t<double> a({s.size() / w, w});
                     std::memcpy(a.mutable_data(), s.data(), s.size() * sizeof(double));
                     sols.append(a);
                 }
                 d["solutions"]         = sols;
                 d["true_rel_residual"] = r.true_rel_residual;
                 d["iterations"]        = r.iterations;
                 d["batch_requests"]    = r.batch_requests;
                 d["batch_rhs"]         = r.batch_rhs;
                 d["joined_running"]    = r.joined_running;
                 d["solve"]             = r.solve;
                 d["arrival"]           = r.arrival;
                 d["admitted"]          = r.admitted;
                 d["finished"]          = r.finished;
                 return d;
             })
        .def("drain", [](serving::Engine& e) {
            py::gil_scoped_release rel;
            e.drain();
        });
}
