// Inverter sources on the host (physics/sources.cpp:35-160).
#include "../include/physics/sources.hpp"

#include <cmath>
#include <cstring>
#include <stdexcept>

namespace physics {

namespace {

/* uniform doubles in [0, 1) from (seed, draw index): splitmix64 finaliser of the sum */
inline double uniform(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + 0x9e3779b97f4a7c15ull * (i + 1);
    z          = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z          = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

/* one complex entry of a random source, draw n of the field (ocl_kernel/random.cl: Z4_complex_number
 * :240-266, gaussianNormalPair :206-236 (Box-Muller, unit variance per real part), Z2_complex_number
 * :270-290; 'one' is the unit entry of set_spinor_cold, operations_spinor.cl:50-58) */
hmc_complex draw(common::sourcecontents c, uint64_t seed, uint64_t n)
{
    const double u1 = uniform(seed, 2 * n), u2 = uniform(seed, 2 * n + 1);
    switch (c) {
        case common::one: return hmc_complex{1., 0.};
        case common::z4: {
            const double h = 1. / std::sqrt(2.);
            return hmc_complex{u1 > 0.5 ? h : -h, u2 > 0.5 ? h : -h};
        }
        case common::gaussian: {
            const double p = std::sqrt(-2. * std::log(1. - u1));
            const double a = 2. * 3.14159265358979323846 * u2;
            return hmc_complex{p * std::cos(a), p * std::sin(a)};
        }
        case common::z2: return hmc_complex{u1 > 0.5 ? 1. : -1., 0.};
    }
    throw std::invalid_argument("unknown source content");
}

struct Extents {
    long ns, nt;
};
Extents extents(const lattices::Spinorfield* sf)
{
    const auto& p = sf->get_system().get_inputparameters();
    return Extents{p.get_nspace(), p.get_ntime()};
}

/* fill the sites selected by keep(x, y, z, t) with the content, the others with zero; entry k of site s is
 * draw 12 s + k, so a source depends on its seed only */
template<typename Keep>
void fill(const lattices::Spinorfield* sf, uint64_t seed, Keep keep)
{
    const Extents e                 = extents(sf);
    const common::sourcecontents c = sf->get_system().get_inputparameters().get_sourcecontent();
    std::vector<spinor> h(sf->get_elements());
    std::memset(h.data(), 0, h.size() * sizeof(spinor));
    for (long t = 0; t < e.nt; ++t)
        for (long z = 0; z < e.ns; ++z)
            for (long y = 0; y < e.ns; ++y)
                for (long x = 0; x < e.ns; ++x) {
                    if (!keep(x, y, z, t)) continue;
                    const long s = x + e.ns * (y + e.ns * (z + e.ns * t));
                    auto* v      = reinterpret_cast<hmc_complex*>(&h[s]);
                    for (int k = 0; k < 12; ++k) v[k] = draw(c, seed, 12 * (uint64_t)s + k);
                }
    sf->import(h.data());
}

}  // namespace

void set_point_source(const lattices::Spinorfield* sf, int k, const SourcesParametersInterface& params)
{
    if (k >= 12 || k < 0) throw std::invalid_argument("k must be within 0..11");
    if (sf->get_system().get_inputparameters().get_sourcecontent() != common::one)
        throw std::invalid_argument("point sources take the source content 'one' only");
    const long ns = params.getNs(), nt = params.getNt();
    const long x = params.getSourceX(), y = params.getSourceY(), z = params.getSourceZ(), t = params.getSourceT();
    if (x >= ns || y >= ns || z >= ns || t >= nt) throw std::invalid_argument("point source outside the lattice");
    std::vector<spinor> h(sf->get_elements());
    std::memset(h.data(), 0, h.size() * sizeof(spinor));
    const int colour = k / 4, spin = k - 4 * colour;
    reinterpret_cast<hmc_complex*>(&h[x + ns * (y + ns * (z + ns * t))])[3 * spin + colour] = hmc_complex{1., 0.};
    sf->import(h.data());
}

void set_volume_source(const lattices::Spinorfield* sf, const PRNG& prng)
{
    fill(sf, prng.next_seed(), [](long, long, long, long) { return true; });
}

void set_timeslice_source(const lattices::Spinorfield* sf, const PRNG& prng, int t)
{
    if (t < 0 || t >= extents(sf).nt) throw std::invalid_argument("time slice outside the lattice");
    fill(sf, prng.next_seed(), [t](long, long, long, long tt) { return tt == t; });
}

void set_zslice_source(const lattices::Spinorfield* sf, const PRNG& prng, int z)
{
    if (z < 0 || z >= extents(sf).ns) throw std::invalid_argument("z-slice outside the lattice");
    fill(sf, prng.next_seed(), [z](long, long, long zz, long) { return zz == z; });
}

std::vector<lattices::Spinorfield*> create_sources(const hardware::System& system, const PRNG& prng, size_t n_sources,
                                                   InterfacesHandler& ih)
{
    const SourcesParametersInterface& p = ih.getSourcesParametersInterface();
    std::vector<lattices::Spinorfield*> out;
    try {
        for (size_t k = 0; k < n_sources; ++k) {
            out.push_back(new lattices::Spinorfield(system, ih.getInterface<lattices::Spinorfield>()));
            switch (p.getSourceType()) {
                case common::point: set_point_source(out.back(), (int)(k % 12), p); break;
                case common::volume: set_volume_source(out.back(), prng); break;
                case common::timeslice: set_timeslice_source(out.back(), prng, (int)p.getSourceT()); break;
                case common::zslice: set_zslice_source(out.back(), prng, (int)p.getSourceZ()); break;
                default: throw std::domain_error("no such sourcetype");
            }
        }
    } catch (...) {
        for (auto* f : out) delete f;
        throw;
    }
    return out;
}

}  // namespace physics
