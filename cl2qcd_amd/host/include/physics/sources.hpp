// Inverter sources (physics/sources.hpp:34-64, physics/sources.cpp:35-160).
//
// Full-lattice spinor fields are host arrays in this build (they only feed convert_to_eoprec), so the
// sources are written on the host.  Random contents come from a counter-based generator keyed by a seed
// from the PRNG, not from the reference's ranluxcl streams, which depend on the device's thread count.
#pragma once

#include <vector>

#include "interfacesHandler.hpp"
#include "lattices/lattices.hpp"

namespace physics {

/* point source k (0..11) at (sourceX, sourceY, sourceZ, sourceT): 1 in spin k % 4, colour k / 4
 * (ocl_kernel/spinorfield_point_source.cl, operations_geometry.cl:476-484); content 'one' only, as the
 * reference's kernel */
void set_point_source(const lattices::Spinorfield* spinorfield, int k, const SourcesParametersInterface& params);
/* every site filled with the source content (ocl_kernel/spinorfield_volume_source.cl) */
void set_volume_source(const lattices::Spinorfield* spinorfield, const PRNG& prng);
/* the sites of time slice t filled, zero elsewhere (ocl_kernel/spinorfield_timeslice_source.cl) */
void set_timeslice_source(const lattices::Spinorfield* spinorfield, const PRNG& prng, int t);
/* the sites of z-slice z filled, zero elsewhere (ocl_kernel/spinorfield_zslice_source.cl) */
void set_zslice_source(const lattices::Spinorfield* spinorfield, const PRNG& prng, int z);

/* the sources of the inverter: for k < n_sources, a point source k % 12 or a fresh random source of the
 * configured type (physics/sources.cpp:108-160 fill_sources); the caller owns the fields */
std::vector<lattices::Spinorfield*> create_sources(const hardware::System& system, const PRNG& prng, size_t n_sources,
                                                   InterfacesHandler& interfacesHandler);

}  // namespace physics
