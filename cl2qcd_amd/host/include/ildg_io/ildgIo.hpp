// ILDG/LIME gauge-configuration files (ildg_io/ildgIo_gaugefield.hpp, ildg_io/checksum.hpp,
// ildg_io/lime/limeFileWriter.hpp): host-side reading and writing of the links, with the SciDAC checksum.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../common_types.hpp"

namespace ildgIo {

/* SciDAC checksum of a configuration (ildg_io/checksum.hpp:29-70): per site of rank
 * ((t NS + z) NS + y) NS + x the CRC-32 of its 4 links as stored in the file, rotated left by rank % 29
 * into suma and by rank % 31 into sumb (XOR). */
struct Checksum {
    uint32_t suma = 0, sumb = 0;
    bool operator==(const Checksum& o) const { return suma == o.suma && sumb == o.sumb; }
    bool operator!=(const Checksum& o) const { return !(*this == o); }
};

/* CRC-32 (IEEE 802.3, reflected, as boost::crc_32_type) of n bytes */
uint32_t crc32(const unsigned char* p, size_t n);

/* checksum of the ildg-binary-data payload of an NS^3 x NT configuration
 * (ildg_io/ildgIo_gaugefield.cpp:104-131); nbytes must be the payload size of the precision */
Checksum calculate_ildg_checksum(const char* buf, size_t nbytes, size_t NT, size_t NS);

/* what the file says besides the links */
struct Sourcefileparameters {
    int lx = 0, ly = 0, lz = 0, lt = 0, precision = 64;
    Checksum checksum;        // the file's scidac-checksum record (zero if absent)
    bool has_checksum = false;
    std::string xlf_info;     // the xlf-info record, as text
};

/* The links of an ILDG file of an ns^3 x nt lattice, in the reference's host order (link mu + 4 site,
 * site = x + ns (y + ns (z + ns t)), internal mu = T,X,Y,Z; the file stores X,Y,Z,T).  Throws
 * std::invalid_argument if the extents differ, std::runtime_error on a malformed file or, unless
 * ignore_checksum, on a payload that does not match the file's checksum
 * (ildg_io/ildgIo_gaugefield.cpp:44-67). */
std::vector<Matrixsu3> readGaugefieldFromSourcefile(const std::string& file, int ns, int nt,
                                                    Sourcefileparameters* info = nullptr,
                                                    bool ignore_checksum = false);

/* Physics parameters recorded in the xlf-info record (ildg_io/sourcefileParameters/SourcefileParameters.cpp:140-159) */
struct XlfInfo {
    double plaquette = 0., beta = 0., kappa = 0., mu = 0., mubar = 0.;
    int trajectory = 0;
};

/* Writes the four records the reference writes (ildg_io/ildgIo_gaugefield.cpp:77-101): xlf-info,
 * ildg-format, ildg-binary-data (precision 64 or 32, big-endian) and scidac-checksum; an existing file
 * of that name is first renamed to name + "_backup" (lime/limeFileWriter.cpp:33-47).  The file is read
 * back and its checksum verified before returning. */
void writeGaugefieldToFile(const std::string& file, const std::vector<Matrixsu3>& links, int ns, int nt,
                           const XlfInfo& info, int precision = 64);

}  // namespace ildgIo
