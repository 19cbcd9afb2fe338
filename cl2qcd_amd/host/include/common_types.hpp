// Host-side value types of the CL2QCD interface (common_header_files/types.hpp:46-206,
// types_fermions.hpp:31-68, globaldefs.hpp:30-49), with the same memory layouts so
// host arrays filled for the reference can be imported unchanged.
#pragma once

#include <cstdint>

typedef double hmc_float;
typedef uint64_t cl_ulong;

struct hmc_complex {
    hmc_float re;
    hmc_float im;
};
constexpr hmc_complex hmc_complex_one{1., 0.};
constexpr hmc_complex hmc_complex_zero{0., 0.};
constexpr hmc_complex hmc_complex_minusone{-1., 0.};

struct su3vec {
    hmc_complex e0, e1, e2;
};
struct spinor {
    su3vec e0, e1, e2, e3;
};
struct Matrixsu3 {
    hmc_complex e00, e01, e02, e10, e11, e12, e20, e21, e22;
};

#ifndef EVEN
#define EVEN 0
#endif
#ifndef ODD
#define ODD 1
#endif

namespace common {
enum action { wilson = 0, clover, twistedmass, tlsym, iwasaki, dbw2, rooted_stagg };
enum solver { cg = 0, bicgstab, bicgstab_save };
enum startcondition { cold_start = 0, hot_start, start_from_source };
/* inverter sources (common_header_files/types.hpp:203-204) */
enum sourcetypes { point = 1, volume, timeslice, zslice };
enum sourcecontents { one = 1, z4, gaussian, z2 };
}  // namespace common
