// def.hpp -- dispatch flags of the MI355X path (reference: def.hpp:7-13,
// prog.hpp:35).
#ifndef TIPL_HIP_DEF_HPP
#define TIPL_HIP_DEF_HPP

#include "../tiplhip.h"

namespace tipl {

// The HIP path is compiled in whenever these headers are used.
constexpr bool use_hip = true;

// tipl::has_gpu: a visible HIP device.
inline bool has_gpu()
{
    int n = 0;
    return tiplhip_device_count(&n) == TIPLHIP_OK && n > 0;
}

// The default cancellation flag of the reference's entry points.
inline bool prog_aborted = false;

}  // namespace tipl

#endif
