// tipl.hpp -- umbrella header of the MI355X registration path: the subset of
// TIPL's API that tipl::reg::linear and tipl::reg::cdm_common need, backed
// by libtiplhip.so (include/tiplhip.h).
#ifndef TIPL_HIP_TIPL_HPP
#define TIPL_HIP_TIPL_HPP

#include "def.hpp"
#include "filter/gaussian.hpp"
#include "hip.hpp"
#include "image.hpp"
#include "io/nifti.hpp"
#include "numerical/dif.hpp"
#include "optimization.hpp"
#include "reg/cdm.hpp"
#include "reg/linear.hpp"
#include "reg/mm_reg.hpp"
#include "transformation.hpp"

#endif
