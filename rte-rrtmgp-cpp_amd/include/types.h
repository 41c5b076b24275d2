// Float / Bool typedefs of the class API (reference include/types.h:13-25): RTE_USE_SP selects
// single precision, RTE_USE_CBOOL makes Bool a signed char (always the case for the HIP kernels).
#ifndef RRTMGP_AMD_TYPES_H
#define RRTMGP_AMD_TYPES_H

#include <cfloat>

#ifdef RTE_USE_CBOOL
using Bool = signed char;
#else
using Bool = signed char;
#endif

#ifdef RTE_USE_SP
using Float = float;
const Float Float_epsilon = FLT_EPSILON;
#else
using Float = double;
const Float Float_epsilon = DBL_EPSILON;
#endif

#endif
