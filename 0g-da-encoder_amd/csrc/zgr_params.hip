// Translation unit of the parameter-file kernels (params_io.h): the powers of tau in G1
// and G2, the G2 vanishes and the conversions to the serialized point forms.  Kept
// apart from zgr.hip so that the two compile in parallel.
#include <hip/hip_runtime.h>

#include "params_io.h"
