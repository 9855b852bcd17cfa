// km_fused_mhd_exact.hip -- fused CT-MHD stage kernels, exact form (see km_fused_mhd.inc).
#include <hip/hip_runtime.h>

#include "km_device.hpp"
#include "km_internal.hpp"

#define KM_FUSED_NS exact
#define KM_FUSED_T double
#include "km_fused_mhd.inc"
