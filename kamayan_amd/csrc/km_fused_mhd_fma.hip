// km_fused_mhd_fma.hip -- fused CT-MHD stage kernels, fma form (see km_fused_mhd.inc).
#include <hip/hip_runtime.h>

#include "km_device.hpp"
#include "km_internal.hpp"

#define KM_FUSED_NS fma
#define KM_FUSED_T double
#define KM_FUSED_FAST 1
#include "km_fused_mhd.inc"
