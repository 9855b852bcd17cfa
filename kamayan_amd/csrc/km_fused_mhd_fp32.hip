// km_fused_mhd_fp32.hip -- fused CT-MHD stage kernels, fp32 form (see km_fused_mhd.inc).
#include <hip/hip_runtime.h>

#include "km_device.hpp"
#include "km_internal.hpp"

#define KM_FUSED_NS fp32
#define KM_FUSED_T float
#include "km_fused_mhd.inc"
