// km_fused_fp32.hip -- fused RK-stage sweeps, fp32 form (see km_fused.inc).
#include <hip/hip_runtime.h>

#include "km_device.hpp"
#include "km_internal.hpp"

#define KM_FUSED_NS fp32
#define KM_FUSED_T float
#include "km_fused.inc"
#include "km_stage3.inc"
#include "km_stage2.inc"
