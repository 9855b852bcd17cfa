// km_fused_fma.hip -- fused RK-stage sweeps, fma form (see km_fused.inc).
#include <hip/hip_runtime.h>

#include "km_device.hpp"
#include "km_internal.hpp"

#define KM_FUSED_NS fma
#define KM_FUSED_T double
#define KM_FUSED_FAST 1
#include "km_fused.inc"
#include "km_stage3.inc"
#include "km_stage2.inc"
