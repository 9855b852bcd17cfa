// km_fused_exact.hip -- fused RK-stage sweeps, exact form (see km_fused.inc).
#include <hip/hip_runtime.h>

#include "km_device.hpp"
#include "km_internal.hpp"

#define KM_FUSED_NS exact
#define KM_FUSED_T double
#include "km_fused.inc"
#include "km_stage3.inc"
#include "km_stage2.inc"
