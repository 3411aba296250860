// engine_internal.h -- engine hooks shared by the batch/graph and autotune
// translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "msosm_c.h"
#include "plan.h"

namespace msosm {
int engine_reset_async(msosm_engine *e);
hipStream_t engine_stream(msosm_engine *e);
float2 *engine_output(msosm_engine *e);
const Plan &engine_plan(msosm_engine *e);
} // namespace msosm
