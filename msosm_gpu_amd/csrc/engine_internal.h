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
// Host-side stream position of an engine: the global index of its next block
// (shared-ring forms) or the product-ring cursor (reference-structured form).
struct EngineCursor {
    long long g0 = 0;
    int current_index = 0;
    bool have_output = false;
};
EngineCursor engine_cursor(const msosm_engine *e);
void engine_set_cursor(msosm_engine *e, const EngineCursor &c);
bool engine_timing(const msosm_engine *e);
} // namespace msosm
