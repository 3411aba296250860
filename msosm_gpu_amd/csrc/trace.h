// trace.h -- rocTX ranges around the C ABI's work entry points, so a
// profiler timeline (rocprofv3 --marker-trace) shows which host call
// enqueued which kernels.  Without a profiler attached a range costs one
// library call.
#pragma once
#include <rocprofiler-sdk-roctx/roctx.h>

namespace msosm {
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};
} // namespace msosm
