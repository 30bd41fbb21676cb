// pmp_trace.h -- optional roctx ranges around the C-ABI's calls, for
// rocprofv3 --marker-trace (host-side ranges: the enqueue of the kernels,
// which themselves appear in the kernel trace).  Turned on by the environment
// variable PMP_ROCTX=1 or by pmp_set_roctx; the ROCm marker library is opened
// at run time, so libpmp has no link-time dependency on it, and with tracing
// off a range costs a few loads and a branch.
#pragma once

namespace pmp {

struct TraceRange {
    explicit TraceRange(const char* name);
    ~TraceRange();
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
    bool on;
};

// Turns the ranges on (true if the marker library could be opened) or off.
bool trace_enable(bool on);

}  // namespace pmp
