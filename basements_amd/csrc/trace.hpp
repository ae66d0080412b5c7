// trace.hpp -- roctx ranges around the library's host entry points and the
// request loop's stages, for rocprofv3 --marker-trace and any other roctx
// consumer.  Without a tool attached a range costs two calls that return at
// once.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace bm {

struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

}  // namespace bm

#define BM_TRACE_CAT2(a, b) a##b
#define BM_TRACE_CAT(a, b) BM_TRACE_CAT2(a, b)
// A range from here to the end of the enclosing scope.
#define BM_TRACE(name) const ::bm::TraceRange BM_TRACE_CAT(bm_trace_range_, __LINE__)(name)
