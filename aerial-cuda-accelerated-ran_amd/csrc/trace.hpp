// Named host ranges around the API's phases, visible to rocprofv3
// --marker-trace (roctx): the counterpart of the reference's NVTX
// PUSH_RANGE / POP_RANGE (cuPHY/src/cuphy/cuphy_internal.h:38-52), which
// cuPHY wraps around the phases of PuschRx::run (pusch_rx.cpp:5982-5998).
//
// Built in by default; -DCUPHY_HIP_NO_ROCTX compiles every range away (as
// USE_NVTX=0 does in the reference).  With no tool attached a range costs
// two calls into the roctx library, so ranges sit around API calls and
// phases, never around per-code-block work.
#pragma once

#ifndef CUPHY_HIP_NO_ROCTX
#include <rocprofiler-sdk-roctx/roctx.h>

namespace nrtrace {

// One roctx range for the lifetime of the object (push in the constructor,
// pop in the destructor), so that every return path closes it.
class range {
public:
    explicit range(const char* name) { roctxRangePushA(name); }
    ~range() { roctxRangePop(); }
    range(const range&) = delete;
    range& operator=(const range&) = delete;
};

}  // namespace nrtrace

#define NRTRACE_CAT2(a, b) a##b
#define NRTRACE_CAT(a, b) NRTRACE_CAT2(a, b)
// A range named `name` (a string literal) from here to the end of the scope.
#define CUPHY_HIP_RANGE(name) ::nrtrace::range NRTRACE_CAT(nrtrace_range_, __LINE__)(name)
#else
#define CUPHY_HIP_RANGE(name) ((void)0)
#endif
