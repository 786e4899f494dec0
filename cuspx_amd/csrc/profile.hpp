// profile.hpp -- optional per-kernel HIP-event timing of the launch
// sequences.  When enabled, every launcher records an event after each
// kernel on the stream it launched on; durations are the gaps between
// consecutive events of one pipeline call.
#pragma once
#include <hip/hip_runtime.h>

namespace spx {
void prof_begin(hipStream_t s, const char* set);  // start of one pipeline call
void prof_mark(hipStream_t s, const char* kernel);  // after a kernel launch
}  // namespace spx
