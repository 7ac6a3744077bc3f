// Event-based per-kernel timing for ffhip_profile_start/stop.  When profiling
// is off the hooks cost one branch.  prof_launch times a launch with events
// that the dispatch itself stamps at the kernel's start and end
// (hipExtLaunchKernel), so the figure is the kernel's execution time without
// the gaps that separate event records would add around it.
#pragma once

#include <hip/hip_ext.h>

#include "common.h"

namespace ffhip {

extern bool g_prof_on;
void prof_begin_impl(const char* name, double bytes, hipStream_t st);
void prof_end_impl(hipStream_t st);

inline void prof_begin(const char* name, double bytes, hipStream_t st) {
  if (g_prof_on) prof_begin_impl(name, bytes, st);
}
inline void prof_end(hipStream_t st) {
  if (g_prof_on) prof_end_impl(st);
}

// a start/stop event pair recorded under `name` (false if none is available)
bool prof_pair(const char* name, double bytes, hipEvent_t* a, hipEvent_t* b);

template <typename F, typename... Args>
inline void prof_launch(const char* name, double bytes, F kern, dim3 grid, dim3 block, size_t sh, hipStream_t st,
                        Args... args) {
  hipEvent_t a = nullptr, b = nullptr;
  if (g_prof_on && prof_pair(name, bytes, &a, &b))
    hipExtLaunchKernelGGL(kern, grid, block, (std::uint32_t)sh, st, a, b, 0u, args...);
  else
    hipLaunchKernelGGL(kern, grid, block, sh, st, args...);
}

}  // namespace ffhip
