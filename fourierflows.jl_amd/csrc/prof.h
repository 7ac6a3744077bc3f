// Event-based per-kernel timing for ffhip_profile_start/stop.  When profiling
// is off the hooks cost one branch.
#pragma once

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

}  // namespace ffhip
