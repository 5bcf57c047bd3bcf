// Kernel timing hooks (see profile.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace dem {
// Returns a slot (>= 0) when profiling is on, -1 otherwise.
int prof_begin(const char* name, hipStream_t st);
void prof_end(int slot, hipStream_t st);
}  // namespace dem

#define DEM_PROF_LAUNCH(name, st, ...)                \
  do {                                                \
    int slot_ = dem::prof_begin(name, st);            \
    __VA_ARGS__;                                      \
    dem::prof_end(slot_, st);                         \
  } while (0)
