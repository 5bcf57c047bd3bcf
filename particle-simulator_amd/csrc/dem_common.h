// Shared helpers of libdemhip: status handling, launch checks, alignment.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <stdio.h>

#include "../../include/dem_hip.h"

namespace dem {

// Records the message for dem_last_error() and returns `code`.
int set_error(int code, const char* fmt, ...);

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Carves consecutive 256-byte-aligned sub-buffers out of a caller workspace.
struct Carver {
  char* base;
  size_t used = 0;
  explicit Carver(void* b) : base(static_cast<char*>(b)) {}
  // With a null base the returned "pointers" are byte offsets (used to size
  // workspaces and to locate named sub-buffers without a device).
  template <class T>
  T* take(size_t count) {
    used = align_up(used, 256);
    T* p = reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(base) + used);
    used += count * sizeof(T);
    return p;
  }
};

inline hipStream_t as_stream(dem_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace dem

#define DEM_CHECK_LAUNCH(name)                                                       \
  do {                                                                               \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess)                                                            \
      return dem::set_error(DEM_ERR_HIP, "%s: %s", name, hipGetErrorString(e_));     \
  } while (0)

#define DEM_CHECK_HIP(call)                                                          \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return dem::set_error(DEM_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_));    \
  } while (0)
