// Pinned host memory shared by the host runtimes (mhmsw.cpp, locassm.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>

// Grown as needed and kept: staging copies go from it at the link's full rate
// and run asynchronously on the stream that issues them.
struct HostPinned {
  uint8_t* ptr = nullptr;
  size_t cap = 0;
  // touch: write every page now (a latency-bound caller reserving ahead),
  // not on a later copy's critical path
  hipError_t ensure(size_t n, bool touch = false) {
    if (n <= cap) return hipSuccess;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(n + n / 8, 4096);
    const hipError_t e = hipHostMalloc((void**)&ptr, want, hipHostMallocDefault);
    if (e == hipSuccess) {
      cap = want;
      if (touch) std::memset(ptr, 0, want);
    }
    return e;
  }
  void release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};
