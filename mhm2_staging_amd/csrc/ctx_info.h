// What an entry point outside mhmsw.cpp needs from a context (mhmsw_ctx is
// defined there): its device, stream, scoring and launch parameters, and its
// last-error string.
#pragma once
#include <hip/hip_runtime_api.h>

#include "../../include/mhmsw.h"
#include "sw_common.h"

namespace mhmsw {

struct CtxInfo {
  int device;
  int cus;
  hipStream_t stream;
  ScoreParams sc;
  int mode;
  mhmsw_tuning tuning;
};

// false for a null context
bool ctx_info(const mhmsw_ctx* c, CtxInfo* out);
// sets the context's last-error message and returns code
int ctx_fail(mhmsw_ctx* c, int code, const char* msg);

// Maps the no-allocation entry point's scratch array on `device` (once per
// device; noalloc.hip).  Returns MHMSW_OK or MHMSW_E_HIP.
int noalloc_setup(int device);

}  // namespace mhmsw
