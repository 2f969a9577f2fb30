// Internal helpers shared by the host setup code and the HIP kernels.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mgpbd.h"

void mgpbd_set_error(const std::string& msg);
int mgpbd_fail(int code, const std::string& msg);

#define MGPBD_HIP_TRY(expr)                                                          \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess)                                                            \
      return mgpbd_fail(MGPBD_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// A launch is checked right after it is queued; kernel faults surface at the next sync.
#define MGPBD_LAUNCH_CHECK(name)                                                     \
  do {                                                                               \
    hipError_t _e = hipGetLastError();                                               \
    if (_e != hipSuccess)                                                            \
      return mgpbd_fail(MGPBD_EHIP, std::string(name) + ": " + hipGetErrorString(_e)); \
  } while (0)

// The stream a context queues its work on (the caller's stream outside a PCG solve).
extern "C" void* mgpbd_internal_stream(mgpbd_ctx* ctx);
