// Host helpers shared by the C-ABI translation units: error recording,
// checked device allocation and launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/clsph_hip.h"
#include "state.h"

namespace clsph {
namespace host {

inline thread_local std::string g_err;  // errors raised before a context exists

inline int fail(clsph_ctx* ctx, int code, const std::string& msg) {
  if (ctx)
    ctx->last_error = msg;
  else
    g_err = msg;
  return code;
}

#define HIPCK(ctx, expr)                                                        \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                       \
      return ::clsph::host::fail((ctx), CLSPH_EDEVICE,                          \
                                 std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dalloc(clsph_ctx* ctx, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(ctx, CLSPH_ENOMEM,
                std::string("hipMalloc of ") + std::to_string(sizeof(T) * count) +
                    " bytes: " + hipGetErrorString(e));
  }
  return CLSPH_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

inline int check_launch(clsph_ctx* ctx) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(ctx, CLSPH_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  return CLSPH_OK;
}

// Reads the device error bits of every scene and turns them into a return
// code (defined in capi.cpp).
int device_errors(clsph_state* st);
// Sets the particle count of a single-scene state: work-group tables and
// device parameters follow (defined in capi.cpp).
int resize_single(clsph_state* st, uint32_t n);

}  // namespace host
}  // namespace clsph
