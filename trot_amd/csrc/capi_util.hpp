// Status, error-message and workspace helpers shared by the C-ABI translation units.
#pragma once

#include <stdio.h>

#include <hip/hip_runtime.h>

#include "../../include/trot_amd.h"

// message of the last failing call on this thread (defined in capi.hip)
extern thread_local char g_trot_err[512];

static inline int fail(int code, const char* msg) {
  snprintf(g_trot_err, sizeof(g_trot_err), "%s", msg);
  return code;
}
static inline int hip_status(hipError_t e, const char* where) {
  if (e == hipSuccess) return TROT_OK;
  snprintf(g_trot_err, sizeof(g_trot_err), "%s: %s", where, hipGetErrorString(e));
  return TROT_ERR_HIP;
}
#define HIPCHK(expr)                                    \
  do {                                                  \
    hipError_t _e = (expr);                             \
    if (_e != hipSuccess) return hip_status(_e, #expr); \
  } while (0)

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// bump allocator over the caller's workspace
struct Carve {
  char* base;
  size_t cap, used;
  template <typename T>
  T* take(size_t count) {
    size_t bytes = align256(count * sizeof(T));
    if (used + bytes > cap) return nullptr;
    T* p = reinterpret_cast<T*>(base + used);
    used += bytes;
    return p;
  }
};
