// Shared plumbing for the cytnx_amd C-ABI: status codes, the thread-local error message,
// HIP error mapping and dtype sizes.  Every entry point returns a status; none exits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/cytnx_amd.h"

namespace cxa {

void set_error(const std::string &msg);
const std::string &get_error();

// element size in bytes for a reference dtype id (include/Type.hpp:177-179); 0 if unknown
inline int elem_size(int dtype) {
  switch (dtype) {
    case CXA_C128: return 16;
    case CXA_C64: return 8;
    case CXA_F64: return 8;
    case CXA_F32: return 4;
    case CXA_I64: return 8;
    case CXA_U64: return 8;
    case CXA_I32: return 4;
    case CXA_U32: return 4;
    case CXA_I16: return 2;
    case CXA_U16: return 2;
    case CXA_BOOL: return 1;
    default: return 0;
  }
}

inline int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
inline int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  set_error(buf);
  return code;
}

// A host-to-device copy of set-up data (plan metadata) that waits only for itself: issued on
// a non-blocking stream of the library's own (one per device, made on first use) and waited
// for there, so it neither waits for nor is ordered after work queued on the caller's streams.
// The data is on the device when it returns.
int upload_now(void *dst, const void *src, size_t bytes);
// The same for clearing device memory (bytes set to zero).
int zero_now(void *dst, size_t bytes);
// That stream, for other set-up work that must not wait for the caller's queued work (the
// caller synchronises it before the results are used).
int upload_stream(hipStream_t *s);

// Device memory of objects that are made and dropped often (a DMRG run makes a chain per bond
// update): pool_release keeps a block for reuse instead of freeing it (hipFree waits for the
// device to drain), pool_alloc hands out a kept block of the same size class (sizes round up
// to quarter octaves), after one device synchronisation if blocks were released since the
// last, so that no launch queued before a release still uses the block.  At most a few GiB
// per device are kept; beyond that, and for blocks the pool did not make, release frees.
int pool_alloc(void **p, size_t bytes);
void pool_release(void *p);

}  // namespace cxa

#define CXA_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return cxa::fail(CXA_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                             \
  } while (0)

#define CXA_RC(call)              \
  do {                            \
    const int rc_ = (call);       \
    if (rc_) return rc_;          \
  } while (0)

#define CXA_CHECK(cond, code, ...)                 \
  do {                                                \
    if (!(cond)) return cxa::fail((code), __VA_ARGS__); \
  } while (0)
