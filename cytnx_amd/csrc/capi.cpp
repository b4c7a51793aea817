// Library-wide C-ABI entry points: error reporting and version; the set-up upload stream.
#include <mutex>
#include <string>

#include "cxa_common.hpp"

namespace cxa {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
const std::string &get_error() { return g_last_error; }

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_upload_mu;
hipStream_t g_upload[kMaxDevices] = {};  // kept until the process ends
}  // namespace

int upload_stream(hipStream_t *s) {
  int dev = 0;
  CXA_HIP(hipGetDevice(&dev));
  CXA_CHECK(dev >= 0 && dev < kMaxDevices, CXA_ERR_INVALID, "upload: device %d", dev);
  std::lock_guard<std::mutex> lk(g_upload_mu);
  if (!g_upload[dev]) CXA_HIP(hipStreamCreateWithFlags(&g_upload[dev], hipStreamNonBlocking));
  *s = g_upload[dev];
  return CXA_OK;
}

int upload_now(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return CXA_OK;
  hipStream_t s;
  const int rc = upload_stream(&s);
  if (rc) return rc;
  CXA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  CXA_HIP(hipStreamSynchronize(s));
  return CXA_OK;
}

int zero_now(void *dst, size_t bytes) {
  if (bytes == 0) return CXA_OK;
  hipStream_t s;
  const int rc = upload_stream(&s);
  if (rc) return rc;
  CXA_HIP(hipMemsetAsync(dst, 0, bytes, s));
  CXA_HIP(hipStreamSynchronize(s));
  return CXA_OK;
}

}  // namespace cxa

extern "C" const char *cxa_last_error(void) { return cxa::get_error().c_str(); }

extern "C" int cxa_abi_version(void) { return 6; }
