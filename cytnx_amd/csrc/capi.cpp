// Library-wide C-ABI entry points: error reporting and version; the set-up upload stream.
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>

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

namespace {
struct Pool {
  std::mutex mu;
  std::multimap<size_t, void *> kept[kMaxDevices];  // size class -> blocks
  std::unordered_map<void *, std::pair<size_t, int>> made;  // block -> (class, device)
  size_t held[kMaxDevices] = {};
  bool dirty[kMaxDevices] = {};  // blocks released since the last device synchronisation
};
Pool &pool() {
  static Pool *p = new Pool();  // lives until the process ends
  return *p;
}
constexpr size_t kPoolCap = (size_t)4 << 30;

size_t size_class(size_t b) {
  if (b <= 4096) return 4096;
  const size_t base = (size_t)1 << (63 - __builtin_clzll(b - 1));  // base < b <= 2 base
  for (int q = 1; q < 4; ++q)
    if (base + (base >> 2) * q >= b) return base + (base >> 2) * q;
  return base << 1;
}

// free every kept block of a device (after it has drained)
int pool_trim(Pool &P, int dev) {
  CXA_HIP(hipDeviceSynchronize());
  for (auto &kv : P.kept[dev]) {
    P.made.erase(kv.second);
    (void)hipFree(kv.second);
  }
  P.kept[dev].clear();
  P.held[dev] = 0;
  P.dirty[dev] = false;
  return CXA_OK;
}
}  // namespace

int pool_alloc(void **p, size_t bytes) {
  *p = nullptr;
  int dev = 0;
  CXA_HIP(hipGetDevice(&dev));
  CXA_CHECK(dev >= 0 && dev < kMaxDevices, CXA_ERR_INVALID, "pool: device %d", dev);
  const size_t cls = size_class(bytes);
  Pool &P = pool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.kept[dev].find(cls);
  if (it != P.kept[dev].end()) {
    if (P.dirty[dev]) {
      CXA_HIP(hipDeviceSynchronize());
      P.dirty[dev] = false;
    }
    *p = it->second;
    P.kept[dev].erase(it);
    P.held[dev] -= cls;
    return CXA_OK;
  }
  if (hipMalloc(p, cls) != hipSuccess) {
    (void)hipGetLastError();
    CXA_RC(pool_trim(P, dev));
    CXA_HIP(hipMalloc(p, cls));
  }
  P.made[*p] = {cls, dev};
  return CXA_OK;
}

void pool_release(void *p) {
  if (!p) return;
  Pool &P = pool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.made.find(p);
  if (it == P.made.end()) {
    (void)hipFree(p);
    return;
  }
  const size_t cls = it->second.first;
  const int dev = it->second.second;
  if (P.held[dev] + cls > kPoolCap) {
    P.made.erase(it);
    (void)hipFree(p);
    return;
  }
  P.kept[dev].emplace(cls, p);
  P.held[dev] += cls;
  P.dirty[dev] = true;
}

}  // namespace cxa

extern "C" const char *cxa_last_error(void) { return cxa::get_error().c_str(); }

extern "C" int cxa_abi_version(void) { return 6; }
