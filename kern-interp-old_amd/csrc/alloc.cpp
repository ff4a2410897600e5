// Caching device allocator behind DBuf (see common.h).
#include <map>
#include <mutex>

#include "common.h"

namespace kigpu {

namespace {
std::mutex g_mu;
std::multimap<size_t, void*> g_free;  // rounded size -> block (current device only)
}  // namespace

size_t dev_round(size_t bytes) {
  size_t r = 256;
  while (r < bytes) r <<= 1;
  return r;
}

void* dev_alloc(size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_free.find(bytes);
    if (it != g_free.end()) {
      void* p = it->second;
      g_free.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    // give cached blocks back and try once more
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& kv : g_free) (void)hipFree(kv.second);
    g_free.clear();
    (void)hipGetLastError();
    e = hipMalloc(&p, bytes);
  }
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    throw HipError{e == hipErrorOutOfMemory ? KIGPU_ERR_NOMEM : KIGPU_ERR_HIP};
  }
  return p;
}

void dev_free(void* p, size_t bytes) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_mu);
  g_free.emplace(bytes, p);
}

}  // namespace kigpu
