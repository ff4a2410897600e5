// Caching device allocator behind DBuf (see common.h).
#include <map>
#include <mutex>
#include <utility>
#include <vector>

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

// Blocks freed on this thread wait here until dev_release_pending(): a
// block can still be read or written by kernels queued before its free, on
// either stream of a context, so it is handed out again only once the
// thread has waited for those streams.
namespace {
struct Pending {
  std::vector<std::pair<size_t, void*>> blocks;
  ~Pending() {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& b : blocks) g_free.emplace(b.first, b.second);
  }
};
thread_local Pending t_pending;
}  // namespace

void dev_free(void* p, size_t bytes) {
  if (!p) return;
  t_pending.blocks.emplace_back(bytes, p);
}

void dev_release_pending() {
  if (t_pending.blocks.empty()) return;
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& b : t_pending.blocks) g_free.emplace(b.first, b.second);
  t_pending.blocks.clear();
}

}  // namespace kigpu
