// Process-wide record of dynamic-LDS limits (no HIP dependency, so that the
// CPU tests can drive it from many threads).
#pragma once

#include <mutex>
#include <utility>
#include <vector>

namespace kigpu {

// Dynamic-LDS limits granted to kernels (hipFuncAttributeMaxDynamicSharedMemorySize).
// The limit belongs to the kernel on a device, for every thread of the
// process, so the record of what was granted is process-wide and a limit is
// only ever raised: a thread that asks for less must not lower the limit
// under a launch of another thread that asked for more (with a record per
// thread, a new thread set its smaller size over the larger one in force).
// `set(f, bytes)` raises the limit; it runs under the record's lock, so two
// threads cannot raise it out of order.
struct LdsGrants {
  std::mutex mu;
  std::vector<std::pair<std::pair<const void*, int>, size_t>> granted;
  template <typename Set>
  void allow(const void* f, int dev, size_t bytes, Set&& set) {
    std::lock_guard<std::mutex> lk(mu);
    for (auto& g : granted)
      if (g.first.first == f && g.first.second == dev) {
        if (bytes <= g.second) return;
        set(f, bytes);
        g.second = bytes;
        return;
      }
    set(f, bytes);
    granted.push_back({{f, dev}, bytes});
  }
};

}  // namespace kigpu
