// Threads of one process asking for dynamic-LDS limits of one kernel, the
// way the factorization's launches do from the threads of a caller; after
// each request the thread "launches" with its size, which needs the limit in
// force to be at least that size.  Compared: the process-wide record of
// kern-interp-old_amd/csrc/lds_grants.h, and a record kept per thread.
// Prints the number of launches that found the limit below their size.
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "lds_grants.h"

namespace {
std::atomic<size_t> g_limit{0};  // the attribute in force for the kernel
std::atomic<long> g_short{0};     // launches that found it below their size
std::atomic<long> g_lowered{0};   // sets that lowered it

void set_limit(const void*, size_t b) {
  if (b < g_limit.load()) ++g_lowered;
  g_limit.store(b);
}

void launch(size_t bytes) {
  if (g_limit.load() < bytes) ++g_short;
}

// the per-thread record: each thread remembers only its own grants
void allow_per_thread(const void* f, size_t bytes) {
  thread_local std::vector<std::pair<const void*, size_t>> granted;
  for (auto& g : granted)
    if (g.first == f) {
      if (bytes <= g.second) return;
      set_limit(f, bytes);
      g.second = bytes;
      return;
    }
  set_limit(f, bytes);
  granted.push_back({f, bytes});
}

// all threads meet here: every thread makes its request of a round before
// any thread launches, whatever the host's scheduling, so the requests of
// different threads interleave in every round
struct Barrier {
  std::mutex mu;
  std::condition_variable cv;
  int count = 0, gen = 0;
  const int n;
  explicit Barrier(int n_) : n(n_) {}
  void wait() {
    std::unique_lock<std::mutex> lk(mu);
    const int g = gen;
    if (++count == n) {
      count = 0;
      ++gen;
      cv.notify_all();
    } else {
      cv.wait(lk, [&] { return gen != g; });
    }
  }
};
}  // namespace

int main(int argc, char** argv) {
  const bool per_thread = argc > 1 && argv[1][0] == 't';
  static int kernel;  // stands for the kernel's address
  kigpu::LdsGrants grants;
  constexpr int kThreads = 8;
  Barrier bar(kThreads);
  std::vector<std::thread> th;
  for (int t = 0; t < kThreads; ++t)
    th.emplace_back([&, t] {
      std::mt19937 rng(t);
      for (int i = 0; i < 2000; ++i) {
        const size_t bytes = 1024 * (1 + rng() % 150);
        if (per_thread) allow_per_thread(&kernel, bytes);
        else grants.allow(&kernel, 0, bytes, set_limit);
        bar.wait();
        launch(bytes);
        bar.wait();
      }
    });
  for (auto& x : th) x.join();
  std::printf("%ld %ld\n", g_short.load(), g_lowered.load());
  return 0;
}
