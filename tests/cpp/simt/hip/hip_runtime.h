// Host emulation of the HIP device built-ins that the Gauss-Jordan routines
// of kern-interp-old_amd/csrc/device_util.h use, for a CPU test under
// AddressSanitizer.  Found first on the include path of that test only, in
// place of the real <hip/hip_runtime.h>; the library never sees it.
//
// One OS thread per GPU thread of one workgroup.  __syncthreads is a
// barrier over the workgroup, and every cross-lane built-in (DPP moves,
// readlane, shuffles, ballots) goes through an exchange slot per lane with
// a barrier over the wavefront before and after the read, so all 64 lanes
// of a wavefront must reach each of them, as on the GPU.
#pragma once

#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
// a variable in LDS: one copy for the process, which runs one workgroup
#define __shared__ static

namespace simt {

struct Dim3 {
  unsigned x = 0, y = 0, z = 0;
};

struct Block {
  Dim3 dim;
  std::unique_ptr<std::barrier<>> bar;
  std::vector<std::unique_ptr<std::barrier<>>> wbar;  // one per wavefront
  std::vector<uint64_t> xch;                          // one exchange slot per thread
};

inline thread_local Dim3 tl_tid;
inline thread_local Block* tl_blk = nullptr;

// Value of lane src_of(lane) of the calling thread's wavefront (own value
// when src_of gives a lane outside the wavefront).
template <class F>
inline uint64_t exchange(uint64_t v, F src_of) {
  const int t = (int)tl_tid.x, lane = t & 63, w = t >> 6;
  tl_blk->xch[t] = v;
  tl_blk->wbar[w]->arrive_and_wait();
  const int s = src_of(lane);
  const uint64_t r = (s >= 0 && s < 64) ? tl_blk->xch[w * 64 + s] : v;
  tl_blk->wbar[w]->arrive_and_wait();
  return r;
}

// run body() on nthreads emulated GPU threads of one workgroup
template <class F>
inline void launch(int nthreads, F body) {
  Block b;
  b.dim.x = (unsigned)nthreads;
  b.dim.y = b.dim.z = 1;
  b.bar = std::make_unique<std::barrier<>>(nthreads);
  for (int w = 0; w < (nthreads + 63) / 64; ++w)
    b.wbar.push_back(std::make_unique<std::barrier<>>(std::min(64, nthreads - 64 * w)));
  b.xch.assign(nthreads, 0);
  std::vector<std::thread> th;
  th.reserve(nthreads);
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&, t] {
      tl_tid.x = (unsigned)t;
      tl_blk = &b;
      body();
    });
  for (auto& x : th) x.join();
}

inline uint64_t bits(int v) { return (uint32_t)v; }

}  // namespace simt

#define threadIdx (simt::tl_tid)
#define blockDim (simt::tl_blk->dim)

inline void __syncthreads() { simt::tl_blk->bar->arrive_and_wait(); }

inline void __builtin_amdgcn_wave_barrier() {
  const int w = (int)simt::tl_tid.x >> 6;
  simt::tl_blk->wbar[w]->arrive_and_wait();
}
inline void __builtin_amdgcn_fence(int, const char*) {}

// DPP: quad_perm (ctrl < 0x100) and row_ror:n (0x121 .. 0x12F), the two
// forms device_util.h uses; lane i reads lane (i - n) mod 16 of its row.
inline int __builtin_amdgcn_update_dpp(int, int v, int ctrl, int, int, bool) {
  return (int)simt::exchange(simt::bits(v), [ctrl](int lane) {
    if (ctrl < 0x100) return (lane & ~3) | ((ctrl >> (2 * (lane & 3))) & 3);
    const int n = ctrl & 0xF;
    return (lane & ~15) | (((lane & 15) - n) & 15);
  });
}
inline int __builtin_amdgcn_readlane(int v, int l) {
  return (int)simt::exchange(simt::bits(v), [l](int) { return l; });
}
inline int __builtin_amdgcn_readfirstlane(int v) { return __builtin_amdgcn_readlane(v, 0); }

template <typename T>
inline T __shfl(T v, int src, int width = 64) {
  static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
  uint64_t u = 0;
  std::memcpy(&u, &v, sizeof(T));
  u = simt::exchange(u, [src, width](int lane) { return (lane / width) * width + src % width; });
  T r;
  std::memcpy(&r, &u, sizeof(T));
  return r;
}
template <typename T>
inline T __shfl_up(T v, unsigned d, int width = 64) {
  uint64_t u = 0;
  std::memcpy(&u, &v, sizeof(T));
  u = simt::exchange(u, [d, width](int lane) {
    const int base = (lane / width) * width;
    return lane - (int)d >= base ? lane - (int)d : lane;
  });
  T r;
  std::memcpy(&r, &u, sizeof(T));
  return r;
}
inline unsigned long long __ballot(int pred) {
  const int t = (int)simt::tl_tid.x, lane = t & 63, w = t >> 6;
  simt::tl_blk->xch[t] = pred ? 1 : 0;
  simt::tl_blk->wbar[w]->arrive_and_wait();
  unsigned long long m = 0;
  for (int i = 0; i < 64; ++i)
    if (simt::tl_blk->xch[w * 64 + i]) m |= 1ull << i;
  simt::tl_blk->wbar[w]->arrive_and_wait();
  (void)lane;
  return m;
}
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }

inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
inline void __builtin_amdgcn_sched_barrier(int) {}
inline float __int_as_float(int v) {
  float f;
  std::memcpy(&f, &v, 4);
  return f;
}
inline unsigned __float_as_uint(float f) {
  unsigned v;
  std::memcpy(&v, &f, 4);
  return v;
}
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline int min(int a, int b) { return a < b ? a : b; }
inline int __float_as_int(float f) {
  int v;
  std::memcpy(&v, &f, 4);
  return v;
}
inline long long __double_as_longlong(double d) {
  long long v;
  std::memcpy(&v, &d, 8);
  return v;
}
inline double __longlong_as_double(long long v) {
  double d;
  std::memcpy(&d, &v, 8);
  return d;
}
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
using std::fabs;
using std::fma;
using std::fmaf;
inline float fabsf(float x) { return std::fabs(x); }
using std::sqrt;
