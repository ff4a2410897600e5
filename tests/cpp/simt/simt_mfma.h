// Host emulation of the matrix-core steps v_mfma_f64_16x16x4f64 and
// v_mfma_f32_16x16x4f32 for the thread-emulation tests
// (tests/cpp/simt/hip/hip_runtime.h); clang only (the operand types are clang
// vectors).  Operand layout as on gfx950: lane l holds A(l % 16, l / 16) and
// B(l / 16, l % 16); accumulator element v of lane l is D((l / 16) + 4 v,
// l % 16) for FP64 and D(4 (l / 16) + v, l % 16) for FP32.  All 64 lanes of
// the wavefront must call them.
#pragma once

#include <hip/hip_runtime.h>

typedef double simt_d4 __attribute__((ext_vector_type(4)));
typedef float simt_f4 __attribute__((ext_vector_type(4)));

namespace simt {
inline std::vector<double>& mfma_slots() {
  static std::vector<double> s(2 * 4096);
  return s;
}
}  // namespace simt

inline simt_d4 __builtin_amdgcn_mfma_f64_16x16x4f64(double a, double b, simt_d4 c, int, int, int) {
  const int t = (int)simt::tl_tid.x, lane = t & 63, w = t >> 6;
  std::vector<double>& s = simt::mfma_slots();
  s[2 * t] = a;
  s[2 * t + 1] = b;
  simt::tl_blk->wbar[w]->arrive_and_wait();
  simt_d4 d = c;
  for (int v = 0; v < 4; ++v) {
    const int m = (lane >> 4) + 4 * v, nn = lane & 15;
    double acc = c[v];
    for (int kk = 0; kk < 4; ++kk)
      acc += s[2 * (w * 64 + m + 16 * kk)] * s[2 * (w * 64 + nn + 16 * kk) + 1];
    d[v] = acc;
  }
  simt::tl_blk->wbar[w]->arrive_and_wait();
  return d;
}

inline simt_f4 __builtin_amdgcn_mfma_f32_16x16x4f32(float a, float b, simt_f4 c, int, int, int) {
  const int t = (int)simt::tl_tid.x, lane = t & 63, w = t >> 6;
  std::vector<double>& s = simt::mfma_slots();
  s[2 * t] = a;
  s[2 * t + 1] = b;
  simt::tl_blk->wbar[w]->arrive_and_wait();
  simt_f4 d = c;
  for (int v = 0; v < 4; ++v) {
    const int m = 4 * (lane >> 4) + v, nn = lane & 15;
    float acc = c[v];
    for (int kk = 0; kk < 4; ++kk)
      acc += (float)s[2 * (w * 64 + m + 16 * kk)] * (float)s[2 * (w * 64 + nn + 16 * kk) + 1];
    d[v] = acc;
  }
  simt::tl_blk->wbar[w]->arrive_and_wait();
  return d;
}
