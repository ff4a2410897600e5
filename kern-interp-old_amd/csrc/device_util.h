// Wavefront / workgroup reductions for wave64 (gfx950).
#pragma once

#include <hip/hip_runtime.h>

namespace kigpu {

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// argmax over (value, index): larger value wins, ties go to the smaller index
__device__ __forceinline__ void wave_argmax(double& v, int& idx) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double v2 = __shfl_xor(v, o, 64);
    int i2 = __shfl_xor(idx, o, 64);
    if (v2 > v || (v2 == v && i2 < idx)) {
      v = v2;
      idx = i2;
    }
  }
}

// sum over a workgroup of NW waves; `red` holds NW doubles of LDS.  Every
// thread gets the total, combined in wave order (deterministic).
template <int NW>
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) t += red[i];
  __syncthreads();
  return t;
}

}  // namespace kigpu

namespace kigpu {

// Solve X u = b (trans = false) or X^T u = b (trans = true) with the LU
// factors of an r x r matrix (X = P L U, LAPACK dgetrf convention: unit lower
// L and U packed in `lu` with leading dimension ld, piv[c] = row swapped with
// row c).  One wavefront, lane i holds u_i, so r <= 64.  All lanes of the wave
// must call it; the result is returned in lane i.
template <typename T>
__device__ __forceinline__ T wave_lu_solve(const T* lu, int ld, const int* piv, int r, T u,
                                           bool trans) {
  const int lane = threadIdx.x & 63;
  if (!trans) {
    for (int c = 0; c < r; ++c) {  // P^T b: interchanges in forward order (dlaswp)
      int p = piv[c];
      if (p != c) {
        T vc = __shfl(u, c, 64), vp = __shfl(u, p, 64);
        if (lane == c) u = vp;
        if (lane == p) u = vc;
      }
    }
    for (int t = 0; t < r - 1; ++t) {  // unit lower
      T ut = __shfl(u, t, 64);
      if (lane > t && lane < r) u -= lu[lane + (size_t)t * ld] * ut;
    }
    for (int t = r - 1; t >= 0; --t) {  // upper
      if (lane == t) u /= lu[t + (size_t)t * ld];
      T ut = __shfl(u, t, 64);
      if (lane < t) u -= lu[lane + (size_t)t * ld] * ut;
    }
    return u;
  }
  for (int t = 0; t < r; ++t) {  // U^T w = b (lower)
    if (lane == t) u /= lu[t + (size_t)t * ld];
    T ut = __shfl(u, t, 64);
    if (lane > t && lane < r) u -= lu[t + (size_t)lane * ld] * ut;
  }
  for (int t = r - 1; t > 0; --t) {  // L^T v = w (unit upper)
    T ut = __shfl(u, t, 64);
    if (lane < t) u -= lu[t + (size_t)lane * ld] * ut;
  }
  for (int c = r - 1; c >= 0; --c) {  // y = P v: interchanges in reverse order
    int p = piv[c];
    if (p != c) {
      T vc = __shfl(u, c, 64), vp = __shfl(u, p, 64);
      if (lane == c) u = vp;
      if (lane == p) u = vc;
    }
  }
  return u;
}

// The same for any r, one thread doing the whole column (used when r > 64).
template <typename T>
__device__ inline void thread_lu_solve(const T* lu, int ld, const int* piv, int r, T* u,
                                       size_t stride, bool trans) {
  auto U_ = [&](int i) -> T& { return u[(size_t)i * stride]; };
  if (!trans) {
    for (int c = 0; c < r; ++c) {
      int p = piv[c];
      if (p != c) {
        T t0 = U_(c);
        U_(c) = U_(p);
        U_(p) = t0;
      }
    }
    for (int i = 0; i < r; ++i) {
      T s = U_(i);
      for (int t = 0; t < i; ++t) s -= lu[i + (size_t)t * ld] * U_(t);
      U_(i) = s;
    }
    for (int i = r - 1; i >= 0; --i) {
      T s = U_(i);
      for (int t = i + 1; t < r; ++t) s -= lu[i + (size_t)t * ld] * U_(t);
      U_(i) = s / lu[i + (size_t)i * ld];
    }
    return;
  }
  for (int i = 0; i < r; ++i) {
    T s = U_(i);
    for (int t = 0; t < i; ++t) s -= lu[t + (size_t)i * ld] * U_(t);
    U_(i) = s / lu[i + (size_t)i * ld];
  }
  for (int i = r - 1; i >= 0; --i) {
    T s = U_(i);
    for (int t = i + 1; t < r; ++t) s -= lu[t + (size_t)i * ld] * U_(t);
    U_(i) = s;
  }
  for (int c = r - 1; c >= 0; --c) {
    int p = piv[c];
    if (p != c) {
      T t0 = U_(c);
      U_(c) = U_(p);
      U_(p) = t0;
    }
  }
}

// solve for `nvec` vectors held in memory (vector v at base + v * vstride,
// element i at + i * estride), one wave per vector
template <typename T, int NW>
__device__ inline void block_lu_solve(const T* lu, int ld, const int* piv, int r, T* base,
                                      size_t vstride, size_t estride, int nvec, bool trans) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (r <= 64) {
    for (int v = wv; v < nvec; v += NW) {
      T* b = base + (size_t)v * vstride;
      T u = lane < r ? b[(size_t)lane * estride] : T(0);
      u = wave_lu_solve(lu, ld, piv, r, u, trans);
      if (lane < r) b[(size_t)lane * estride] = u;
    }
  } else {
    for (int v = threadIdx.x; v < nvec; v += NW * 64)
      thread_lu_solve(lu, ld, piv, r, base + (size_t)v * vstride, estride, trans);
  }
}

}  // namespace kigpu
