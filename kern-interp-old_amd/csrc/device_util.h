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

namespace kigpu {

// Make LDS writes of this wavefront visible to all its lanes.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// In-place Gauss-Jordan inverse of the n x n matrix A (LDS, column-major,
// leading dimension n) with partial row pivoting: the pivot of step c is the
// first largest |A(i, c)|, i >= c, of the current matrix -- the same choice
// as dgetrf's -- and piv[c] records it.  Executed by ONE wavefront; prow and
// fcol are n-element LDS scratch vectors.  Returns false on a zero pivot.
template <typename T>
__device__ bool wave_gj_inverse(T* A, int n, int* piv, T* prow, T* fcol) {
  const int lane = threadIdx.x & 63;
  bool ok = true;
  for (int c = 0; c < n; ++c) {
    double best = -1.0;
    int bi = n;
    for (int i = c + lane; i < n; i += 64) {
      double v = fabs((double)A[i + (size_t)c * n]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    wave_argmax(best, bi);
    const int p = bi;
    if (lane == 0) piv[c] = p;
    if (p != c)
      for (int j = lane; j < n; j += 64) {
        T t0 = A[c + (size_t)j * n];
        A[c + (size_t)j * n] = A[p + (size_t)j * n];
        A[p + (size_t)j * n] = t0;
      }
    wave_sync();
    const T d = A[c + (size_t)c * n];
    if (d == T(0)) ok = false;
    const T inv = T(1) / d;
    for (int j = lane; j < n; j += 64) prow[j] = j == c ? inv : A[c + (size_t)j * n] * inv;
    for (int i = lane; i < n; i += 64) fcol[i] = A[i + (size_t)c * n];
    wave_sync();
    for (int j = 0; j < n; ++j) {
      const T pj = prow[j];
      for (int i = lane; i < n; i += 64) {
        T* a = A + i + (size_t)j * n;
        if (i == c) *a = pj;
        else if (j == c) *a = -fcol[i] * inv;
        else *a -= fcol[i] * pj;
      }
    }
    wave_sync();
  }
  for (int c = n - 1; c >= 0; --c) {
    const int p = piv[c];
    if (p != c)
      for (int i = lane; i < n; i += 64) {
        T t0 = A[i + (size_t)c * n];
        A[i + (size_t)c * n] = A[i + (size_t)p * n];
        A[i + (size_t)p * n] = t0;
      }
    wave_sync();
  }
  return ok;
}

// The same executed by a whole workgroup (block barriers); for larger n.
template <typename T>
__device__ bool block_gj_inverse(T* A, int n, int* piv, T* prow, T* fcol, int* s_p) {
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6;
  bool ok = true;
  for (int c = 0; c < n; ++c) {
    if (wv == 0) {
      double best = -1.0;
      int bi = n;
      for (int i = c + lane; i < n; i += 64) {
        double v = fabs((double)A[i + (size_t)c * n]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      wave_argmax(best, bi);
      if (lane == 0) {
        *s_p = bi;
        piv[c] = bi;
      }
    }
    __syncthreads();
    const int p = *s_p;
    if (p != c)
      for (int j = tid; j < n; j += nt) {
        T t0 = A[c + (size_t)j * n];
        A[c + (size_t)j * n] = A[p + (size_t)j * n];
        A[p + (size_t)j * n] = t0;
      }
    __syncthreads();
    const T d = A[c + (size_t)c * n];
    if (d == T(0)) ok = false;
    const T inv = T(1) / d;
    for (int j = tid; j < n; j += nt) prow[j] = j == c ? inv : A[c + (size_t)j * n] * inv;
    for (int i = tid; i < n; i += nt) fcol[i] = A[i + (size_t)c * n];
    __syncthreads();
    const int nw = nt >> 6;
    for (int j = wv; j < n; j += nw) {
      const T pj = prow[j];
      T* col = A + (size_t)j * n;
      for (int i = lane; i < n; i += 64) {
        if (i == c) col[i] = pj;
        else if (j == c) col[i] = -fcol[i] * inv;
        else col[i] -= fcol[i] * pj;
      }
    }
    __syncthreads();
  }
  for (int c = n - 1; c >= 0; --c) {
    const int p = piv[c];
    if (p != c)
      for (int i = tid; i < n; i += nt) {
        T t0 = A[i + (size_t)c * n];
        A[i + (size_t)c * n] = A[i + (size_t)p * n];
        A[i + (size_t)p * n] = t0;
      }
    __syncthreads();
  }
  return ok;
}

}  // namespace kigpu

namespace kigpu {

// f(i, j) over a rows x cols index space: wavefront w takes columns
// w, w + NW, ...; its lanes take rows.  Avoids integer division by runtime
// sizes in the hot loops.
template <int NW, typename F>
__device__ __forceinline__ void for2d(int rows, int cols, F&& f) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j = wv; j < cols; j += NW)
    for (int i = lane; i < rows; i += 64) f(i, j);
}

}  // namespace kigpu
