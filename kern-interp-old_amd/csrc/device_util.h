// Wavefront / workgroup reductions for wave64 (gfx950).
#pragma once

#include <hip/hip_runtime.h>

namespace kigpu {

// Cross-lane steps with DPP (no LDS round trip): quad_perm xor 1 and xor 2,
// then row_ror 4 and 8 give every lane the result of its 16-lane row; the
// four rows are combined through readlane, so every lane gets the same,
// deterministically ordered result.  All 64 lanes must be active.
namespace dpp {
constexpr int kXor1 = 0xB1;   // quad_perm [1, 0, 3, 2]
constexpr int kXor2 = 0x4E;   // quad_perm [2, 3, 0, 1]
constexpr int kRor4 = 0x124;  // row_ror:4
constexpr int kRor8 = 0x128;  // row_ror:8

template <int CTRL>
__device__ __forceinline__ int mov(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float mov(float v) {
  return __int_as_float(mov<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = mov<CTRL>((int)b), hi = mov<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ int lane_value(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_value(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
}  // namespace dpp

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += dpp::mov<dpp::kXor1>(v);
  v += dpp::mov<dpp::kXor2>(v);
  v += dpp::mov<dpp::kRor4>(v);
  v += dpp::mov<dpp::kRor8>(v);
  return (dpp::lane_value(v, 0) + dpp::lane_value(v, 16)) +
         (dpp::lane_value(v, 32) + dpp::lane_value(v, 48));
}

// 1/d for d > 0 with one Newton step: relative error below 2.2e-15 over
// d in [1e-20, 1e20] (measured on gfx950 against 1/d in long double on 4M
// log-uniform samples: v_rcp_f64 alone 4.6e-8, one step 2.2e-15), which is
// below the rounding of an N-term sum of the refinement residual.
__device__ __forceinline__ double recip1(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return fma(r, fma(-d, r, 1.0), r);
}

// argmax over (value, index): larger value wins, ties go to the smaller index
__device__ __forceinline__ void argmax_step(double& v, int& idx, double v2, int i2) {
  if (v2 > v || (v2 == v && i2 < idx)) {
    v = v2;
    idx = i2;
  }
}

__device__ __forceinline__ void wave_argmax(double& v, int& idx) {
  argmax_step(v, idx, dpp::mov<dpp::kXor1>(v), dpp::mov<dpp::kXor1>(idx));
  argmax_step(v, idx, dpp::mov<dpp::kXor2>(v), dpp::mov<dpp::kXor2>(idx));
  argmax_step(v, idx, dpp::mov<dpp::kRor4>(v), dpp::mov<dpp::kRor4>(idx));
  argmax_step(v, idx, dpp::mov<dpp::kRor8>(v), dpp::mov<dpp::kRor8>(idx));
  double bv = dpp::lane_value(v, 0);
  int bi = dpp::lane_value(idx, 0);
  argmax_step(bv, bi, dpp::lane_value(v, 16), dpp::lane_value(idx, 16));
  argmax_step(bv, bi, dpp::lane_value(v, 32), dpp::lane_value(idx, 32));
  argmax_step(bv, bi, dpp::lane_value(v, 48), dpp::lane_value(idx, 48));
  v = bv;
  idx = bi;
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

// A pivot the elimination can divide by: nonzero and finite (false for NaN).
// The pivot searches below start from the first candidate row, so a column
// without any finite candidate keeps the diagonal as its pivot: the row
// index stays inside the matrix, and the zero, infinite or NaN pivot is
// reported through the return value instead.
template <typename T>
__host__ __device__ __forceinline__ bool pivot_usable(T d) {
  const double a = fabs((double)d);
  return a > 0.0 && a <= 1.7976931348623157e308;
}


// Gauss-Jordan inverse with two barriers per column.  The row interchange
// of step c is folded into the elimination (the new pivot row is read from
// row p, the old row c is captured in orow and becomes row p), and the wave
// that eliminates column c + 1 also finds the next pivot, so a step is:
// capture [barrier] eliminate + next pivot [barrier].  Pivots follow
// dgetrf: first largest |A(i, c)|, i >= c, of the current matrix; piv[c]
// records them.  The column interchanges that finish the inverse are not
// applied: on return inv(A)(i, j) = A[i + sigma[j] * n].  prow, orow, fcol
// are n-element LDS scratch vectors, piv and sigma n-element LDS int arrays.
// Returns false on a zero or non-finite pivot.
template <typename T>
__device__ bool block_gj_inverse_perm(T* __restrict__ A, int n, int* __restrict__ piv,
                                      int* __restrict__ sigma, T* __restrict__ prow,
                                      T* __restrict__ orow, T* __restrict__ fcol, int* s_p) {
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6;
  const int nw = nt >> 6;
  bool ok = true;
  if (wv == 0) {  // pivot of column 0
    double best = -1.0;
    int bi = 0;  // stays 0 when no candidate is finite (NaN compares false)
    for (int i = lane; i < n; i += 64) {
      double v = fabs((double)A[i]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    wave_argmax(best, bi);
    if (lane == 0) *s_p = bi;
  }
  __syncthreads();
  for (int c = 0; c < n; ++c) {
    const int p = *s_p;
    const T d = A[p + (size_t)c * n];
    if (!pivot_usable(d)) ok = false;
    const T inv = T(1) / d;
    for (int j = tid; j < n; j += nt) {
      prow[j] = j == c ? inv : A[p + (size_t)j * n] * inv;
      orow[j] = A[c + (size_t)j * n];
    }
    for (int i = tid; i < n; i += nt) fcol[i] = A[i + (size_t)c * n];
    if (tid == 0) piv[c] = p;
    __syncthreads();
    const T fp = orow[c];  // A(c, c) before the interchange: row p's multiplier
    // Eliminate.  Wave w takes columns j(jj) = (c + 1 + jj) mod n for
    // jj = w, w + nw, ...; its lanes take rows.  Loads of a batch of four
    // columns are issued before any store so that they overlap.  The wave
    // with jj = 0 (column c + 1) also finds the next pivot.
    double best = -1.0;
    int bi = c + 1;  // the first candidate of column c + 1; kept when none is finite
    for (int i = lane; i < n; i += 64) {
      const bool rp = i == p, rc = i == c;
      const T f = rp ? fp : fcol[i];
      for (int jb = wv; jb < n; jb += 4 * nw) {
        T src[4], pj[4];
        int jx[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int jj = jb + u * nw;
          int j = c + 1 + jj;
          if (j >= n) j -= n;
          jx[u] = jj < n ? j : -1;
          if (jj < n) {
            pj[u] = prow[j];
            src[u] = rp ? orow[j] : A[i + (size_t)j * n];
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int j = jx[u];
          if (j < 0) continue;
          const T v = rc ? pj[u] : (j == c ? -f * inv : src[u] - f * pj[u]);
          A[i + (size_t)j * n] = v;
          if (jb + u * nw == 0 && i > c) {  // column c + 1
            const double a = fabs((double)v);
            if (a > best) {
              best = a;
              bi = i;
            }
          }
        }
      }
    }
    if (wv == 0 && c + 1 < n) {
      wave_argmax(best, bi);
      if (lane == 0) *s_p = bi;
    }
    __syncthreads();
  }
  if (tid == 0) {  // the column interchanges, in reverse order, as a permutation
    for (int j = 0; j < n; ++j) sigma[j] = j;
    for (int c = n - 1; c >= 0; --c) {
      const int q = piv[c];
      const int t0 = sigma[c];
      sigma[c] = sigma[q];
      sigma[q] = t0;
    }
  }
  __syncthreads();
  return ok;
}

// argmax over (value, index) of a wavefront in the value's own precision:
// larger value wins, ties go to the smaller index; every lane gets the result
template <typename V>
__device__ __forceinline__ void wave_argmax_t(V& v, int& idx) {
  auto step = [&](V v2, int i2) {
    if (v2 > v || (v2 == v && i2 < idx)) {
      v = v2;
      idx = i2;
    }
  };
  step(dpp::mov<dpp::kXor1>(v), dpp::mov<dpp::kXor1>(idx));
  step(dpp::mov<dpp::kXor2>(v), dpp::mov<dpp::kXor2>(idx));
  step(dpp::mov<dpp::kRor4>(v), dpp::mov<dpp::kRor4>(idx));
  step(dpp::mov<dpp::kRor8>(v), dpp::mov<dpp::kRor8>(idx));
  V bv = dpp::lane_value(v, 0);
  int bi = dpp::lane_value(idx, 0);
#pragma unroll
  for (int l = 16; l < 64; l += 16) {
    const V v2 = dpp::lane_value(v, l);
    const int i2 = dpp::lane_value(idx, l);
    if (v2 > bv || (v2 == bv && i2 < bi)) {
      bv = v2;
      bi = i2;
    }
  }
  v = bv;
  idx = bi;
}

// 1/d by the hardware reciprocal and Newton steps (one for FP32, two for
// FP64: within an ulp or two of the correctly rounded quotient, measured
// for FP64 on gfx950), for the pivots of the register Gauss-Jordan.
__device__ __forceinline__ float fast_recip(float d) {
  const float r = __builtin_amdgcn_rcpf(d);
  return fmaf(r, fmaf(-d, r, 1.0f), r);
}
__device__ __forceinline__ double fast_recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  return fma(r, fma(-d, r, 1.0), r);
}

// X^{-1} of the r x r matrix X (LDS, column-major, ld r), r <= R <= 64, by
// Gauss-Jordan exchange steps with partial pivoting.  Lane i holds row i in
// registers (rows r .. R-1 are an identity border that never pivots while
// a real row is left); the current column is always register 0: each step
// writes the updated column j into register j - 1 and the pivot column's new
// values into register R - 1, so after r steps register j holds column
// (j + r) mod R.  Rows are not interchanged: step c picks the largest
// |X(i, c)| over the rows not yet used (near-ties to 2^-17 go to the first
// physical row, so the order can differ from dgetrf's interchanged one),
// and X^{-1}(c', p(c)) ends up in the row of lane p(c') (p = the step's
// pivot row).  Z (r x r, ld r) receives X^{-1}; piv_s (LDS, >= r ints)
// the pivot row of each step; prow (LDS, R values) carries the pivot row
// from its lane to the others.  Returns false on a zero or non-finite
// pivot.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return fma(a, b, c); }

template <typename T, int R>
__device__ bool wave_gj_reg(const T* X, int r, T* Z, int* piv_s, T* prow) {
  const int lane = threadIdx.x & 63;
  T a[R];
#pragma unroll
  for (int j = 0; j < R; ++j)
    a[j] = (lane < r && j < r) ? X[lane + (size_t)j * r] : (lane == j ? T(1) : T(0));
  bool used = lane >= R;  // lanes past the border hold no row
  bool ok = true;
  int pmine = -1;  // the step at which this lane's row pivoted
  for (int c = 0; c < r; ++c) {
    // pivot search on 32-bit keys: the FP32 magnitude's bits with the low 6
    // replaced by 63 - lane, so one unsigned wave maximum finds the largest
    // candidate and, among candidates within 2^-17 of each other, the first
    // row (partial pivoting needs no more than that); used rows key 0, a
    // NaN candidate keys below every number
    unsigned key = 0;
    if (!used) {
      const float m = fabsf((float)a[0]);
      key = (m == m ? (__float_as_uint(m) & ~63u) : 0u) | (unsigned)(63 - lane);
    }
    key = max(key, (unsigned)dpp::mov<dpp::kXor1>((int)key));
    key = max(key, (unsigned)dpp::mov<dpp::kXor2>((int)key));
    key = max(key, (unsigned)dpp::mov<dpp::kRor4>((int)key));
    key = max(key, (unsigned)dpp::mov<dpp::kRor8>((int)key));
    const unsigned k0 = (unsigned)__builtin_amdgcn_readlane((int)key, 0);
    const unsigned k1 = (unsigned)__builtin_amdgcn_readlane((int)key, 16);
    const unsigned k2 = (unsigned)__builtin_amdgcn_readlane((int)key, 32);
    const unsigned k3 = (unsigned)__builtin_amdgcn_readlane((int)key, 48);
    const int p = 63 - (int)(max(max(k0, k1), max(k2, k3)) & 63u);
    if (lane == p) {
#pragma unroll
      for (int j = 0; j < R; ++j) prow[j] = a[j];
    }
    wave_sync();
    const T pv = prow[0];
    if (!pivot_usable(pv)) ok = false;
    const T inv = fast_recip(pv);
    const T f = a[0];
    const T tc = lane == p ? inv - T(1) : -f * inv;
    // the pivot row read 16 values at a time (the scheduling barriers keep
    // all R of them from being held in registers at once)
#pragma unroll
    for (int j0 = 1; j0 < R; j0 += 16) {
#pragma unroll
      for (int j = j0; j < j0 + 16 && j < R; ++j) a[j - 1] = fma_t(tc, prow[j], a[j]);
      __builtin_amdgcn_sched_barrier(0);
    }
    a[R - 1] = lane == p ? inv : -f * inv;
    if (lane == p) {
      used = true;
      pmine = c;
    }
    if (lane == 0) piv_s[c] = p;
    wave_sync();  // every lane has read prow before the next step rewrites it
  }
  // (a border row pivots only when a column has no usable entry; the
  // result is then flagged, and nothing is written outside Z)
  if (pmine >= 0 && lane < r) {
#pragma unroll
    for (int j = 0; j < R; ++j) {
      int col = j + r;
      if (col >= R) col -= R;
      if (col < r) {
        const int q = piv_s[col];
        if (q < r) Z[pmine + (size_t)q * r] = a[j];
      }
    }
  }
  wave_sync();
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
