// Device code of the blocked Gauss-Jordan inverse (dense_inv.hip): one
// 256-thread workgroup per matrix.  Kept in a header so that a CPU test can
// run the same routines on emulated GPU threads.
#pragma once

#include <climits>

#include "device_util.h"

namespace kigpu {

namespace gj {

constexpr int kGjTB = 256;
constexpr int kGjWave = 64;  // gfx950 wavefront
constexpr int kGjNW = kGjTB / kGjWave;
constexpr int kPB = 16;  // panel width (= the MFMA tile width)

typedef double gj_d4 __attribute__((ext_vector_type(4)));

// Pivot key of a row: 0 for a row that cannot pivot, else the bits of |v|
// (a NaN as 0) with the low 11 replaced by 2047 - row.  The largest key is
// the largest |v|, and among values equal to 2^-41 relative the first row:
// dgetrf's choice except for such near-ties.  Rows are below 1024, so every
// candidate's key is nonzero.
__device__ __forceinline__ uint64_t gj_key(double v, bool can, int row) {
  uint64_t b = (uint64_t)__double_as_longlong(v) & 0x7FFFFFFFFFFFFFFFull;
  b = b > 0x7FF0000000000000ull ? 0ull : b;
  return ((b & ~0x7FFull) | (uint64_t)(2047 - row)) & (0ull - (uint64_t)can);
}
__device__ __forceinline__ int gj_key_row(uint64_t key) { return 2047 - (int)(key & 0x7FFull); }

template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  const uint32_t lo = (uint32_t)dpp::mov<CTRL>((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)dpp::mov<CTRL>((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t max_u64(uint64_t a, uint64_t b) { return a > b ? a : b; }
// largest key of the wavefront, in every lane (all 64 lanes active)
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t k) {
  k = max_u64(k, dpp_u64<dpp::kXor1>(k));
  k = max_u64(k, dpp_u64<dpp::kXor2>(k));
  k = max_u64(k, dpp_u64<dpp::kRor4>(k));
  k = max_u64(k, dpp_u64<dpp::kRor8>(k));
  return max_u64(max_u64(lane_u64(k, 0), lane_u64(k, 16)),
                 max_u64(lane_u64(k, 32), lane_u64(k, 48)));
}

// one step's candidates: each wave's best key and that row's panel entries
// (register layout of the step)
struct GjCand {
  uint64_t key[kGjNW];
  double row[kGjNW][kPB];
};

struct GjMat {
  int n;
  double* B;     // working matrix, ld n
  double* R;     // saved pivot rows of a panel: R(t, j) at R[j + n * t]
  int* pr;       // pivot row of each step
  int* pstep;    // step at which a row pivoted, -1 before
};

// Exchange steps on panel columns [k, k + bw) of B; writes the panel back,
// records the steps in pr / pstep and saves the pivot rows of the other
// columns in R.  All threads of the workgroup call it.  ONE: wave 0 alone
// holds the panel, rows lane + 64 u (n <= 64 U), and a step needs no
// workgroup barrier (the pivot row passes through LDS within the wave);
// else the four waves hold rows tid + 256 u (n <= 256 U) and meet once per
// step.  Returns false (in thread 0) if a pivot was zero or not finite.
template <int U, bool ONE>
__device__ bool gj_panel(const GjMat& M, int k, int bw, GjCand* cand) {
  const int tid = threadIdx.x, lane = tid & (kGjWave - 1), wv = tid >> 6;
  constexpr int kStride = ONE ? kGjWave : kGjTB;
  const int base = ONE ? lane : tid;
  const int n = M.n;
  bool ok = true;
  if (!ONE || wv == 0) {
    double a[U][kPB];
    bool can[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int row = u * kStride + base;
      const bool valid = row < n;
      can[u] = valid && M.pstep[row] < 0;
#pragma unroll
      for (int j = 0; j < kPB; ++j)
        a[u][j] = (valid && j < bw) ? M.B[row + (size_t)(k + j) * n] : 0.0;
    }
    int myp = 0;
    for (int t = 0; t < bw; ++t) {
      uint64_t key = 0;
#pragma unroll
      for (int u = 0; u < U; ++u) key = max_u64(key, gj_key(a[u][0], can[u], u * kStride + base));
      key = wave_max_u64(key);
      double sp[kPB];
      int p;
      if (ONE) {
        p = gj_key_row(key);  // some row is always left unpivoted: p < n
        if (lane == (p & (kGjWave - 1))) {
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (u == p / kGjWave) {
#pragma unroll
              for (int j = 0; j < kPB; ++j) cand[0].row[0][j] = a[u][j];
            }
        }
        wave_sync();
#pragma unroll
        for (int j = 0; j < kPB; ++j) sp[j] = cand[0].row[0][j];
        wave_sync();  // read before the next step's owner writes
      } else {
        GjCand& cd = cand[t & 1];
        const int pw = gj_key_row(key);  // this wave's best row (if key != 0)
        if (key != 0 && tid == (pw & (kGjTB - 1))) {
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (u == pw / kGjTB) {
#pragma unroll
              for (int j = 0; j < kPB; ++j) cd.row[wv][j] = a[u][j];
            }
        }
        if (lane == 0) cd.key[wv] = key;
        __syncthreads();
        uint64_t kb = cd.key[0];
        int wb = 0;
#pragma unroll
        for (int w = 1; w < kGjNW; ++w)
          if (cd.key[w] > kb) {
            kb = cd.key[w];
            wb = w;
          }
        p = gj_key_row(kb);
#pragma unroll
        for (int j = 0; j < kPB; ++j) sp[j] = cd.row[wb][j];
      }
      const double d = sp[0];
      if (!pivot_usable(d)) ok = false;
      const double inv = 1.0 / d;
#pragma unroll
      for (int j = 1; j < kPB; ++j) sp[j] *= inv;
      // the step: register j - 1 takes column j, register 15 the pivot
      // column; every row as a non-pivot row, then the pivot row's lane
      // overwrites its slot (p is the same in every lane)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const double f = a[u][0];
#pragma unroll
        for (int j = 1; j < kPB; ++j) a[u][j - 1] = fma(-f, sp[j], a[u][j]);
        a[u][kPB - 1] = -f * inv;
        can[u] = can[u] && u * kStride + base != p;
      }
      if (base == p % kStride) {
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (u == p / kStride) {
#pragma unroll
            for (int j = 1; j < kPB; ++j) a[u][j - 1] = sp[j];
            a[u][kPB - 1] = inv;
          }
      }
      if (lane == t) myp = p;  // lane t of each wave keeps step t's pivot row
    }
    if (wv == 0 && lane < bw) {
      M.pr[k + lane] = myp;
      M.pstep[myp] = k + lane;
    }
    // register j now holds panel column (j + bw) mod 16
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int row = u * kStride + base;
      if (row < n) {
#pragma unroll
        for (int j = 0; j < kPB; ++j) {
          int col = j + bw;
          if (col >= kPB) col -= kPB;
          if (col < bw) M.B[row + (size_t)(k + col) * n] = a[u][j];
        }
      }
    }
  }
  __syncthreads();  // the panel and pr of the last step
  // R(t, j) = B(pr[k + t], j) for the columns outside the panel; rows
  // t >= bw of R are zero (the last panel), so the tile products need no
  // mask in t
  for (int t = 0; t < kPB; ++t) {
    const int p = t < bw ? M.pr[k + t] : 0;
    for (int j = tid; j < n; j += kGjTB)
      if (j < k || j >= k + bw) M.R[j + (size_t)n * t] = t < bw ? M.B[p + (size_t)j * n] : 0.0;
  }
  __syncthreads();
  return ok;
}

// B(i, J) = [i not in P] B(i, J) + W(i, :) R(:, J) for every 16 x 16 tile
// outside the panel's column block; tiles q0, q0 + qs, ... of the
// ceil(n/16)^2 grid (row blocks fastest), each by one wave.  The product is
// formed transposed, D(j, i) = sum_t R(t, j) W(i, t), so that the 16 lanes
// of a tile row cover 16 consecutive rows i of B.  Indices past n are
// clamped for the loads (their results are not stored), and R's rows past
// bw are zero, so every operand load is unconditional and the eight of a
// tile are in flight together.
__device__ void gj_update_tiles(const GjMat& M, int k, int bw, int q0, int qs) {
  const int lane = threadIdx.x & (kGjWave - 1);
  const int l16 = lane & 15, l4 = lane >> 4;
  const int n = M.n, nb = (n + 15) >> 4, kb = k >> 4;
  int ibk = q0, jbk = 0;
  while (ibk >= nb) {
    ibk -= nb;
    ++jbk;
  }
  for (; jbk < nb;) {
    if (jbk != kb) {
      const int ia = ibk * 16 + l16;
      const int iac = min(ia, n - 1);
      const int j0 = jbk * 16, jac = min(j0 + l16, n - 1);
      double av[4], bv[4], cv[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int tt = 4 * s + l4;
        av[s] = M.R[jac + (size_t)n * tt];                       // R(t, j0 + l16)
        bv[s] = M.B[iac + (size_t)(k + min(tt, bw - 1)) * n];    // W(i0 + l16, t)
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) cv[v] = M.B[iac + (size_t)min(j0 + l4 + 4 * v, n - 1) * n];
      const int ps = M.pstep[iac];
      gj_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], bv[s], acc, 0, 0, 0);
      const bool inP = ps >= k && ps < k + bw;
      if (ia < n) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int j = j0 + l4 + 4 * v;
          if (j < n) M.B[ia + (size_t)j * n] = (inP ? 0.0 : cv[v]) + acc[v];
        }
      }
    }
    ibk += qs;
    while (ibk >= nb) {
      ibk -= nb;
      ++jbk;
    }
  }
}

// dynamic LDS: pr, pstep (n ints each), then B and R when in LDS
__host__ __device__ inline size_t gj_lds_ints(int n) {
  return (2 * (size_t)n * sizeof(int) + 15) & ~(size_t)15;
}
__host__ __device__ inline size_t gj_lds_bytes(int n, bool in_lds) {
  return gj_lds_ints(n) + (in_lds ? sizeof(double) * ((size_t)n * n + (size_t)kPB * n) : 0);
}

// The whole inverse of matrix x by the calling workgroup; smem: the
// dynamic LDS (gj_lds_bytes), cand: two GjCand in LDS.
template <int U, bool ONE, bool LDS>
__device__ void dense_gj_block(int x, unsigned char* smem, GjCand* cand, const int* n_arr,
                               const int64_t* doff, const int64_t* moff, int n0,
                               const double* A_all, double* Ainv_all, double* ws, double* rws,
                               int* piv_all, int* status, int bit) {
  const int tid = threadIdx.x, wv = tid >> 6;
  const int n = n_arr ? n_arr[x] : n0;
  if (n == 0) return;
  const size_t mo = n_arr ? (size_t)moff[x] : 0, dof = n_arr ? (size_t)doff[x] : 0;
  const double* A = A_all + mo;
  double* Ainv = Ainv_all + mo;
  GjMat M;
  M.n = n;
  M.pr = reinterpret_cast<int*>(smem);
  M.pstep = M.pr + n;
  if (LDS) {
    M.B = reinterpret_cast<double*>(smem + gj_lds_ints(n));
    M.R = M.B + (size_t)n * n;
  } else {
    M.B = ws + mo;
    M.R = rws + (size_t)kPB * dof;
  }
  {  // copy in, four loads in flight per thread
    const int nn = n * n;
    int e = tid;
    for (; e + 3 * kGjTB < nn; e += 4 * kGjTB) {
      const double v0 = A[e], v1 = A[e + kGjTB], v2 = A[e + 2 * kGjTB], v3 = A[e + 3 * kGjTB];
      M.B[e] = v0;
      M.B[e + kGjTB] = v1;
      M.B[e + 2 * kGjTB] = v2;
      M.B[e + 3 * kGjTB] = v3;
    }
    for (; e < nn; e += kGjTB) M.B[e] = A[e];
  }
  for (int i = tid; i < n; i += kGjTB) M.pstep[i] = -1;
  __syncthreads();
  bool ok = true;
  for (int k = 0; k < n; k += kPB) {
    const int bw = min(kPB, n - k);
    ok = gj_panel<U, ONE>(M, k, bw, cand) && ok;
    gj_update_tiles(M, k, bw, wv, kGjNW);
    __syncthreads();
  }
  if (!ok && tid == 0) atomicOr(status, bit);
  // inv(A)(i, j) = B(pr[i], pstep[j])
  for (int j = wv; j < n; j += kGjNW) {
    const double* col = M.B + (size_t)M.pstep[j] * n;
    for (int i = tid & (kGjWave - 1); i < n; i += kGjWave) Ainv[i + (size_t)j * n] = col[M.pr[i]];
  }
  if (piv_all)
    for (int i = tid; i < n; i += kGjTB) piv_all[dof + i] = M.pr[i];
}

}  // namespace gj
}  // namespace kigpu
