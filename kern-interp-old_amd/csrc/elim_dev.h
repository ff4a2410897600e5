// Device routines of the elimination (decouple, skel_factorization.cpp:55-91)
// and of the ID matrix generation shared by factor.hip's kernels: staged
// geometry, kernel entries from it, the matrix-core tile products of one
// wavefront, the children's update layout, and the elimination of one box
// by one wavefront (the body of k_decouple_wave).  Needs only the HIP device
// built-ins, so that a CPU test can run them on emulated threads.
#pragma once

#include "device_util.h"
#include "entries.h"
#include "geom.h"

namespace kigpu {
namespace elim {

__device__ __forceinline__ void flag(int* status, int bit) { atomicOr(status, bit); }

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

// Geometry of a DOF staged in LDS: point x, y, normal x, y, weight, curvature.
constexpr int kGeo = 6;

__device__ __forceinline__ void stage_geo(const GeomView& g, int dof, double* dst) {
  const int p = dof / g.sd;
  dst[0] = g.px[p];
  dst[1] = g.py[p];
  dst[2] = g.nx[p];
  dst[3] = g.ny[p];
  dst[4] = g.w[p];
  dst[5] = g.kappa[p];
}

// K(t, s) from staged geometry (same operations as kernel_entry)
__device__ __forceinline__ double staged_entry(int pde, int tdof, const double* tg, int sdof,
                                               const double* sg) {
#pragma clang fp contract(off)
  if (pde == KIGPU_STOKES)
    return stokes_entry(tdof, sdof, tg[0], tg[1], tg[2], tg[3], sg[0], sg[1], sg[2], sg[3], sg[4],
                        sg[5]);
  return laplace_entry(pde, tg[0] - sg[0], tg[1] - sg[1], sg[4], sg[5], sg[2], sg[3], tg[2],
                       tg[3], false);
}

constexpr int kWdecMaxN = 64;

typedef float wd_f4 __attribute__((ext_vector_type(4)));
typedef double wd_d4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  typedef wd_f4 acc;
  __device__ static acc step(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // row of accumulator element v of lane l (column l % 16)
  __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <>
struct Mfma<double> {
  typedef wd_d4 acc;
  __device__ static acc step(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

// out(i, j, sum_t A(i, t) B(t, j)) for every (i, j) of an M x N product by
// one wavefront, in 16 x 16 tiles with 4-deep MFMA steps.  The A operand of
// lane l is A(i0 + l % 16, t0 + l / 16), the B operand B(t0 + l / 16,
// j0 + l % 16); entries past M, N or K enter as zeros (read at a clamped
// index and replaced, so the loads need no branch).  The next step's
// operands are loaded before the current MFMA is issued.
template <typename T, typename FA, typename FB, typename FO>
__device__ __forceinline__ void wave_gemm(int M, int N, int K, FA&& A, FB&& B, FO&& out) {
  const int lane = threadIdx.x & (kWave - 1);
  const int l16 = lane & 15, l4 = lane >> 4;
  for (int i0 = 0; i0 < M; i0 += 16)
    for (int j0 = 0; j0 < N; j0 += 16) {
      const int ia = i0 + l16, jb = j0 + l16;
      const int iac = min(ia, M - 1), jbc = min(jb, N - 1);
      auto ld = [&](int t0, T& a, T& b) {
        const int t = t0 + l4, tc = min(t, K - 1);
        const T av = A(iac, tc), bv = B(tc, jbc);
        a = (ia < M && t < K) ? av : T(0);
        b = (jb < N && t < K) ? bv : T(0);
      };
      typename Mfma<T>::acc c = {T(0), T(0), T(0), T(0)};
      T a, b;
      ld(0, a, b);
      for (int t0 = 0; t0 < K; t0 += 4) {
        T an, bn;
        ld(min(t0 + 4, K - 1), an, bn);
        c = Mfma<T>::step(a, b, c);
        a = an;
        b = bn;
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + Mfma<T>::row(lane, v);
        if (i < M && jb < N) out(i, jb, c[v]);
      }
    }
}


// LDS of one k_decouple_wave box with |active_box| = n and rank k:
// staged geometry, the [skel | redundant] DOFs, their active positions and
// the inverse map, pivots, the Gauss-Jordan's pivot row, then
//   Kp (n x n): K_BN; after X_rr is formed it holds X_rr^{-1}, and at the
//               end the Schur update;
//   X (r x r):  X_rr, kept for the refinement step of L;
//   Tm (k x r): T; once X_rr is formed it holds L;
//   K_r_sn (r x k), K_sn_r (k x r).
struct WdecLds {
  size_t geo, sr, pos, ipos, piv, gjv, kp, x, tm, krsn, ksnr, total;
};
__host__ __device__ inline WdecLds wdec_layout(int n, int k, size_t tsz) {
  WdecLds L;
  const int r = n - k;
  size_t o = 0;
  L.geo = o;  o = align16(o + sizeof(double) * kGeo * (size_t)n);
  L.sr = o;   o = align16(o + sizeof(int) * (size_t)n);
  L.pos = o;  o = align16(o + sizeof(int) * (size_t)n);
  L.ipos = o; o = align16(o + sizeof(int) * (size_t)n);
  L.piv = o;  o = align16(o + sizeof(int) * (size_t)kWdecMaxN);
  L.gjv = o;  o = align16(o + tsz * (size_t)kWdecMaxN);
  L.kp = o;   o = align16(o + tsz * (size_t)n * n);
  L.x = o;    o = align16(o + tsz * (size_t)r * r);
  L.tm = o;   o = align16(o + tsz * (size_t)k * r);
  L.krsn = o; o = align16(o + tsz * (size_t)k * r);
  L.ksnr = o; o = align16(o + tsz * (size_t)k * r);
  L.total = o;
  return L;
}
// the most any box of size <= n needs (over every rank k)
__host__ inline size_t wdec_lds_bound(int n, size_t tsz) {
  size_t best = wdec_layout(n, n, tsz).total;
  for (int m = 1; m <= n; ++m)
    for (int k = 1; k < m; ++k) best = std::max(best, wdec_layout(m, k, tsz).total);
  return best;
}

// The children's accumulated updates of a box, flattened: child c's
// osize x osize block holds entries [eo[c], eo[c + 1]) in column-major
// order, entry (a, b) at active positions (co[c] + a, co[c] + b).
struct ChildUpd {
  int co[5], eo[5], cld[4];
  uint64_t cacc[4];
};

// element c (0..3) of a 4-entry array held in registers (a runtime index
// into the array itself would put it in scratch memory)
template <typename V>
__device__ __forceinline__ V pick4(const V* a, int c) {
  return c == 0 ? a[0] : c == 1 ? a[1] : c == 2 ? a[2] : a[3];
}

// f(a', b', v) for the entries e = e0 + lane + 64 u (u < 8) of the
// flattened update; the 8 loads of a lane are issued before any f runs.
template <typename T, typename F>
__device__ __forceinline__ void child_update_batch(const ChildUpd& cu, int e0, F&& f) {
  const int lane = threadIdx.x & (kWave - 1);
  T v[8];
  int pa[8], pb[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int e = e0 + u * kWave + lane;
    pa[u] = -1;
    pb[u] = 0;
    v[u] = T(0);
    if (e < cu.eo[4]) {
      const int c = (e >= cu.eo[1]) + (e >= cu.eo[2]) + (e >= cu.eo[3]);
      const int c0 = pick4(cu.co, c), e0c = pick4(cu.eo, c);
      const int sz = pick4(cu.co + 1, c) - c0;
      const int loc = e - e0c;
      int b = (int)(((float)loc + 0.5f) / (float)sz);  // exact: loc < 2^12
      int a = loc - b * sz;
      if (a < 0) { --b; a += sz; }
      if (a >= sz) { ++b; a -= sz; }
      const T* acc = reinterpret_cast<const T*>(pick4(cu.cacc, c));
      if (acc) {
        v[u] = acc[a + (size_t)b * pick4(cu.cld, c)];
        pa[u] = c0 + a;
        pb[u] = c0 + b;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < 8; ++u)
    if (pa[u] >= 0) f(pa[u], pb[u], v[u]);
}

// ---------------------------------------------------------------------------
// elimination of the redundant DOFs (decouple, skel_factorization.cpp:55-91)
// ---------------------------------------------------------------------------

struct DecLds {
  size_t uact, kp, x, krsn, ksnr, tm, lm, sr, pos, piv, sig, total;
};

__host__ __device__ inline DecLds dec_layout(int n, size_t tsz) {
  DecLds L;
  size_t o = 0;
  const size_t nn = (size_t)n * n, q4 = nn / 4 + 1;
  L.uact = o; o = align16(o + nn * tsz);
  L.kp = o;   o = align16(o + nn * tsz);
  L.x = o;    o = align16(o + nn * tsz);
  L.krsn = o; o = align16(o + q4 * tsz);
  L.ksnr = o; o = align16(o + q4 * tsz);
  L.tm = o;   o = align16(o + q4 * tsz);
  L.lm = o;   o = align16(o + q4 * tsz);
  L.sr = o;   o = align16(o + sizeof(int) * (size_t)n);
  L.pos = o;  o = align16(o + sizeof(int) * (size_t)n);
  L.piv = o;  o = align16(o + sizeof(int) * (size_t)n);
  L.sig = o;  o = align16(o + sizeof(int) * (size_t)n);
  L.total = o;
  return L;
}

// Factor slab of a compressed node (k skeleton, r redundant DOFs):
//   T (k x r) | L (k x r) | U (r x k) | X_rr^{-1} (r x r) | Acc (k x k)
// all column-major with leading dimension = height.
__host__ __device__ inline size_t slab_elems(int n) { return ((size_t)n * n * 5 + 3) / 4 + 4; }

constexpr int kDecTB = 1024;
constexpr int kDecNW = kDecTB / kWave;

// The elimination of box x by the calling 1024-thread workgroup: the steps of
// decouple with the products as dot products over the box's working arrays
// and X_rr^{-1} by a block Gauss-Jordan in dgetrf's pivot order.  base: the
// working arrays (dec_layout of the box; the launch's dynamic LDS, or the
// block's slice of a global workspace); s_p: one int of LDS.
template <typename T>
__device__ __forceinline__ void decouple_block_box(
    int x, unsigned char* base, int* s_p, GeomView g, int pde, const int* is_leaf,
    const int* child, const int* osize_in, const uint64_t* acc_ptr_in, const int* acc_ld_in,
    const int* nact, const int* rank, const int64_t* act_off, const int* sr_pool,
    const int* pos_pool, int* piv_pool, const uint64_t* fac_ptr, int* osize, uint64_t* acc_ptr,
    int* acc_ld, int* status) {
  const int n = nact[x];
  const int k = rank[x];
  const int tid = threadIdx.x;
  if (n == 0) {
    if (tid == 0) {
      osize[x] = 0;
      acc_ptr[x] = 0;
      acc_ld[x] = 0;
    }
    return;
  }
  const DecLds Ly = dec_layout(n, sizeof(T));
  T* Uact = reinterpret_cast<T*>(base + Ly.uact);
  T* slab = reinterpret_cast<T*>(fac_ptr[x]);
  const bool leaf = is_leaf[x];

  // accumulated updates of the children, block diagonal in active order
  if (!leaf) {
    for (int i = tid; i < n * n; i += kDecTB) Uact[i] = T(0);
    __syncthreads();
    int o = 0;
    for (int s = 0; s < 4; ++s) {
      int y = child[4 * x + s];
      int sz = osize_in[y];
      const T* acc = reinterpret_cast<const T*>(acc_ptr_in[y]);
      if (acc) {
        int ld = acc_ld_in[y];
        for2d<kDecNW>(sz, sz, [&](int i, int j) {
          Uact[(o + i) + (size_t)(o + j) * n] = acc[i + (size_t)j * ld];
        });
      }
      o += sz;
    }
    __syncthreads();
  }
  if (k < 0) {  // not compressed: hand the whole active list and its update up
    if (!leaf) {
      for (int i = tid; i < n * n; i += kDecTB) slab[i] = Uact[i];
    }
    if (tid == 0) {
      osize[x] = n;
      acc_ptr[x] = leaf ? 0 : reinterpret_cast<uint64_t>(slab);
      acc_ld[x] = n;
    }
    return;
  }
  const int r = n - k;
  T* Kp = reinterpret_cast<T*>(base + Ly.kp);
  T* X = reinterpret_cast<T*>(base + Ly.x);
  T* Krsn = reinterpret_cast<T*>(base + Ly.krsn);
  T* Ksnr = reinterpret_cast<T*>(base + Ly.ksnr);
  T* Ts = reinterpret_cast<T*>(base + Ly.tm);
  T* Ls = reinterpret_cast<T*>(base + Ly.lm);
  int* sr = reinterpret_cast<int*>(base + Ly.sr);
  int* pos = reinterpret_cast<int*>(base + Ly.pos);
  int* piv = reinterpret_cast<int*>(base + Ly.piv);
  int* sig = reinterpret_cast<int*>(base + Ly.sig);
  const int64_t aoff = act_off[x];
  for (int j = tid; j < n; j += kDecTB) {
    sr[j] = sr_pool[aoff + j];
    pos[j] = pos_pool[aoff + j];
  }
  for (int e = tid; e < k * r; e += kDecTB) Ts[e] = slab[e];  // T from the ID
  __syncthreads();
  // K_BN = K(BN, BN) - update, permuted to [s r] (skel_factorization.cpp:66-78)
  for2d<kDecNW>(n, n, [&](int i, int j) {
    double v = kernel_entry(g, pde, sr[i], sr[j]);
    if (!leaf) v -= (double)Uact[pos[i] + (size_t)pos[j] * n];
    Kp[i + (size_t)j * n] = (T)v;
  });
  __syncthreads();
  T* Lm = slab + (size_t)k * r;             // k x r
  T* Um = slab + 2 * (size_t)k * r;         // r x k
  T* Xinv = slab + 3 * (size_t)k * r;       // r x r
  T* Acc = Xinv + (size_t)r * r;            // k x k
  auto KP = [&](int i, int j) -> T { return Kp[i + (size_t)j * n]; };
  // K_r_sn = K_rs - T^T K_ss (r x k); K_sn_r = K_sr - K_ss T (k x r)
  for2d<kDecNW>(r, k, [&](int i, int j) {
    T s = KP(k + i, j);
    for (int t = 0; t < k; ++t) s -= Ts[t + (size_t)i * k] * KP(t, j);
    Krsn[i + (size_t)j * r] = s;
  });
  for2d<kDecNW>(k, r, [&](int i, int j) {
    T s = KP(i, k + j);
    for (int t = 0; t < k; ++t) s -= KP(i, t) * Ts[t + (size_t)j * k];
    Ksnr[i + (size_t)j * k] = s;
  });
  __syncthreads();
  // X_rr = K_rr - T^T K_sr - K_r_sn T (r x r)
  for2d<kDecNW>(r, r, [&](int i, int j) {
    T s = KP(k + i, k + j);
    for (int t = 0; t < k; ++t) s -= Ts[t + (size_t)i * k] * KP(t, k + j);
    for (int t = 0; t < k; ++t) s -= Krsn[i + (size_t)t * r] * Ts[t + (size_t)j * k];
    X[i + (size_t)j * r] = s;
  });
  __syncthreads();
  // K_BN and T are dead from here on: the K_BN area holds a copy of X_rr and
  // the two pivot scratch vectors (r^2 + 2r <= n^2 since k >= 1)
  T* Xc = Kp;
  for (int e = tid; e < r * r; e += kDecTB) Xc[e] = X[e];
  __syncthreads();
  // X_rr^{-1} in place (Gauss-Jordan with dgetrf's pivot order), columns
  // left permuted: X_rr^{-1}(i, j) = Z[i + sig[j] r].  Scratch: two vectors
  // after the copy of X_rr, one after X_rr (r^2 + r <= n^2).
  T* Z = X;
  if (!block_gj_inverse_perm(X, r, piv, sig, Xc + (size_t)r * r, X + (size_t)r * r,
                             Xc + (size_t)r * r + r, s_p) &&
      tid == 0)
    flag(status, 4);
  // L0 = K_sn_r X_rr^{-1} (k x r), U = X_rr^{-1} K_r_sn (r x k).  The right
  // residual of an explicit inverse is small but its left residual is not:
  // multiplied from the right, X_rr^{-1} loses digits when X_rr is
  // ill-conditioned (tight id_tol), so L takes one refinement step against
  // X_rr below, which brings it to the accuracy of the dgetrs solve of
  // skel_factorization.cpp:98-103.
  for2d<kDecNW>(k, r, [&](int i, int j) {
    const T* zj = Z + (size_t)sig[j] * r;
    T s = T(0);
    for (int t = 0; t < r; ++t) s += Ksnr[i + (size_t)t * k] * zj[t];
    Ls[i + (size_t)j * k] = s;
  });
  for2d<kDecNW>(r, k, [&](int i, int j) {
    T s = T(0);
    for (int t = 0; t < r; ++t) s += Z[i + (size_t)sig[t] * r] * Krsn[t + (size_t)j * r];
    Um[i + (size_t)j * r] = s;
  });
  for2d<kDecNW>(r, r, [&](int i, int j) { Xinv[i + (size_t)j * r] = Z[i + (size_t)sig[j] * r]; });
  for (int c = tid; c < r; c += kDecTB) piv_pool[aoff + c] = piv[c];
  __syncthreads();
  // R = K_sn_r - L0 X_rr in place of K_sn_r (each entry reads only itself)
  for2d<kDecNW>(k, r, [&](int i, int j) {
    T s = Ksnr[i + (size_t)j * k];
    for (int t = 0; t < r; ++t) s -= Ls[i + (size_t)t * k] * Xc[t + (size_t)j * r];
    Ksnr[i + (size_t)j * k] = s;
  });
  __syncthreads();
  // L = L0 + R X_rr^{-1}
  for2d<kDecNW>(k, r, [&](int i, int j) {
    const T* zj = Z + (size_t)sig[j] * r;
    T s = Ls[i + (size_t)j * k];
    for (int t = 0; t < r; ++t) s += Ksnr[i + (size_t)t * k] * zj[t];
    Ls[i + (size_t)j * k] = s;
    Lm[i + (size_t)j * k] = s;
  });
  __syncthreads();
  // schur_update = L K_r_sn (k x k), plus the descendants' update on the skeleton
  for2d<kDecNW>(k, k, [&](int i, int j) {
    T s = T(0);
    for (int t = 0; t < r; ++t) s += Ls[i + (size_t)t * k] * Krsn[t + (size_t)j * r];
    if (!leaf) s += Uact[pos[i] + (size_t)pos[j] * n];
    Acc[i + (size_t)j * k] = s;
  });
  if (tid == 0) {
    osize[x] = k;
    acc_ptr[x] = reinterpret_cast<uint64_t>(Acc);
    acc_ld[x] = k;
  }
}

// ---------------------------------------------------------------------------
// elimination with one wavefront per box (|active_box| <= 64): the same
// steps as k_decouple, with the products on the matrix cores
// (v_mfma_f32_16x16x4f32 for the FP32 factor, v_mfma_f64_16x16x4f64 for
// FP64) and X_rr^{-1} by a Gauss-Jordan held in registers, row i in lane i,
// so no step waits on a workgroup barrier.
// ---------------------------------------------------------------------------

// The elimination of box x by the calling wavefront; smem: its dynamic LDS
// (wdec_lds_bound of the launch's largest box).  RMAX: the largest r the
// launch's boxes can have (n <= RMAX + 1); the register Gauss-Jordan is
// instantiated up to that width only.
template <typename T, int RMAX>
__device__ __forceinline__ void decouple_wave_box(
    int x, unsigned char* smem, GeomView g, int pde, const int* is_leaf, const int* child,
    const int* osize_in, const uint64_t* acc_ptr_in, const int* acc_ld_in, const int* nact,
    const int* rank, const int64_t* act_off, const int* sr_pool, const int* pos_pool,
    int* piv_pool, const uint64_t* fac_ptr, int* osize, uint64_t* acc_ptr, int* acc_ld,
    int* status) {
  const int lane = threadIdx.x;  // a 64-thread block: the lane
  // loads that need only x, issued together
  const int n = nact[x];
  const int kr = rank[x];
  const bool leaf = is_leaf[x];
  const int y = lane < 4 ? child[4 * x + lane] : -1;  // -1 for a leaf's slots
  const int64_t aoff = act_off[x];
  T* slab = reinterpret_cast<T*>(fac_ptr[x]);
  if (n == 0) {
    if (lane == 0) {
      osize[x] = 0;
      acc_ptr[x] = 0;
      acc_ld[x] = 0;
    }
    return;
  }
  const int k = kr < 0 ? n : kr, r = n - k;
  // the second round: children (lanes 0..3), the [skel | redundant] lists
  // (one DOF per lane, n <= 64) and T (k x r, 8 values per lane per batch)
  const int csz = (!leaf && y >= 0) ? osize_in[y] : 0;
  const uint64_t cap = (!leaf && y >= 0) ? acc_ptr_in[y] : 0;
  const int cldl = (!leaf && y >= 0) ? acc_ld_in[y] : 0;
  int md = 0, ps = 0;
  if (kr >= 0 && lane < n) {
    md = sr_pool[aoff + lane];
    ps = pos_pool[aoff + lane];
  }
  const int kt = kr >= 0 ? k * r : 0;
  T tv[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int e = u * kWave + lane;
    tv[u] = e < kt ? slab[e] : T(0);
  }
  ChildUpd cu;
  cu.co[0] = cu.eo[0] = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int sz = __shfl(csz, c, kWave);
    cu.co[c + 1] = cu.co[c] + sz;
    cu.eo[c + 1] = cu.eo[c] + sz * sz;
    cu.cacc[c] = (uint64_t)__shfl((long long)cap, c, kWave);
    cu.cld[c] = __shfl(cldl, c, kWave);
  }
  if (kr < 0) {  // not compressed: hand the whole active list and its update up
    if (!leaf) {  // lane a writes row a, each entry once
      const int ca = (lane >= cu.co[1]) + (lane >= cu.co[2]) + (lane >= cu.co[3]);
      const T* acca = reinterpret_cast<const T*>(pick4(cu.cacc, ca));
      const int coa = pick4(cu.co, ca), lda = pick4(cu.cld, ca);
      for (int b0 = 0; b0 < n; b0 += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int b = b0 + u;
          const int cb = (b >= cu.co[1]) + (b >= cu.co[2]) + (b >= cu.co[3]);
          const bool in = lane < n && b < n && cb == ca && b < cu.co[4] && lane < cu.co[4] &&
                          acca;
          v[u] = in ? acca[(lane - coa) + (size_t)(b - coa) * lda] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (lane < n && b0 + u < n) slab[lane + (size_t)(b0 + u) * n] = v[u];
      }
    }
    if (lane == 0) {
      osize[x] = n;
      acc_ptr[x] = leaf ? 0 : reinterpret_cast<uint64_t>(slab);
      acc_ld[x] = n;
    }
    return;
  }
  const WdecLds Ly = wdec_layout(n, k, sizeof(T));
  double* geo = reinterpret_cast<double*>(smem + Ly.geo);
  int* sr = reinterpret_cast<int*>(smem + Ly.sr);
  int* pos = reinterpret_cast<int*>(smem + Ly.pos);
  int* ipos = reinterpret_cast<int*>(smem + Ly.ipos);
  int* piv = reinterpret_cast<int*>(smem + Ly.piv);
  T* gjv = reinterpret_cast<T*>(smem + Ly.gjv);  // pivot row of the Gauss-Jordan
  T* Kp = reinterpret_cast<T*>(smem + Ly.kp);
  T* X = reinterpret_cast<T*>(smem + Ly.x);
  T* Ts = reinterpret_cast<T*>(smem + Ly.tm);
  T* Krsn = reinterpret_cast<T*>(smem + Ly.krsn);
  T* Ksnr = reinterpret_cast<T*>(smem + Ly.ksnr);
  // the third round: this lane's DOF geometry (kept in registers for its row)
  double mg[kGeo] = {0, 0, 0, 0, 0, 0};
  if (lane < n) stage_geo(g, md, mg);
  if (lane < n) {
    sr[lane] = md;
    pos[lane] = ps;
    ipos[ps] = lane;
#pragma unroll
    for (int q = 0; q < kGeo; ++q) geo[kGeo * lane + q] = mg[q];
  }
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int e = u * kWave + lane;
    if (e < kt) Ts[e] = tv[u];
  }
  for (int e0 = 8 * kWave; e0 < kt; e0 += kWave) {  // T past the first batch
    const int e = e0 + lane;
    if (e < kt) Ts[e] = slab[e];
  }
  wave_sync();
  // K_BN = K(BN, BN) - update, permuted to [s r] (skel_factorization.cpp:66-78):
  // 64 / n lanes per row, lane l on row l % n and every (64 / n)-th column
  // from l / n, four columns in flight
  {
    const int npl = kWave / n;
    if (lane < npl * n) {
      const int i = lane % n;
      double tg[kGeo];
#pragma unroll
      for (int q = 0; q < kGeo; ++q) tg[q] = geo[kGeo * i + q];
      const int ti = sr[i];
#pragma unroll 4
      for (int j = lane / n; j < n; j += npl)
        Kp[i + (size_t)j * n] = (T)staged_entry(pde, ti, tg, sr[j], geo + kGeo * j);
    }
  }
  wave_sync();
  if (!leaf) {
    for (int e0 = 0; e0 < cu.eo[4]; e0 += 8 * kWave)
      child_update_batch<T>(cu, e0, [&](int a, int b, T v) {
        if (a >= n || b >= n) return;
        const size_t e = ipos[a] + (size_t)ipos[b] * n;
        Kp[e] = (T)((double)Kp[e] - (double)v);
      });
    wave_sync();
  }
  auto KP = [&](int i, int j) -> T { return Kp[i + (size_t)j * n]; };
  auto TT = [&](int i, int t) -> T { return Ts[t + (size_t)i * k]; };  // T^T (r x k)
  auto TM = [&](int t, int j) -> T { return Ts[t + (size_t)j * k]; };  // T (k x r)
  // K_r_sn = K_rs - T^T K_ss (r x k); K_sn_r = K_sr - K_ss T (k x r)
  wave_gemm<T>(r, k, k, TT, [&](int t, int j) { return KP(t, j); },
               [&](int i, int j, T v) { Krsn[i + (size_t)j * r] = KP(k + i, j) - v; });
  wave_gemm<T>(k, r, k, [&](int i, int t) { return KP(i, t); }, TM,
               [&](int i, int j, T v) { Ksnr[i + (size_t)j * k] = KP(i, k + j) - v; });
  // X_rr = K_rr - T^T K_sr - K_r_sn T (r x r), in two products
  wave_gemm<T>(r, r, k, TT, [&](int t, int j) { return KP(t, k + j); },
               [&](int i, int j, T v) { X[i + (size_t)j * r] = KP(k + i, k + j) - v; });
  wave_sync();
  wave_gemm<T>(r, r, k, [&](int i, int t) { return Krsn[i + (size_t)t * r]; }, TM,
               [&](int i, int j, T v) { X[i + (size_t)j * r] -= v; });
  wave_sync();
  // K_BN and T are dead: X_rr^{-1} goes where K_BN was, L where T was
  T* Z = Kp;
  T* Ls = Ts;
  bool ok;
  if (r <= 16) ok = wave_gj_reg<T, 16>(X, r, Z, piv, gjv);
  else if (RMAX <= 32 || r <= 32) ok = wave_gj_reg<T, 32>(X, r, Z, piv, gjv);
  else ok = wave_gj_reg<T, (RMAX > 32 ? 64 : 32)>(X, r, Z, piv, gjv);
  if (!ok && lane == 0) flag(status, 4);
  T* Lm = slab + (size_t)k * r;        // k x r
  T* Um = slab + 2 * (size_t)k * r;    // r x k
  T* Xinv = slab + 3 * (size_t)k * r;  // r x r
  T* Acc = Xinv + (size_t)r * r;       // k x k
  auto ZZ = [&](int i, int j) -> T { return Z[i + (size_t)j * r]; };
  // L0 = K_sn_r X^{-1} (k x r) into Ls; U = X^{-1} K_r_sn (r x k) into the slab
  wave_gemm<T>(k, r, r, [&](int i, int t) { return Ksnr[i + (size_t)t * k]; }, ZZ,
               [&](int i, int j, T v) { Ls[i + (size_t)j * k] = v; });
  wave_gemm<T>(r, k, r, ZZ, [&](int t, int j) { return Krsn[t + (size_t)j * r]; },
               [&](int i, int j, T v) { Um[i + (size_t)j * r] = v; });
  for (int e = lane; e < r * r; e += kWave) Xinv[e] = Z[e];
  if (lane < r) piv_pool[aoff + lane] = piv[lane];
  wave_sync();
  // one refinement step of L against X_rr (see k_decouple): R = K_sn_r - L0 X_rr
  // in place of K_sn_r, then L = L0 + R X_rr^{-1}
  wave_gemm<T>(k, r, r, [&](int i, int t) { return Ls[i + (size_t)t * k]; },
               [&](int t, int j) { return X[t + (size_t)j * r]; },
               [&](int i, int j, T v) { Ksnr[i + (size_t)j * k] -= v; });
  wave_sync();
  wave_gemm<T>(k, r, r, [&](int i, int t) { return Ksnr[i + (size_t)t * k]; }, ZZ,
               [&](int i, int j, T v) {
                 const T l = Ls[i + (size_t)j * k] + v;
                 Ls[i + (size_t)j * k] = l;
                 Lm[i + (size_t)j * k] = l;
               });
  wave_sync();
  // schur_update = L K_r_sn (k x k) where X_rr^{-1} was, plus the
  // descendants' update on the skeleton, then out to the slab
  T* Sa = Kp;
  wave_gemm<T>(k, k, r, [&](int i, int t) { return Ls[i + (size_t)t * k]; },
               [&](int t, int j) { return Krsn[t + (size_t)j * r]; },
               [&](int i, int j, T v) { Sa[i + (size_t)j * k] = v; });
  wave_sync();
  if (!leaf) {
    for (int e0 = 0; e0 < cu.eo[4]; e0 += 8 * kWave)
      child_update_batch<T>(cu, e0, [&](int a, int b, T v) {
        if (a >= n || b >= n) return;
        const int i = ipos[a], j = ipos[b];
        if (i < k && j < k) Sa[i + (size_t)j * k] += v;
      });
    wave_sync();
  }
  for (int e = lane; e < k * k; e += kWave) Acc[e] = Sa[e];
  if (lane == 0) {
    osize[x] = k;
    acc_ptr[x] = reinterpret_cast<uint64_t>(Acc);
    acc_ld[x] = k;
  }
}

}  // namespace elim
}  // namespace kigpu
