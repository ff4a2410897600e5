// Device routines of the interpolative decomposition (id_compress,
// skel_factorization.cpp:32-52, ki_mat.cpp:505-541, kernel.cpp:153-334)
// behind factor.hip's ID kernels: the ID matrix generated from staged
// geometry, the pivoted Householder QR in LDS, global memory or registers
// (id_box), and the ID from the Gram matrix (id_gram_box).  Needs only the
// HIP device built-ins, so that a CPU test can run them on emulated threads.
#pragma once

#include "elim_dev.h"

namespace kigpu {
namespace elim {

// ---------------------------------------------------------------------------
// ID: generate the matrix, pivoted QR to the rank-revealing step, T solve
// ---------------------------------------------------------------------------

struct IdLds {
  size_t a, rgeo, cgeo, rows, act, vn1, vn2, perm, vb, red, total;
};

constexpr int kIdTB = 1024;
constexpr int kIdNW = kIdTB / kWave;
constexpr int kIdRC = 3;  // columns per wave of the register-resident QR
constexpr int kIdRR = 8;  // rows per lane of the register-resident QR (row lane + 64 u)

// Boxes whose ID matrix fits in the registers of the block: column j in wave
// j % 16, slot j / 16; row r in lane r % 64, slot r / 64.
__host__ __device__ inline bool id_reg_box(int m, int n) {
  return n <= kIdRC * kIdNW && m <= kIdRR * kWave;
}

// The wider boxes, up to 64 columns and 512 rows or 48 columns and 640
// rows: one 512-thread block each (k_id_wide, two waves per SIMD, so twice
// the registers per lane), column j in wave j % 8, slot j / 8, rows as
// above in 8 or 10 slots.
constexpr int kIdWideTB = 512;
constexpr int kIdWideNW = kIdWideTB / kWave;
__host__ __device__ inline bool id_wide_box(int m, int n) {
  return !id_reg_box(m, n) &&
         ((n <= 8 * kIdWideNW && m <= 8 * kWave) || (n <= 6 * kIdWideNW && m <= 10 * kWave));
}



__host__ __device__ inline IdLds id_lds_layout(int m, int n, int q, bool a_in_lds, size_t tsz) {
  IdLds L;
  size_t o = 0;
  L.a = o;
  if (a_in_lds) o = align16(o + (size_t)m * n * tsz);
  L.rgeo = o;
  o = align16(o + sizeof(double) * kGeo * (size_t)q);
  L.cgeo = o;
  o = align16(o + sizeof(double) * kGeo * (size_t)n);
  L.rows = o;
  o = align16(o + sizeof(int) * (size_t)q);
  L.act = o;
  o = align16(o + sizeof(int) * (size_t)n);
  L.vn1 = o;
  o = align16(o + sizeof(double) * (size_t)n);
  L.vn2 = o;
  o = align16(o + sizeof(double) * (size_t)n);
  L.perm = o;
  o = align16(o + sizeof(int) * (size_t)n);
  L.vb = o;  // Householder vector, beta and tau of the register-resident QR
  if (id_reg_box(m, n) || id_wide_box(m, n)) o = align16(o + sizeof(double) * ((size_t)m + 2));
  L.red = o;
  o = align16(o + sizeof(double) * 32);
  L.total = o;
  return L;
}

// proxy-block entry against a box DOF with staged geometry bg (see proxy_entry)
__device__ __forceinline__ double staged_proxy_entry(int pde, int row, int bdof, const double* bg,
                                                     double cx, double cy, double rad, double pw,
                                                     const double* ring_c, const double* ring_s) {
#pragma clang fp contract(off)
  if (pde != KIGPU_STOKES) {
    int half = row >= kProxyPoints;
    int i = half ? row - kProxyPoints : row;
    double pc = ring_c[i], ps = ring_s[i];
    double r1 = (cx + rad * pc) - bg[0];
    double r2 = (cy + rad * ps) - bg[1];
    if (!half) return laplace_entry(pde, r1, r2, bg[4], 0.0, bg[2], bg[3], pc, ps, false);
    return laplace_entry(pde, -r1, -r2, pw, 0.0, pc, ps, bg[2], bg[3], false);
  }
  int half = row >= 2 * kProxyPoints;
  int rr_ = half ? row - 2 * kProxyPoints : row;
  int i = rr_ >> 1, comp = rr_ & 1;
  double pc = ring_c[i], ps = ring_s[i];
  const double r[2] = {(cx + rad * pc) - bg[0], (cy + rad * ps) - bg[1]};
  const double sn[2] = {bg[2], bg[3]};
  const double tn[2] = {pc, ps};
  const double sw = bg[4];
  double rr = r[0] * r[0] + r[1] * r[1];
  int sc = bdof % 2;
  if (!half) {
    double potential = sw * kInvPi * (r[0] * sn[0] + r[1] * sn[1]) / (rr * rr);
    return potential * r[comp] * r[sc] + sw * tn[comp] * sn[sc];
  }
  double potential = -pw * kInvPi * (r[0] * tn[0] + r[1] * tn[1]) / (rr * rr);
  return potential * r[comp] * r[sc] + pw * sn[sc] * tn[comp];
}

// Entry (i, j) of a box's ID matrix [K(rows, B); K(B, rows)^T; proxy]
// (kernel.cpp:153-235) from the geometry staged in LDS.
__device__ __forceinline__ double id_entry(const IdWork& w, int pde, int i, int j,
                                           const int* s_rows, const double* rgeo,
                                           const int* s_act, const double* cgeo,
                                           const double* ring_c, const double* ring_s) {
  const int q = w.q, bd = s_act[j];
  const double* bg = cgeo + kGeo * j;
  if (i < q) return staged_entry(pde, s_rows[i], rgeo + kGeo * i, bd, bg);
  if (i < 2 * q) return staged_entry(pde, bd, bg, s_rows[i - q], rgeo + kGeo * (i - q));
  return staged_proxy_entry(pde, i - 2 * q, bd, bg, w.cx, w.cy, w.rad, w.pw, ring_c, ring_s);
}

constexpr int kIdCR = 8;  // rows per lane of the register-cached trailing update

// One workgroup per box: generate the ID matrix, Householder QR with column
// pivoting (the steps of LAPACK dlaqp2: first-maximum pivot on the partial
// column norms, norm downdating with recomputation when
// (1 - (|a_ij|/vn1)^2)(vn1/vn2)^2 <= sqrt(eps)) until the first i >= 1 with
// |R_ii| < |tol R_00| (ki_mat.cpp:516-528), then T = R11^{-1} R12.
// T: arithmetic of the QR (FP64: an ID at tol ~1e-6 needs it, since
// T = R11^{-1} R12 inherits cond(R11) ~ 1/tol); TS: storage type of the
// interpolation matrix in the factor slab.
// Outputs of a pivoted QR that stopped at rank k, with logical position j
// held by physical column lp[j]: the pivots, the permuted active list, and
// T = R11^{-1} R12 (dtrtrs) into the factor slab.
template <typename T, typename TS, int TB>
__device__ __forceinline__ void id_finish(const IdWork& w, T* A, int lda, int n, int k,
                                          const int* lp, const int* s_act,
                                          const uint64_t* fac_ptr, int* rank, int* sr_pool,
                                          int* pos_pool, int64_t aoff) {
  const int tid = threadIdx.x;
  for (int j = tid; j < n; j += TB) {
    pos_pool[aoff + j] = lp[j];
    sr_pool[aoff + j] = s_act[lp[j]];
  }
  const int r = n - k;
  TS* Tm = reinterpret_cast<TS*>(fac_ptr[w.node]);
  for (int c = tid; c < r; c += TB) {  // back substitution in place of R12's column
    T* rc = A + (size_t)lp[k + c] * lda;
    for (int row = k - 1; row >= 0; --row) {
      T s2 = rc[row];
      for (int j = row + 1; j < k; ++j) s2 -= A[row + (size_t)lp[j] * lda] * rc[j];
      rc[row] = s2 / A[row + (size_t)lp[row] * lda];
    }
    TS* tc = Tm + (size_t)c * k;
    for (int row = 0; row < k; ++row) tc[row] = (TS)rc[row];
  }
  if (tid == 0) rank[w.node] = k;
}

// Householder QR with column pivoting, the same dlaqp2 steps as id_box below,
// with the ID matrix held in registers for the whole factorization: wave w
// keeps columns w, w + 16, w + 32 (physical order), lane l rows l, l + 64,
// ..., and each wave keeps its columns' partial norms.  Per step:
//   * every wave finds the pivot (first maximum of the partial norms over
//     logical positions i..n-1) and swaps the logical-to-physical map, which
//     lane j of every wave holds for position j;
//   * the wave that holds the pivot column runs dlarfg on it and puts the
//     scaled Householder vector, beta and tau in LDS;
//   * after a block barrier every wave applies H(i)^T to its trailing
//     columns in registers and downdates their norms.
// R(0..k-1, :) goes back to A for the back substitution.
template <typename T, typename TS, int TB, int RC, int RR>
__device__ __forceinline__ void id_qr_reg(const IdWork& w, T* A, int lda, int m, int n, double tol,
                                          const int* s_act, double* vn1, const double* vn2,
                                          int* perm, double* vb, const uint64_t* fac_ptr,
                                          int* rank, int* sr_pool, int* pos_pool, int64_t aoff,
                                          int* status) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const double tol3z = sqrt(2.220446049250313e-16);
  T y[RC][RR];
  double n1[RC], n2[RC];
  unsigned live = 0;  // slots holding a trailing column
#pragma unroll
  for (int c = 0; c < RC; ++c) {
    const int pc = wv + (TB / kWave) * c;
    const bool on = pc < n;
    if (on) live |= 1u << c;
#pragma unroll
    for (int u = 0; u < RR; ++u) {
      const int r = lane + kWave * u;
      y[c][u] = on && r < m ? A[r + (size_t)pc * lda] : T(0);
    }
    n1[c] = on ? vn1[pc] : 0.0;
    n2[c] = on ? vn2[pc] : 0.0;
  }
  int lp = lane;  // physical column at logical position `lane` (lanes < n)
  double thresh = 0;
  int k = 0;
  for (int i = 0; i < n; ++i) {
    if (i >= m) {  // the reference reads R(i, i) past the last row here
      if (tid == 0) flag(status, 2);
      k = 0;
      break;
    }
    double best = -1.0;
    int bi = n;
    if (lane >= i && lane < n) {
      const double v = vn1[lp];
      if (v > best) {
        best = v;
        bi = lane;
      }
    }
    wave_argmax(best, bi);
    int pi = __builtin_amdgcn_readfirstlane(bi);
    if (pi >= n) pi = i;  // NaN norms: keep the order
    const int p = dpp::lane_value(lp, pi), pc_i = dpp::lane_value(lp, i);
    lp = lane == i ? p : (lane == pi ? pc_i : lp);
    const int ow = p % (TB / kWave), oc = p / (TB / kWave);
#pragma unroll
    for (int c = 0; c < RC; ++c) {
      if (wv != ow || c != oc) continue;
      // dlarfg on the pivot column below row i (row i < n <= 64: slot 0)
      const double alpha = dpp::lane_value((double)y[c][0], i);
      double s = 0;
#pragma unroll
      for (int u = 0; u < RR; ++u)
        if (lane + kWave * u > i) s += (double)y[c][u] * (double)y[c][u];
      const double xnorm = sqrt(wave_sum(s));
      double beta, tau;
      if (xnorm == 0.0) {
        beta = alpha;
        tau = 0.0;
      } else {
        beta = -copysign(hypot(alpha, xnorm), alpha);
        tau = (beta - alpha) / beta;
      }
      const T scal = xnorm == 0.0 ? T(0) : (T)(1.0 / (alpha - beta));
#pragma unroll
      for (int u = 0; u < RR; ++u) {
        const int r = lane + kWave * u;
        if (r < m) vb[r] = r > i ? (double)(y[c][u] * scal) : 0.0;
      }
      if (lane == 0) {
        vb[m] = beta;
        vb[m + 1] = tau;
      }
      if (lane == i) y[c][0] = (T)beta;  // R(i, i); the column retires
      live &= ~(1u << c);
    }
    __syncthreads();
    const double beta = vb[m], tau = vb[m + 1];
    if (i == 0) {
      thresh = fabs(tol * beta);
    } else if (fabs(beta) < thresh) {
      k = i;
      break;
    }
    if (i == n - 1) {  // no |R_ii| fell below the threshold: not compressed
      k = 0;
      break;
    }
    T vv[RR];
#pragma unroll
    for (int u = 0; u < RR; ++u) {
      const int r = lane + kWave * u;
      vv[u] = r < m ? (T)vb[r] : T(0);
    }
    const T ttau = (T)tau;
#pragma unroll
    for (int c = 0; c < RC; ++c) {
      if (!(live & (1u << c))) continue;
      T d = lane == i ? y[c][0] : T(0);
#pragma unroll
      for (int u = 0; u < RR; ++u) d += vv[u] * y[c][u];
      d = wave_sum(d);
      const T f = ttau * d;
#pragma unroll
      for (int u = 0; u < RR; ++u) y[c][u] -= f * vv[u];
      if (lane == i) y[c][0] -= f;
      const double newi = dpp::lane_value((double)y[c][0], i);
      const double v1 = n1[c];
      if (v1 != 0.0) {
        double tt = fabs(newi) / v1;
        double temp = fmax(1.0 - tt * tt, 0.0);
        double rr = v1 / n2[c];
        double temp2 = temp * rr * rr;
        if (temp2 <= tol3z) {
          double s2 = 0;
#pragma unroll
          for (int u = 0; u < RR; ++u)
            if (lane + kWave * u > i) s2 += (double)y[c][u] * (double)y[c][u];
          s2 = wave_sum(s2);
          const double nv = i + 1 < m ? sqrt(s2) : 0.0;
          n1[c] = nv;
          n2[c] = nv;
        } else {
          n1[c] = v1 * sqrt(temp);
        }
      }
      if (lane == 0) vn1[wv + (TB / kWave) * c] = n1[c];
    }
    __syncthreads();
  }
  __syncthreads();
  if (k == 0) {
    if (tid == 0) rank[w.node] = -1;
    return;
  }
  // R(0..k-1, :) (k < n <= 64: slot 0) and the final column map
#pragma unroll
  for (int c = 0; c < RC; ++c) {
    const int pc = wv + (TB / kWave) * c;
    if (pc < n && lane < k) A[lane + (size_t)pc * lda] = y[c][0];
  }
  if (wv == 0 && lane < n) perm[lane] = lp;
  __syncthreads();
  id_finish<T, TS, TB>(w, A, lda, n, k, perm, s_act, fac_ptr, rank, sr_pool, pos_pool, aoff);
}

// GIVEN: the ID matrix is read from `given` (m x n, column-major) instead of
// generated, for the ID of an arbitrary matrix (ki_Mat::id, ki_mat.cpp:505-541).
template <typename T, typename TS, bool LDS, bool GIVEN = false, int TB = kIdTB>
__device__ __forceinline__ void id_box(const IdWork& w, unsigned char* smem, GeomView g, int pde,
                                       double tol, const int* rows_pool, const int* act_pool,
                                       const int64_t* act_off, const double* ring, T* ws,
                                       int* rank, int* sr_pool, int* pos_pool,
                                       const uint64_t* fac_ptr, int* status,
                                       const double* given = nullptr) {
  const int n = w.n, q = w.q, m = w.m;
  const IdLds Ly = id_lds_layout(m, n, q, LDS, sizeof(T));
  T* A = LDS ? reinterpret_cast<T*>(smem + Ly.a) : ws + w.a_off;
  double* rgeo = reinterpret_cast<double*>(smem + Ly.rgeo);
  double* cgeo = reinterpret_cast<double*>(smem + Ly.cgeo);
  int* s_rows = reinterpret_cast<int*>(smem + Ly.rows);
  int* s_act = reinterpret_cast<int*>(smem + Ly.act);
  double* vn1 = reinterpret_cast<double*>(smem + Ly.vn1);
  double* vn2 = reinterpret_cast<double*>(smem + Ly.vn2);
  int* perm = reinterpret_cast<int*>(smem + Ly.perm);
  double* red = reinterpret_cast<double*>(smem + Ly.red);
  __shared__ int s_pvt;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int lda = m;
  const int64_t aoff = act_off[w.node];

  if (GIVEN) {
    for (int j = tid; j < n; j += TB) {
      s_act[j] = act_pool[aoff + j];
      perm[j] = j;
    }
    if (LDS) for2d<(TB / kWave)>(m, n, [&](int i, int j) { A[i + (size_t)j * lda] = (T)given[i + (size_t)j * m]; });
    __syncthreads();
  } else {
    for (int i = tid; i < q; i += TB) {
      const int d = rows_pool[w.rows_off + i];
      s_rows[i] = d;
      stage_geo(g, d, rgeo + kGeo * i);
    }
    for (int j = tid; j < n; j += TB) {
      const int d = act_pool[aoff + j];
      s_act[j] = d;
      perm[j] = j;
      stage_geo(g, d, cgeo + kGeo * j);
    }
    __syncthreads();

    // ---- ID matrix [K(rows, B); K(B, rows)^T; proxy] (kernel.cpp:153-235)
    const double* ring_c = ring;
    const double* ring_s = ring + kProxyPoints;
    for2d<(TB / kWave)>(m, n, [&](int i, int j) {
      A[i + (size_t)j * lda] = (T)id_entry(w, pde, i, j, s_rows, rgeo, s_act, cgeo, ring_c, ring_s);
    });
    __syncthreads();
  }

  // ---- column norms (dgeqp3: DNRM2 of each column), FP64
  for (int j = wv; j < n; j += (TB / kWave)) {
    const T* col = A + (size_t)j * lda;
    double s = 0;
    for (int i = lane; i < m; i += kWave) {
      double v = (double)col[i];
      s += v * v;
    }
    s = wave_sum(s);
    if (lane == 0) {
      vn1[j] = sqrt(s);
      vn2[j] = vn1[j];
    }
  }
  __syncthreads();
  double* vb = reinterpret_cast<double*>(smem + Ly.vb);
  if constexpr (TB == kIdWideTB) {  // k_id_wide takes wide boxes only
    if (m <= 8 * kWave)
      id_qr_reg<T, TS, TB, 8, 8>(w, A, lda, m, n, tol, s_act, vn1, vn2, perm, vb, fac_ptr, rank,
                                 sr_pool, pos_pool, aoff, status);
    else
      id_qr_reg<T, TS, TB, 6, 10>(w, A, lda, m, n, tol, s_act, vn1, vn2, perm, vb, fac_ptr, rank,
                                  sr_pool, pos_pool, aoff, status);
    return;
  } else {
    if (id_reg_box(m, n)) {
      id_qr_reg<T, TS, TB, kIdRC, kIdRR>(w, A, lda, m, n, tol, s_act, vn1, vn2, perm, vb, fac_ptr,
                                         rank, sr_pool, pos_pool, aoff, status);
      return;
    }

    const double tol3z = sqrt(2.220446049250313e-16);
    double thresh = 0;
    int k = 0;
    for (int i = 0; i < n; ++i) {
      if (i >= m) {  // the reference reads R(i, i) past the last row here
        if (tid == 0) flag(status, 2);
        k = 0;
        break;
      }
      if (wv == 0) {
        double best = -1.0;
        int bi = i;  // NaN norms compare false: the column order is kept
        for (int j = i + lane; j < n; j += kWave) {
          double v = vn1[j];
          if (v > best) {
            best = v;
            bi = j;
          }
        }
        wave_argmax(best, bi);
        if (lane == 0) s_pvt = bi;
      }
      __syncthreads();
      const int p = s_pvt;
      if (p != i) {
        T* ci = A + (size_t)i * lda;
        T* cp = A + (size_t)p * lda;
        for (int r = tid; r < m; r += kIdTB) {
          T t0 = ci[r];
          ci[r] = cp[r];
          cp[r] = t0;
        }
        if (tid == 0) {
          int t1 = perm[i];
          perm[i] = perm[p];
          perm[p] = t1;
          vn1[p] = vn1[i];
          vn2[p] = vn2[i];
        }
      }
      __syncthreads();
      // dlarfg on A(i:m, i)
      T* ci = A + (size_t)i * lda;
      double part = 0;
      for (int r = i + 1 + tid; r < m; r += kIdTB) {
        double v = (double)ci[r];
        part += v * v;
      }
      const double xnorm = sqrt(block_sum<kIdNW>(part, red));
      const double alpha = (double)ci[i];
      double beta, tau;
      if (xnorm == 0.0) {
        beta = alpha;
        tau = 0.0;
      } else {
        beta = -copysign(hypot(alpha, xnorm), alpha);
        tau = (beta - alpha) / beta;
      }
      if (i == 0) {
        thresh = fabs(tol * beta);
      } else if (fabs(beta) < thresh) {
        k = i;
        break;
      }
      if (i == n - 1) {  // no |R_ii| fell below the threshold: not compressed
        k = 0;
        break;
      }
      const T scal = xnorm == 0.0 ? T(0) : (T)(1.0 / (alpha - beta));
      for (int r = i + 1 + tid; r < m; r += kIdTB) ci[r] *= scal;
      __syncthreads();
      if (tid == 0) ci[i] = (T)beta;
      // apply H(i)^T to A(i:m, i+1:n) and downdate the partial column norms
      const T ttau = (T)tau;
      if (m - (i + 1) <= kIdCR * kWave) {
        // Each wave keeps the Householder vector's rows in registers for the
        // whole step and a column's rows while it works on it: one LDS read and
        // one write per entry instead of five accesses (the update is bound by
        // LDS bandwidth).  Same operations in the same order as below.
        T cv[kIdCR];
  #pragma unroll
        for (int u = 0; u < kIdCR; ++u) {
          const int r = i + 1 + lane + u * kWave;
          cv[u] = r < m ? ci[r] : T(0);
        }
        for (int j = i + 1 + wv; j < n; j += kIdNW) {
          T* cj = A + (size_t)j * lda;
          const T aij = cj[i];
          T y[kIdCR];
  #pragma unroll
          for (int u = 0; u < kIdCR; ++u) {
            const int r = i + 1 + lane + u * kWave;
            y[u] = r < m ? cj[r] : T(0);
          }
          T d = lane == 0 ? aij : T(0);
  #pragma unroll
          for (int u = 0; u < kIdCR; ++u)
            if (i + 1 + lane + u * kWave < m) d += cv[u] * y[u];
          // every lane reads the norms before the reduction: lane 0 rewrites
          // them after it
          const double v1 = vn1[j], v2 = vn2[j];
          d = wave_sum(d);
          const T f = ttau * d;
  #pragma unroll
          for (int u = 0; u < kIdCR; ++u) {
            const int r = i + 1 + lane + u * kWave;
            if (r < m) {
              y[u] -= f * cv[u];
              cj[r] = y[u];
            }
          }
          const T newi = aij - f;
          if (lane == 0) cj[i] = newi;
          if (v1 != 0.0) {
            double tt = fabs((double)newi) / v1;
            double temp = fmax(1.0 - tt * tt, 0.0);
            double rr = v1 / v2;
            double temp2 = temp * rr * rr;
            if (temp2 <= tol3z) {
              double s2 = 0;
  #pragma unroll
              for (int u = 0; u < kIdCR; ++u) {
                if (i + 1 + lane + u * kWave < m) {
                  double v = (double)y[u];
                  s2 += v * v;
                }
              }
              s2 = wave_sum(s2);
              if (lane == 0) {
                double nv = i + 1 < m ? sqrt(s2) : 0.0;
                vn1[j] = nv;
                vn2[j] = nv;
              }
            } else if (lane == 0) {
              vn1[j] = v1 * sqrt(temp);
            }
          }
        }
        __syncthreads();
        continue;
      }
      for (int j = i + 1 + wv; j < n; j += kIdNW) {
        T* cj = A + (size_t)j * lda;
        const T aij = cj[i];
        T d = lane == 0 ? aij : T(0);
        for (int r = i + 1 + lane; r < m; r += kWave) d += ci[r] * cj[r];
        const double v1 = vn1[j], v2 = vn2[j];  // read before lane 0 rewrites them
        d = wave_sum(d);
        const T f = ttau * d;
        for (int r = i + 1 + lane; r < m; r += kWave) cj[r] -= f * ci[r];
        const T newi = aij - f;
        if (lane == 0) cj[i] = newi;
        if (v1 != 0.0) {
          double tt = fabs((double)newi) / v1;
          double temp = fmax(1.0 - tt * tt, 0.0);
          double rr = v1 / v2;
          double temp2 = temp * rr * rr;
          if (temp2 <= tol3z) {
            double s2 = 0;
            for (int r = i + 1 + lane; r < m; r += kWave) {
              double v = (double)cj[r];
              s2 += v * v;
            }
            s2 = wave_sum(s2);
            if (lane == 0) {
              double nv = i + 1 < m ? sqrt(s2) : 0.0;
              vn1[j] = nv;
              vn2[j] = nv;
            }
          } else if (lane == 0) {
            vn1[j] = v1 * sqrt(temp);
          }
        }
      }
      __syncthreads();
    }
    __syncthreads();
    if (k == 0) {
      if (tid == 0) rank[w.node] = -1;
      return;
    }
    // ---- outputs: pivots, permuted active list, T = R11^{-1} R12 (dtrtrs)
    for (int j = tid; j < n; j += kIdTB) {
      pos_pool[aoff + j] = perm[j];
      sr_pool[aoff + j] = s_act[perm[j]];
    }
    const int r = n - k;
    TS* Tm = reinterpret_cast<TS*>(fac_ptr[w.node]);
    for (int c = tid; c < r; c += kIdTB) {  // back substitution in place of R12's column
      T* rc = A + (size_t)(k + c) * lda;
      for (int row = k - 1; row >= 0; --row) {
        T s2 = rc[row];
        for (int j = row + 1; j < k; ++j) s2 -= A[row + (size_t)j * lda] * rc[j];
        rc[row] = s2 / A[row + (size_t)row * lda];
      }
      TS* tc = Tm + (size_t)c * k;
      for (int row = 0; row < k; ++row) tc[row] = (TS)rc[row];
    }
    if (tid == 0) rank[w.node] = k;
  }
}

// ---------------------------------------------------------------------------
// ID from the Gram matrix (the FP32 factorization's ID)
// ---------------------------------------------------------------------------
// The pivots, the rank and R of a Householder QR with column pivoting depend
// on A only through A^T A.  Per box (one 256-thread block, n <= 64):
//  * G = A^T A on the FP64 matrix cores (v_mfma_f64_16x16x4f64), the entries
//    of A generated as they are consumed, four rows at a time: wave w takes
//    the row groups w, w + 4, ...; lane l generates A(4 s + l / 16, 16 c +
//    l % 16) for the column groups c, which is both operands' layout, so
//    every 16 x 16 tile (c1 <= c2) of G is one 4-deep step per row group;
//    the four waves' partial tiles are added in wave order (deterministic);
//  * a pivoted Cholesky of G by one wave (lane = logical column position):
//    the same pivot rule as the QR (first largest remaining column norm,
//    here its square, the Schur complement's diagonal), the same rank rule
//    (first i >= 1 with R_ii < |tol R_00|), the same column interchanges;
//    row i of R overwrites row p_i of G, which no later step reads;
//  * T = R11^{-1} R12 by back substitution (dtrtrs), one lane per column.
// In exact arithmetic this is the QR's ID.  In FP64 the Schur complement's
// diagonal carries an absolute error ~ eps * G_jj, so pivots whose remaining
// norms agree to that level (near ties), and ranks with |R_kk| that close
// to the threshold, can come out differently, and T differs from the QR's
// in the directions R11 nearly annihilates; the compressed operator agrees
// to ~1e-10 relative, far below the ID tolerance.  The FP32 factorization
// uses it (its solution is refined against the FP64 kernel apply); the
// FP64 factorization keeps the Householder QR.
constexpr int kGramTB = 256;
constexpr int kGramNW = kGramTB / kWave;
constexpr int kGramMaxN = 64;
constexpr int kGramU = 2;  // row groups generated together
// The rank rule compares remaining norms squared with tol^2 |R_00|^2, and G
// resolves them only to ~eps |A_j|^2: below this tolerance the default
// keeps the pivoted QR.
constexpr double kGramMinTol = 1e-7;
// A remaining norm squared at most this times (step + 1) |A_p|^2 is rounding
// (measured: the Schur complement's diagonal differs from the QR's R_ii^2
// by at most 6.1 (i + 1) eps |A_p|^2 on the star's boxes), so the box's
// rank stops there whatever the tolerance.
constexpr double kGramFloor = 16 * 2.220446049250313e-16;

struct GramLds {
  size_t rgeo, cgeo, rows, act, g, t, pv, lp, total;
};
__host__ __device__ inline GramLds gram_layout(int n, int q) {
  GramLds L;
  size_t o = 0;
  L.rgeo = o; o = align16(o + sizeof(double) * kGeo * (size_t)q);
  L.cgeo = o; o = align16(o + sizeof(double) * kGeo * (size_t)n);
  L.rows = o; o = align16(o + sizeof(int) * (size_t)q);
  L.act = o;  o = align16(o + sizeof(int) * (size_t)n);
  L.g = o;    o = align16(o + sizeof(double) * (size_t)n * n);
  L.t = o;    o = align16(o + sizeof(double) * ((size_t)n * n / 4 + 1));  // k (n - k) <= n^2 / 4
  L.pv = o;   o = align16(o + sizeof(int) * (size_t)n);
  L.lp = o;   o = align16(o + sizeof(int) * (size_t)n);
  L.total = o;
  return L;
}

typedef double gram_d4 __attribute__((ext_vector_type(4)));

template <typename TS, int PDE>
__device__ __forceinline__ void id_gram_box(const IdWork& w, unsigned char* smem, GeomView g,
                                            double tol, const int* rows_pool,
                                            const int* act_pool, const int64_t* act_off,
                                            const double* ring, int* rank, int* sr_pool,
                                            int* pos_pool, const uint64_t* fac_ptr,
                                            int* status) {
  const int n = w.n, q = w.q, m = w.m;
  const GramLds Ly = gram_layout(n, q);
  double* rgeo = reinterpret_cast<double*>(smem + Ly.rgeo);
  double* cgeo = reinterpret_cast<double*>(smem + Ly.cgeo);
  int* s_rows = reinterpret_cast<int*>(smem + Ly.rows);
  int* s_act = reinterpret_cast<int*>(smem + Ly.act);
  double* G = reinterpret_cast<double*>(smem + Ly.g);  // n x n, ld n
  double* Ts = reinterpret_cast<double*>(smem + Ly.t);
  int* s_pv = reinterpret_cast<int*>(smem + Ly.pv);  // physical pivot column of each step
  int* s_lp = reinterpret_cast<int*>(smem + Ly.lp);  // logical position -> physical column
  __shared__ int s_k;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int64_t aoff = act_off[w.node];
  for (int i = tid; i < q; i += kGramTB) {
    const int d = rows_pool[w.rows_off + i];
    s_rows[i] = d;
    stage_geo(g, d, rgeo + kGeo * i);
  }
  for (int j = tid; j < n; j += kGramTB) {
    const int d = act_pool[aoff + j];
    s_act[j] = d;
    stage_geo(g, d, cgeo + kGeo * j);
  }
  __syncthreads();

  // ---- G = A^T A: the three row blocks of the ID matrix one after the
  // other (their rows are added to G in any grouping), so that a row group's
  // entries come from one formula; each lane keeps its four columns'
  // geometry in registers
  const double* ring_c = ring;
  const double* ring_s = ring + kProxyPoints;
  const int nc = (n + 15) >> 4, l16 = lane & 15, l4 = lane >> 4;
  double cg[4][kGeo];
  int cd[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int jc = min(16 * c + l16, n - 1);
    cd[c] = s_act[jc];
#pragma unroll
    for (int u = 0; u < kGeo; ++u) cg[c][u] = cgeo[kGeo * jc + u];
  }
  gram_d4 acc[10];
#pragma unroll
  for (int t = 0; t < 10; ++t) acc[t] = gram_d4{0.0, 0.0, 0.0, 0.0};
  // rows [0, nrow) of one block; f(i, c) is entry (i, column group c)
  auto block_rows = [&](int nrow, auto&& f) {
    for (int s0 = 4 * wv; s0 < nrow; s0 += 4 * kGramNW * kGramU) {
      double e[kGramU][4];
#pragma unroll
      for (int h = 0; h < kGramU; ++h) {
        const int i = s0 + 4 * kGramNW * h + l4;
        const int ic = min(i, nrow - 1);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const double v = c < nc ? f(ic, c) : 0.0;
          e[h][c] = (i < nrow && 16 * c + l16 < n) ? v : 0.0;
        }
      }
#pragma unroll
      for (int h = 0; h < kGramU; ++h) {
        int t = 0;
#pragma unroll
        for (int c1 = 0; c1 < 4; ++c1)
#pragma unroll
          for (int c2 = c1; c2 < 4; ++c2, ++t)
            if (c2 < nc)
              acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(e[h][c1], e[h][c2], acc[t], 0, 0, 0);
      }
    }
  };
  // K(rows, B) and K(B, rows)^T
  block_rows(q, [&](int i, int c) {
    return staged_entry(PDE, s_rows[i], rgeo + kGeo * i, cd[c], cg[c]);
  });
  block_rows(q, [&](int i, int c) {
    return staged_entry(PDE, cd[c], cg[c], s_rows[i], rgeo + kGeo * i);
  });
  // the proxy block
  block_rows(m - 2 * q, [&](int i, int c) {
    return staged_proxy_entry(PDE, i, cd[c], cg[c], w.cx, w.cy, w.rad, w.pw, ring_c, ring_s);
  });
  for (int wr = 0; wr < kGramNW; ++wr) {
    if (wv == wr) {
      int t = 0;
#pragma unroll
      for (int c1 = 0; c1 < 4; ++c1)
#pragma unroll
        for (int c2 = c1; c2 < 4; ++c2, ++t) {
          if (c2 >= nc) continue;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int gi = 16 * c1 + l4 + 4 * v, gj = 16 * c2 + l16;
            if (gi < n && gj < n) {
              double x = acc[t][v];
              if (wr > 0) x += G[gi + (size_t)gj * n];
              G[gi + (size_t)gj * n] = x;
              if (c1 != c2) G[gj + (size_t)gi * n] = x;
            }
          }
        }
    }
    __syncthreads();
  }

  // ---- pivoted Cholesky of G by wave 0; lane j: logical position j
  if (wv == 0) {
    int lp = lane;  // physical column at this logical position
    double d = lane < n ? G[lane + (size_t)lane * n] : 0.0;  // its remaining norm squared
    double gd = d;                                             // its norm squared
    int mypv = 0;  // lane l: the physical pivot column of step l
    double thresh = 0;
    int k = 0;
    for (int i = 0; i < n; ++i) {
      if (i >= m) {  // the reference reads R(i, i) past the last row here
        if (lane == 0) flag(status, 2);
        k = 0;
        break;
      }
      double best = (lane >= i && lane < n) ? d : -1.0;
      int bi = lane;
      wave_argmax(best, bi);
      int pi = __builtin_amdgcn_readfirstlane(bi);
      if (pi < i || pi >= n) pi = i;  // no comparable candidate (NaN): keep the order
      const int p = dpp::lane_value(lp, pi), lp_i = dpp::lane_value(lp, i);
      const double dp = dpp::lane_value(d, pi), d_i = dpp::lane_value(d, i);
      const double gp = dpp::lane_value(gd, pi), gd_i = dpp::lane_value(gd, i);
      lp = lane == i ? p : (lane == pi ? lp_i : lp);
      d = lane == i ? dp : (lane == pi ? d_i : d);
      gd = lane == i ? gp : (lane == pi ? gd_i : gd);
      const double rii = sqrt(fmax(dp, 0.0));
      if (i == 0) {
        thresh = fabs(tol * rii);
      } else if (rii < thresh || dp <= kGramFloor * (i + 1) * gp) {
        // (or the largest remaining norm is at G's rounding level: the rest
        // of the box is numerically dependent on the pivots so far)
        k = i;
        break;
      }
      if (i == n - 1) {  // no R_ii fell below the threshold: not compressed
        k = 0;
        break;
      }
      // row i of R: R(i, c) = (G(p, c) - sum_{l < i} R(l, p) R(l, c)) / R_ii
      // for the columns c at the logical positions after i; R(l, :) is row
      // s_pv[l] of G
      // Every lane runs the loop (the whole wavefront takes part in the
      // readlane); lanes at other positions work on column p and drop the
      // result, so the loop has no branch and every index is in G.
      const bool after = lane > i && lane < n;
      const int lc = after ? lp : p;
      double r = G[p + (size_t)lc * n];
#pragma unroll 4
      for (int l = 0; l < i; ++l) {
        const int pl = __builtin_amdgcn_readlane(mypv, l);
        r -= G[pl + (size_t)p * n] * G[pl + (size_t)lc * n];
      }
      if (after) {
        r /= rii;
        G[p + (size_t)lp * n] = r;
        d -= r * r;
      }
      if (lane == i) mypv = p;
      if (lane == 0) G[p + (size_t)p * n] = rii;
      wave_sync();
    }
    if (lane < n) s_lp[lane] = lp;
    if (lane < k) s_pv[lane] = mypv;
    if (lane == 0) s_k = k;
  }
  __syncthreads();
  const int k = s_k;
  if (k == 0) {
    if (tid == 0) rank[w.node] = -1;
    return;
  }
  // ---- outputs: pivots, permuted active list, T = R11^{-1} R12 (dtrtrs)
  for (int j = tid; j < n; j += kGramTB) {
    pos_pool[aoff + j] = s_lp[j];
    sr_pool[aoff + j] = s_act[s_lp[j]];
  }
  const int r = n - k;
  TS* Tm = reinterpret_cast<TS*>(fac_ptr[w.node]);
  for (int c = tid; c < r; c += kGramTB) {
    const int pc = s_lp[k + c];
    for (int row = k - 1; row >= 0; --row) {
      const int prow = s_pv[row];
      double v = G[prow + (size_t)pc * n];  // R(row, c)
#pragma unroll 4
      for (int j = row + 1; j < k; ++j) v -= G[prow + (size_t)s_pv[j] * n] * Ts[j + (size_t)c * k];
      v /= G[prow + (size_t)prow * n];
      Ts[row + (size_t)c * k] = v;
      Tm[row + (size_t)c * k] = (TS)v;
    }
  }
  if (tid == 0) rank[w.node] = k;
}

}  // namespace elim
}  // namespace kigpu
