// Recursive-skeletonization factorization on gfx950: one batched launch per
// tree level for each phase (index lists, ID, elimination), then the root LU.
//
// Follows kern_interp/skel_factorization/skel_factorization.cpp:94-243
// (skeletonize), :32-52 (id_compress), :55-91 (decouple), the ID of
// kern_interp/ki_mat.cpp:505-541 and the ID matrix of
// kern_interp/kernel/kernel.cpp:153-235.
//
// Differences of form, not of result:
//  * the pivoted QR stops at the rank-revealing step instead of running to
//    min(m, n) (LAPACK dgeqp3 does the full factorization); the skeleton
//    columns and their order, the rank and R11^{-1} R12 are the same, the
//    redundant DOFs keep the column order reached at that step;
//  * the Schur updates of descendants (skel_factorization.cpp:246-294) are
//    propagated child -> parent as one accumulated block per box instead of
//    rescanning every descendant; the sum is the same, the order of the
//    floating-point additions differs.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_util.h"
#include "entries.h"
#include "factor.h"

namespace kigpu {

namespace {

constexpr int kTB = 256;
constexpr int kNW = kTB / kWave;

__device__ __forceinline__ void flag(int* status, int bit) { atomicOr(status, bit); }

// ---------------------------------------------------------------------------
// index lists
// ---------------------------------------------------------------------------

// original_box DOF lists of the leaves (quadtree.cpp:167-170, :256-280)
__global__ void k_obox(int nnodes, const int* is_leaf, const int* pcount, const int64_t* pstart,
                       const int* perm, const int64_t* obox_off, int sd, int* obox) {
  int x = blockIdx.x;
  if (x >= nnodes || !is_leaf[x]) return;
  const int* pp = perm + pstart[x];
  int* out = obox + obox_off[x];
  for (int j = threadIdx.x; j < pcount[x]; j += blockDim.x)
    for (int c = 0; c < sd; ++c) out[sd * j + c] = sd * pp[j] + c;
}

__global__ void k_leaf_dofs(int nnodes, const int* is_leaf, const int* pcount, int sd, int* cnt) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < nnodes) cnt[x] = is_leaf[x] ? pcount[x] * sd : 0;
}

// |active_box| (remove_inactive_dofs_at_box, quadtree.cpp:694-720)
__global__ void k_active_size(int a, int b, const int* is_leaf, const int* child,
                              const int* pcount, const int* osize, int sd, int* nact,
                              int* rank) {
  int x = a + blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= b) return;
  int n = 0;
  if (is_leaf[x]) {
    n = pcount[x] * sd;
  } else {
    for (int s = 0; s < 4; ++s) n += osize[child[4 * x + s]];
  }
  nact[x] = n;
  rank[x] = -1;
}

// the list a node hands to its parent: skel (compressed) or active_box
__device__ __forceinline__ const int* out_list(int y, const int* rank, const int64_t* act_off,
                                               const int* act_pool, const int* sr_pool) {
  return rank[y] >= 0 ? sr_pool + act_off[y] : act_pool + act_off[y];
}

__global__ void k_active_fill(int a, const int* is_leaf, const int* child, const int64_t* obox_off,
                              const int* obox, const int* osize, const int* rank,
                              const int64_t* act_off, const int* nact, int* act_pool,
                              const int* sr_pool) {
  int x = a + blockIdx.x;
  int* out = act_pool + act_off[x];
  if (is_leaf[x]) {
    const int* src = obox + obox_off[x];
    for (int j = threadIdx.x; j < nact[x]; j += blockDim.x) out[j] = src[j];
    return;
  }
  int o = 0;
  for (int s = 0; s < 4; ++s) {
    int y = child[4 * x + s];
    const int* src = out_list(y, rank, act_off, act_pool, sr_pool);
    int cnt = osize[y];
    for (int j = threadIdx.x; j < cnt; j += blockDim.x) out[o + j] = src[j];
    o += cnt;
  }
}

struct NearSrc {
  const int* p;
  int n;
};

// near-list segment of neighbor y: original_box for leaves, else active_box
// (remove_inactive_dofs_at_level, quadtree.cpp:672-690)
__device__ __forceinline__ NearSrc near_segment(int y, const int* is_leaf, const int64_t* obox_off,
                                                const int* obox, const int* pcount, int sd,
                                                const int64_t* act_off, const int* nact,
                                                const int* act_pool) {
  if (is_leaf[y]) return {obox + obox_off[y], pcount[y] * sd};
  return {act_pool + act_off[y], nact[y]};
}

// inner-circle rows of the ID matrix (kernel.cpp:209-218) for levels >= 3;
// one wave per node.  pass 0 counts into nrows, pass 1 writes rows_pool.
__global__ void k_inner_rows(int a, int b, int min_dofs, GeomView g, const int* is_leaf,
                             const int64_t* obox_off, const int* obox, const int* pcount,
                             const int64_t* act_off, const int* nact, const int* act_pool,
                             const int* sl_nbr, const int* sl_cnt, const int64_t* cp_off,
                             const int* cp_nbr, const double* corners, const double* side,
                             int* nrows, const int64_t* rows_off, int* rows_pool, int pass) {
#pragma clang fp contract(off)
  int x = a + blockIdx.x;
  if (x >= b) return;
  const int lane = threadIdx.x;
  if (nact[x] < min_dofs) {
    if (!pass && lane == 0) nrows[x] = 0;
    return;
  }
  const double cx = corners[8 * x] + side[x] / 2.0;
  const double cy = corners[8 * x + 1] + side[x] / 2.0;
  const double lim = kRadiusRatio * side[x];
  int* out = pass ? rows_pool + rows_off[x] : nullptr;
  int cnt = 0;
  int ncp = (int)(cp_off[x + 1] - cp_off[x]);
  int nsl = sl_cnt[x];
  for (int e = 0; e < ncp + nsl; ++e) {
    int y = e < ncp ? cp_nbr[cp_off[x] + e] : sl_nbr[kMaxNbrs * x + (e - ncp)];
    NearSrc src = near_segment(y, is_leaf, obox_off, obox, pcount, g.sd, act_off, nact, act_pool);
    for (int j0 = 0; j0 < src.n; j0 += kWave) {
      int j = j0 + lane;
      bool keep = false;
      int d = 0;
      if (j < src.n) {
        d = src.p[j];
        int pt = d / g.sd;
        double dx = cx - g.px[pt];
        double dy = cy - g.py[pt];
        keep = sqrt(dx * dx + dy * dy) < lim;
      }
      unsigned long long m = __ballot(keep);
      if (pass && keep) out[cnt + __popcll(m & ((1ull << lane) - 1))] = d;
      cnt += __popcll(m);
    }
  }
  if (!pass && lane == 0) nrows[x] = cnt;
}

// outside rows for levels 1 and 2 (kernel.cpp:164-207): every other node's
// active_box at this level, then (level 2) the level-1 leaves' original_box.
__global__ void k_outside_rows(const IdWork* work, int a, int b, int level, int l1a, int l1b,
                               const int* is_leaf, const int64_t* obox_off, const int* obox,
                               const int* pcount, int sd, const int64_t* act_off, const int* nact,
                               const int* act_pool, int* rows_pool) {
  const IdWork w = work[blockIdx.x];
  int* out = rows_pool + w.rows_off;
  int o = 0;
  for (int y = a; y < b; ++y) {
    if (y == w.node) continue;
    const int* src = act_pool + act_off[y];
    for (int j = threadIdx.x; j < nact[y]; j += blockDim.x) out[o + j] = src[j];
    o += nact[y];
  }
  if (level == 2) {
    for (int y = l1a; y < l1b; ++y) {
      if (!is_leaf[y]) continue;
      const int* src = obox + obox_off[y];
      int n = pcount[y] * sd;
      for (int j = threadIdx.x; j < n; j += blockDim.x) out[o + j] = src[j];
      o += n;
    }
  }
}

// ---------------------------------------------------------------------------
// ID: generate the matrix, pivoted QR to the rank-revealing step, T solve
// ---------------------------------------------------------------------------

struct IdLds {
  size_t a, rows, act, vn1, vn2, perm, red, total;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

__host__ __device__ inline IdLds id_lds_layout(int m, int n, int q, bool a_in_lds, size_t tsz) {
  IdLds L;
  size_t o = 0;
  L.a = o;
  if (a_in_lds) o = align16(o + (size_t)m * n * tsz);
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
  L.red = o;
  o = align16(o + sizeof(double) * 16);
  L.total = o;
  return L;
}

// T: arithmetic of the QR (FP64: an ID at tol ~1e-6 needs it, since
// T = R11^{-1} R12 inherits cond(R11) ~ 1/tol); TS: storage type of the
// interpolation matrix in the factor slab.
template <typename T, typename TS, bool LDS>
__global__ __launch_bounds__(kTB) void k_id(const IdWork* work, GeomView g, int pde, double tol,
                                            const int* rows_pool, const int* act_pool,
                                            const int64_t* act_off, const double* ring, T* ws,
                                            int* rank, int* sr_pool, int* pos_pool,
                                            const uint64_t* fac_ptr, int* status) {
  extern __shared__ __align__(16) unsigned char smem[];
  const IdWork w = work[blockIdx.x];
  const int n = w.n, q = w.q, m = w.m;
  const IdLds Ly = id_lds_layout(m, n, q, LDS, sizeof(T));
  T* A = LDS ? reinterpret_cast<T*>(smem + Ly.a) : ws + w.a_off;
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

  for (int i = tid; i < q; i += kTB) s_rows[i] = rows_pool[w.rows_off + i];
  for (int j = tid; j < n; j += kTB) {
    s_act[j] = act_pool[aoff + j];
    perm[j] = j;
  }
  __syncthreads();

  // ---- ID matrix [K(rows, B); K(B, rows)^T; proxy] (kernel.cpp:153-235)
  const double* ring_c = ring;
  const double* ring_s = ring + kProxyPoints;
  for (int j = 0; j < n; ++j) {
    const int bd = s_act[j];
    T* col = A + (size_t)j * lda;
    for (int i = tid; i < m; i += kTB) {
      double v;
      if (i < q) v = kernel_entry(g, pde, s_rows[i], bd);
      else if (i < 2 * q) v = kernel_entry(g, pde, bd, s_rows[i - q]);
      else v = proxy_entry(g, pde, i - 2 * q, bd, w.cx, w.cy, w.rad, w.pw, ring_c, ring_s);
      col[i] = (T)v;
    }
  }
  __syncthreads();

  // ---- column norms (dgeqp3: DNRM2 of each column), FP64
  for (int j = wv; j < n; j += kNW) {
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

  // ---- Householder QR with column pivoting (LAPACK dlaqp2 steps), stopping
  // at the first i >= 1 with |R_ii| < |tol R_00| (ki_mat.cpp:516-528)
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
      int bi = n;
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
      for (int r = tid; r < m; r += kTB) {
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
    for (int r = i + 1 + tid; r < m; r += kTB) {
      double v = (double)ci[r];
      part += v * v;
    }
    const double xnorm = sqrt(block_sum<kNW>(part, red));
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
    for (int r = i + 1 + tid; r < m; r += kTB) ci[r] *= scal;
    __syncthreads();
    if (tid == 0) ci[i] = (T)beta;
    // apply H(i)^T to A(i:m, i+1:n) and downdate the partial column norms
    const T ttau = (T)tau;
    for (int j = i + 1 + wv; j < n; j += kNW) {
      T* cj = A + (size_t)j * lda;
      const T aij = cj[i];
      T d = lane == 0 ? aij : T(0);
      for (int r = i + 1 + lane; r < m; r += kWave) d += ci[r] * cj[r];
      d = wave_sum(d);
      const T f = ttau * d;
      for (int r = i + 1 + lane; r < m; r += kWave) cj[r] -= f * ci[r];
      const T newi = aij - f;
      if (lane == 0) cj[i] = newi;
      const double v1 = vn1[j];
      if (v1 != 0.0) {
        double tt = fabs((double)newi) / v1;
        double temp = fmax(1.0 - tt * tt, 0.0);
        double rr = v1 / vn2[j];
        double temp2 = temp * rr * rr;
        if (temp2 <= tol3z) {
          double s = 0;
          for (int r = i + 1 + lane; r < m; r += kWave) {
            double v = (double)cj[r];
            s += v * v;
          }
          s = wave_sum(s);
          if (lane == 0) {
            double nv = i + 1 < m ? sqrt(s) : 0.0;
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
  for (int j = tid; j < n; j += kTB) {
    pos_pool[aoff + j] = perm[j];
    sr_pool[aoff + j] = s_act[perm[j]];
  }
  const int r = n - k;
  TS* Tm = reinterpret_cast<TS*>(fac_ptr[w.node]);
  for (int c = tid; c < r; c += kTB) {  // back substitution in place of R12's column
    T* rc = A + (size_t)(k + c) * lda;
    for (int row = k - 1; row >= 0; --row) {
      T s = rc[row];
      for (int j = row + 1; j < k; ++j) s -= A[row + (size_t)j * lda] * rc[j];
      rc[row] = s / A[row + (size_t)row * lda];
    }
    TS* tc = Tm + (size_t)c * k;
    for (int row = 0; row < k; ++row) tc[row] = (TS)rc[row];
  }
  if (tid == 0) rank[w.node] = k;
}

// ---------------------------------------------------------------------------
// elimination of the redundant DOFs (decouple, skel_factorization.cpp:55-91)
// ---------------------------------------------------------------------------

struct DecLds {
  size_t uact, kp, x, krsn, ksnr, sr, pos, piv, total;
};

__host__ __device__ inline DecLds dec_layout(int n, size_t tsz) {
  DecLds L;
  size_t o = 0;
  size_t nn = (size_t)n * n;
  L.uact = o; o = align16(o + nn * tsz);
  L.kp = o;   o = align16(o + nn * tsz);
  L.x = o;    o = align16(o + nn * tsz);
  L.krsn = o; o = align16(o + nn * tsz / 4 + tsz);
  L.ksnr = o; o = align16(o + nn * tsz / 4 + tsz);
  L.sr = o;   o = align16(o + sizeof(int) * (size_t)n);
  L.pos = o;  o = align16(o + sizeof(int) * (size_t)n);
  L.piv = o;  o = align16(o + sizeof(int) * (size_t)n);
  L.total = o;
  return L;
}

// Factor slab of a compressed node (k skeleton, r redundant DOFs):
//   T (k x r) | L (k x r) | U (r x k) | X_rr LU (r x r) | Acc (k x k)
// all column-major with leading dimension = height.
__host__ __device__ inline size_t slab_elems(int n) { return ((size_t)n * n * 5 + 3) / 4 + 4; }

template <typename T, bool LDS>
__global__ __launch_bounds__(kTB) void k_decouple(int a, GeomView g, int pde, const int* is_leaf,
                                                  const int* child, const int* osize_in,
                                                  const uint64_t* acc_ptr_in, const int* acc_ld_in,
                                                  const int* nact, const int* rank,
                                                  const int64_t* act_off, const int* act_pool,
                                                  const int* sr_pool, const int* pos_pool,
                                                  int* piv_pool, const uint64_t* fac_ptr,
                                                  int* osize, uint64_t* acc_ptr, int* acc_ld,
                                                  T* ws, size_t ws_stride, int* status) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int x = a + blockIdx.x;
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
  unsigned char* base = LDS ? smem : reinterpret_cast<unsigned char*>(ws + ws_stride * blockIdx.x);
  T* Uact = reinterpret_cast<T*>(base + Ly.uact);
  T* slab = reinterpret_cast<T*>(fac_ptr[x]);
  const bool leaf = is_leaf[x];

  // accumulated updates of the children, block diagonal in active order
  if (!leaf) {
    for (int i = tid; i < n * n; i += kTB) Uact[i] = T(0);
    __syncthreads();
    int o = 0;
    for (int s = 0; s < 4; ++s) {
      int y = child[4 * x + s];
      int sz = osize_in[y];
      const T* acc = reinterpret_cast<const T*>(acc_ptr_in[y]);
      if (acc) {
        int ld = acc_ld_in[y];
        for (int e = tid; e < sz * sz; e += kTB) {
          int i = e % sz, j = e / sz;
          Uact[(o + i) + (size_t)(o + j) * n] = acc[i + (size_t)j * ld];
        }
      }
      o += sz;
    }
    __syncthreads();
  }
  if (k < 0) {  // not compressed: hand the whole active list and its update up
    if (!leaf) {
      for (int i = tid; i < n * n; i += kTB) slab[i] = Uact[i];
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
  int* sr = reinterpret_cast<int*>(base + Ly.sr);
  int* pos = reinterpret_cast<int*>(base + Ly.pos);
  int* piv = reinterpret_cast<int*>(base + Ly.piv);
  const int64_t aoff = act_off[x];
  for (int j = tid; j < n; j += kTB) {
    sr[j] = sr_pool[aoff + j];
    pos[j] = pos_pool[aoff + j];
  }
  __syncthreads();
  // K_BN = K(BN, BN) - update, permuted to [s r] (skel_factorization.cpp:66-78)
  for (int e = tid; e < n * n; e += kTB) {
    int i = e % n, j = e / n;
    double v = kernel_entry(g, pde, sr[i], sr[j]);
    if (!leaf) v -= (double)Uact[pos[i] + (size_t)pos[j] * n];
    Kp[e] = (T)v;
  }
  __syncthreads();
  const T* Tm = slab;                       // k x r from the ID
  T* Lm = slab + (size_t)k * r;             // k x r
  T* Um = slab + 2 * (size_t)k * r;         // r x k
  T* Xlu = slab + 3 * (size_t)k * r;        // r x r
  T* Acc = Xlu + (size_t)r * r;             // k x k
  auto KP = [&](int i, int j) -> T { return Kp[i + (size_t)j * n]; };
  // K_r_sn = K_rs - T^T K_ss (r x k); K_sn_r = K_sr - K_ss T (k x r)
  for (int e = tid; e < r * k; e += kTB) {
    int i = e % r, j = e / r;
    T s = KP(k + i, j);
    for (int t = 0; t < k; ++t) s -= Tm[t + (size_t)i * k] * KP(t, j);
    Krsn[i + (size_t)j * r] = s;
  }
  for (int e = tid; e < k * r; e += kTB) {
    int i = e % k, j = e / k;
    T s = KP(i, k + j);
    for (int t = 0; t < k; ++t) s -= KP(i, t) * Tm[t + (size_t)j * k];
    Ksnr[i + (size_t)j * k] = s;
  }
  __syncthreads();
  // X_rr = K_rr - T^T K_sr - K_r_sn T (r x r)
  for (int e = tid; e < r * r; e += kTB) {
    int i = e % r, j = e / r;
    T s = KP(k + i, k + j);
    for (int t = 0; t < k; ++t) s -= Tm[t + (size_t)i * k] * KP(t, k + j);
    for (int t = 0; t < k; ++t) s -= Krsn[i + (size_t)t * r] * Tm[t + (size_t)j * k];
    X[i + (size_t)j * r] = s;
  }
  __syncthreads();
  // LU with partial pivoting (dgetrf): first largest |.| in the column
  const int lane = tid & (kWave - 1), wv = tid / kWave;
  __shared__ int s_p;
  for (int c = 0; c < r; ++c) {
    if (wv == 0) {
      double best = -1.0;
      int bi = r;
      for (int i = c + lane; i < r; i += kWave) {
        double v = fabs((double)X[i + (size_t)c * r]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      wave_argmax(best, bi);
      if (lane == 0) {
        s_p = bi;
        piv[c] = bi;
      }
    }
    __syncthreads();
    const int p = s_p;
    if (p != c) {
      for (int j = tid; j < r; j += kTB) {
        T t0 = X[c + (size_t)j * r];
        X[c + (size_t)j * r] = X[p + (size_t)j * r];
        X[p + (size_t)j * r] = t0;
      }
    }
    __syncthreads();
    const T d = X[c + (size_t)c * r];
    if (d == T(0)) {
      if (tid == 0) flag(status, 4);  // singular X_rr: the reference leaves it unchecked
    } else {
      const T inv = T(1) / d;
      for (int i = c + 1 + tid; i < r; i += kTB) X[i + (size_t)c * r] *= inv;
    }
    __syncthreads();
    const int rem = r - c - 1;
    for (int e = tid; e < rem * rem; e += kTB) {
      int i = c + 1 + e % rem, j = c + 1 + e / rem;
      X[i + (size_t)j * r] -= X[i + (size_t)c * r] * X[c + (size_t)j * r];
    }
    __syncthreads();
  }
  // L = K_sn_r X_rr^{-1}: rows z with X^T z^T = K_sn_r(row, :)^T (dgetrs 'T')
  for (int e = tid; e < k * r; e += kTB) Lm[e] = Ksnr[e];
  // U = X_rr^{-1} K_r_sn (dgetrs 'N')
  for (int e = tid; e < r * k; e += kTB) Um[e] = Krsn[e];
  __syncthreads();
  block_lu_solve<T, kNW>(X, r, piv, r, Lm, 1, k, k, true);
  block_lu_solve<T, kNW>(X, r, piv, r, Um, r, 1, k, false);
  __syncthreads();
  for (int e = tid; e < r * r; e += kTB) Xlu[e] = X[e];
  for (int c = tid; c < r; c += kTB) piv_pool[aoff + c] = piv[c];
  __syncthreads();
  // schur_update = L K_r_sn (k x k), plus the descendants' update on the skeleton
  for (int e = tid; e < k * k; e += kTB) {
    int i = e % k, j = e / k;
    T s = T(0);
    for (int t = 0; t < r; ++t) s += Lm[i + (size_t)t * k] * Krsn[t + (size_t)j * r];
    if (!leaf) s += Uact[pos[i] + (size_t)pos[j] * n];
    Acc[e] = s;
  }
  if (tid == 0) {
    osize[x] = k;
    acc_ptr[x] = reinterpret_cast<uint64_t>(Acc);
    acc_ld[x] = k;
  }
}

// ---------------------------------------------------------------------------
// root: allskel_mat = K(allskel, allskel) - updates, FP64 LU (skel_factorization.cpp:116-128)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(1024) void k_root(int x, GeomView g, int pde, const int* child,
                                               const int* osize, const uint64_t* acc_ptr,
                                               const int* acc_ld, const int* dofs, int n,
                                               double* A, int* piv, int* status) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < n * n; e += nt) {
    int i = e % n, j = e / n;
    A[e] = kernel_entry(g, pde, dofs[i], dofs[j]);
  }
  __syncthreads();
  int o = 0;
  for (int s = 0; s < 4; ++s) {
    int y = child[4 * x + s];
    if (y < 0) break;
    int sz = osize[y];
    const T* acc = reinterpret_cast<const T*>(acc_ptr[y]);
    if (acc) {
      int ld = acc_ld[y];
      for (int e = tid; e < sz * sz; e += nt) {
        int i = e % sz, j = e / sz;
        A[(o + i) + (size_t)(o + j) * n] -= (double)acc[i + (size_t)j * ld];
      }
    }
    o += sz;
  }
  __syncthreads();
  __shared__ int s_p;
  const int lane = tid & (kWave - 1), wv = tid / kWave;
  for (int c = 0; c < n; ++c) {
    if (wv == 0) {
      double best = -1.0;
      int bi = n;
      for (int i = c + lane; i < n; i += kWave) {
        double v = fabs(A[i + (size_t)c * n]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      wave_argmax(best, bi);
      if (lane == 0) {
        s_p = bi;
        piv[c] = bi;
      }
    }
    __syncthreads();
    const int p = s_p;
    if (p != c)
      for (int j = tid; j < n; j += nt) {
        double t0 = A[c + (size_t)j * n];
        A[c + (size_t)j * n] = A[p + (size_t)j * n];
        A[p + (size_t)j * n] = t0;
      }
    __syncthreads();
    const double d = A[c + (size_t)c * n];
    if (d == 0.0) {
      if (tid == 0) flag(status, 8);
    } else {
      const double inv = 1.0 / d;
      for (int i = c + 1 + tid; i < n; i += nt) A[i + (size_t)c * n] *= inv;
    }
    __syncthreads();
    const int rem = n - c - 1;
    for (int e = tid; e < rem * rem; e += nt) {
      int i = c + 1 + e % rem, j = c + 1 + e / rem;
      A[i + (size_t)j * n] -= A[i + (size_t)c * n] * A[c + (size_t)j * n];
    }
    __syncthreads();
  }
}

__global__ void k_add_i64(int64_t* p, int n, int64_t add) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] += add;
}

__global__ void k_set_ptr(uint64_t* dst, int a, int b, const int* nact, uint64_t base,
                          const int64_t* offs, size_t tsz) {
  int x = a + blockIdx.x * blockDim.x + threadIdx.x;
  if (x < b) dst[x] = base + (uint64_t)offs[x - a] * tsz;
}

}  // namespace

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------

template <typename T>
static void factor_levels(Context& c, TreeData& t, const kigpu_factor_opts& o) {
  hipStream_t s = c.stream;
  const int nn = t.nnodes;
  const GeomView g{t.px.p, t.py.p, t.nx.p, t.ny.p, t.w.p, t.kappa.p, t.npts, t.sd};
  const int pde = o.pde;
  const int prow = pde == KIGPU_STOKES ? 4 * kProxyPoints : 2 * kProxyPoints;

  // leaves' original_box DOF lists
  DBuf<int> cnt;
  cnt.alloc(nn);
  t.obox_off.alloc(nn);
  k_leaf_dofs<<<ceil_div(nn, kTB), kTB, 0, s>>>(nn, t.is_leaf.p, t.pcount.p, t.sd, cnt.p);
  exclusive_scan_i32_to_i64(cnt.p, t.obox_off.p, nn, s);
  t.obox_pool.alloc(t.ndof);
  k_obox<<<nn, 64, 0, s>>>(nn, t.is_leaf.p, t.pcount.p, t.pstart.p, t.perm_all.p, t.obox_off.p,
                           t.sd, t.obox_pool.p);

  t.act_off.alloc(nn);
  t.nact.alloc(nn);
  t.nrows.alloc(nn);
  t.rank.alloc(nn);
  t.osize.alloc(nn);
  t.fac_ptr.alloc(nn);
  t.acc_ptr.alloc(nn);
  t.acc_ld.alloc(nn);
  t.fac_ptr.zero(s);
  t.acc_ptr.zero(s);
  t.osize.zero(s);
  t.status.alloc(1);
  t.status.zero(s);
  const int64_t pool = (int64_t)t.ndof * t.nlevels + 64;
  t.act_pool.alloc(pool);
  t.sr_pool.alloc(pool);
  t.pos_pool.alloc(pool);
  t.piv_pool.alloc(pool);
  t.act_pool_size = pool;
  t.work.clear();
  t.work.resize(t.nlevels);
  DBuf<int64_t> loc;
  loc.alloc(nn + 1);
  DBuf<IdWork> dwork;
  std::vector<int> h_n, h_q;
  std::vector<int64_t> h_off;
  const size_t lds_cap = c.lds_max;
  int64_t pool_base = 0;

  for (int l = t.nlevels - 1; l >= 0; --l) {
    const int a = (int)t.level_off[l], b = (int)t.level_off[l + 1], cn = b - a;
    auto lw = std::make_unique<LevelWork>();
    lw->level = l;
    lw->a = a;
    lw->b = b;
    hipEventRecord(c.ev[0], s);
    // ---- active lists
    k_active_size<<<ceil_div(cn, kTB), kTB, 0, s>>>(a, b, t.is_leaf.p, t.child.p, t.pcount.p,
                                                   t.osize.p, t.sd, t.nact.p, t.rank.p);
    exclusive_scan_i32_to_i64(t.nact.p + a, t.act_off.p + a, cn, s);
    k_add_i64<<<ceil_div(cn, kTB), kTB, 0, s>>>(t.act_off.p + a, cn, pool_base);
    k_active_fill<<<cn, 64, 0, s>>>(a, t.is_leaf.p, t.child.p, t.obox_off.p, t.obox_pool.p,
                                    t.osize.p, t.rank.p, t.act_off.p, t.nact.p, t.act_pool.p,
                                    t.sr_pool.p);
    if (l == 0) {
      hipEventRecord(c.ev[1], s);
      break;
    }
    if (l >= 3)
      k_inner_rows<<<cn, kWave, 0, s>>>(a, b, o.min_dofs, g, t.is_leaf.p, t.obox_off.p,
                                        t.obox_pool.p, t.pcount.p, t.act_off.p, t.nact.p,
                                        t.act_pool.p, t.sl_nbr.p, t.sl_cnt.p, t.cp_off.p,
                                        t.cp_nbr.p, t.corners.p, t.side.p, t.nrows.p, nullptr,
                                        nullptr, 0);
    h_n.resize(cn);
    h_q.assign(cn, 0);
    KI_HIP(hipMemcpyAsync(h_n.data(), t.nact.p + a, sizeof(int) * cn, hipMemcpyDeviceToHost, s));
    if (l >= 3)
      KI_HIP(hipMemcpyAsync(h_q.data(), t.nrows.p + a, sizeof(int) * cn, hipMemcpyDeviceToHost,
                            s));
    KI_HIP(hipStreamSynchronize(s));
    ++c.tim.syncs;
    if (l <= 2) {
      int64_t tot = 0;
      for (int i = 0; i < cn; ++i) tot += h_n[i];
      int64_t extra = 0;
      if (l == 2) {
        for (int y = (int)t.level_off[1]; y < (int)t.level_off[2]; ++y)
          if (t.h_is_leaf[y]) extra += (int64_t)t.h_pcount[y] * t.sd;
      }
      for (int i = 0; i < cn; ++i) h_q[i] = (int)(tot - h_n[i] + extra);
    }
    // ---- work list: boxes with >= min_dofs active DOFs and a non-empty ID matrix
    std::vector<IdWork> wl;
    std::vector<int64_t> rows_off(cn, 0);
    int64_t rtot = 0, atot = 0;
    int nmax = 0;
    for (int i = 0; i < cn; ++i) {
      nmax = std::max(nmax, h_n[i]);
      if (h_n[i] < o.min_dofs) continue;
      const int x = a + i;
      IdWork w{};
      w.node = x;
      w.n = h_n[i];
      w.q = h_q[i];
      w.proxy = l >= 3;
      w.m = 2 * w.q + (w.proxy ? prow : 0);
      if (w.m == 0) continue;
      w.rows_off = rtot;
      rows_off[i] = rtot;
      rtot += w.q;
      const double* cc = &t.h_corners[8 * (size_t)x];
      w.cx = cc[0] + t.h_side[x] / 2.0;
      w.cy = cc[1] + t.h_side[x] / 2.0;
      w.rad = t.h_side[x] * kRadiusRatio;
      w.pw = 2.0 * M_PI * w.rad / kProxyPoints;
      IdLds L = id_lds_layout(w.m, w.n, w.q, true, sizeof(double));
      w.in_lds = L.total <= lds_cap;
      w.a_off = atot;
      if (!w.in_lds) atot += (int64_t)w.m * w.n;
      wl.push_back(w);
    }
    lw->nmax = nmax;
    // ---- factor slabs for every node of the level
    std::vector<int64_t> soff(cn);
    int64_t stot = 0;
    for (int i = 0; i < cn; ++i) {
      soff[i] = stot;
      stot += (int64_t)slab_elems(h_n[i]);
    }
    DBuf<int64_t> dsoff;
    dsoff.upload(soff, s);
    if (sizeof(T) == 8) {
      lw->slab64.alloc(stot);
      k_set_ptr<<<ceil_div(cn, kTB), kTB, 0, s>>>(t.fac_ptr.p, a, b, t.nact.p,
                                                 reinterpret_cast<uint64_t>(lw->slab64.p),
                                                 dsoff.p, sizeof(T));
    } else {
      lw->slab32.alloc(stot);
      k_set_ptr<<<ceil_div(cn, kTB), kTB, 0, s>>>(t.fac_ptr.p, a, b, t.nact.p,
                                                 reinterpret_cast<uint64_t>(lw->slab32.p),
                                                 dsoff.p, sizeof(T));
    }
    // ---- ID rows
    t.rows_pool.alloc(std::max<int64_t>(rtot, 1));
    if (!wl.empty()) {
      dwork.upload(wl, s);
      if (l >= 3) {
        DBuf<int64_t> droff;
        droff.upload(rows_off, s);
        // rows_off is indexed by node - a inside the kernel through this shifted pointer
        k_inner_rows<<<cn, kWave, 0, s>>>(a, b, o.min_dofs, g, t.is_leaf.p, t.obox_off.p,
                                          t.obox_pool.p, t.pcount.p, t.act_off.p, t.nact.p,
                                          t.act_pool.p, t.sl_nbr.p, t.sl_cnt.p, t.cp_off.p,
                                          t.cp_nbr.p, t.corners.p, t.side.p, t.nrows.p,
                                          droff.p - a, t.rows_pool.p, 1);
        KI_HIP(hipStreamSynchronize(s));
      } else {
        k_outside_rows<<<(int)wl.size(), kTB, 0, s>>>(
            dwork.p, a, b, l, (int)t.level_off[1], (int)t.level_off[2], t.is_leaf.p,
            t.obox_off.p, t.obox_pool.p, t.pcount.p, t.sd, t.act_off.p, t.nact.p, t.act_pool.p,
            t.rows_pool.p);
      }
    }
    hipEventRecord(c.ev[1], s);
    // ---- ID
    if (!wl.empty()) {
      std::vector<IdWork> wl_lds, wl_glb;
      size_t lds_bytes = 0, glds = 0;
      for (auto& w : wl) {
        if (w.in_lds) {
          wl_lds.push_back(w);
          lds_bytes = std::max(lds_bytes, id_lds_layout(w.m, w.n, w.q, true, sizeof(double)).total);
        } else {
          wl_glb.push_back(w);
          glds = std::max(glds, id_lds_layout(w.m, w.n, w.q, false, sizeof(double)).total);
        }
      }
      double* ws = nullptr;
      if (atot) {
        t.ws64.alloc(atot);
        ws = t.ws64.p;
      }
      DBuf<IdWork> d1, d2;
      if (!wl_lds.empty()) {
        d1.upload(wl_lds, s);
        allow_lds(k_id<double, T, true>, lds_bytes);
        KI_HIP(hipEventRecord(c.ev[12], s));
        k_id<double, T, true><<<(int)wl_lds.size(), kTB, lds_bytes, s>>>(
            d1.p, g, pde, o.id_tol, t.rows_pool.p, t.act_pool.p, t.act_off.p, c.ring.p, ws,
            t.rank.p, t.sr_pool.p, t.pos_pool.p, t.fac_ptr.p, t.status.p);
        KI_HIP(hipGetLastError());
        KI_HIP(hipEventRecord(c.ev[13], s));
        KI_HIP(hipEventSynchronize(c.ev[13]));
        float kms = 0;
        KI_HIP(hipEventElapsedTime(&kms, c.ev[12], c.ev[13]));
        c.tim.id_kernel_s += kms * 1e-3;
        c.tim.id_launches += 1;
        lw->id_lds = wl_lds;
      }
      if (!wl_glb.empty()) {
        d2.upload(wl_glb, s);
        allow_lds(k_id<double, T, false>, glds);
        k_id<double, T, false><<<(int)wl_glb.size(), kTB, glds, s>>>(
            d2.p, g, pde, o.id_tol, t.rows_pool.p, t.act_pool.p, t.act_off.p, c.ring.p, ws,
            t.rank.p, t.sr_pool.p, t.pos_pool.p, t.fac_ptr.p, t.status.p);
        KI_HIP(hipGetLastError());
      }
      KI_HIP(hipStreamSynchronize(s));
    }
    hipEventRecord(c.ev[2], s);
    // ---- elimination for every node of the level
    {
      DecLds L = dec_layout(std::max(nmax, 1), sizeof(T));
      KI_HIP(hipEventRecord(c.ev[14], s));
      if (L.total <= lds_cap) {
        allow_lds(k_decouple<T, true>, L.total);
        k_decouple<T, true><<<cn, kTB, L.total, s>>>(
            a, g, pde, t.is_leaf.p, t.child.p, t.osize.p, t.acc_ptr.p, t.acc_ld.p, t.nact.p,
            t.rank.p, t.act_off.p, t.act_pool.p, t.sr_pool.p, t.pos_pool.p, t.piv_pool.p,
            t.fac_ptr.p, t.osize.p, t.acc_ptr.p, t.acc_ld.p, (T*)nullptr, 0, t.status.p);
      } else {
        size_t stride = (L.total + sizeof(T) - 1) / sizeof(T);
        T* ws;
        if (sizeof(T) == 8) {
          t.ws64.alloc(stride * cn);
          ws = reinterpret_cast<T*>(t.ws64.p);
        } else {
          t.ws32.alloc(stride * cn);
          ws = reinterpret_cast<T*>(t.ws32.p);
        }
        k_decouple<T, false><<<cn, kTB, 0, s>>>(
            a, g, pde, t.is_leaf.p, t.child.p, t.osize.p, t.acc_ptr.p, t.acc_ld.p, t.nact.p,
            t.rank.p, t.act_off.p, t.act_pool.p, t.sr_pool.p, t.pos_pool.p, t.piv_pool.p,
            t.fac_ptr.p, t.osize.p, t.acc_ptr.p, t.acc_ld.p, ws, stride, t.status.p);
      }
      KI_HIP(hipGetLastError());
      KI_HIP(hipEventRecord(c.ev[15], s));
      KI_HIP(hipEventSynchronize(c.ev[15]));
      float kms = 0;
      KI_HIP(hipEventElapsedTime(&kms, c.ev[14], c.ev[15]));
      c.tim.dec_kernel_s += kms * 1e-3;
      c.tim.dec_launches += 1;
    }
    hipEventRecord(c.ev[3], s);
    KI_HIP(hipStreamSynchronize(s));
    float ms = 0;
    hipEventElapsedTime(&ms, c.ev[0], c.ev[1]);
    c.tim.t_lists += ms * 1e-3;
    hipEventElapsedTime(&ms, c.ev[1], c.ev[2]);
    c.tim.t_id += ms * 1e-3;
    hipEventElapsedTime(&ms, c.ev[2], c.ev[3]);
    c.tim.t_decouple += ms * 1e-3;
    pool_base += 0;
    (void)h_off;
    t.work[l] = std::move(lw);
    // next level's lists start after this level's
    int64_t lastoff = 0;
    KI_HIP(hipMemcpyAsync(&lastoff, t.act_off.p + b - 1, sizeof(int64_t), hipMemcpyDeviceToHost,
                          s));
    KI_HIP(hipStreamSynchronize(s));
    pool_base = lastoff + h_n[cn - 1];
  }
  (void)loc;

  // ---- root (level 0)
  hipEventRecord(c.ev[4], s);
  const int root = 0;
  int n0 = 0;
  int64_t r_off = 0;
  KI_HIP(hipMemcpyAsync(&n0, t.nact.p + root, sizeof(int), hipMemcpyDeviceToHost, s));
  KI_HIP(hipMemcpyAsync(&r_off, t.act_off.p + root, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  KI_HIP(hipStreamSynchronize(s));
  t.nroot = n0;
  if (n0 > 0) {
    t.root_dofs.alloc(n0);
    KI_HIP(hipMemcpyAsync(t.root_dofs.p, t.act_pool.p + r_off, sizeof(int) * n0,
                          hipMemcpyDeviceToDevice, s));
    t.root_lu.alloc((size_t)n0 * n0);
    t.root_piv.alloc(n0);
    k_root<T><<<1, 1024, 0, s>>>(root, g, pde, t.child.p, t.osize.p, t.acc_ptr.p, t.acc_ld.p,
                                 t.root_dofs.p, n0, t.root_lu.p, t.root_piv.p, t.status.p);
    KI_HIP(hipGetLastError());
  }
  hipEventRecord(c.ev[5], s);
  KI_HIP(hipStreamSynchronize(s));
  float ms = 0;
  hipEventElapsedTime(&ms, c.ev[4], c.ev[5]);
  c.tim.t_root += ms * 1e-3;
}

// Algorithmic flops of the ID kernel for one box, as executed: entry
// generation (c_gen flops per entry), initial column norms, the QR steps up to
// the rank-revealing one (all n steps when the box is not compressed) and the
// T = R11^{-1} R12 solve.
static double id_flops(int m, int n, int k, double c_gen) {
  double f = c_gen * m * n + 2.0 * m * n;
  const int steps = k > 0 ? k + 1 : n;
  for (int i = 0; i < steps; ++i) {
    f += 2.0 * (m - i);
    if (i + 1 < steps) f += 4.0 * (m - i) * (n - i - 1);
  }
  if (k > 0) f += (double)k * k * (n - k);
  return f;
}

// elimination: K_BN generation plus 6k^2r + 8kr^2 + 2/3 r^3
static double dec_flops(int n, int k, double c_gen) {
  if (k <= 0) return 0;
  const double r = n - k;
  return c_gen * n * n + 6.0 * k * k * r + 8.0 * k * r * r + 2.0 / 3.0 * r * r * r;
}

void factor_tree(Context& c, TreeData& t, const kigpu_factor_opts& o) {
  c.tim.t_lists = c.tim.t_id = c.tim.t_decouple = c.tim.t_root = 0;
  c.tim.id_launches = c.tim.dec_launches = 0;
  c.tim.id_kernel_s = c.tim.id_flops = c.tim.dec_kernel_s = c.tim.dec_flops = 0;
  if (o.precision == KIGPU_FP32_REFINE) factor_levels<float>(c, t, o);
  else factor_levels<double>(c, t, o);
  int st = 0;
  t.status.download(&st, 1, c.stream);
  KI_HIP(hipStreamSynchronize(c.stream));
  if (st & 2) KI_FAIL(KIGPU_ERR_RANK, "ID matrix with fewer rows than columns");
  // host copies of the per-node results
  const int nn = t.nnodes;
  t.h_rank.resize(nn);
  t.h_nact.resize(nn);
  t.h_act_off.resize(nn);
  t.rank.download(t.h_rank.data(), nn, c.stream);
  t.nact.download(t.h_nact.data(), nn, c.stream);
  t.act_off.download(t.h_act_off.data(), nn, c.stream);
  t.h_sr.resize(t.act_pool_size);
  t.h_act.resize(t.act_pool_size);
  t.sr_pool.download(t.h_sr.data(), t.act_pool_size, c.stream);
  t.act_pool.download(t.h_act.data(), t.act_pool_size, c.stream);
  KI_HIP(hipStreamSynchronize(c.stream));
  const double c_gen = o.pde == KIGPU_STOKES ? 20.0 : 11.0;
  for (auto& lw : t.work) {
    if (!lw) continue;
    for (const IdWork& w : lw->id_lds) c.tim.id_flops += id_flops(w.m, w.n, t.h_rank[w.node], c_gen);
    for (int x = lw->a; x < lw->b; ++x) c.tim.dec_flops += dec_flops(t.h_nact[x], t.h_rank[x], c_gen);
  }
  t.opts = o;
  t.factored = true;
}

}  // namespace kigpu
