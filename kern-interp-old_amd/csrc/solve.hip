// Solve sweeps of the skeletonized factorization (skel_factorization.cpp:
// 341-403), the FP64 matrix-free kernel apply used for iterative refinement
// and the boundary -> domain evaluation (kernel.cpp:337-357).
//
// The upward pass applies, per compressed box and level (finest first),
//   x_r -= T^T x_s ;  x_s -= L x_r ;  x_r <- X_rr^{-1} x_r
// where the third step is the reference's separate D^{-1} pass fused in: a
// box's redundant DOFs are touched by no later (coarser) box, so the order of
// the operations on each DOF is unchanged.  Then the root solve and the
// downward pass, levels coarse to fine:  x_r -= U x_s ;  x_s -= T x_r.
//
// The transposed solve (adjoint, x = F^{-T} b for the same factorization F)
// reverses the product: per box, finest level first,
//   x_r -= T^T x_s ;  x_s -= U^T x_r ;  x_r <- X_rr^{-T} x_r
// then allskel_mat^{-T} on the root skeleton, then coarse to fine
//   x_r -= L^T x_s ;  x_s -= T x_r.
#include <algorithm>
#include <cmath>

#include <mutex>

#include "device_util.h"
#include "entries.h"
#include "factor.h"

namespace kigpu {

namespace {

constexpr int kTB = 256;
constexpr int kNW = kTB / kWave;
constexpr int kRC = 16;  // right-hand sides per block

// The threads that work on one node: the whole block, or one wavefront
// (four children of a node side by side)
struct BlockGroup {
  int tid;
  static constexpr int nt = kTB;
  __device__ void sync() const { __syncthreads(); }
};
struct WaveGroup {
  int tid;
  static constexpr int nt = kWave;
  __device__ void sync() const { wave_sync(); }
};

// Factor records the sweeps read, per node
struct SweepArgs {
  const int* rank;
  const int* nact;
  const int64_t* act_off;
  const int* sr_pool;
  const uint64_t* fac_ptr;
};

// One node of the upward pass on right-hand sides [c0, c0 + nc); v: LDS for
// 2 n nc values.
// mode 0: x_r -= T^T x_s; x_s -= L x_r        (sweeps of U in skeletonize, :169-185)
// mode 1: the same, then x_r <- X_rr^{-1} x_r   (solve's upward pass with D^{-1})
// mode 2: x_r -= T^T x_s; x_s -= U^T x_r       (sweeps of Psi^T, :187-204)
// mode 3: mode 2, then x_r <- X_rr^{-T} x_r     (transposed solve's upward pass)
template <typename T, typename G>
__device__ __forceinline__ void up_node(const G& g, const SweepArgs& f, int node, T* x, int ldx,
                                        int c0, int nc, int mode, T* v) {
  const int k = f.rank[node];
  if (k < 0) return;
  const int n = f.nact[node], r = n - k;
  const int* dofs = f.sr_pool + f.act_off[node];
  const T* Tm = reinterpret_cast<const T*>(f.fac_ptr[node]);
  const T* Lm = Tm + (size_t)k * r;
  const T* Xinv = Tm + 3 * (size_t)k * r;
  T* tmp = v + (size_t)n * nc;
  const int tid = g.tid;
  for (int e = tid; e < n * nc; e += G::nt) {
    int i = e % n, c = e / n;
    v[e] = x[dofs[i] + (size_t)(c0 + c) * ldx];
  }
  g.sync();
  for (int e = tid; e < r * nc; e += G::nt) {  // x_r -= T^T x_s
    int i = e % r, c = e / r;
    const T* vs = v + (size_t)c * n;
    T s = vs[k + i];
    for (int t = 0; t < k; ++t) s -= Tm[t + (size_t)i * k] * vs[t];
    v[k + i + (size_t)c * n] = s;
  }
  g.sync();
  if (mode >= 2) {
    const T* Um = Tm + 2 * (size_t)k * r;  // r x k
    for (int e = tid; e < k * nc; e += G::nt) {  // x_s -= U^T x_r
      int t = e % k, c = e / k;
      const T* vs = v + (size_t)c * n;
      T s = vs[t];
      for (int i = 0; i < r; ++i) s -= Um[i + (size_t)t * r] * vs[k + i];
      v[t + (size_t)c * n] = s;
    }
  } else {
    for (int e = tid; e < k * nc; e += G::nt) {  // x_s -= L x_r
      int t = e % k, c = e / k;
      const T* vs = v + (size_t)c * n;
      T s = vs[t];
      for (int i = 0; i < r; ++i) s -= Lm[t + (size_t)i * k] * vs[k + i];
      v[t + (size_t)c * n] = s;
    }
  }
  for (int e = tid; e < r * nc; e += G::nt) {  // x_r <- X_rr^{-1} x_r (mode 1, ^{-T} mode 3)
    int i = e % r, c = e / r;
    const T* vs = v + (size_t)c * n + k;
    T s;
    if (mode == 1) {
      s = T(0);
      for (int j = 0; j < r; ++j) s += Xinv[i + (size_t)j * r] * vs[j];
    } else if (mode == 3) {
      s = T(0);
      for (int j = 0; j < r; ++j) s += Xinv[j + (size_t)i * r] * vs[j];
    } else {
      s = vs[i];
    }
    tmp[e] = s;
  }
  g.sync();
  for (int e = tid; e < k * nc; e += G::nt) {
    int t = e % k, c = e / k;
    x[dofs[t] + (size_t)(c0 + c) * ldx] = v[t + (size_t)c * n];
  }
  for (int e = tid; e < r * nc; e += G::nt) {
    int i = e % r, c = e / r;
    x[dofs[k + i] + (size_t)(c0 + c) * ldx] = tmp[e];
  }
}

// One node of the downward pass; v: LDS for n nc values.
// trans = 0: x_r -= U x_s; x_s -= T x_r.  trans = 1: x_r -= L^T x_s; x_s -= T x_r.
template <typename T, typename G>
__device__ __forceinline__ void down_node(const G& g, const SweepArgs& f, int node, T* x, int ldx,
                                          int c0, int nc, int trans, T* v) {
  const int k = f.rank[node];
  if (k < 0) return;
  const int n = f.nact[node], r = n - k;
  const int* dofs = f.sr_pool + f.act_off[node];
  const T* Tm = reinterpret_cast<const T*>(f.fac_ptr[node]);
  const T* Lm = Tm + (size_t)k * r;
  const T* Um = Tm + 2 * (size_t)k * r;
  const int tid = g.tid;
  for (int e = tid; e < n * nc; e += G::nt) {
    int i = e % n, c = e / n;
    v[e] = x[dofs[i] + (size_t)(c0 + c) * ldx];
  }
  g.sync();
  for (int e = tid; e < r * nc; e += G::nt) {  // x_r -= U x_s  (or L^T x_s)
    int i = e % r, c = e / r;
    const T* vs = v + (size_t)c * n;
    T s = vs[k + i];
    if (trans)
      for (int t = 0; t < k; ++t) s -= Lm[t + (size_t)i * k] * vs[t];
    else
      for (int t = 0; t < k; ++t) s -= Um[i + (size_t)t * r] * vs[t];
    v[k + i + (size_t)c * n] = s;
  }
  g.sync();
  for (int e = tid; e < k * nc; e += G::nt) {  // x_s -= T x_r
    int t = e % k, c = e / k;
    const T* vs = v + (size_t)c * n;
    T s = vs[t];
    for (int i = 0; i < r; ++i) s -= Tm[t + (size_t)i * k] * vs[k + i];
    v[t + (size_t)c * n] = s;
  }
  g.sync();
  for (int e = tid; e < r * nc; e += G::nt) {  // only the redundant part changed last
    int i = e % r, c = e / r;
    x[dofs[k + i] + (size_t)(c0 + c) * ldx] = v[k + i + (size_t)c * n];
  }
  for (int e = tid; e < k * nc; e += G::nt) {
    int t = e % k, c = e / k;
    x[dofs[t] + (size_t)(c0 + c) * ldx] = v[t + (size_t)c * n];
  }
}

// one level: a block per node of [a, a + gridDim.x), kRC right-hand sides
// per blockIdx.y
template <typename T>
__global__ __launch_bounds__(kTB) void k_up(int a, SweepArgs f, T* x, int ldx, int nrhs,
                                            int mode) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int c0 = blockIdx.y * kRC;
  up_node<T>(BlockGroup{(int)threadIdx.x}, f, a + blockIdx.x, x, ldx, c0, min(kRC, nrhs - c0),
             mode, reinterpret_cast<T*>(smem));
}

template <typename T>
__global__ __launch_bounds__(kTB) void k_down(int a, SweepArgs f, T* x, int ldx, int nrhs,
                                              int trans) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int c0 = blockIdx.y * kRC;
  down_node<T>(BlockGroup{(int)threadIdx.x}, f, a + blockIdx.x, x, ldx, c0, min(kRC, nrhs - c0),
               trans, reinterpret_cast<T*>(smem));
}

// Two levels in one launch: a block per node of the coarser level [a, ...),
// its four children one wavefront each (LDS regions of `wstride` values),
// then the node itself with the whole block.  A node's sweep touches only
// its own active DOFs, which lie among its children's, so a block needs no
// other block's results.  Upward: children first; downward: node first.
static_assert(kTB == 4 * kWave, "one wavefront per child");
template <typename T>
__global__ __launch_bounds__(kTB) void k_up_pair(int a, const int* child, SweepArgs f, T* x,
                                                 int ldx, int nrhs, int mode, int wstride) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* lds = reinterpret_cast<T*>(smem);
  const int node = a + blockIdx.x;
  const int c0 = blockIdx.y * kRC, nc = min(kRC, nrhs - c0);
  const int w = threadIdx.x / kWave;
  const int ch = child[4 * node + w];
  if (ch >= 0)
    up_node<T>(WaveGroup{(int)threadIdx.x % kWave}, f, ch, x, ldx, c0, nc, mode,
               lds + (size_t)w * wstride);
  __syncthreads();
  up_node<T>(BlockGroup{(int)threadIdx.x}, f, node, x, ldx, c0, nc, mode, lds);
}

template <typename T>
__global__ __launch_bounds__(kTB) void k_down_pair(int a, const int* child, SweepArgs f, T* x,
                                                   int ldx, int nrhs, int trans, int wstride) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* lds = reinterpret_cast<T*>(smem);
  const int node = a + blockIdx.x;
  const int c0 = blockIdx.y * kRC, nc = min(kRC, nrhs - c0);
  down_node<T>(BlockGroup{(int)threadIdx.x}, f, node, x, ldx, c0, nc, trans, lds);
  __syncthreads();
  const int w = threadIdx.x / kWave;
  const int ch = child[4 * node + w];
  if (ch >= 0)
    down_node<T>(WaveGroup{(int)threadIdx.x % kWave}, f, ch, x, ldx, c0, nc, trans,
                 lds + (size_t)w * wstride);
}

// x[allskel] <- allskel_mat^{-1} x[allskel] (trans: ^{-T}) in FP64 with the
// explicit inverse; one block per (right-hand side, instance), one thread
// per row.
template <typename T>
__global__ __launch_bounds__(kTB) void k_root_apply(const int* root_n, const int64_t* root_doff,
                                                    const int64_t* root_moff, const int* dofs_all,
                                                    const double* Ainv_all, T* x, int ldx,
                                                    int trans) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.y;
  const int n = root_n[b];
  if (n == 0) return;
  const int* dofs = dofs_all + root_doff[b];
  const double* Ainv = Ainv_all + root_moff[b];
  double* v = reinterpret_cast<double*>(smem);
  T* xc = x + (size_t)blockIdx.x * ldx;
  for (int i = threadIdx.x; i < n; i += kTB) v[i] = (double)xc[dofs[i]];
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kTB) {
    double s = 0;
    if (trans)
      for (int j = 0; j < n; ++j) s += Ainv[j + (size_t)i * n] * v[j];
    else
      for (int j = 0; j < n; ++j) s += Ainv[i + (size_t)j * n] * v[j];
    xc[dofs[i]] = (T)s;
  }
}

// ---- FP64 matrix-free apply -----------------------------------------------
constexpr int kAT = 256;  // targets per block = sources per tile
constexpr int kAR = 4;    // right-hand sides per pass

struct Targets {
  const double* x;
  const double* y;
  const double* nx;  // null for domain targets (forward)
  const double* ny;
  int n;             // target points
};

// Block-diagonal interaction of independent instances.  Tile b of kAT "own"
// points is [t0[b], t1[b]) of instance inst[b]; the "other" points it meets
// are [opt[inst[b]], opt[inst[b] + 1]), cut into gridDim.y equal splits
// (rounded to kAT).  A single instance is one segment.
struct Tiles {
  const int* inst;
  const int* t0;
  const int* t1;
  const int64_t* opt;
};

__device__ __forceinline__ void split_range(const Tiles& tl, int* o0, int* o1) {
  const int b = tl.inst[blockIdx.x];
  const int a = (int)tl.opt[b], e = (int)tl.opt[b + 1];
  int per = (e - a + (int)gridDim.y - 1) / (int)gridDim.y;
  per = (per + kAT - 1) / kAT * kAT;
  *o0 = min(e, a + (int)blockIdx.y * per);
  *o1 = min(e, *o0 + per);
}

// y_part[split] = sum over this split's sources; one thread per target point
template <int PDE, bool FWD>
__global__ __launch_bounds__(kAT) void k_apply(Targets tg, Tiles tl, GeomView g, const double* mu,
                                               int ldm, int nrhs, double* part, int ldp) {
#pragma clang fp contract(off)
  __shared__ double sx[kAT], sy[kAT], snx[kAT], sny[kAT], sw[kAT], sk[kAT];
  __shared__ double smu[kAT * 2 * kAR];
  constexpr int SD = PDE == KIGPU_STOKES ? 2 : 1;
  const int ti = tl.t0[blockIdx.x] + threadIdx.x;
  const bool valid = ti < tl.t1[blockIdx.x];
  const double tx = valid ? tg.x[ti] : 0, ty = valid ? tg.y[ti] : 0;
  const double tnx = (valid && !FWD) ? tg.nx[ti] : 0, tny = (valid && !FWD) ? tg.ny[ti] : 0;
  int s0, s1;
  split_range(tl, &s0, &s1);
  for (int c0 = 0; c0 < nrhs; c0 += kAR) {
    const int nc = min(kAR, nrhs - c0);
    double acc[kAR][SD];
#pragma unroll
    for (int c = 0; c < kAR; ++c)
#pragma unroll
      for (int a = 0; a < SD; ++a) acc[c][a] = 0;
    for (int b0 = s0; b0 < s1; b0 += kAT) {
      const int nb = min(kAT, s1 - b0);
      __syncthreads();
      if (threadIdx.x < nb) {
        int sp = b0 + threadIdx.x;
        sx[threadIdx.x] = g.px[sp];
        sy[threadIdx.x] = g.py[sp];
        snx[threadIdx.x] = g.nx[sp];
        sny[threadIdx.x] = g.ny[sp];
        sw[threadIdx.x] = g.w[sp];
        sk[threadIdx.x] = g.kappa[sp];
        for (int c = 0; c < nc; ++c)
          for (int b = 0; b < SD; ++b)
            smu[(threadIdx.x * SD + b) * kAR + c] = mu[SD * sp + b + (size_t)(c0 + c) * ldm];
      }
      __syncthreads();
      if (!valid) continue;
      for (int j = 0; j < nb; ++j) {
        const double r1 = tx - sx[j], r2 = ty - sy[j];
        const double w = sw[j];
        if (PDE != KIGPU_STOKES) {
          double kv;
          if (FWD && PDE == KIGPU_LAPLACE_NEUMANN) {
            kv = w * kInv2Pi * log(sqrt(r1 * r1 + r2 * r2));
          } else {
            kv = laplace_entry(PDE, r1, r2, w, sk[j], snx[j], sny[j], tnx, tny, FWD);
          }
#pragma unroll
          for (int c = 0; c < kAR; ++c)
            if (c < nc) acc[c][0] += kv * smu[j * kAR + c];
        } else {
#pragma unroll
          for (int ta = 0; ta < SD; ++ta)
#pragma unroll
            for (int sb = 0; sb < SD; ++sb) {
              double kv = stokes_entry(ta, sb, tx, ty, tnx, tny, sx[j], sy[j], snx[j], sny[j], w,
                                       sk[j]);
#pragma unroll
              for (int c = 0; c < kAR; ++c)
                if (c < nc) acc[c][ta] += kv * smu[(j * SD + sb) * kAR + c];
            }
        }
      }
    }
    if (valid) {
      for (int c = 0; c < nc; ++c)
        for (int a = 0; a < SD; ++a)
          part[(size_t)blockIdx.y * ldp * nrhs + SD * ti + a + (size_t)(c0 + c) * ldp] =
              acc[c][a];
    }
  }
}

// Transposed apply: y_part[split](i) = sum over this split's rows j of
// K(j, i) lam(j), one thread per column (source point) i.  Rows are the
// boundary points (K^T) or domain targets without normals (forward^T).
template <int PDE, bool FWD>
__global__ __launch_bounds__(kAT) void k_apply_t(Targets rows, Tiles tl, GeomView g,
                                                 const double* lam, int ldl, int nrhs,
                                                 double* part, int ldp) {
#pragma clang fp contract(off)
  __shared__ double sx[kAT], sy[kAT], snx[kAT], sny[kAT];
  __shared__ double slam[kAT * 2 * kAR];
  constexpr int SD = PDE == KIGPU_STOKES ? 2 : 1;
  const int si = tl.t0[blockIdx.x] + threadIdx.x;
  const bool valid = si < tl.t1[blockIdx.x];
  const double ox = valid ? g.px[si] : 0, oy = valid ? g.py[si] : 0;
  const double onx = valid ? g.nx[si] : 0, ony = valid ? g.ny[si] : 0;
  const double ow = valid ? g.w[si] : 0, okap = valid ? g.kappa[si] : 0;
  int j0, j1;
  split_range(tl, &j0, &j1);
  for (int c0 = 0; c0 < nrhs; c0 += kAR) {
    const int nc = min(kAR, nrhs - c0);
    double acc[kAR][SD];
#pragma unroll
    for (int c = 0; c < kAR; ++c)
#pragma unroll
      for (int a = 0; a < SD; ++a) acc[c][a] = 0;
    for (int b0 = j0; b0 < j1; b0 += kAT) {
      const int nb = min(kAT, j1 - b0);
      __syncthreads();
      if (threadIdx.x < nb) {
        int rp = b0 + threadIdx.x;
        sx[threadIdx.x] = rows.x[rp];
        sy[threadIdx.x] = rows.y[rp];
        snx[threadIdx.x] = rows.nx ? rows.nx[rp] : 0.0;
        sny[threadIdx.x] = rows.ny ? rows.ny[rp] : 0.0;
        for (int c = 0; c < nc; ++c)
          for (int b = 0; b < SD; ++b)
            slam[(threadIdx.x * SD + b) * kAR + c] = lam[SD * rp + b + (size_t)(c0 + c) * ldl];
      }
      __syncthreads();
      if (!valid) continue;
      for (int j = 0; j < nb; ++j) {
        const double r1 = sx[j] - ox, r2 = sy[j] - oy;  // target (row) - source (own)
        if (PDE != KIGPU_STOKES) {
          double kv;
          if (FWD && PDE == KIGPU_LAPLACE_NEUMANN) {
            kv = ow * kInv2Pi * log(sqrt(r1 * r1 + r2 * r2));
          } else {
            kv = laplace_entry(PDE, r1, r2, ow, okap, onx, ony, snx[j], sny[j], FWD);
          }
#pragma unroll
          for (int c = 0; c < kAR; ++c)
            if (c < nc) acc[c][0] += kv * slam[j * kAR + c];
        } else {
#pragma unroll
          for (int ta = 0; ta < SD; ++ta)
#pragma unroll
            for (int sb = 0; sb < SD; ++sb) {
              double kv = stokes_entry(ta, sb, sx[j], sy[j], snx[j], sny[j], ox, oy, onx, ony, ow,
                                       okap);
#pragma unroll
              for (int c = 0; c < kAR; ++c)
                if (c < nc) acc[c][sb] += kv * slam[(j * SD + ta) * kAR + c];
            }
        }
      }
    }
    if (valid) {
      for (int c = 0; c < nc; ++c)
        for (int a = 0; a < SD; ++a)
          part[(size_t)blockIdx.y * ldp * nrhs + SD * si + a + (size_t)(c0 + c) * ldp] =
              acc[c][a];
    }
  }
}

// 1/d to full FP64 precision for d > 0: the hardware reciprocal (about 2^-23
// relative error) and two Newton steps r <- r + r (1 - d r), each squaring the
// relative error.
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}


// Laplace double layer, one right-hand side (the refinement residual and the
// interior evaluation).  Per source s the LDS tile holds (x, y), the scaled
// normal a = -w n / (2 pi) and (diagonal value, density), so a pair costs
// three 16-byte LDS reads and
//   q = (r . a) / |r|^2,  or the diagonal value where r is exactly zero
// (the entries of kernel.cpp:11-36 up to rounding).
// TR = false: own points are targets, other points sources, in = mu.
// TR = true:  own points are sources, other points targets, in = lam (rows).
template <bool TR>
__global__ __launch_bounds__(kAT) void k_apply_dl(Targets other_or_tg, Tiles tl, GeomView g,
                                                  const double* in, double* part, int ldp) {
  __shared__ double2 s_p[kAT], s_a[kAT], s_d[kAT];
  const int oi = tl.t0[blockIdx.x] + threadIdx.x;
  const bool valid = oi < tl.t1[blockIdx.x];
  // own point
  double ox = 0, oy = 0, oax = 0, oay = 0, od = 0;
  if (valid) {
    if (TR) {
      const double cw = -g.w[oi] * kInv2Pi;
      ox = g.px[oi];
      oy = g.py[oi];
      oax = cw * g.nx[oi];
      oay = cw * g.ny[oi];
      od = 0.5 + 0.5 * g.kappa[oi] * g.w[oi] * kInv2Pi;
    } else {
      ox = other_or_tg.x[oi];
      oy = other_or_tg.y[oi];
    }
  }
  int j0, j1;
  split_range(tl, &j0, &j1);
  double acc = 0;
  for (int b0 = j0; b0 < j1; b0 += kAT) {
    const int nb = min(kAT, j1 - b0);
    __syncthreads();
    if (threadIdx.x < nb) {
      const int j = b0 + threadIdx.x;
      if (TR) {
        s_p[threadIdx.x] = make_double2(other_or_tg.x[j], other_or_tg.y[j]);
        s_d[threadIdx.x] = make_double2(0.0, in[j]);
      } else {
        const double w = g.w[j];
        const double cw = -w * kInv2Pi;
        s_p[threadIdx.x] = make_double2(g.px[j], g.py[j]);
        s_a[threadIdx.x] = make_double2(cw * g.nx[j], cw * g.ny[j]);
        s_d[threadIdx.x] = make_double2(0.5 + 0.5 * g.kappa[j] * w * kInv2Pi, in[j]);
      }
    }
    __syncthreads();
    if (!valid) continue;
#pragma unroll 4
    for (int jj = 0; jj < nb; ++jj) {
      const double2 p = s_p[jj];
      const double2 dm = s_d[jj];
      double dx, dy, ax, ay, dv;
      if (TR) {  // target = row point jj, source = own
        dx = p.x - ox;
        dy = p.y - oy;
        ax = oax;
        ay = oay;
        dv = od;
      } else {   // target = own, source = jj
        const double2 a = s_a[jj];
        dx = ox - p.x;
        dy = oy - p.y;
        ax = a.x;
        ay = a.y;
        dv = dm.x;
      }
      const double den = fma(dx, dx, dy * dy);
      const double num = fma(dx, ax, dy * ay);
      const double q = den != 0.0 ? num * recip(den) : dv;
      acc = fma(q, dm.y, acc);
    }
  }
  if (valid) part[(size_t)blockIdx.y * ldp + oi] = acc;
}

// K mu for the Laplace double layer when no two nodes coincide (the
// refinement residual), by unordered pairs.  With r = x_i - x_j and
// a_j = -w_j mu_j n_j / (2 pi) the two entries of a pair share |r|^2:
//   K(i, j) mu_j = (r . a_j) / |r|^2,    K(j, i) mu_i = -(r . a_i) / |r|^2,
// so one reciprocal serves both directions and the pairs of two distinct
// blocks are evaluated once.  Each instance's nodes are cut into blocks of
// kSymT; one workgroup takes one block pair (u, v), u <= v:
//   u < v: lane l of wave w owns the kSymP targets w*kSymT/2 + l + 64p of
//          block u; in step s of a 64-source chunk it meets source
//          (l + s) mod 64 of block v, so in every step each source is met by
//          one lane.  A source's running sum moves one lane down per step
//          (a cross-lane shuffle), collecting each lane's targets once, and
//          lands in the wave's LDS row after 64 steps (no atomics, a fixed
//          summation order);
//   u = v: ordered pairs of the block, sources broadcast, |r|^2 clamped away
//          from zero for the node itself (0 * 1e300 = 0) and the diagonal
//          1/2 + kappa w / (4 pi) added once.
// The contribution of pair (u, v) to block u goes to partial slot v, the one
// to block v to slot u, so every node gets each of its instance's slots once
// and k_sum_sym adds them.  Padding lanes and sources sit at x = +-1e30 with
// a = 0: their pairs with real nodes add exactly 0 and |r|^2 stays finite.
constexpr int kSymT = 512;              // nodes per block
constexpr int kSymP = 4;                // targets per lane
constexpr int kSymTB = kSymT / kSymP;   // threads per workgroup
constexpr int kSymNW = kSymTB / kWave;  // waves per workgroup
static_assert(kSymP == 4, "the off-diagonal loop shares one reciprocal between four pairs");
constexpr double kFar = 1e30;
constexpr double kTinyR2 = 1e-300;

struct SymPair {
  int t0, tn;      // target block [t0, t0 + tn)
  int s0, sn;      // source block
  int slot_t;      // partial slot of the target block's sums
  int slot_s;      // partial slot of the source block's sums; -1: diagonal block
};

__global__ __launch_bounds__(kSymTB) void k_apply_sym(const SymPair* pairs, GeomView g,
                                                      const double* mu, double* part, int ldp) {
  __shared__ double2 s_xy[kSymT], s_a[kSymT];
  // after the pair loop s_xy holds the waves' source sums, kSymNW rows of kSymT
  static_assert(sizeof(double2) >= sizeof(double) * kSymNW, "source sums fit in s_xy");
  double* s_acc = reinterpret_cast<double*>(s_xy);
  const SymPair pr = pairs[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  for (int j = tid; j < kSymT; j += kSymTB) {
    if (j < pr.sn) {
      const int p = pr.s0 + j;
      const double c = -g.w[p] * mu[p] * kInv2Pi;
      s_xy[j] = make_double2(g.px[p], g.py[p]);
      s_a[j] = make_double2(c * g.nx[p], c * g.ny[p]);
    } else {
      s_xy[j] = make_double2(-kFar, 0.0);
      s_a[j] = make_double2(0.0, 0.0);
    }
  }
  double tx[kSymP], ty[kSymP], tax[kSymP], tay[kSymP], ta[kSymP];
#pragma unroll
  for (int p = 0; p < kSymP; ++p) {
    const int i = wv * (kWave * kSymP) + lane + kWave * p;
    ta[p] = 0;
    if (i < pr.tn) {
      const int q = pr.t0 + i;
      const double c = -g.w[q] * mu[q] * kInv2Pi;
      tx[p] = g.px[q];
      ty[p] = g.py[q];
      tax[p] = c * g.nx[q];
      tay[p] = c * g.ny[q];
    } else {
      tx[p] = kFar;
      ty[p] = 0;
      tax[p] = 0;
      tay[p] = 0;
    }
  }
  __syncthreads();
  if (pr.slot_s < 0) {
#pragma unroll 2
    for (int j = 0; j < pr.sn; ++j) {
      const double2 xy = s_xy[j], a = s_a[j];
#pragma unroll
      for (int p = 0; p < kSymP; ++p) {
        const double dx = tx[p] - xy.x, dy = ty[p] - xy.y;
        const double d = fmax(fma(dx, dx, dy * dy), kTinyR2);
        const double n = fma(dx, a.x, dy * a.y);
        ta[p] = fma(n, recip1(d), ta[p]);
      }
    }
#pragma unroll
    for (int p = 0; p < kSymP; ++p) {
      const int i = wv * (kWave * kSymP) + lane + kWave * p;
      if (i < pr.tn) {
        const int q = pr.t0 + i;
        const double dv = 0.5 + 0.5 * g.kappa[q] * g.w[q] * kInv2Pi;
        part[(size_t)pr.slot_t * ldp + q] = fma(dv, mu[q], ta[p]);
      }
    }
    return;
  }
  const int ns = (pr.sn + kWave - 1) / kWave * kWave;
  const int next = (lane + 1) & (kWave - 1);
  // source sums stay in registers until every wave is done with s_xy
  double jsv[kSymT / kWave];
#pragma unroll
  for (int bb = 0; bb < kSymT / kWave; ++bb) {
    const int base = bb * kWave;
    jsv[bb] = 0;
    if (base >= ns) continue;
    const double2* cxy = s_xy + base;
    const double2* ca = s_a + base;
    // js: running sum of the source this lane meets in the current step; the
    // source met in step s was met in step s - 1 by lane + 1, whose sum moves
    // here before this lane adds its targets' part
    double js = 0;
#pragma unroll 2
    for (int s = 0; s < kWave; ++s) {
      const int j = (lane + s) & (kWave - 1);
      const double2 xy = cxy[j], a = ca[j];
      double dx[kSymP], dy[kSymP], d[kSymP];
#pragma unroll
      for (int p = 0; p < kSymP; ++p) {
        dx[p] = tx[p] - xy.x;
        dy[p] = ty[p] - xy.y;
        d[p] = fma(dx[p], dx[p], dy[p] * dy[p]);
      }
      // one reciprocal for the four pairs (v_rcp_f64 issues at about a
      // quarter of the FMA rate): 1/(d0 d1 d2 d3) with one Newton step,
      // then 1/d0 = d1 (d2 d3) / (d0 d1 d2 d3) and so on; all d lie in
      // [1e-20, 4e60] (distinct nodes, padding at 1e30), so no product
      // overflows or underflows
      const double p01 = d[0] * d[1], p23 = d[2] * d[3];
      const double rr = recip1(p01 * p23);
      const double r01 = p23 * rr, r23 = p01 * rr;
      const double inv[kSymP] = {d[1] * r01, d[0] * r01, d[3] * r23, d[2] * r23};
      double cs = 0;
#pragma unroll
      for (int p = 0; p < kSymP; ++p) {
        const double ni = fma(dx[p], a.x, dy[p] * a.y);
        const double nj = fma(dx[p], tax[p], dy[p] * tay[p]);
        ta[p] = fma(ni, inv[p], ta[p]);
        cs = fma(nj, inv[p], cs);
      }
      js = __shfl(js, next) + cs;  // all lanes hold 0 before the first step
    }
    // after the last step this lane holds the whole wave's sum of source lane - 1
    jsv[bb] = js;
  }
#pragma unroll
  for (int p = 0; p < kSymP; ++p) {
    const int i = wv * (kWave * kSymP) + lane + kWave * p;
    if (i < pr.tn) part[(size_t)pr.slot_t * ldp + pr.t0 + i] = ta[p];
  }
  __syncthreads();
#pragma unroll
  for (int bb = 0; bb < kSymT / kWave; ++bb)
    s_acc[wv * kSymT + bb * kWave + ((lane + kWave - 1) & (kWave - 1))] = jsv[bb];
  __syncthreads();
  for (int j = tid; j < pr.sn; j += kSymTB) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < kSymNW; ++w) s += s_acc[w * kSymT + j];
    part[(size_t)pr.slot_s * ldp + pr.s0 + j] = -s;
  }
}

// Apply to many right-hand sides as a product on the FP64 matrix cores:
// y_part[split] = K(tile rows, split's sources) * mu(split's sources, :).
// A block of 4 waves owns 128 target DOFs (32 per wave, two 16-row blocks)
// and 16 * NCB right-hand sides per pass.  Sources go through LDS in chunks
// of 32 DOFs (per-source values and the density block), the next chunk
// loaded into registers while the current one is multiplied.  Each lane
// computes the kernel entries of its own A-operand slots of
// v_mfma_f64_16x16x4f64 (row lane % 16, source lane / 16 of each 4-wide
// step) and reads its B-operand slot (source lane / 16, column lane % 16)
// from LDS; accumulator element v of lane l is row l / 16 + 4 v, column
// l % 16 of its 16 x 16 block.  Rows past the tile are computed and not
// stored (a row of the product depends only on its own row of K); sources
// past the split are placed far away with zero weight and zero density, so
// their entries are 0.
//   Laplace: per source (x, y), a = -w n / (2 pi) and the diagonal value;
//   the entry is (r . a) / |r|^2 (one Newton step on v_rcp_f64, as in
//   k_apply_sym), or the diagonal value where r is exactly zero.
//   Laplace-Neumann and Stokes: the entries of laplace_entry / stokes_entry.
typedef double mm_d4 __attribute__((ext_vector_type(4)));
constexpr int kMmTB = 256;
constexpr int kMmRows = 128;  // target DOFs per block
constexpr int kMmS = 32;      // source DOFs per LDS chunk
constexpr int kMmLd = 72;     // LDS row stride of the density chunk (64 columns + 8)

template <int PDE, int NCB>
__global__ __launch_bounds__(kMmTB) void k_apply_mm(Targets tg, Tiles tl, GeomView g,
                                                    const double* mu, int ldm, int nrhs,
                                                    double* part, int ldp) {
  constexpr int SD = PDE == KIGPU_STOKES ? 2 : 1;
  constexpr int SP = kMmS / SD;                    // source points per chunk
  constexpr int NM = kMmS * 16 * NCB / kMmTB;      // density values per thread and chunk
  static_assert(16 * NCB <= kMmLd && NM >= 1, "density chunk");
  __shared__ double s_src[6][SP];
  __shared__ double s_mu[kMmS * kMmLd];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int tp0 = tl.t0[blockIdx.x], tp1 = tl.t1[blockIdx.x];
  double tx[2], ty[2], tnx[2], tny[2];
  int tc[2];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    const int row = 32 * wv + 16 * rb + (lane & 15);
    const int pt = min(tp0 + row / SD, tp1 - 1);  // rows past the tile repeat the last point
    tc[rb] = row % SD;
    tx[rb] = tg.x[pt];
    ty[rb] = tg.y[pt];
    tnx[rb] = tg.nx[pt];
    tny[rb] = tg.ny[pt];
  }
  // this split's source points, in whole chunks
  const int b = tl.inst[blockIdx.x];
  const int oa = (int)tl.opt[b], oe = (int)tl.opt[b + 1];
  int per = (oe - oa + (int)gridDim.y - 1) / (int)gridDim.y;
  per = (per + SP - 1) / SP * SP;
  const int s0 = min(oe, oa + (int)blockIdx.y * per), s1 = min(oe, s0 + per);
  const size_t pbase = (size_t)blockIdx.y * ldp * nrhs;
  for (int c0 = 0; c0 < nrhs; c0 += 16 * NCB) {
    mm_d4 acc[2][NCB];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) acc[rb][cb] = mm_d4{0.0, 0.0, 0.0, 0.0};
    // registers holding the next chunk
    double rs[6];
    double rm[NM];
    auto fetch = [&](int b0) {
      const int nb = min(SP, s1 - b0);
      if (tid < SP) {
        const int sp = b0 + tid;
        if (tid < nb) {
          const double w = g.w[sp], kap = g.kappa[sp];
          rs[0] = g.px[sp];
          rs[1] = g.py[sp];
          if (PDE == KIGPU_LAPLACE) {
            rs[2] = -w * kInv2Pi * g.nx[sp];
            rs[3] = -w * kInv2Pi * g.ny[sp];
            rs[4] = 0.5 + 0.5 * kap * w * kInv2Pi;
            rs[5] = 0.0;
          } else {
            rs[2] = g.nx[sp];
            rs[3] = g.ny[sp];
            rs[4] = w;
            rs[5] = kap;
          }
        } else {
          rs[0] = kFar;
          rs[1] = kFar;
          rs[2] = rs[3] = rs[4] = rs[5] = 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < NM; ++u) {
        const int e = tid + kMmTB * u;
        const int cc = e / kMmS, d = e - cc * kMmS;
        const int col = c0 + cc;
        rm[u] = (d < SD * nb && col < nrhs) ? mu[(size_t)SD * b0 + d + (size_t)col * ldm] : 0.0;
      }
    };
    if (s0 < s1) fetch(s0);
    for (int b0 = s0; b0 < s1; b0 += SP) {
      __syncthreads();
      if (tid < SP)
#pragma unroll
        for (int q = 0; q < 6; ++q) s_src[q][tid] = rs[q];
#pragma unroll
      for (int u = 0; u < NM; ++u) {
        const int e = tid + kMmTB * u;
        const int cc = e / kMmS, d = e - cc * kMmS;
        s_mu[d * kMmLd + cc] = rm[u];
      }
      __syncthreads();
      if (b0 + SP < s1) fetch(b0 + SP);
#pragma unroll 2
      for (int ks = 0; ks < kMmS / 4; ++ks) {
        const int d = 4 * ks + (lane >> 4);  // source DOF in the chunk
        const int sp = d / SD, sc = d % SD;
        const double sx = s_src[0][sp], sy = s_src[1][sp];
        double a[2];
        if (PDE == KIGPU_LAPLACE) {
          const double ax = s_src[2][sp], ay = s_src[3][sp], dg = s_src[4][sp];
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) {
            const double r1 = tx[rb] - sx, r2 = ty[rb] - sy;
            const double rr = fma(r1, r1, r2 * r2);
            const double v = fma(r1, ax, r2 * ay) * recip1(rr);
            a[rb] = rr == 0.0 ? dg : v;
          }
        } else {
          const double snx = s_src[2][sp], sny = s_src[3][sp], sw = s_src[4][sp],
                       sk = s_src[5][sp];
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) {
            double v;
            if (PDE == KIGPU_STOKES)
              v = stokes_entry(tc[rb], sc, tx[rb], ty[rb], tnx[rb], tny[rb], sx, sy, snx, sny,
                               sw, sk);
            else
              v = laplace_entry(PDE, tx[rb] - sx, ty[rb] - sy, sw, sk, snx, sny, tnx[rb],
                                tny[rb], false);
            a[rb] = sx == kFar ? 0.0 : v;  // padding sources
          }
        }
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
          const double bv = s_mu[d * kMmLd + 16 * cb + (lane & 15)];
#pragma unroll
          for (int rb = 0; rb < 2; ++rb)
            acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], bv, acc[rb][cb], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 32 * wv + 16 * rb + (lane >> 4) + 4 * v;
          const int col = c0 + 16 * cb + (lane & 15);
          if (tp0 + row / SD < tp1 && col < nrhs)
            part[pbase + (size_t)SD * tp0 + row + (size_t)col * ldp] = acc[rb][cb][v];
        }
  }
}

// y over the blocks (3 ints per block: start, length, slots of its
// instance): node i of block blockIdx.x adds its slots; 64 nodes x 4 slot
// phases per workgroup, the four phase sums combined in a fixed order
__global__ __launch_bounds__(256) void k_sum_sym(const int* blocks, const double* part, int ldp,
                                                 double* y) {
  __shared__ double red[4][kWave];
  const int* b = blocks + 3 * blockIdx.x;
  const int p0 = b[0], n = b[1], ns = b[2];
  const int li = threadIdx.x & (kWave - 1), ph = threadIdx.x / kWave;
  const int i = blockIdx.y * kWave + li;
  double s = 0;
  if (i < n)
    for (int k = ph; k < ns; k += 4) s += part[(size_t)k * ldp + p0 + i];
  red[ph][li] = s;
  __syncthreads();
  if (ph == 0 && i < n) y[p0 + i] = (red[0][li] + red[1][li]) + (red[2][li] + red[3][li]);
}

// flag = 1 if two points of one leaf have identical coordinates (coincident
// points always share a leaf); one 64-thread block per node
__global__ void k_dup_points(int nnodes, const int* is_leaf, const int64_t* pstart,
                             const int* pcount, const int* perm, const double* px,
                             const double* py, int* flag) {
  const int x = blockIdx.x;
  if (x >= nnodes || !is_leaf[x]) return;
  const int c = pcount[x];
  const int* pp = perm + pstart[x];
  for (int e = threadIdx.x; e < c * c; e += blockDim.x) {
    const int a = e / c, b = e - a * c;
    if (a < b && px[pp[a]] == px[pp[b]] && py[pp[a]] == py[pp[b]]) *flag = 1;  // all write 1
  }
}

__global__ void k_sum_parts(const double* part, int nsplit, size_t len, double* y) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double s = 0;
  for (int p = 0; p < nsplit; ++p) s += part[(size_t)p * len + i];
  y[i] = s;
}

// r = f - y (FP64); partial sums of squares of r and f per (instance
// blockIdx.z, column blockIdx.y) over the instance's DOFs [doff[z], doff[z+1])
__global__ void k_residual(const double* f, const double* y, double* r, int n, int nrhs,
                           const int64_t* doff, double* sums) {
  __shared__ double red[2 * kNW];
  const int c = blockIdx.y;
  const int i0 = (int)doff[blockIdx.z], i1 = (int)doff[blockIdx.z + 1];
  double sr = 0, sf = 0;
  for (int i = i0 + blockIdx.x * blockDim.x + threadIdx.x; i < i1; i += gridDim.x * blockDim.x) {
    double fv = f[i + (size_t)c * n];
    double rv = fv - y[i + (size_t)c * n];
    r[i + (size_t)c * n] = rv;
    sr += rv * rv;
    sf += fv * fv;
  }
  sr = wave_sum(sr);
  sf = wave_sum(sf);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[wv] = sr;
    red[kNW + wv] = sf;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int i = 0; i < kNW; ++i) {
      a += red[i];
      b += red[kNW + i];
    }
    const size_t o = ((size_t)blockIdx.z * nrhs + c) * gridDim.x + blockIdx.x;
    sums[2 * o] = a;
    sums[2 * o + 1] = b;
  }
}

template <typename A, typename B>
__global__ void k_convert(const A* in, B* out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (B)in[i];
}

__global__ void k_axpy_f32(double* x, const float* d, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += (double)d[i];
}

}  // namespace

// Levels l and l + 1 of the sweeps in one launch (k_up_pair / k_down_pair)
// when both exist and four children's LDS regions plus the node's fit; the
// LDS bytes of that launch and each child's region (values) come back.
template <typename T>
static bool pair_fits(const Context& c, const TreeData& t, int l, int rcol, int per_n,
                      size_t* lds, int* wstride) {
  if (!c.tune.pair_sweeps) return false;
  if (l < 1 || l + 1 >= t.nlevels || !t.work[l] || !t.work[l + 1]) return false;
  const LevelWork& wp = *t.work[l];
  const LevelWork& wc = *t.work[l + 1];
  *wstride = per_n * wc.nmax * rcol;
  const size_t vals = std::max<size_t>(4 * (size_t)*wstride, (size_t)per_n * wp.nmax * rcol);
  *lds = sizeof(T) * vals;
  return *lds > 0 && *lds <= c.lds_max;
}

// upward pass over all levels, finest first (see up_node for the modes)
template <typename T>
void up_sweeps(Context& c, TreeData& t, T* x, int ldx, int nrhs, int mode) {
  hipStream_t s = c.stream;
  const int ncb = ceil_div(nrhs, kRC);
  const int rcol = std::min(nrhs, kRC);
  const SweepArgs f{t.rank.p, t.nact.p, t.act_off.p, t.sr_pool.p, t.fac_ptr.p};
  for (int l = t.nlevels - 1; l >= 1;) {
    size_t lds = 0;
    int ws = 0;
    if (pair_fits<T>(c, t, l - 1, rcol, 2, &lds, &ws)) {  // levels l - 1 and l
      const LevelWork& w = *t.work[l - 1];
      allow_lds(k_up_pair<T>, lds);
      k_up_pair<T><<<dim3(w.b - w.a, ncb), kTB, lds, s>>>(w.a, t.child.p, f, x, ldx, nrhs, mode,
                                                          ws);
      l -= 2;
      continue;
    }
    LevelWork& w = *t.work[l];
    if (w.nmax > 0) {
      lds = 2 * sizeof(T) * (size_t)w.nmax * rcol;
      allow_lds(k_up<T>, lds);
      k_up<T><<<dim3(w.b - w.a, ncb), kTB, lds, s>>>(w.a, f, x, ldx, nrhs, mode);
    }
    l -= 1;
  }
}
template void up_sweeps<double>(Context&, TreeData&, double*, int, int, int);
template void up_sweeps<float>(Context&, TreeData&, float*, int, int, int);

// downward pass, coarsest first (trans: with L^T in place of U)
template <typename T>
void down_sweeps(Context& c, TreeData& t, T* x, int ldx, int nrhs, int trans = 0) {
  hipStream_t s = c.stream;
  const int ncb = ceil_div(nrhs, kRC);
  const int rcol = std::min(nrhs, kRC);
  const SweepArgs f{t.rank.p, t.nact.p, t.act_off.p, t.sr_pool.p, t.fac_ptr.p};
  for (int l = 1; l < t.nlevels;) {
    size_t lds = 0;
    int ws = 0;
    if (pair_fits<T>(c, t, l, rcol, 1, &lds, &ws)) {  // levels l and l + 1
      const LevelWork& w = *t.work[l];
      allow_lds(k_down_pair<T>, lds);
      k_down_pair<T><<<dim3(w.b - w.a, ncb), kTB, lds, s>>>(w.a, t.child.p, f, x, ldx, nrhs,
                                                            trans, ws);
      l += 2;
      continue;
    }
    LevelWork& w = *t.work[l];
    if (w.nmax > 0) {
      lds = sizeof(T) * (size_t)w.nmax * rcol;
      allow_lds(k_down<T>, lds);
      k_down<T><<<dim3(w.b - w.a, ncb), kTB, lds, s>>>(w.a, f, x, ldx, nrhs, trans);
    }
    l += 1;
  }
  KI_HIP(hipGetLastError());
}

namespace {

// y(:, c) += U alpha(:, c)  (U: nd x h, FP64)
__global__ void k_add_ualpha(int nd, int h, const double* U, const double* alpha, double* y) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (d >= nd) return;
  double s = 0;
  for (int j = 0; j < h; ++j) s += U[d + (size_t)j * nd] * alpha[j + (size_t)c * h];
  y[d + (size_t)c * nd] += s;
}

// rb(i, c) = alpha(i, c) - sum_d Psi^T(d, i) x(d, c)
__global__ __launch_bounds__(kTB) void k_border_resid(int nd, int h, const double* PsiT,
                                                      const double* x, const double* alpha,
                                                      double* rb) {
  __shared__ double red[kNW];
  const int i = blockIdx.x, c = blockIdx.y;
  double s = 0;
  for (int d = threadIdx.x; d < nd; d += kTB) s += PsiT[d + (size_t)i * nd] * x[d + (size_t)c * nd];
  s = block_sum<kNW>(s, red);
  if (threadIdx.x == 0) rb[i + (size_t)c * h] = alpha[i + (size_t)c * h] - s;
}

__global__ void k_axpy_f64(double* x, const double* d, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += d[i];
}

template <typename T>
void sweeps_launch(Context& c, TreeData& t, T* x, int ldx, int nrhs, const double* g, bool trans);

// x <- F^{-1} x, or F^{-T} x with trans (no border then).  With graphs on,
// the launches of a sweep are captured on its second call with the same
// vector, shape and border vector, and replayed from then on while the
// bordered step's scratch stays where the capture found it: the first call
// runs them directly (it sets the LDS limits and sizes the border scratch, so
// the captured launches allocate nothing), and the sweeps hold no memset,
// host wait or allocation.
template <typename T>
void sweeps(Context& c, TreeData& t, T* x, int ldx, int nrhs, const double* g = nullptr,
            bool trans = false) {
  if (!c.graphs) {
    sweeps_launch<T>(c, t, x, ldx, nrhs, g, trans);
    return;
  }
  hipStream_t s = c.stream;
  TreeData::SweepGraph* e = nullptr;
  for (auto& q : t.sweep_graphs)
    if (q.tsz == (int)sizeof(T) && q.x == x && q.ldx == ldx && q.nrhs == nrhs && q.g == g &&
        q.trans == trans)
      e = &q;
  if (!e) {
    if (t.sweep_graphs.size() >= 8) {  // keep a few: the oldest goes
      if (t.sweep_graphs.front().exec) (void)hipGraphExecDestroy(t.sweep_graphs.front().exec);
      t.sweep_graphs.erase(t.sweep_graphs.begin());
    }
    t.sweep_graphs.push_back({(int)sizeof(T), x, ldx, nrhs, g, trans, nullptr, {}});
    sweeps_launch<T>(c, t, x, ldx, nrhs, g, trans);
    return;
  }
  const void* scr[3] = {t.bgram.p, t.bM.p, t.alpha_dev.p};
  if (e->exec && (e->scr[0] != scr[0] || e->scr[1] != scr[1] || e->scr[2] != scr[2])) {
    // the border scratch moved since the capture: the graph would write the
    // freed blocks.  Run directly and capture again on the next call.
    (void)hipGraphExecDestroy(e->exec);
    e->exec = nullptr;
    sweeps_launch<T>(c, t, x, ldx, nrhs, g, trans);
    return;
  }
  if (!e->exec) {
    hipGraph_t graph = nullptr;
    KI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      sweeps_launch<T>(c, t, x, ldx, nrhs, g, trans);
    } catch (...) {
      (void)hipStreamEndCapture(s, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    KI_HIP(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    KI_HIP(ie);
    e->exec = exec;
    e->scr[0] = t.bgram.p;
    e->scr[1] = t.bM.p;
    e->scr[2] = t.alpha_dev.p;
    ++c.tim.graph_captures;
  }
  KI_HIP(hipGraphLaunch(e->exec, s));
  ++c.tim.graph_replays;
}

template <typename T>
void sweeps_launch(Context& c, TreeData& t, T* x, int ldx, int nrhs, const double* g, bool trans) {
  hipStream_t s = c.stream;
  up_sweeps<T>(c, t, x, ldx, nrhs, trans ? 3 : 1);
  if (t.nborder > 0) {
    border_step<T>(c, t, x, ldx, nrhs, g);
  } else if (t.nroot > 0) {
    k_root_apply<T><<<dim3(nrhs, t.ninst), kTB, sizeof(double) * t.nroot_max, s>>>(
        t.root_n_p, t.root_doff_p, t.root_moff_p, t.root_dofs.p, t.root_lu.p, x, ldx, trans);
  }
  down_sweeps<T>(c, t, x, ldx, nrhs, trans);
}

}  // namespace

namespace {
__global__ void k_split_xy(const double* xy, int n, double* x, double* y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    x[i] = xy[2 * i];
    y[i] = xy[2 * i + 1];
  }
}

// tiles of kAT own points per segment [own[i], own[i+1]); other segments opt
struct TileBuf {
  DBuf<int> inst, t0, t1;
  DBuf<int64_t> opt;
  std::vector<int64_t> own, other;  // the segments it was built from
  int ntiles = 0;
  int maxlen = 0;  // longest other-segment
  void build(const std::vector<int64_t>& own_, const std::vector<int64_t>& other_, hipStream_t s,
             int tile = kAT) {
    own = own_;
    other = other_;
    std::vector<int> hi, h0, h1;
    maxlen = 0;
    for (size_t b = 0; b + 1 < own.size(); ++b) {
      for (int64_t p = own[b]; p < own[b + 1]; p += tile) {
        hi.push_back((int)b);
        h0.push_back((int)p);
        h1.push_back((int)std::min<int64_t>(own[b + 1], p + tile));
      }
      maxlen = std::max(maxlen, (int)(other[b + 1] - other[b]));
    }
    ntiles = (int)hi.size();
    inst.upload(hi, s);
    t0.upload(h0, s);
    t1.upload(h1, s);
    opt.upload(other, s);
  }
  Tiles view() const { return Tiles{inst.p, t0.p, t1.p, opt.p}; }
  // splits of the other range: enough blocks to fill the chip, each split at
  // least kAT points
  int splits(int cap) const {
    const int by_len = std::max(1, ceil_div(maxlen, kAT));
    return std::max(1, std::min({cap, ceil_div(2048, std::max(ntiles, 1)), by_len}));
  }
};
}  // namespace

// interleaved [x0, y0, x1, ...] device points -> separate x and y arrays
void split_points(Context& c, const double* d_xy, int n, DBuf<double>* x, DBuf<double>* y) {
  x->alloc(n);
  y->alloc(n);
  if (n > 0) k_split_xy<<<ceil_div(n, 256), 256, 0, c.stream>>>(d_xy, n, x->p, y->p);
  KI_HIP(hipGetLastError());
}

namespace {

// K(all, all) restricted to each instance's diagonal block (FWD = false), or
// the forward matrix at domain targets (FWD = true, one instance)
template <bool TRANS>
void run_apply(Context& c, const TreeData& t, int pde, bool fwd, const Targets& other,
               const TileBuf& tb, const double* d_in, int ldi, int nrhs, double* d_out, int ldo) {
  hipStream_t s = c.stream;
  const GeomView g{t.px.p, t.py.p, t.nx.p, t.ny.p, t.w.p, t.kappa.p, t.npts, t.sd};
  const size_t len = (size_t)ldo * nrhs;
  DBuf<double> part;
  if (!TRANS && !fwd && nrhs >= 2) {
    // many right-hand sides: the product on the matrix cores, tiles of
    // kMmRows target DOFs
    const int sd = pde == KIGPU_STOKES ? 2 : 1;
    TileBuf mt;
    mt.build(tb.own, tb.other, s, kMmRows / sd);
    const int msplit = std::max(1, std::min(ceil_div(2048, std::max(mt.ntiles, 1)),
                                            ceil_div(mt.maxlen, kMmS / sd * 8)));
    part.alloc(len * msplit);
    const dim3 mgrid(mt.ntiles, msplit);
    const Tiles mtl = mt.view();
#define KI_MM(P)                                                                              \
  do {                                                                                        \
    if (nrhs <= 16)                                                                           \
      k_apply_mm<P, 1><<<mgrid, kMmTB, 0, s>>>(other, mtl, g, d_in, ldi, nrhs, part.p, ldo);  \
    else if (nrhs <= 32)                                                                      \
      k_apply_mm<P, 2><<<mgrid, kMmTB, 0, s>>>(other, mtl, g, d_in, ldi, nrhs, part.p, ldo);  \
    else                                                                                      \
      k_apply_mm<P, 4><<<mgrid, kMmTB, 0, s>>>(other, mtl, g, d_in, ldi, nrhs, part.p, ldo);  \
  } while (0)
    if (pde == KIGPU_LAPLACE) KI_MM(KIGPU_LAPLACE);
    else if (pde == KIGPU_LAPLACE_NEUMANN) KI_MM(KIGPU_LAPLACE_NEUMANN);
    else KI_MM(KIGPU_STOKES);
#undef KI_MM
    k_sum_parts<<<ceil_div((long)len, 256), 256, 0, s>>>(part.p, msplit, len, d_out);
    KI_HIP(hipGetLastError());
    return;
  }
  const int split = tb.splits(TRANS || fwd ? c.tune.fwd_split_cap : c.tune.apply_split_cap);
  part.alloc(len * split);
  dim3 grid(tb.ntiles, split);
  const Tiles tl = tb.view();
  if (pde == KIGPU_LAPLACE && nrhs == 1) {
    if (TRANS)
      k_apply_dl<true><<<grid, kAT, 0, s>>>(other, tl, g, d_in, part.p, ldo);
    else
      k_apply_dl<false><<<grid, kAT, 0, s>>>(other, tl, g, d_in, part.p, ldo);
    k_sum_parts<<<ceil_div((long)len, 256), 256, 0, s>>>(part.p, split, len, d_out);
    KI_HIP(hipGetLastError());
    return;
  }
#define KI_RUN(P, F)                                                                         \
  do {                                                                                       \
    if (TRANS)                                                                               \
      k_apply_t<P, F><<<grid, kAT, 0, s>>>(other, tl, g, d_in, ldi, nrhs, part.p, ldo);      \
    else                                                                                     \
      k_apply<P, F><<<grid, kAT, 0, s>>>(other, tl, g, d_in, ldi, nrhs, part.p, ldo);        \
  } while (0)
  if (pde == KIGPU_LAPLACE) {
    if (fwd) KI_RUN(KIGPU_LAPLACE, true); else KI_RUN(KIGPU_LAPLACE, false);
  } else if (pde == KIGPU_LAPLACE_NEUMANN) {
    if (fwd) KI_RUN(KIGPU_LAPLACE_NEUMANN, true); else KI_RUN(KIGPU_LAPLACE_NEUMANN, false);
  } else {
    if (fwd) KI_RUN(KIGPU_STOKES, true); else KI_RUN(KIGPU_STOKES, false);
  }
#undef KI_RUN
  k_sum_parts<<<ceil_div((long)len, 256), 256, 0, s>>>(part.p, split, len, d_out);
  KI_HIP(hipGetLastError());
}

}  // namespace

// whether two nodes coincide (checked once per tree)
static bool has_coincident_points(Context& c, const TreeData& t) {
  if (t.dup_state < 0) {
    DBuf<int> flag;
    flag.alloc(1);
    flag.zero(c.stream);
    k_dup_points<<<t.nnodes, 64, 0, c.stream>>>(t.nnodes, t.is_leaf.p, t.pstart.p, t.pcount.p,
                                                t.perm_all.p, t.px.p, t.py.p, flag.p);
    int h = 0;
    flag.download(&h, 1, c.stream);
    stream_wait(c.stream);
    t.dup_state = h;
  }
  return t.dup_state != 0;
}

// block pairs of the symmetric apply, built once per tree: per instance,
// blocks of kSymT nodes and every pair u <= v of them
static void ensure_sym_plan(Context& c, const TreeData& t) {
  if (t.sym_npairs > 0) return;
  std::vector<int> pairs, blocks;
  int slots = 0;
  for (int i = 0; i < t.ninst; ++i) {
    const int p0 = (int)t.inst_pt[i], n = (int)(t.inst_pt[i + 1] - t.inst_pt[i]);
    const int nb = ceil_div(n, kSymT);
    slots = std::max(slots, nb);
    for (int u = 0; u < nb; ++u) {
      const int u0 = p0 + u * kSymT, un = std::min(kSymT, n - u * kSymT);
      blocks.insert(blocks.end(), {u0, un, nb});
      for (int v = u; v < nb; ++v) {
        const int v0 = p0 + v * kSymT, vn = std::min(kSymT, n - v * kSymT);
        pairs.insert(pairs.end(), {u0, un, v0, vn, v, u == v ? -1 : u});
      }
    }
  }
  static_assert(sizeof(SymPair) == 6 * sizeof(int), "SymPair is six ints");
  t.sym_pairs.upload(pairs, c.stream);
  t.sym_blocks.upload(blocks, c.stream);
  t.sym_npairs = (int)(pairs.size() / 6);
  t.sym_nblocks = (int)(blocks.size() / 3);
  t.sym_slots = slots;
}

void kernel_apply(Context& c, const TreeData& t, int pde, const double* d_mu, int nrhs,
                  double* d_y) {
  Targets tg{t.px.p, t.py.p, t.nx.p, t.ny.p, t.npts};
  TileBuf tb;
  if (pde == KIGPU_LAPLACE && nrhs == 1 && !has_coincident_points(c, t)) {
    hipStream_t s = c.stream;
    const GeomView g{t.px.p, t.py.p, t.nx.p, t.ny.p, t.w.p, t.kappa.p, t.npts, t.sd};
    ensure_sym_plan(c, t);
    DBuf<double> part;
    part.alloc((size_t)t.ndof * t.sym_slots);
    KI_HIP(hipEventRecord(c.ev[12], s));
    k_apply_sym<<<t.sym_npairs, kSymTB, 0, s>>>(
        reinterpret_cast<const SymPair*>(t.sym_pairs.p), g, d_mu, part.p, t.ndof);
    KI_HIP(hipEventRecord(c.ev[13], s));
    k_sum_sym<<<dim3(t.sym_nblocks, kSymT / kWave), 256, 0, s>>>(t.sym_blocks.p, part.p, t.ndof,
                                                                d_y);
    KI_HIP(hipGetLastError());
    stream_wait(s);
    float ms = 0;
    KI_HIP(hipEventElapsedTime(&ms, c.ev[12], c.ev[13]));
    c.tim.apply_kernel_s += ms * 1e-3;
    c.tim.apply_launches += 1;
    c.tim.apply_pairs += [&] {
      double p = 0;
      for (int i = 0; i < t.ninst; ++i) {
        const double ni = (double)(t.inst_pt[i + 1] - t.inst_pt[i]);
        p += ni * ni;
      }
      return p;
    }();
    return;
  }
  tb.build(t.inst_pt, t.inst_pt, c.stream);
  run_apply<false>(c, t, pde, false, tg, tb, d_mu, t.ndof, nrhs, d_y, t.ndof);
  stream_wait(c.stream);
}

void kernel_apply_transpose(Context& c, const TreeData& t, int pde, const double* d_lam, int nrhs,
                            double* d_y) {
  Targets rows{t.px.p, t.py.p, t.nx.p, t.ny.p, t.npts};
  TileBuf tb;
  tb.build(t.inst_pt, t.inst_pt, c.stream);
  run_apply<true>(c, t, pde, false, rows, tb, d_lam, t.ndof, nrhs, d_y, t.ndof);
  stream_wait(c.stream);
}

void forward_apply(Context& c, const TreeData& t, int pde, const double* d_tgt, int ntgt,
                   const double* d_mu, int nrhs, double* d_u) {
  DBuf<double> tx, ty;
  split_points(c, d_tgt, ntgt, &tx, &ty);
  Targets tg{tx.p, ty.p, nullptr, nullptr, ntgt};
  const int sd = pde == KIGPU_STOKES ? 2 : 1;
  TileBuf tb;  // own = targets, other = all boundary points
  tb.build({0, (int64_t)ntgt}, {0, (int64_t)t.npts}, c.stream);
  run_apply<false>(c, t, pde, true, tg, tb, d_mu, t.ndof, nrhs, d_u, sd * ntgt);
  stream_wait(c.stream);
}

void forward_apply_transpose(Context& c, const TreeData& t, int pde, const double* d_tgt, int ntgt,
                             const double* d_v, int nrhs, double* d_y) {
  DBuf<double> tx, ty;
  split_points(c, d_tgt, ntgt, &tx, &ty);
  Targets rows{tx.p, ty.p, nullptr, nullptr, ntgt};
  const int sd = pde == KIGPU_STOKES ? 2 : 1;
  TileBuf tb;  // own = boundary points, other = targets
  tb.build({0, (int64_t)t.npts}, {0, (int64_t)ntgt}, c.stream);
  run_apply<true>(c, t, pde, true, rows, tb, d_v, sd * ntgt, nrhs, d_y, t.ndof);
  stream_wait(c.stream);
}

// r = f - y written out; per (instance, column) sums of r^2 and f^2
static void residual_sums(Context& c, const TreeData& t, const double* f, const double* y,
                          double* r, int nrhs, std::vector<double>* rr, std::vector<double>* ff) {
  hipStream_t s = c.stream;
  const int ni = t.ninst, n = t.ndof;
  const int nb = ni == 1 ? 64 : 4;
  if (t.d_dof_off.n != (size_t)ni + 1) {  // instance DOF offsets, uploaded once per tree
    std::vector<int64_t> doff(ni + 1);
    for (int i = 0; i <= ni; ++i) doff[i] = t.inst_pt[i] * t.sd;
    t.d_dof_off.upload(doff, s);
  }
  // the partial sums go straight to page-locked memory
  const size_t nh = 2 * (size_t)nb * nrhs * ni;
  double* h = c.pin_down.get<double>(nh, s);
  k_residual<<<dim3(nb, nrhs, ni), kTB, 0, s>>>(f, y, r, n, nrhs, t.d_dof_off.p,
                                                c.pin_down.dget<double>());
  stream_wait(s);
  rr->assign((size_t)nrhs * ni, 0);
  ff->assign((size_t)nrhs * ni, 0);
  for (size_t q = 0; q < (size_t)nrhs * ni; ++q)
    for (int i = 0; i < nb; ++i) {
      (*rr)[q] += h[2 * (q * nb + i)];
      (*ff)[q] += h[2 * (q * nb + i) + 1];
    }
}

// ---------------------------------------------------------------------------
// solve without device allocation
// ---------------------------------------------------------------------------
// FP32 staging copy of the right-hand sides for solve_tree_noalloc, one per
// GPU, held until the process ends.  In host code `scratch` names a host
// stand-in, not device memory: kernels get the address hipGetSymbolAddress
// returns for the current device (noalloc_init).
static __device__ __align__(256) float scratch[kNoAllocMaxEntries];

namespace {
std::mutex g_noalloc_mu;
std::vector<float*> g_noalloc_ptr;  // per device, set by noalloc_init
}  // namespace

void noalloc_init(int device) {
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if ((int)g_noalloc_ptr.size() <= device) g_noalloc_ptr.resize(device + 1, nullptr);
  if (g_noalloc_ptr[device]) return;
  void* p = nullptr;
  KI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_noalloc_ptr[device] = static_cast<float*>(p);
}

// x = F^{-1} b with the factorization's own arithmetic and no refinement,
// the same kernels in the same order as solve_tree with max_iters = 0, but
// no device memory is allocated: the FP32 factor's staging copy of x lives
// in `scratch` (at most kNoAllocMaxEntries = ndof * nrhs values), the FP64
// factor works in place in d_x.  Trees with a border are rejected (the
// bordered step keeps per-call buffers).  Calls that overlap on one GPU
// share its scratch and race.
void solve_tree_noalloc(Context& c, TreeData& t, const double* d_b, int nrhs, double* d_x) {
  hipStream_t s = c.stream;
  if (t.nborder > 0) KI_FAIL(KIGPU_ERR_NOTSUPP, "the no-allocation solve takes no border");
  const size_t len = (size_t)t.ndof * nrhs;
  if (t.opts.precision == KIGPU_FP64) {
    if (d_x != d_b)
      KI_HIP(hipMemcpyAsync(d_x, d_b, sizeof(double) * len, hipMemcpyDeviceToDevice, s));
    sweeps<double>(c, t, d_x, t.ndof, nrhs, nullptr, false);
    stream_wait(s);
    return;
  }
  if (len > (size_t)kNoAllocMaxEntries)
    KI_FAIL(KIGPU_ERR_ARG, "ndof * nrhs exceeds the no-allocation scratch (kNoAllocMaxEntries)");
  float* x32 = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (c.device < (int)g_noalloc_ptr.size()) x32 = g_noalloc_ptr[c.device];
  }
  if (!x32) KI_FAIL(KIGPU_ERR_STATE, "no-allocation scratch not set up on this device");
  const int g = ceil_div((long)len, 256);
  k_convert<double, float><<<g, 256, 0, s>>>(d_b, x32, len);
  sweeps<float>(c, t, x32, t.ndof, nrhs, nullptr, false);
  k_convert<float, double><<<g, 256, 0, s>>>(x32, d_x, len);
  KI_HIP(hipGetLastError());
  stream_wait(s);
}

void solve_tree(Context& c, TreeData& t, const double* d_b, int nrhs, double* d_x, double* d_alpha,
                const kigpu_refine_opts* ro, kigpu_solve_info* info, bool trans) {
  hipStream_t s = c.stream;
  if (trans && t.nborder > 0)
    KI_FAIL(KIGPU_ERR_NOTSUPP, "transposed solve of a bordered (multiply connected) system");
  c.tim.apply_launches = 0;
  c.tim.apply_kernel_s = c.tim.apply_pairs = 0;
  c.tim.tc_launches = 0;
  c.tim.tc_kernel_s = c.tim.tc_eval_s = c.tim.tc_series_pairs = c.tim.tc_direct_pairs = 0;
  c.tim.tc_series_terms = 0;
  c.tim.graph_replays = c.tim.graph_captures = 0;
  const size_t len = (size_t)t.ndof * nrhs;
  const int h = t.nborder;
  const size_t alen = (size_t)h * nrhs;
  if (info) {
    info->iterations = 0;
    info->rel_residual = -1;
  }
  KI_HIP(hipEventRecord(c.ev[6], s));
  if (t.opts.precision == KIGPU_FP64) {
    if (c.graphs) {
      // the captured sweeps run on the tree's own vector, whose address
      // stays the same from solve to solve
      t.gx.alloc(len);
      KI_HIP(hipMemcpyAsync(t.gx.p, d_b, sizeof(double) * len, hipMemcpyDeviceToDevice, s));
      sweeps<double>(c, t, t.gx.p, t.ndof, nrhs, nullptr, trans);
      KI_HIP(hipMemcpyAsync(d_x, t.gx.p, sizeof(double) * len, hipMemcpyDeviceToDevice, s));
    } else {
      if (d_x != d_b)
        KI_HIP(hipMemcpyAsync(d_x, d_b, sizeof(double) * len, hipMemcpyDeviceToDevice, s));
      sweeps<double>(c, t, d_x, t.ndof, nrhs, nullptr, trans);
    }
    if (h && d_alpha)
      KI_HIP(hipMemcpyAsync(d_alpha, t.alpha_dev.p, sizeof(double) * alen,
                            hipMemcpyDeviceToDevice, s));
    KI_HIP(hipEventRecord(c.ev[7], s));
    stream_wait(s);
    float ms = 0;
    KI_HIP(hipEventElapsedTime(&ms, c.ev[6], c.ev[7]));
    c.tim.t_solve = ms * 1e-3;
    c.tim.t_refine_apply = 0;
    return;
  }
  // FP32 factor: solve in FP32, then FP64 refinement
  //   r = [f - K mu - U alpha; alpha - Psi mu],  [mu; alpha] += F^{-1} r
  // The host waits only for each residual's norms; every phase is timed by
  // events read after the last wait.
  const int iters = ro ? ro->max_iters : 0;
  // the residual apply: the treecode for the Laplace double layer with no
  // coincident nodes (unless the context asks for the direct sum), else the
  // direct kernels
  const bool tc_try = iters > 0 && !trans && c.refine_apply == 0 &&
                      t.opts.pde == KIGPU_LAPLACE && treecode_usable(t);
  int* dup_h = nullptr;
  if (tc_try && t.dup_state < 0) {
    // coincident-node test on the side stream, its flag written straight to
    // page-locked memory and read once the first FP32 solve is queued
    dup_h = c.pin_down.get<int>(1, s);
    *dup_h = 0;
    KI_HIP(hipEventRecord(c.ev_fork, s));
    KI_HIP(hipStreamWaitEvent(c.side, c.ev_fork, 0));
    k_dup_points<<<t.nnodes, 64, 0, c.side>>>(t.nnodes, t.is_leaf.p, t.pstart.p, t.pcount.p,
                                               t.perm_all.p, t.px.p, t.py.p,
                                               c.pin_down.dget<int>());
    KI_HIP(hipEventRecord(c.ev[5], c.side));
  }
  t.x32.alloc(len);
  const int g = ceil_div((long)len, 256);
  k_convert<double, float><<<g, 256, 0, s>>>(d_b, t.x32.p, len);
  sweeps<float>(c, t, t.x32.p, t.ndof, nrhs, nullptr, trans);
  k_convert<float, double><<<g, 256, 0, s>>>(t.x32.p, d_x, len);
  if (h) {
    t.alpha64.alloc(alen);
    KI_HIP(hipMemcpyAsync(t.alpha64.p, t.alpha_dev.p, sizeof(double) * alen,
                          hipMemcpyDeviceToDevice, s));
  }
  KI_HIP(hipEventRecord(c.ev[7], s));
  if (dup_h) {
    event_wait(c.ev[5]);
    t.dup_state = *dup_h;
  }
  const bool tc = tc_try && t.dup_state == 0;
  // the treecode's lists are built on the side stream while the first FP32
  // solve runs (the first apply waits for them)
  if (tc) treecode_prepare(c, t, c.side);
  // event pairs of the refinement's applies and corrections, and the
  // treecode's three events per apply
  c.ensure_events(7 * (size_t)(iters + 1) + 1);
  int evn = 0;
  auto mark = [&]() {
    KI_HIP(hipEventRecord(c.evpool[evn], s));
    return evn++;
  };
  std::vector<int> seg_apply, seg_solve, seg_tc;
  if (iters > 0) {
    t.r64.alloc(len);
    t.f64.alloc(len);
    if (h) t.rb64.alloc(alen);
    std::vector<double> rr, ff, hb(alen);
    for (int it = 0; it <= iters; ++it) {
      seg_apply.push_back(mark());
      if (trans) {
        kernel_apply_transpose(c, t, t.opts.pde, d_x, nrhs, t.f64.p);
      } else if (tc) {
        seg_tc.push_back(evn);
        evn += 3;
        treecode_apply(c, t, d_x, nrhs, t.f64.p, &c.evpool[seg_tc.back()]);
      } else {
        kernel_apply(c, t, t.opts.pde, d_x, nrhs, t.f64.p);
      }
      if (h) {
        k_add_ualpha<<<dim3(ceil_div(t.ndof, 256), nrhs), 256, 0, s>>>(t.ndof, h, t.bU64in.p,
                                                                      t.alpha64.p, t.f64.p);
        k_border_resid<<<dim3(h, nrhs), kTB, 0, s>>>(t.ndof, h, t.bP64in.p, d_x, t.alpha64.p,
                                                     t.rb64.p);
      }
      residual_sums(c, t, d_b, t.f64.p, t.r64.p, nrhs, &rr, &ff);
      if (h) {
        t.rb64.download(hb.data(), alen, s);
        stream_wait(s);
        for (int col = 0; col < nrhs; ++col)
          for (int i = 0; i < h; ++i) rr[col] += hb[i + (size_t)col * h] * hb[i + (size_t)col * h];
      }
      double rel = 0;  // worst column of the worst instance
      for (size_t q = 0; q < rr.size(); ++q)
        rel = std::max(rel, ff[q] > 0 ? std::sqrt(rr[q] / ff[q]) : std::sqrt(rr[q]));
      seg_apply.push_back(mark());
      if (info) {
        info->rel_residual = rel;
        info->iterations = it;
      }
      if (rel <= ro->rel_residual || it == iters) break;
      seg_solve.push_back(mark());
      k_convert<double, float><<<g, 256, 0, s>>>(t.r64.p, t.x32.p, len);
      sweeps<float>(c, t, t.x32.p, t.ndof, nrhs, h ? t.rb64.p : nullptr, trans);
      k_axpy_f32<<<g, 256, 0, s>>>(d_x, t.x32.p, len);
      if (h) k_axpy_f64<<<ceil_div((long)alen, 256), 256, 0, s>>>(t.alpha64.p, t.alpha_dev.p, alen);
      seg_solve.push_back(mark());
    }
  }
  if (h && d_alpha)
    KI_HIP(hipMemcpyAsync(d_alpha, t.alpha64.p, sizeof(double) * alen, hipMemcpyDeviceToDevice, s));
  const unsigned long long* hw = nullptr;
  if (tc) {
    hw = c.pin_down.get<unsigned long long>(3, s);
    copy_to_host(c.pin_down, s, {{t.tc_work.p, 0, 3 * sizeof(unsigned long long)}});
  }
  stream_wait(s);
  if (tc) {
    c.tim.tc_series_pairs = (double)hw[0];
    c.tim.tc_direct_pairs = (double)hw[1];
    c.tim.tc_series_terms = (double)hw[2];
  }
  auto el = [&](hipEvent_t e0, hipEvent_t e1) {
    float ms = 0;
    KI_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms * 1e-3;
  };
  c.tim.t_solve = el(c.ev[6], c.ev[7]);
  for (size_t i = 0; i + 1 < seg_solve.size(); i += 2)
    c.tim.t_solve += el(c.evpool[seg_solve[i]], c.evpool[seg_solve[i + 1]]);
  c.tim.t_refine_apply = 0;
  for (size_t i = 0; i + 1 < seg_apply.size(); i += 2)
    c.tim.t_refine_apply += el(c.evpool[seg_apply[i]], c.evpool[seg_apply[i + 1]]);
  for (int e : seg_tc) treecode_timing(c, &c.evpool[e]);
}

}  // namespace kigpu
