// Treecode apply of the Laplace double layer for the FP64 refinement
// residual: y = K mu with the entries of kernel.cpp:11-36, one density.
//
// With z = x_i - x_j as a complex number and a_j = -w_j mu_j (n_j) / (2 pi)
// (also complex), the off-diagonal entry times the density is
//   K(i, j) mu_j = (r . a_j) / |r|^2 = Re(a_j / z),
// so the far field of a box C with centre c and radius rho (half its
// diagonal) is the Cauchy series
//   sum_{j in C} a_j / (x - x_j) = (1 / (x - c)) sum_k M_k w^k,
//   M_k = sum_{j in C} a_j ((x_j - c) / rho)^k,   w = rho / (x - c).
// A box is used through its series when rho <= theta |x - c| for every target
// x of the leaf being evaluated (theta = 1/2, tested against the distance
// from c to the leaf's box); the first kTcP terms then leave a relative
// truncation below theta^P / (1 - theta) = 7e-15.  Everything else is summed
// directly, with the diagonal value 1/2 + kappa w / (4 pi) where r is exactly
// zero, as in k_apply_sym.  Boxes of levels 0 and 1 are always opened.
//
// Three kernels: k_tc_coef (centres for the acceptance test, series of every
// box of level >= 2 with at least kTcMinPts points; per apply), k_tc_plan
// (interaction lists of every leaf; once per tree), k_tc_eval (per apply).
#include <algorithm>
#include <cmath>

#include "device_util.h"
#include "entries.h"
#include "factor.h"

namespace kigpu {

namespace {

constexpr int kTcP = 48;              // series terms
constexpr double kTcTheta2 = 0.25;    // acceptance: rho^2 <= theta^2 dist^2
constexpr int kTcMinLevel = 2;        // boxes of coarser levels are always opened
constexpr int kTcMinPts = 16;         // smaller accepted boxes are summed directly
constexpr int kTcFar = 768;           // list capacities per leaf
constexpr int kTcNear = 256;
constexpr int kTcFront = 1024;        // breadth-first frontier of the list build
constexpr int kTcTB = 256;            // evaluation: 4 waves per leaf
constexpr int kTcNW = kTcTB / kWave;
constexpr double kTcFar2 = 1e30;      // padding sources, as in k_apply_sym

// Centre (x, y), radius and its inverse of every box of level >= kTcMinLevel,
// and the series of those with at least kTcMinPts points.  One block per box;
// each thread takes every 256th point, 16 powers per pass, wave sums then a
// sum over the four waves in a fixed order.
__global__ __launch_bounds__(256) void k_tc_coef(int base, const int64_t* pstart,
                                                 const int* pcount, const int* perm,
                                                 const double* corners, GeomView g,
                                                 const double* mu, double4* ctr, double2* coef) {
  __shared__ double red[4][32];
  const int node = base + blockIdx.x;
  const int cnt = pcount[node];
  const double* cc = corners + 8 * (size_t)node;
  const double hx = 0.5 * (cc[6] - cc[0]), hy = 0.5 * (cc[3] - cc[1]);
  const double cx = cc[0] + hx, cy = cc[1] + hy;
  const double rho = sqrt(hx * hx + hy * hy);
  const double irho = 1.0 / rho;
  if (threadIdx.x == 0) ctr[node] = make_double4(cx, cy, rho, irho);
  if (cnt < kTcMinPts) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int64_t p0 = pstart[node];
  for (int q = 0; q < kTcP / 16; ++q) {
    double sr[16], si[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) sr[k] = si[k] = 0.0;
    for (int j = tid; j < cnt; j += 256) {
      const int p = perm[p0 + j];
      const double s = -g.w[p] * mu[p] * kInv2Pi;
      double ar = s * g.nx[p], ai = s * g.ny[p];
      const double zr = (g.px[p] - cx) * irho, zi = (g.py[p] - cy) * irho;
      if (q > 0) {  // a zeta^(16 q)
        double er = zr, ei = zi;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const double nr = er * er - ei * ei, ni = 2.0 * er * ei;
          er = nr;
          ei = ni;
        }
        for (int t = 0; t < q; ++t) {
          const double nr = ar * er - ai * ei, ni = ar * ei + ai * er;
          ar = nr;
          ai = ni;
        }
      }
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        sr[k] += ar;
        si[k] += ai;
        const double nr = ar * zr - ai * zi, ni = ar * zi + ai * zr;
        ar = nr;
        ai = ni;
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const double a = wave_sum(sr[k]), b = wave_sum(si[k]);
      if (lane == 0) {
        red[wv][2 * k] = a;
        red[wv][2 * k + 1] = b;
      }
    }
    __syncthreads();
    if (tid < 32) {
      const double s = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
      reinterpret_cast<double*>(coef + (size_t)node * kTcP + 16 * q)[tid] = s;
    }
    __syncthreads();
  }
}

// Interaction lists of leaf B (one wave): breadth first from B's instance
// root; a box is accepted (series, or direct when small), summed directly
// when it is a leaf, or else opened.  Counts -1: a list overflowed, and the
// evaluation sums the whole instance directly for B.  work[0], work[1]:
// totals of (target, series) and (target, source) pairs over all leaves.
__global__ __launch_bounds__(64) void k_tc_plan(const int* is_leaf, const int* pcount,
                                                const int* level, const int* child,
                                                const int* inst, const double* corners,
                                                const double4* ctr, int* far_list,
                                                int* near_list, int* nfar, int* nnear,
                                                unsigned long long* work) {
  __shared__ int fr[2][kTcFront];
  const int B = blockIdx.x, lane = threadIdx.x;
  const int cntB = pcount[B];
  if (!is_leaf[B] || cntB == 0) return;
  const double* cc = corners + 8 * (size_t)B;
  const double hbx = 0.5 * (cc[6] - cc[0]), hby = 0.5 * (cc[3] - cc[1]);
  const double bx = cc[0] + hbx, by = cc[1] + hby;
  int* fl = far_list + (size_t)B * kTcFar;
  int* nl = near_list + (size_t)B * kTcNear;
  if (lane == 0) fr[0][0] = inst[B];  // the instance's root is node inst[B]
  __syncthreads();
  int nf = 0, nn = 0, nfr = 1, cur = 0;
  long long npts_near = 0;
  bool ovf = false;
  const uint64_t below = (1ull << lane) - 1;
  while (nfr > 0) {
    int nnext = 0;
    for (int i0 = 0; i0 < nfr; i0 += kWave) {
      const int i = i0 + lane;
      int kind = 0, C = -1, cnt = 0;
      if (i < nfr) {
        C = fr[cur][i];
        cnt = C >= 0 ? pcount[C] : 0;
        if (cnt > 0) {
          bool acc = false;
          if (level[C] >= kTcMinLevel) {
            const double4 q = ctr[C];
            const double dx = fmax(fabs(q.x - bx) - hbx, 0.0);
            const double dy = fmax(fabs(q.y - by) - hby, 0.0);
            acc = q.z * q.z <= kTcTheta2 * (dx * dx + dy * dy);
          }
          if (acc) kind = cnt >= kTcMinPts ? 1 : 2;
          else kind = is_leaf[C] ? 2 : 3;
        }
      }
      const uint64_t mf = __ballot(kind == 1), mn = __ballot(kind == 2), me = __ballot(kind == 3);
      if (kind == 1) {
        const int pos = nf + __popcll(mf & below);
        if (pos < kTcFar) fl[pos] = C;
      } else if (kind == 2) {
        const int pos = nn + __popcll(mn & below);
        if (pos < kTcNear) nl[pos] = C;
        npts_near += cnt;
      } else if (kind == 3) {
        const int pos = nnext + 4 * __popcll(me & below);
        if (pos + 4 <= kTcFront)
          for (int q = 0; q < 4; ++q) fr[cur ^ 1][pos + q] = child[4 * (size_t)C + q];
      }
      nf += __popcll(mf);
      nn += __popcll(mn);
      nnext += 4 * __popcll(me);
    }
    if (nf > kTcFar || nn > kTcNear || nnext > kTcFront) {
      ovf = true;
      break;
    }
    __syncthreads();
    cur ^= 1;
    nfr = nnext;
  }
  npts_near = wave_sum((double)npts_near);
  if (lane == 0) {
    nfar[B] = ovf ? -1 : nf;
    nnear[B] = ovf ? -1 : nn;
    if (!ovf) {
      atomicAdd(work, (unsigned long long)cntB * (unsigned long long)nf);
      atomicAdd(work + 1, (unsigned long long)cntB * (unsigned long long)npts_near);
    }
  }
}

// per-wave LDS handshake: the wave's lanes see each other's LDS writes
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct TcSrc {
  double x[kTcNW][kWave], y[kTcNW][kWave], ax[kTcNW][kWave], ay[kTcNW][kWave],
      dg[kTcNW][kWave];
};

// u += sum over n sources (point index src(j)) of K(target, j) mu_j, the
// sources staged 64 at a time in the wave's LDS rows
template <typename F>
__device__ __forceinline__ void tc_direct(TcSrc& S, int wv, int lane, const GeomView& g,
                                          const double* mu, double zx, double zy, int n, F src,
                                          double& u) {
  for (int j0 = 0; j0 < n; j0 += kWave) {
    const int j = j0 + lane;
    wave_sync();
    if (j < n) {
      const int p = src(j);
      const double w = g.w[p], m = mu[p];
      const double s = -w * m * kInv2Pi;
      S.x[wv][lane] = g.px[p];
      S.y[wv][lane] = g.py[p];
      S.ax[wv][lane] = s * g.nx[p];
      S.ay[wv][lane] = s * g.ny[p];
      S.dg[wv][lane] = (0.5 + 0.5 * g.kappa[p] * w * kInv2Pi) * m;
    } else {
      S.x[wv][lane] = kTcFar2;
      S.y[wv][lane] = kTcFar2;
      S.ax[wv][lane] = 0.0;
      S.ay[wv][lane] = 0.0;
      S.dg[wv][lane] = 0.0;
    }
    wave_sync();
    const int m = min(kWave, n - j0);
    for (int k = 0; k < m; ++k) {
      const double rx = zx - S.x[wv][k], ry = zy - S.y[wv][k];
      const double rr = fma(rx, rx, ry * ry);
      const double v = fma(rx, S.ax[wv][k], ry * S.ay[wv][k]) * recip1(rr);
      u += rr == 0.0 ? S.dg[wv][k] : v;
    }
  }
}

// y over the points of leaf B (one block of 4 waves, a target per lane):
// wave w takes every 4th series box and every 4th direct box of B's lists;
// the four partial sums are added in wave order.
__global__ __launch_bounds__(kTcTB) void k_tc_eval(const int* is_leaf, const int* pcount,
                                                   const int64_t* pstart, const int* perm,
                                                   const int* inst, const int64_t* inst_pt,
                                                   GeomView g, const double* mu,
                                                   const double4* ctr, const double2* coef,
                                                   const int* far_list, const int* near_list,
                                                   const int* nfar, const int* nnear, double* y) {
  __shared__ double2 s_cf[kTcNW][kTcP];
  __shared__ TcSrc S;
  __shared__ double s_u[kTcNW][kWave];
  const int B = blockIdx.x;
  const int cntB = pcount[B];
  if (!is_leaf[B] || cntB == 0) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int64_t p0B = pstart[B];
  const int nf = nfar[B], nn = nnear[B];
  const int* fl = far_list + (size_t)B * kTcFar;
  const int* nl = near_list + (size_t)B * kTcNear;
  for (int t0 = 0; t0 < cntB; t0 += kWave) {
    const int ti = t0 + lane;
    const bool tv = ti < cntB;
    const int tp = perm[p0B + (tv ? ti : 0)];
    const double zx = g.px[tp], zy = g.py[tp];
    double u = 0;
    if (nf >= 0) {
      for (int e = wv; e < nf; e += kTcNW) {
        const int C = fl[e];
        const double4 q = ctr[C];
        wave_sync();
        if (lane < kTcP) s_cf[wv][lane] = coef[(size_t)C * kTcP + lane];
        wave_sync();
        const double dx = zx - q.x, dy = zy - q.y;
        const double inv = recip1(fma(dx, dx, dy * dy));
        const double wr = q.z * dx * inv, wi = -q.z * dy * inv;
        double sr = s_cf[wv][kTcP - 1].x, si = s_cf[wv][kTcP - 1].y;
        for (int k = kTcP - 2; k >= 0; --k) {
          const double2 m = s_cf[wv][k];
          const double nr = fma(sr, wr, fma(-si, wi, m.x));
          const double ni = fma(sr, wi, fma(si, wr, m.y));
          sr = nr;
          si = ni;
        }
        u = fma(sr * wr - si * wi, q.w, u);
      }
      for (int e = wv; e < nn; e += kTcNW) {
        const int D = nl[e];
        const int64_t pd = pstart[D];
        tc_direct(S, wv, lane, g, mu, zx, zy, pcount[D], [&](int j) { return perm[pd + j]; }, u);
      }
    } else {  // lists overflowed: the whole instance, directly
      const int b = inst[B];
      const int a0 = (int)inst_pt[b], n = (int)(inst_pt[b + 1] - inst_pt[b]);
      const int per = (n + kTcNW - 1) / kTcNW;
      const int lo = min(n, wv * per), hi = min(n, lo + per);
      tc_direct(S, wv, lane, g, mu, zx, zy, hi - lo, [&](int j) { return a0 + lo + j; }, u);
    }
    s_u[wv][lane] = u;
    __syncthreads();
    if (wv == 0 && tv) y[tp] = (s_u[0][lane] + s_u[1][lane]) + (s_u[2][lane] + s_u[3][lane]);
    __syncthreads();
  }
}

}  // namespace

bool treecode_usable(const TreeData& t) { return t.nnodes > 0 && t.sd == 1; }

void treecode_apply(Context& c, const TreeData& t, const double* d_mu, double* d_y) {
  hipStream_t s = c.stream;
  const int nn = t.nnodes;
  const int base = t.nlevels > kTcMinLevel ? (int)t.level_off[kTcMinLevel] : nn;
  const GeomView g{t.px.p, t.py.p, t.nx.p, t.ny.p, t.w.p, t.kappa.p, t.npts, t.sd};
  const bool plan = !t.tc_ready;
  if (plan) {
    t.tc_ctr.alloc((size_t)nn);
    t.tc_coef.alloc((size_t)nn * kTcP);
    t.tc_far.alloc((size_t)nn * kTcFar);
    t.tc_near.alloc((size_t)nn * kTcNear);
    t.tc_nfar.alloc(nn);
    t.tc_nnear.alloc(nn);
    t.tc_work.alloc(2);
    t.tc_work.zero(s);
  }
  KI_HIP(hipEventRecord(c.ev[12], s));
  if (nn > base)
    k_tc_coef<<<nn - base, 256, 0, s>>>(base, t.pstart.p, t.pcount.p, t.perm_all.p,
                                        t.corners.p, g, d_mu, t.tc_ctr.p, t.tc_coef.p);
  if (plan) {
    k_tc_plan<<<nn, kWave, 0, s>>>(t.is_leaf.p, t.pcount.p, t.level.p, t.child.p, t.inst.p,
                                   t.corners.p, t.tc_ctr.p, t.tc_far.p, t.tc_near.p,
                                   t.tc_nfar.p, t.tc_nnear.p, t.tc_work.p);
    t.tc_ready = true;
  }
  KI_HIP(hipEventRecord(c.ev[14], s));
  k_tc_eval<<<nn, kTcTB, 0, s>>>(t.is_leaf.p, t.pcount.p, t.pstart.p, t.perm_all.p, t.inst.p,
                                 t.d_inst_pt.p, g, d_mu, t.tc_ctr.p, t.tc_coef.p, t.tc_far.p,
                                 t.tc_near.p, t.tc_nfar.p, t.tc_nnear.p, d_y);
  KI_HIP(hipEventRecord(c.ev[13], s));
  KI_HIP(hipGetLastError());
  stream_wait(s);
  float ms = 0, ms_eval = 0;
  KI_HIP(hipEventElapsedTime(&ms, c.ev[12], c.ev[13]));
  KI_HIP(hipEventElapsedTime(&ms_eval, c.ev[14], c.ev[13]));
  c.tim.tc_launches += 1;
  c.tim.tc_kernel_s += ms * 1e-3;
  c.tim.tc_eval_s += ms_eval * 1e-3;
}

}  // namespace kigpu
