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
// zero, as in k_apply_sym.  Boxes of levels 0 to 2 are always opened.
//
// Kernels: k_tc_ctr (box centres and radii; once per tree), k_tc_coef
// (series of every box of level >= 3 with at least kTcMinPts points; per
// apply), k_tc_plan (interaction lists of every leaf; once per tree),
// k_tc_eval (per apply).
#include <algorithm>
#include <cmath>

#include "device_util.h"
#include "entries.h"
#include "factor.h"

namespace kigpu {

namespace {

constexpr int kTcP = 48;              // series terms
constexpr double kTcTheta2 = 0.25;    // acceptance: rho^2 <= theta^2 dist^2

// Terms a box's series needs for a leaf: with r = rho / dist(centre, leaf
// box) <= 1/2 the tail after P terms is below r^P / (1 - r) <= 2 r^P of
// sum|a_j| / |x - c|.  P = 48 at the acceptance limit gives 2^-47; for
// smaller r, P = 96 / log2(1 / r^2) (rounded up to even, with a margin for
// the single-precision logarithm) keeps 2 r^P <= 2^-47 as well.
__device__ __forceinline__ int tc_nterms(double rho2, double g2) {
  if (!(g2 > 0.0)) return kTcP;
  const float x = -__log2f((float)(rho2 / g2));  // log2(1 / r^2) >= 2
  if (!(x > 2.0f)) return kTcP;
  int P = (int)ceilf(96.0f * 1.001f / x);
  return min(kTcP, max(2, P + (P & 1)));
}
constexpr int kTcMinLevel = 3;        // boxes of coarser levels are always opened
constexpr int kTcMinPts = 16;         // smaller accepted boxes are summed directly
constexpr int kTcFar = 768;           // list capacities per leaf
constexpr int kTcNear = 256;
constexpr int kTcFront = 1024;        // breadth-first frontier of the list build
constexpr int kTcTB = 256;            // evaluation: 4 waves per leaf
constexpr int kTcNW = kTcTB / kWave;
constexpr double kTcFar2 = 1e30;      // padding sources, as in k_apply_sym

// per-wave LDS handshake: the wave's lanes see each other's LDS writes
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Centre (x, y), radius and its inverse of every box (once per tree).
__global__ void k_tc_ctr(int nn, const double* corners, double4* ctr) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nn) return;
  const double* cc = corners + 8 * (size_t)node;
  const double hx = 0.5 * (cc[6] - cc[0]), hy = 0.5 * (cc[3] - cc[1]);
  const double rho = sqrt(hx * hx + hy * hy);
  ctr[node] = make_double4(cc[0] + hx, cc[1] + hy, rho, 1.0 / rho);
}

// Series of the boxes in the work list (boxes of level >= kTcMinLevel with at
// least kTcMinPts points).  A block takes four boxes of at most kTcSmall
// points, one per wave, or one larger box over its four waves (entry .y ==
// -2).  Each lane sums its points' a_j zeta_j^k for all k in registers (one
// pass over the points, loads kept in flight two points at a time); the 64
// partial sums of each power go through LDS, lanes 4k .. 4k + 3 add a
// quarter each and a quad exchange completes the sum; a large box adds its
// four waves in wave order.  Fixed summation order throughout.
constexpr int kTcSmall = 256;

__global__ __launch_bounds__(kTcTB) void k_tc_coef(const int4* work, const int64_t* pstart,
                                                   const int* pcount, const int* perm,
                                                   GeomView g, const double* mu, int ldm,
                                                   const double4* ctr, double2* coef,
                                                   size_t cstride) {
  __shared__ double2 part[kTcNW][16][kWave];
  __shared__ double2 wsum[kTcNW][kTcP];
  mu += (size_t)blockIdx.y * ldm;  // one density per grid row
  coef += (size_t)blockIdx.y * cstride;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int4 wk = work[blockIdx.x];
  const bool big = wk.y == -2;
  const int node = big ? wk.x : (wv == 0 ? wk.x : wv == 1 ? wk.y : wv == 2 ? wk.z : wk.w);
  const int first = big ? tid : lane, stride = big ? kTcTB : kWave;
  double sr[kTcP], si[kTcP];
#pragma unroll
  for (int k = 0; k < kTcP; ++k) sr[k] = si[k] = 0.0;
  if (node >= 0) {
    const int cnt = pcount[node];
    const int64_t p0 = pstart[node];
    const double4 q = ctr[node];
    for (int j = first; j < cnt; j += 2 * stride) {
      const int j2 = j + stride;
      const int pa = perm[p0 + j], pb = perm[p0 + (j2 < cnt ? j2 : j)];
      const double wa = g.w[pa], ma = mu[pa], xa = g.px[pa], ya = g.py[pa], na = g.nx[pa],
                   nb_ = g.ny[pa];
      const double wb = g.w[pb], mb = mu[pb], xb = g.px[pb], yb = g.py[pb], nxb = g.nx[pb],
                   nyb = g.ny[pb];
      for (int h = 0; h < 2; ++h) {
        if (h == 1 && j2 >= cnt) break;
        const double s = -(h ? wb * mb : wa * ma) * kInv2Pi;
        double ar = s * (h ? nxb : na), ai = s * (h ? nyb : nb_);
        const double zr = ((h ? xb : xa) - q.x) * q.w, zi = ((h ? yb : ya) - q.y) * q.w;
#pragma unroll
        for (int k = 0; k < kTcP; ++k) {
          sr[k] += ar;
          si[k] += ai;
          const double nr = ar * zr - ai * zi, ni = ar * zi + ai * zr;
          ar = nr;
          ai = ni;
        }
      }
    }
  }
  // per wave: 64 partials of each power -> one sum, 16 powers at a time
#pragma unroll
  for (int c = 0; c < kTcP / 16; ++c) {
    wave_sync();
#pragma unroll
    for (int k = 0; k < 16; ++k) part[wv][k][lane] = make_double2(sr[16 * c + k], si[16 * c + k]);
    wave_sync();
    const int k = lane >> 2, r = lane & 3;
    double tr = 0, ti = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const double2 v = part[wv][k][16 * r + i];
      tr += v.x;
      ti += v.y;
    }
    tr += dpp::mov<dpp::kXor1>(tr);
    ti += dpp::mov<dpp::kXor1>(ti);
    tr += dpp::mov<dpp::kXor2>(tr);
    ti += dpp::mov<dpp::kXor2>(ti);
    if (r == 0) {
      if (big) wsum[wv][16 * c + k] = make_double2(tr, ti);
      else if (node >= 0) coef[(size_t)node * kTcP + 16 * c + k] = make_double2(tr, ti);
    }
  }
  if (big) {
    __syncthreads();
    if (tid < kTcP) {
      const double2 a = wsum[0][tid], b = wsum[1][tid], c = wsum[2][tid], d = wsum[3][tid];
      coef[(size_t)node * kTcP + tid] = make_double2((a.x + b.x) + (c.x + d.x),
                                                      (a.y + b.y) + (c.y + d.y));
    }
  }
}

// Interaction lists of leaf B (one wave): breadth first from B's instance
// root; a box is accepted (series, or direct when small), summed directly
// when it is a leaf, or else opened.  Counts -1: a list overflowed, and the
// evaluation sums the whole instance directly for B.  work[0], work[1]:
// totals of (target, series) and (target, source) pairs over all leaves;
// work[2]: the series terms those pairs sum.
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
  long long npts_near = 0, nterms = 0;
  bool ovf = false;
  const uint64_t below = (1ull << lane) - 1;
  while (nfr > 0) {
    int nnext = 0;
    for (int i0 = 0; i0 < nfr; i0 += kWave) {
      const int i = i0 + lane;
      int kind = 0, C = -1, cnt = 0, pterms = 0;
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
            if (acc) pterms = tc_nterms(q.z * q.z, dx * dx + dy * dy);
          }
          if (acc) kind = cnt >= kTcMinPts ? 1 : 2;
          else kind = is_leaf[C] ? 2 : 3;
        }
      }
      const uint64_t mf = __ballot(kind == 1), mn = __ballot(kind == 2), me = __ballot(kind == 3);
      if (kind == 1) {
        const int pos = nf + __popcll(mf & below);
        if (pos < kTcFar) fl[pos] = C;
        nterms += pterms;
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
  nterms = wave_sum((double)nterms);
  if (lane == 0) {
    nfar[B] = ovf ? -1 : nf;
    nnear[B] = ovf ? -1 : nn;
    if (!ovf) {
      atomicAdd(work, (unsigned long long)cntB * (unsigned long long)nf);
      atomicAdd(work + 1, (unsigned long long)cntB * (unsigned long long)npts_near);
      atomicAdd(work + 2, (unsigned long long)cntB * (unsigned long long)nterms);
    }
  }
}

// near-field sources in LDS: all of a leaf's near boxes at once when their
// points fit, else 64 per wave at a time
constexpr int kTcNearPts = 512;
struct TcSrc {
  double x[kTcNearPts], y[kTcNearPts], ax[kTcNearPts], ay[kTcNearPts], dg[kTcNearPts];
};

// source point p into slot i of the LDS store
__device__ __forceinline__ void tc_stage(TcSrc& S, int i, const GeomView& g, const double* mu,
                                         int p) {
  const double w = g.w[p], m = mu[p];
  const double s = -w * m * kInv2Pi;
  S.x[i] = g.px[p];
  S.y[i] = g.py[p];
  S.ax[i] = s * g.nx[p];
  S.ay[i] = s * g.ny[p];
  S.dg[i] = (0.5 + 0.5 * g.kappa[p] * w * kInv2Pi) * m;
}

// u += sum over the LDS sources [lo, hi) of K(target, j) mu_j
__device__ __forceinline__ void tc_sum(const TcSrc& S, int lo, int hi, double zx, double zy,
                                       double& u) {
  for (int k = lo; k < hi; ++k) {
    const double rx = zx - S.x[k], ry = zy - S.y[k];
    const double rr = fma(rx, rx, ry * ry);
    const double v = fma(rx, S.ax[k], ry * S.ay[k]) * recip1(rr);
    u += rr == 0.0 ? S.dg[k] : v;
  }
}

// u += sum over n sources (point index src(j)) of K(target, j) mu_j, the
// sources staged 64 at a time in the wave's LDS slots [64 wv, 64 wv + 64)
template <typename F>
__device__ __forceinline__ void tc_direct(TcSrc& S, int wv, int lane, const GeomView& g,
                                          const double* mu, double zx, double zy, int n, F src,
                                          double& u) {
  const int o = kWave * wv;
  for (int j0 = 0; j0 < n; j0 += kWave) {
    const int j = j0 + lane;
    wave_sync();
    if (j < n) tc_stage(S, o + lane, g, mu, src(j));
    wave_sync();
    tc_sum(S, o, o + min(kWave, n - j0), zx, zy, u);
  }
}

// y over the points of leaf B (one block of 4 waves, a target per lane):
// wave w takes every 4th series box of B's list and a quarter of the points
// of its near boxes; the four partial sums are added in wave order.
__global__ __launch_bounds__(kTcTB) void k_tc_eval(const int* is_leaf, const int* pcount,
                                                   const int64_t* pstart, const int* perm,
                                                   const int* inst, const int64_t* inst_pt,
                                                   GeomView g, const double* mu, int ldm,
                                                   const double4* ctr, const double2* coef,
                                                   size_t cstride, const int* far_list,
                                                   const int* near_list, const int* nfar,
                                                   const int* nnear, double* y, int ldy,
                                                   const double* corners) {
  __shared__ double2 s_cf[kTcNW][kTcP];
  __shared__ TcSrc S;
  __shared__ double s_u[kTcNW][kWave];
  __shared__ int fl[kTcFar], nl[kTcNear];
  const int B = blockIdx.x;
  const int cntB = pcount[B];
  if (!is_leaf[B] || cntB == 0) return;
  mu += (size_t)blockIdx.y * ldm;  // one density per grid row
  coef += (size_t)blockIdx.y * cstride;
  y += (size_t)blockIdx.y * ldy;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int64_t p0B = pstart[B];
  const int nf = nfar[B], nn = nnear[B];
  // the leaf's box, for each series box's distance to it
  const double* cb = corners + 8 * (size_t)B;
  const double hbx = 0.5 * (cb[6] - cb[0]), hby = 0.5 * (cb[3] - cb[1]);
  const double bx = cb[0] + hbx, by = cb[1] + hby;
  __shared__ int s_off[kWave + 1];
  for (int e = tid; e < nf; e += kTcTB) fl[e] = far_list[(size_t)B * kTcFar + e];
  for (int e = tid; e < nn; e += kTcTB) nl[e] = near_list[(size_t)B * kTcNear + e];
  __syncthreads();
  // the near boxes' points, staged once for every target of B when at most
  // 64 boxes with at most kTcNearPts points in all
  int nt = -1;
  if (nf >= 0 && nn <= kWave) {
    if (wv == 0) {
      int inc = lane < nn ? pcount[nl[lane]] : 0;
      for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(inc, o);
        if (lane >= o) inc += t;
      }
      s_off[lane + 1] = inc;
      if (lane == 0) s_off[0] = 0;
    }
    __syncthreads();
    if (s_off[nn] <= kTcNearPts) {
      nt = s_off[nn];
      for (int idx = tid; idx < nt; idx += kTcTB) {
        int lo = 0, hi = nn;  // s_off[lo] <= idx < s_off[hi]
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_off[mid] <= idx) lo = mid;
          else hi = mid;
        }
        tc_stage(S, idx, g, mu, perm[pstart[nl[lo]] + (idx - s_off[lo])]);
      }
      __syncthreads();
    }
  }
  for (int t0 = 0; t0 < cntB; t0 += kWave) {
    const int ti = t0 + lane;
    const bool tv = ti < cntB;
    const int tp = perm[p0B + (tv ? ti : 0)];
    const double zx = g.px[tp], zy = g.py[tp];
    double u = 0;
    if (nf >= 0) {
      // the next box's centre and series are loaded while this one is summed
      double4 qn = make_double4(0.0, 0.0, 0.0, 0.0);
      double2 cn = make_double2(0.0, 0.0);
      if (wv < nf) {
        qn = ctr[fl[wv]];
        if (lane < kTcP) cn = coef[(size_t)fl[wv] * kTcP + lane];
      }
      for (int e = wv; e < nf; e += kTcNW) {
        const double4 q = qn;
        wave_sync();
        if (lane < kTcP) s_cf[wv][lane] = cn;
        wave_sync();
        if (e + kTcNW < nf) {
          const int C2 = fl[e + kTcNW];
          qn = ctr[C2];
          if (lane < kTcP) cn = coef[(size_t)C2 * kTcP + lane];
        }
        // terms this box needs for this leaf (the same for the whole wave)
        const double gx = fmax(fabs(q.x - bx) - hbx, 0.0), gy = fmax(fabs(q.y - by) - hby, 0.0);
        // the same for every lane (box and leaf are): made scalar, so that
        // the Horner loop below runs as a uniform loop
        const int P = __builtin_amdgcn_readfirstlane(tc_nterms(q.z * q.z, gx * gx + gy * gy));
        const double dx = zx - q.x, dy = zy - q.y;
        const double inv = recip1(fma(dx, dx, dy * dy));
        const double wr = q.z * dx * inv, wi = -q.z * dy * inv;
        // sum_{k<P} M_k w^k = E(w^2) + w O(w^2): two Horner chains of half
        // the length over the even and the odd terms
        const double vr = wr * wr - wi * wi, vi = 2.0 * wr * wi;
        double er = s_cf[wv][P - 2].x, ei = s_cf[wv][P - 2].y;
        double orr = s_cf[wv][P - 1].x, oi = s_cf[wv][P - 1].y;
#pragma unroll 4
        for (int k = P - 4; k >= 0; k -= 2) {
          const double2 me = s_cf[wv][k], mo = s_cf[wv][k + 1];
          const double ner = fma(er, vr, fma(-ei, vi, me.x));
          const double nei = fma(er, vi, fma(ei, vr, me.y));
          const double nor = fma(orr, vr, fma(-oi, vi, mo.x));
          const double noi = fma(orr, vi, fma(oi, vr, mo.y));
          er = ner;
          ei = nei;
          orr = nor;
          oi = noi;
        }
        const double sr = er + (orr * wr - oi * wi), si = ei + (orr * wi + oi * wr);
        u = fma(sr * wr - si * wi, q.w, u);
      }
      if (nt >= 0) {  // staged: wave w sums a quarter of the points
        const int per = (nt + kTcNW - 1) / kTcNW;
        const int lo = min(nt, wv * per), hi = min(nt, lo + per);
        tc_sum(S, lo, hi, zx, zy, u);
      } else {
        for (int e = wv; e < nn; e += kTcNW) {
          const int D = nl[e];
          const int64_t pd = pstart[D];
          tc_direct(S, wv, lane, g, mu, zx, zy, pcount[D], [&](int j) { return perm[pd + j]; },
                    u);
        }
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

// Everything the applies reuse for the tree's lifetime: box centres, the
// series work list and the interaction lists, on stream st; ev[15] marks
// their completion.
void treecode_prepare(Context& c, const TreeData& t, hipStream_t st) {
  if (t.tc_ready) return;
  const int nn = t.nnodes;
  const int base = t.nlevels > kTcMinLevel ? (int)t.level_off[kTcMinLevel] : nn;
  t.tc_ctr.alloc((size_t)nn);
  t.tc_far.alloc((size_t)nn * kTcFar);
  t.tc_near.alloc((size_t)nn * kTcNear);
  t.tc_nfar.alloc(nn);
  t.tc_nnear.alloc(nn);
  t.tc_work.alloc(3);
  t.tc_work.zero(st);
  k_tc_ctr<<<ceil_div(nn, 256), 256, 0, st>>>(nn, t.corners.p, t.tc_ctr.p);
  // series work list: boxes of at most kTcSmall points four to a block,
  // larger ones a block each (the node's point counts are mirrored on the
  // host by the tree build)
  std::vector<int4> wl;
  int4 cur = make_int4(-1, -1, -1, -1);
  int fill = 0;
  for (int x = base; x < nn; ++x) {
    const int cnt = t.h_pcount[x];
    if (cnt < kTcMinPts) continue;
    if (cnt > kTcSmall) {
      wl.push_back(make_int4(x, -2, -1, -1));
      continue;
    }
    int* slot = fill == 0 ? &cur.x : fill == 1 ? &cur.y : fill == 2 ? &cur.z : &cur.w;
    *slot = x;
    if (++fill == 4) {
      wl.push_back(cur);
      cur = make_int4(-1, -1, -1, -1);
      fill = 0;
    }
  }
  if (fill > 0) wl.push_back(cur);
  t.tc_nwork = (int)wl.size();
  t.tc_wlist.upload(wl, st);
  k_tc_plan<<<nn, kWave, 0, st>>>(t.is_leaf.p, t.pcount.p, t.level.p, t.child.p, t.inst.p,
                                  t.corners.p, t.tc_ctr.p, t.tc_far.p, t.tc_near.p, t.tc_nfar.p,
                                  t.tc_nnear.p, t.tc_work.p);
  KI_HIP(hipEventRecord(c.ev[15], st));
  KI_HIP(hipGetLastError());
  t.tc_ready = true;
}

void treecode_apply(Context& c, const TreeData& t, const double* d_mu, int nrhs, double* d_y,
                    const hipEvent_t* tev) {
  hipStream_t s = c.stream;
  const int nn = t.nnodes;
  const GeomView g{t.px.p, t.py.p, t.nx.p, t.ny.p, t.w.p, t.kappa.p, t.npts, t.sd};
  if (!t.tc_ready) treecode_prepare(c, t, s);
  const size_t cstride = (size_t)nn * kTcP;
  t.tc_coef.alloc(cstride * nrhs);
  KI_HIP(hipStreamWaitEvent(s, c.ev[15], 0));  // the lists may come from the side stream
  if (tev) KI_HIP(hipEventRecord(tev[0], s));
  if (t.tc_nwork > 0)
    k_tc_coef<<<dim3(t.tc_nwork, nrhs), kTcTB, 0, s>>>(t.tc_wlist.p, t.pstart.p, t.pcount.p,
                                                      t.perm_all.p, g, d_mu, t.ndof, t.tc_ctr.p,
                                                      t.tc_coef.p, cstride);
  if (tev) KI_HIP(hipEventRecord(tev[1], s));
  k_tc_eval<<<dim3(nn, nrhs), kTcTB, 0, s>>>(t.is_leaf.p, t.pcount.p, t.pstart.p, t.perm_all.p,
                                             t.inst.p, t.d_inst_pt.p, g, d_mu, t.ndof, t.tc_ctr.p,
                                             t.tc_coef.p, cstride, t.tc_far.p, t.tc_near.p,
                                             t.tc_nfar.p, t.tc_nnear.p, d_y, t.ndof, t.corners.p);
  if (tev) KI_HIP(hipEventRecord(tev[2], s));
  KI_HIP(hipGetLastError());
  c.tim.tc_launches += 1;
}

void treecode_timing(Context& c, const hipEvent_t* tev) {
  float ms = 0, ms_eval = 0;
  KI_HIP(hipEventElapsedTime(&ms, tev[0], tev[2]));
  KI_HIP(hipEventElapsedTime(&ms_eval, tev[1], tev[2]));
  c.tim.tc_kernel_s += ms * 1e-3;
  c.tim.tc_eval_s += ms_eval * 1e-3;
}

}  // namespace kigpu
