// Quadtree construction on the GPU, bit-exact with the reference's
// sequential point insertion (kern_interp/quadtree/quadtree.cpp:72-285) and
// neighbor lists (:27-69, :119-161).
//
// The reference inserts points one at a time; a leaf splits as soon as it
// holds more than leaf_dofs DOFs and a new child that already overflows
// splits at once, depth first.  Children are appended to their level when
// their parent splits.  The same tree and node order come out of a
// level-synchronous construction:
//   * a box splits iff its final DOF count exceeds leaf_dofs;
//   * it splits while inserting point T_eff = max(T, T_eff(parent)), where T
//     is the index of its (floor(leaf_dofs/sd)+1)-th point (points of a box
//     are kept in ascending index order);
//   * boxes that split during the same insertion do so in depth-first
//     pre-order, i.e. in lexicographic order of their child-slot paths;
// so the nodes of level l+1 are the 4 children (slot order BL, TL, TR, BR) of
// the splitting nodes of level l sorted by (T_eff, path).
// Every coordinate is produced by the same IEEE double operations as the
// reference: midpoint ((c6 - c0) / 2.0) + c0, side / 2.0, and the neighbor
// distance sqrt(dx*dx + dy*dy) without contraction.
#include <algorithm>
#include <cmath>

#include <climits>

#include "device_util.h"
#include "tree.h"

namespace kigpu {

namespace {

constexpr int kTB = 256;

// ---- root box: min/max pooled over x and y (quadtree.cpp:80-88) ----------
constexpr int kMMBlocks = 120;

__global__ void k_minmax_part(const double* px, const double* py, int n, double* part) {
  __shared__ double smin[kTB], smax[kTB];
  double mn = px[0], mx = px[0];
  for (int i = blockIdx.x * kTB + threadIdx.x; i < n; i += gridDim.x * kTB) {
    double a = px[i], b = py[i];
    mn = fmin(mn, fmin(a, b));
    mx = fmax(mx, fmax(a, b));
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int st = kTB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + st]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = smin[0];
    part[2 * blockIdx.x + 1] = smax[0];
  }
}

// one wave: min and max are exact, so the order of the reduction is free
__global__ void k_minmax_final(const double* part, int np, double* out) {
  const int l = threadIdx.x;
  double mn = part[0], mx = part[1];
  for (int i = l; i < np; i += kWave) {
    mn = fmin(mn, part[2 * i]);
    mx = fmax(mx, part[2 * i + 1]);
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    mn = fmin(mn, __shfl_xor(mn, o));
    mx = fmax(mx, __shfl_xor(mx, o));
  }
  if (l == 0) {
    out[0] = mn;
    out[1] = mx;
  }
}

__global__ void k_iota(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

__global__ void k_fill_i32(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// per-instance root boxes: block i reduces points [pt_off[i], pt_off[i+1])
__global__ void k_minmax_seg(const double* px, const double* py, const int64_t* pt_off,
                             double* out) {
  __shared__ double smin[kTB], smax[kTB];
  const int64_t a = pt_off[blockIdx.x], b = pt_off[blockIdx.x + 1];
  double mn = px[a], mx = px[a];
  for (int64_t i = a + threadIdx.x; i < b; i += kTB) {
    double u = px[i], v = py[i];
    mn = fmin(mn, fmin(u, v));
    mx = fmax(mx, fmax(u, v));
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int st = kTB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + st]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = smin[0];
    out[2 * blockIdx.x + 1] = smax[0];
  }
}

// pnode of every point = its instance (the instance's root is node i)
__global__ void k_fill_inst(const int64_t* pt_off, int* pnode) {
  const int64_t a = pt_off[blockIdx.x], b = pt_off[blockIdx.x + 1];
  for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x) pnode[i] = blockIdx.x;
}

__device__ __forceinline__ void node_mid(const double* c, double* midx, double* midy) {
#pragma clang fp contract(off)
  *midx = ((c[6] - c[0]) / 2.0) + c[0];
  *midy = ((c[3] - c[1]) / 2.0) + c[1];
}

__device__ __forceinline__ int quadrant(double x, double y, double midx, double midy) {
  // quadtree.cpp:177-185: BL=0, TL=1, TR=2, BR=3 with strict '<'
  if (x < midx) return y < midy ? 0 : 1;
  return y < midy ? 3 : 2;
}

// per node of level l: split flag and effective split time
__global__ void k_split_info(int c, const int* pcount, const int64_t* pstart, const int* perm,
                             const int* parent_local, const int* parent_st, int sd, int leaf,
                             int* split, int* st) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c) return;
  int cnt = pcount[i];
  int sp = (long)cnt * sd > leaf;
  split[i] = sp;
  if (!sp) {
    st[i] = -1;
    return;
  }
  int t = perm[pstart[i] + leaf / sd];
  if (parent_st) t = max(t, parent_st[parent_local[i]]);
  st[i] = t;
}

// Rank of each splitting node among the splitting nodes of its instance by
// (split point, child-slot path).  The nodes of a level are grouped by
// instance (segment b = nodes [seg[b], seg[b+1])); block b sorts its
// segment with a bitonic sort of (st << 40 | path) keys.  Needs segments of
// at most 4096 nodes, paths of at most 20 levels and split points below
// 2^24; k_rank_pairs covers the rest.  seg_cnt[b] receives the segment's
// number of splitting nodes.
constexpr int kSortMax = 4096;

// The block first computes its segment's split flags and split times (the
// work of k_split_info); with a single instance (one block) it also does the
// segment scan of k_seg_scan at the end.
__global__ __launch_bounds__(1024) void k_rank_sort(const int64_t* seg, const int* pcount,
                                                    const int64_t* pstart, const int* perm,
                                                    const int* parent_local,
                                                    const int* parent_st, int sd, int leaf,
                                                    int* split, int* st, const uint64_t* path,
                                                    int* rank, int* seg_cnt, int* seg_split,
                                                    int64_t* next_seg, int* xfer,
                                                    int64_t* rank_pstart) {
  __shared__ uint64_t key[kSortMax];
  __shared__ int val[kSortMax];
  __shared__ int s_wtot[1024 / kWave];
  __shared__ int s_cnt;
  const int base = (int)seg[blockIdx.x];
  const int c = (int)seg[blockIdx.x + 1] - base;
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  // split flags and times; the splitting nodes' keys compacted in node order
  for (int i0 = 0; i0 < c; i0 += 1024) {
    const int i = i0 + threadIdx.x;
    const int x = base + i;
    int sp = 0, t = -1;
    if (i < c) {
      sp = (long)pcount[x] * sd > leaf;
      if (sp) {
        t = perm[pstart[x] + leaf / sd];
        if (parent_st) t = max(t, parent_st[parent_local[x]]);
      }
      split[x] = sp;
      st[x] = t;
      if (!sp) rank[x] = -1;
    }
    const unsigned long long m = __ballot(sp);
    if (lane == 0) s_wtot[wv] = __popcll(m);
    __syncthreads();
    int o = s_cnt;
    for (int w = 0; w < wv; ++w) o += s_wtot[w];
    if (sp) {
      const int q = o + __popcll(m & ((1ull << lane) - 1));
      key[q] = ((uint64_t)t << 40) | path[x];
      val[q] = i;
    }
    __syncthreads();
    if (threadIdx.x == 1023) s_cnt = o + __popcll(m);
    __syncthreads();
  }
  const int ns = s_cnt;
  int P = 1;
  while (P < ns) P <<= 1;
  for (int i = ns + threadIdx.x; i < P; i += 1024) {
    key[i] = ~0ull;
    val[i] = -1;
  }
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < P; i += 1024) {
        int j = i ^ stride;
        if (j > i) {
          bool up = (i & size) == 0;
          uint64_t a = key[i], b = key[j];
          if ((a > b) == up) {
            key[i] = b;
            key[j] = a;
            int t = val[i];
            val[i] = val[j];
            val[j] = t;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < ns; i += 1024) rank[base + val[i]] = i;
  if (gridDim.x == 1 && rank_pstart) {
    // one instance: the offset of each splitting node's points among the
    // next level's, i.e. the prefix of the point counts in rank order
    int64_t carry = 0;
    for (int i0 = 0; i0 < ns; i0 += 1024) {
      const int i = i0 + threadIdx.x;
      const int v = i < ns ? pcount[base + val[i]] : 0;
      int incl = v;
      for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += t;
      }
      if (lane == kWave - 1) s_wtot[wv] = incl;
      __syncthreads();
      int64_t o = carry, tot = 0;
      for (int w = 0; w < 1024 / kWave; ++w) {
        if (w < wv) o += s_wtot[w];
        tot += s_wtot[w];
      }
      if (i < ns) rank_pstart[base + val[i]] = o + incl - v;
      __syncthreads();
      carry += tot;
    }
  }
  if (threadIdx.x != 0) return;
  if (gridDim.x > 1) {
    seg_cnt[blockIdx.x] = s_cnt;
    return;
  }
  const int run = s_cnt;  // k_seg_scan for one instance
  seg_split[0] = 0;
  xfer[0] = 0;
  next_seg[0] = 0;
  seg_split[1] = run;
  xfer[1] = run;
  next_seg[1] = 4 * (int64_t)run;
}

// the same by counting, one thread per node, for segments the sort cannot take
__global__ void k_rank_pairs(int c, const int* inst, const int64_t* seg, const int* split,
                             const int* st, const uint64_t* path, int* rank, int* seg_cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c) return;
  if (!split[i]) {
    rank[i] = -1;
    return;
  }
  const int b = inst[i];
  const int a0 = (int)seg[b], a1 = (int)seg[b + 1];
  int si = st[i];
  uint64_t pi = path[i];
  int r = 0;
  for (int j = a0; j < a1; ++j) {
    if (!split[j]) continue;
    int sj = st[j];
    if (sj < si || (sj == si && path[j] < pi)) ++r;
  }
  rank[i] = r;
  atomicAdd(seg_cnt + b, 1);
}

// exclusive scan of the per-instance split counts (ninst + 1 entries)
__global__ void k_seg_scan(int ninst, int* seg_cnt, int* seg_split, int64_t* next_seg,
                           int* xfer) {
  // also: the next level's node offsets (4 children per split), the offsets
  // for the host (xfer), and seg_cnt cleared for the next level
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int run = 0;
  for (int b = 0; b < ninst; ++b) {
    seg_split[b] = run;
    xfer[b] = run;
    next_seg[b] = 4 * (int64_t)run;
    run += seg_cnt[b];
    seg_cnt[b] = 0;
  }
  seg_split[ninst] = run;
  xfer[ninst] = run;
  next_seg[ninst] = 4 * (int64_t)run;
}

// instance-local rank -> rank within the level
__global__ void k_rank_offset(int c, const int64_t* dc, const int* inst, const int* seg_split,
                              int* rank) {
  if (dc) c = (int)*dc;  // the level's node count, when only the device has it
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < c && rank[i] >= 0) rank[i] += seg_split[inst[i]];
}

// records of the 4 children of splitting node i (node_subdivide, quadtree.cpp:203-253)
__device__ __forceinline__ void child_records(int i, const int* rank, const double* corners,
                                              const double* side, const uint64_t* path,
                                              const int* inst, int* n_parent, int* n_slot,
                                              uint64_t* n_path, double* n_corners,
                                              double* n_side, int* n_inst) {
#pragma clang fp contract(off)
  const double* cc = corners + 8 * i;
  double midx, midy;
  node_mid(cc, &midx, &midy);
  double hs = side[i] / 2.0;
  for (int s = 0; s < 4; ++s) {
    int o = 4 * rank[i] + s;
    double k[8];
    for (int j = 0; j < 8; ++j) k[j] = cc[j];
    if (s == 0) {
      k[3] = midy; k[4] = midx; k[5] = midy; k[6] = midx;
    } else if (s == 1) {
      k[1] = midy; k[4] = midx; k[6] = midx; k[7] = midy;
    } else if (s == 2) {
      k[0] = midx; k[1] = midy; k[2] = midx; k[7] = midy;
    } else {
      k[0] = midx; k[2] = midx; k[3] = midy; k[5] = midy;
    }
    for (int j = 0; j < 8; ++j) n_corners[8 * o + j] = k[j];
    n_side[o] = hs;
    n_parent[o] = i;
    n_slot[o] = s;
    n_path[o] = path[i] * 4 + s;
    n_inst[o] = inst[i];
  }
}

// ---- stable 4-way partition of the points of splitting nodes -----------
struct Q4 {
  int v[4];
};

constexpr int kScanTB = 1024;

// quadrant of each point position (-1: its node does not split; to qcode),
// then per quadrant the number of earlier positions of the block with that
// quadrant (wave ballots, then the waves' totals); block totals out
__global__ __launch_bounds__(kScanTB) void k_q4_block(int np, const int* dnp, const int* perm,
                                                      const int* pnode,
                                                      const int* rank, const double* corners,
                                                      const double* px, const double* py,
                                                      int* qcode, Q4* pre, Q4* btot) {
  __shared__ int wtot[kScanTB / kWave][4];
  if (dnp) np = *dnp;  // the level's point count, when only the device has it
  const int p = blockIdx.x * kScanTB + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  int q = -1;
  if (p < np) {
    const int nd = pnode[p];
    if (rank[nd] >= 0) {
      double midx, midy;
      node_mid(corners + 8 * nd, &midx, &midy);
      const int pt = perm[p];
      q = quadrant(px[pt], py[pt], midx, midy);
    }
    qcode[p] = q;
  }
  // per-quadrant counts: lanes below in the wave (ballots), then the waves
  // below in the block
  const unsigned long long below = (1ull << lane) - 1;
  Q4 e;
  for (int k = 0; k < 4; ++k) {
    const unsigned long long m = __ballot(q == k);
    e.v[k] = __popcll(m & below);
    if (lane == 0) wtot[wv][k] = __popcll(m);
  }
  __syncthreads();
  for (int w = 0; w < wv; ++w)
    for (int k = 0; k < 4; ++k) e.v[k] += wtot[w][k];
  if (p < np) pre[p] = e;  // exclusive within the block
  if (threadIdx.x == kScanTB - 1) {
    Q4 tot = e;
    if (q >= 0) tot.v[q] += 1;
    btot[blockIdx.x] = tot;
  }
}

// pre += the exclusive prefix of the block totals (each block of kTB = 256
// positions lies in one scan block b and sums btot[0 .. b-1] itself, in wave
// order); the block holding position np also writes the level total
__global__ void k_q4_apply(int np, const int* dnp, int nb, Q4* pre, const Q4* btot, Q4* total,
                           int* np_next) {
  __shared__ int s_add[8];
  if (dnp) np = *dnp;
  const int p0 = blockIdx.x * blockDim.x;
  const int p = p0 + threadIdx.x;
  const int b = p0 / kScanTB;
  const bool last = np >= p0 && np < p0 + (int)blockDim.x;
  if (threadIdx.x < kWave) {
    int a[4] = {0, 0, 0, 0}, t[4] = {0, 0, 0, 0};
    for (int i = threadIdx.x; i < nb; i += kWave) {
      const Q4 v = btot[i];
      for (int q = 0; q < 4; ++q) {
        if (i < b) a[q] += v.v[q];
        t[q] += v.v[q];
      }
    }
    for (int q = 0; q < 4; ++q) {
      a[q] = wave_sum(a[q]);
      t[q] = wave_sum(t[q]);
    }
    if (threadIdx.x == 0)
      for (int q = 0; q < 4; ++q) {
        s_add[q] = a[q];
        s_add[4 + q] = t[q];
      }
  }
  __syncthreads();
  if (p < np) {
    Q4 v = pre[p];
    for (int q = 0; q < 4; ++q) v.v[q] += s_add[q];
    pre[p] = v;
  }
  if (last && p == np) {
    Q4 v;
    for (int q = 0; q < 4; ++q) v.v[q] = s_add[4 + q];
    pre[np] = v;
    *total = v;
    if (np_next) *np_next = v.v[0] + v.v[1] + v.v[2] + v.v[3];
  }
}

// records of the 4 children of every splitting node and its points per child
// (counts between the node's first and one-past-last position)
// (counts between the node's first and one-past-last position).  With
// rank_pstart (the node's first point among the next level's) the children's
// first points go to n_pstart as well.
__global__ void k_children_counts(int c, const int64_t* dc, const int* rank, const double* corners,
                                  const double* side, const uint64_t* path, const int* inst,
                                  const int64_t* pstart, const int* pcount, const Q4* pre,
                                  int* n_parent, int* n_slot, uint64_t* n_path,
                                  double* n_corners, double* n_side, int* n_inst, int* ccnt,
                                  const int64_t* rank_pstart, int64_t* n_pstart) {
  if (dc) c = (int)*dc;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c || rank[i] < 0) return;
  child_records(i, rank, corners, side, path, inst, n_parent, n_slot, n_path, n_corners, n_side,
                n_inst);
  const Q4 a = pre[pstart[i]], b = pre[pstart[i] + pcount[i]];
  int64_t o = rank_pstart ? rank_pstart[i] : 0;
  for (int q = 0; q < 4; ++q) {
    const int cq = b.v[q] - a.v[q];
    ccnt[4 * rank[i] + q] = cq;
    if (rank_pstart) {
      n_pstart[4 * rank[i] + q] = o;
      o += cq;
    }
  }
}

__global__ void k_scatter(int np, const int* dnp, const int* perm, const int* pnode,
                          const int* qcode, const int* rank, const int64_t* pstart, const Q4* pre,
                          const int64_t* cstart, int* nperm_out, int* npnode_out) {
  if (dnp) np = *dnp;
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int q = qcode[p];
  if (q < 0) return;
  const int nd = pnode[p];
  const int ch = 4 * rank[nd] + q;
  const int64_t dst = cstart[ch] + (pre[p].v[q] - pre[pstart[nd]].v[q]);
  nperm_out[dst] = perm[p];
  npnode_out[dst] = ch;
}

// get_descendent_neighbors' wall test (quadtree.cpp:121-155) for big box B and
// smaller box d, including the left/right test `y < top && y < bottom` (:143).
__device__ __forceinline__ bool wall_touch(const double* bc, const double* dc) {
  double top = bc[3], bottom = bc[1], left = bc[0], right = bc[4];
  for (int i = 0; i < 8; i += 2) {
    double x = dc[i], y = dc[i + 1];
    if (x > left && x < right) {
      if (fabs(y - top) < 1e-14) return true;
      if (fabs(y - bottom) < 1e-14) return true;
    }
    if (y < top && y < bottom) {
      if (fabs(x - left) < 1e-14) return true;
      if (fabs(x - right) < 1e-14) return true;
    }
  }
  return false;
}

// Coarse leaves pushed into node d by get_descendent_neighbors: for each
// ancestor level lb (ascending) the leaves among anc(d, lb)'s same-level
// neighbors in level order, if the wall test passes.  pass 0 counts, pass 1
// writes at cp_off.
// ancestors of every node, kMaxDepth per node (anc[d][k] = ancestor at level
// k, anc[d][level] = d), filled level by level in k_flatten
constexpr int kMaxDepth = 32;

// the coarser leaves next to node d (get_descendent_neighbors,
// quadtree.cpp:119-161): for each ancestor level lb = 1 .. l-1, the leaves
// among the ancestor's same-level neighbors in ascending node id (the
// reference's list order), kept if they touch d.  pass 0 counts, pass 1
// writes.  Everything stays in registers: the up-to-8 candidates are
// ordered by a sorting network.
// kMaxDepth lanes per node d: lane lb handles ancestor level lb, and a
// prefix sum over the lanes keeps the ascending-lb order.
static_assert(kMaxDepth == 32 && kWave % kMaxDepth == 0, "32 lanes per node");
__global__ void k_coarse(int nnodes, const int* anc, const int* level, const int* is_leaf,
                         const double* corners, const int* sl_nbr, const int* sl_cnt,
                         const int64_t* cp_off, int* cp_cnt, int* cp_nbr, int pass) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int d = gid / kMaxDepth;
  const int lb = gid % kMaxDepth;
  const int l = d < nnodes ? level[d] : 0;
  const double* dc = corners + 8 * (size_t)(d < nnodes ? d : 0);
  int c[kMaxNbrs];
  bool touch[kMaxNbrs] = {};
  int cnt = 0;
  if (lb >= 1 && lb < l) {
    const int nd = anc[(size_t)d * kMaxDepth + lb];
    const int ns = sl_cnt[nd];
#pragma unroll
    for (int j = 0; j < kMaxNbrs; ++j) {
      const int bb = j < ns ? sl_nbr[kMaxNbrs * (size_t)nd + j] : -1;
      c[j] = (bb >= 0 && is_leaf[bb]) ? bb : INT_MAX;
    }
    // Batcher's odd-even merge network for 8 keys
#define KI_CE(a, b)                     \
  {                                     \
    const int lo_ = min(c[a], c[b]);    \
    const int hi_ = max(c[a], c[b]);    \
    c[a] = lo_;                         \
    c[b] = hi_;                         \
  }
    KI_CE(0, 1) KI_CE(2, 3) KI_CE(4, 5) KI_CE(6, 7)
    KI_CE(0, 2) KI_CE(1, 3) KI_CE(4, 6) KI_CE(5, 7)
    KI_CE(1, 2) KI_CE(5, 6)
    KI_CE(0, 4) KI_CE(1, 5) KI_CE(2, 6) KI_CE(3, 7)
    KI_CE(2, 4) KI_CE(3, 5)
    KI_CE(1, 2) KI_CE(3, 4) KI_CE(5, 6)
#undef KI_CE
#pragma unroll
    for (int i = 0; i < kMaxNbrs; ++i) {
      touch[i] = c[i] != INT_MAX && wall_touch(corners + 8 * (size_t)c[i], dc);
      cnt += touch[i];
    }
  }
  // exclusive prefix of cnt over the node's lanes
  int incl = cnt;
  for (int k = 1; k < kMaxDepth; k <<= 1) {
    const int t = __shfl_up(incl, k, kMaxDepth);
    if (lb >= k) incl += t;
  }
  if (d >= nnodes) return;
  if (!pass) {
    if (lb == kMaxDepth - 1) cp_cnt[d] = incl;
    return;
  }
  if (cnt == 0) return;
  int64_t o = cp_off[d] + (incl - cnt);
#pragma unroll
  for (int i = 0; i < kMaxNbrs; ++i)
    if (touch[i]) cp_nbr[o++] = c[i];
}

struct LevelBuf {
  int count = 0;
  std::vector<int64_t> seg;  // node offsets of the instances (ninst + 1)
  DBuf<int64_t> dseg;
  DBuf<int> inst;
  DBuf<int> parent_local, slot, split, st, rank, pcount;
  DBuf<int64_t> pstart;
  DBuf<uint64_t> path;
  DBuf<double> corners, side;
  DBuf<int> perm, pnode;
  int nperm = 0;
};

// one launch per level: copy the level's records into the flat node arrays
struct FlatArgs {
  int c, base, parent_base, next_base, level, has_next;
  int64_t pbase;
  int nperm;               // this level's point permutation, copied to perm_all + pbase
  const int* perm;
  int* perm_all;
  int* anc;                // ancestor table, kMaxDepth per node
  const int *parent_local, *slot, *split, *rank, *pcount, *inst;
  const int64_t* pstart;
  const double *corners, *side;
  int *parent, *child, *is_leaf, *levelv, *slotv, *pcountv, *instv;
  int64_t* pstartv;
  double *cornersv, *sidev;
};

// Same-level neighbor lists (compute_neighbor_lists, quadtree.cpp:33-50):
// siblings, then cousins.  kSlLanes lanes per node: lane j tests the
// children of the parent's j-th neighbor, and a prefix sum over the lanes
// puts the cousins in the serial loop's (j, s) order.  The parent level's
// flat records (child, is_leaf, lists) come from the previous launch; this
// level's own records are read from the level's arrays, and a child of a
// parent-level node is always on this level.
constexpr int kSlLanes = kMaxNbrs;
static_assert(kSlLanes == 8 && kWave % kSlLanes == 0, "8 lanes per node");
__device__ __forceinline__ void same_level_lane(const FlatArgs& f, int i, int j, int* sl_nbr,
                                                int* sl_cnt, int* status) {
#pragma clang fp contract(off)
  const int p = i < f.c && f.parent_local[i] >= 0 ? f.parent_base + f.parent_local[i] : -1;
  int found[4];
  int nf = 0;
  if (p >= 0 && j < sl_cnt[p]) {
    const int y = sl_nbr[kMaxNbrs * p + j];
    if (!f.is_leaf[y]) {
      const double thr = f.side[i] * sqrt(2.0) + 1e-5;
      const double x0 = f.corners[8 * i], y0 = f.corners[8 * i + 1];
      for (int s = 0; s < 4; ++s) {
        const int z = f.child[4 * y + s];
        if (z < 0) continue;
        const int zl = z - f.base;
        const double dx = x0 - f.corners[8 * zl];
        const double dy = y0 - f.corners[8 * zl + 1];
        if (sqrt(dx * dx + dy * dy) < thr) found[nf++] = z;
      }
    }
  }
  // exclusive prefix of nf over the node's lanes
  int incl = nf;
  for (int d = 1; d < kSlLanes; d <<= 1) {
    const int t = __shfl_up(incl, d, kSlLanes);
    if (j >= d) incl += t;
  }
  const int total = __shfl(incl, kSlLanes - 1, kSlLanes);
  if (i >= f.c) return;
  const int x = f.base + i;
  if (p < 0) {
    if (j == 0) sl_cnt[x] = 0;
    return;
  }
  int* out = sl_nbr + kMaxNbrs * x;
  if (j < 3) {  // the siblings first
    const int sx = f.slot[i];
    const int s = j < sx ? j : j + 1;
    out[j] = f.child[4 * p + s];
  }
  const int o = 3 + incl - nf;
  for (int k = 0; k < nf; ++k) {
    if (o + k < kMaxNbrs) out[o + k] = found[k];
    else atomicOr(status, 1);
  }
  if (j == 0) sl_cnt[x] = min(3 + total, kMaxNbrs);
}

// one launch per level: the level's point permutation and records into the
// flat arrays, its ancestor rows, and its same-level neighbor lists
__global__ void k_flatten(FlatArgs f, int* sl_nbr, int* sl_cnt, int* status) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  same_level_lane(f, g / kSlLanes, g % kSlLanes, sl_nbr, sl_cnt, status);
  const int i = g;
  if (i < f.nperm) f.perm_all[f.pbase + i] = f.perm[i];
  if (i >= f.c) return;
  const int x = f.base + i;
  {
    int* ax = f.anc + (size_t)x * kMaxDepth;
    if (f.parent_local[i] >= 0) {
      const int* ap = f.anc + (size_t)(f.parent_base + f.parent_local[i]) * kMaxDepth;
      for (int k2 = 0; k2 < f.level; ++k2) ax[k2] = ap[k2];
    }
    ax[f.level] = x;
  }
  f.parent[x] = f.parent_local[i] < 0 ? -1 : f.parent_base + f.parent_local[i];
  for (int s = 0; s < 4; ++s)
    f.child[4 * x + s] = (f.has_next && f.rank[i] >= 0) ? f.next_base + 4 * f.rank[i] + s : -1;
  f.is_leaf[x] = f.split[i] ? 0 : 1;
  f.levelv[x] = f.level;
  f.slotv[x] = f.slot[i];
  f.instv[x] = f.inst[i];
  f.pcountv[x] = f.pcount[i];
  f.pstartv[x] = f.pstart[i] + f.pbase;
  for (int j = 0; j < 8; ++j) f.cornersv[8 * x + j] = f.corners[8 * i + j];
  f.sidev[x] = f.side[i];
}

// level 0: one root box per instance around its points' min/max
// (quadtree.cpp:80-88); also the level's instance offsets
__global__ void k_root_level(int ni, const int64_t* inst_pt, const double* mm, int* parent_local,
                             int* slot, int* pcount, int64_t* pstart, uint64_t* path,
                             double* corners, double* side, int* inst, int64_t* seg) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > ni) return;
  seg[i] = i;
  if (i == ni) return;
  const double lo = mm[2 * i] - 0.01, hi = mm[2 * i + 1] + 0.01;
  double* c = corners + 8 * (size_t)i;
  c[0] = lo;
  c[1] = lo;
  c[2] = lo;
  c[3] = hi;
  c[4] = hi;
  c[5] = hi;
  c[6] = hi;
  c[7] = lo;
  side[i] = hi - lo;
  pcount[i] = (int)(inst_pt[i + 1] - inst_pt[i]);
  pstart[i] = inst_pt[i];
  parent_local[i] = -1;
  slot[i] = 0;
  path[i] = 0;
  inst[i] = i;
}

// number of coarse-leaf entries (last offset + last count) into cp_off[nn],
// and with the status word into xfer for the host
__global__ void k_cp_total(int nn, int64_t* cp_off, const int* cnt, const int* status,
                           int64_t* xfer) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t tot = nn ? cp_off[nn - 1] + cnt[nn - 1] : 0;
  cp_off[nn] = tot;
  xfer[0] = tot;
  xfer[1] = *status;
}

// the host mirrors of the factorization into mapped host memory: doubles
// corners (8 nn) | side (nn) | min/max (2 ni), then ints inst | is_leaf | pcount
__global__ void k_mirror_out(int nn, int ni, const double* corners, const double* side,
                             const double* mm, const int* inst, const int* is_leaf,
                             const int* pcount, double* hd, int* hi) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < 8 * (int64_t)nn) hd[e] = corners[e];
  if (e < nn) {
    hd[8 * (int64_t)nn + e] = side[e];
    hi[e] = inst[e];
    hi[nn + e] = is_leaf[e];
    hi[2 * (int64_t)nn + e] = pcount[e];
  }
  if (e < 2 * ni) hd[9 * (int64_t)nn + e] = mm[e];
}

}  // namespace

void build_tree(TreeData& t, hipStream_t s, int* syncs, PinnedHost& pin,
                std::vector<hipEvent_t>& evs) {
  const int n = t.npts;
  const int ni = t.ninst;
  if (t.inst_pt.empty()) t.inst_pt = {0, (int64_t)n};
  if ((int)t.inst_pt.size() != ni + 1 || t.inst_pt[ni] != n)
    KI_FAIL(KIGPU_ERR_ARG, "instance point offsets do not cover the points");
  t.d_inst_pt.upload(t.inst_pt, s);
  // root boxes: min/max pooled over x and y (quadtree.cpp:80-88), per instance
  DBuf<double> mm;
  mm.alloc(2 * (size_t)ni);
  if (ni == 1) {
    DBuf<double> mpart;
    mpart.alloc(2 * kMMBlocks);
    k_minmax_part<<<kMMBlocks, kTB, 0, s>>>(t.px.p, t.py.p, n, mpart.p);
    k_minmax_final<<<1, 64, 0, s>>>(mpart.p, kMMBlocks, mm.p);
  } else {
    k_minmax_seg<<<ni, kTB, 0, s>>>(t.px.p, t.py.p, t.d_inst_pt.p, mm.p);
  }

  std::vector<std::unique_ptr<LevelBuf>> lv, spare;
  // Pipelined build: a splitting node holds at least q1 points, so below the
  // root an instance has at most 4 floor(n_i / q1) nodes per level and a
  // level at most `cap` nodes and n points.  When every instance's bound fits
  // the sort, level l + 1 is queued with those bounds and the counts the
  // kernels leave on the device, and the host reads level l's counts while it
  // runs; a level queued past the last one finds no nodes and does nothing.
  // Levels deeper than the sort takes (paths of 20 levels) restart the build
  // with a wait per level.
  const int q1 = t.leaf_dofs / t.sd + 1;
  int64_t inst_bound = 1;
  for (int i = 0; i < ni; ++i)
    inst_bound = std::max<int64_t>(inst_bound, 4 * ((t.inst_pt[i + 1] - t.inst_pt[i]) / q1));
  const int cap = (int)std::max<int64_t>(4 * ((int64_t)n / q1), 1);
  bool pipe = inst_bound <= kSortMax && (int64_t)n < (1ll << 24);
  const int slot = ni + 5;  // per level: seg_split (ni + 1) | quadrant totals (Q4)
  int* hx = pin.get<int>((size_t)slot * kMaxDepth, s);
  int* dx = pin.dget<int>();
  if (pipe) ensure_events(evs, kMaxDepth);
  DBuf<int> seg_cnt, seg_split, qcode, d_np;
  DBuf<int64_t> rank_pstart;
  DBuf<Q4> pre, btot;
  seg_split.alloc(ni + 1);
  d_np.alloc(kMaxDepth + 1);  // points per level (pipelined)
  for (int attempt = 0;; ++attempt) {
    lv.clear();
    seg_cnt.alloc(ni);
    seg_cnt.zero(s);
    {
      auto r = std::make_unique<LevelBuf>();
      r->count = ni;
      r->seg.resize(ni + 1);
      for (int i = 0; i <= ni; ++i) r->seg[i] = i;
      r->parent_local.alloc(ni);
      r->slot.alloc(ni);
      r->pcount.alloc(ni);
      r->pstart.alloc(ni);
      r->path.alloc(ni);
      r->corners.alloc(8 * (size_t)ni);
      r->side.alloc(ni);
      r->inst.alloc(ni);
      r->dseg.alloc(ni + 1);
      k_root_level<<<ceil_div(ni + 1, kTB), kTB, 0, s>>>(
          ni, t.d_inst_pt.p, mm.p, r->parent_local.p, r->slot.p, r->pcount.p, r->pstart.p,
          r->path.p, r->corners.p, r->side.p, r->inst.p, r->dseg.p);
      r->perm.alloc(n);
      r->pnode.alloc(n);
      r->nperm = n;
      k_iota<<<ceil_div(n, kTB), kTB, 0, s>>>(r->perm.p, n);
      k_fill_inst<<<ni, kTB, 0, s>>>(t.d_inst_pt.p, r->pnode.p);
      lv.push_back(std::move(r));
    }
    bool restart = false;
    // host copy of level l's counts (its slot): next level's segments and points
    auto read_slot = [&](int l, LevelBuf& N) {
      const int* h = hx + (size_t)slot * l;
      N.count = 4 * h[ni];
      N.seg.resize(ni + 1);
      for (int i = 0; i <= ni; ++i) N.seg[i] = 4 * (int64_t)h[i];
      N.nperm = h[ni + 1] + h[ni + 2] + h[ni + 3] + h[ni + 4];
      return h[ni];
    };
    for (int l = 0;; ++l) {
      LevelBuf& L = *lv[l];
      if (pipe && l > 20) {  // beyond the sort's path length
        stream_wait(s);
        restart = true;
        break;
      }
      if (l >= kMaxDepth - 1) KI_FAIL(KIGPU_ERR_NOTSUPP, "quadtree deeper than 31 levels");
      const bool dev = pipe && l > 0;  // this level's counts are on the device only
      const int c = dev ? cap : L.count;
      const int np = dev ? n : L.nperm;
      const int* dnp = dev ? d_np.p + l : nullptr;
      const int64_t* dc = dev ? L.dseg.p + ni : nullptr;
      int* xf = dx + (size_t)slot * l;
      Q4* qtot = reinterpret_cast<Q4*>(xf + ni + 1);
      L.split.alloc(c);
      L.st.alloc(c);
      L.rank.alloc(c);
      auto nx = std::make_unique<LevelBuf>();
      nx->dseg.alloc(ni + 1);
      const int* pst = l ? lv[l - 1]->st.p : nullptr;
      bool sorted = pipe;
      if (!pipe) {
        int64_t segmax = 0;
        for (int i = 0; i < ni; ++i) segmax = std::max(segmax, L.seg[i + 1] - L.seg[i]);
        sorted = segmax <= kSortMax && l <= 20 && (int64_t)n < (1ll << 24);
      }
      // one sorted instance: the segment scan and the children's point
      // offsets come out of the sort
      const bool one = sorted && ni == 1;
      if (one) rank_pstart.alloc(c);
      if (sorted) {  // split flags and times inside
        k_rank_sort<<<ni, 1024, 0, s>>>(L.dseg.p, L.pcount.p, L.pstart.p, L.perm.p,
                                        L.parent_local.p, pst, t.sd, t.leaf_dofs, L.split.p,
                                        L.st.p, L.path.p, L.rank.p, seg_cnt.p, seg_split.p,
                                        nx->dseg.p, xf, one ? rank_pstart.p : nullptr);
      } else {
        k_split_info<<<ceil_div(c, kTB), kTB, 0, s>>>(c, L.pcount.p, L.pstart.p, L.perm.p,
                                                      L.parent_local.p, pst, t.sd, t.leaf_dofs,
                                                      L.split.p, L.st.p);
        k_rank_pairs<<<ceil_div(c, kTB), kTB, 0, s>>>(c, L.inst.p, L.dseg.p, L.split.p, L.st.p,
                                                      L.path.p, L.rank.p, seg_cnt.p);
      }
      if (!one) k_seg_scan<<<1, 64, 0, s>>>(ni, seg_cnt.p, seg_split.p, nx->dseg.p, xf);
      if (ni > 1)
        k_rank_offset<<<ceil_div(c, kTB), kTB, 0, s>>>(c, dc, L.inst.p, seg_split.p, L.rank.p);
      // classify and scan the points of splitting nodes
      const int nb = ceil_div(std::max(np, 1), kScanTB);
      qcode.alloc(np);
      pre.alloc((size_t)np + 1);
      btot.alloc(nb);
      k_q4_block<<<nb, kScanTB, 0, s>>>(np, dnp, L.perm.p, L.pnode.p, L.rank.p, L.corners.p,
                                        t.px.p, t.py.p, qcode.p, pre.p, btot.p);
      k_q4_apply<<<ceil_div(np + 1, kTB), kTB, 0, s>>>(np, dnp, nb, pre.p, btot.p, qtot,
                                                       pipe ? d_np.p + l + 1 : nullptr);
      LevelBuf& N = *nx;
      if (!pipe) {
        stream_wait(s);
        ++*syncs;
        if (read_slot(l, N) == 0) break;
      } else {
        N.count = cap;  // bounds until the host reads the level's slot
        N.nperm = n;
      }
      N.parent_local.alloc(N.count);
      N.slot.alloc(N.count);
      N.path.alloc(N.count);
      N.corners.alloc(8 * (size_t)N.count);
      N.side.alloc(N.count);
      N.pcount.alloc(N.count);
      N.pstart.alloc(N.count);
      N.inst.alloc(N.count);
      k_children_counts<<<ceil_div(c, kTB), kTB, 0, s>>>(
          c, dc, L.rank.p, L.corners.p, L.side.p, L.path.p, L.inst.p, L.pstart.p, L.pcount.p,
          pre.p, N.parent_local.p, N.slot.p, N.path.p, N.corners.p, N.side.p, N.inst.p,
          N.pcount.p, one ? rank_pstart.p : nullptr, N.pstart.p);
      if (!one) {
        if (pipe) exclusive_scan_i32_to_i64_dev(N.pcount.p, N.pstart.p, N.dseg.p + ni, s);
        else exclusive_scan_i32_to_i64(N.pcount.p, N.pstart.p, N.count, s);
      }
      N.perm.alloc(std::max(N.nperm, 1));
      N.pnode.alloc(std::max(N.nperm, 1));
      k_scatter<<<ceil_div(np, kTB), kTB, 0, s>>>(np, dnp, L.perm.p, L.pnode.p, qcode.p,
                                                  L.rank.p, L.pstart.p, pre.p, N.pstart.p,
                                                  N.perm.p, N.pnode.p);
      lv.push_back(std::move(nx));
      if (pipe) {
        KI_HIP(hipEventRecord(evs[l], s));
        if (l > 0) {  // level l - 1's counts, while level l runs
          event_wait(evs[l - 1]);
          ++*syncs;
          if (read_slot(l - 1, *lv[l]) == 0) {  // level l has no nodes
            // its buffers stay alive until the queued no-op kernels are done
            for (size_t k = l; k < lv.size(); ++k) spare.push_back(std::move(lv[k]));
            lv.resize(l);
            break;
          }
        }
      }
    }
    if (!restart) break;
    pipe = false;
  }

  // ---- flatten to global node ids ----
  t.nlevels = (int)lv.size();
  t.level_off.assign(t.nlevels + 1, 0);
  for (int l = 0; l < t.nlevels; ++l) t.level_off[l + 1] = t.level_off[l] + lv[l]->count;
  t.nnodes = (int)t.level_off[t.nlevels];
  const int nn = t.nnodes;
  t.parent.alloc(nn);
  t.child.alloc(4 * (size_t)nn);
  t.is_leaf.alloc(nn);
  t.level.alloc(nn);
  t.slot.alloc(nn);
  t.corners.alloc(8 * (size_t)nn);
  t.side.alloc(nn);
  t.pstart.alloc(nn);
  t.pcount.alloc(nn);
  t.inst.alloc(nn);
  int64_t ptot = 0;
  for (auto& L : lv) ptot += L->nperm;
  t.perm_all.alloc(std::max<int64_t>(ptot, 1));
  DBuf<int> anc;
  anc.alloc((size_t)nn * kMaxDepth);
  // neighbor lists, filled level by level with the flat records
  t.sl_nbr.alloc(kMaxNbrs * (size_t)nn);
  t.sl_cnt.alloc(nn);
  t.status.alloc(1);
  t.status.zero(s);
  int64_t pbase = 0;
  for (int l = 0; l < t.nlevels; ++l) {
    LevelBuf& L = *lv[l];
    FlatArgs f;
    f.c = L.count;
    f.base = (int)t.level_off[l];
    f.parent_base = l ? (int)t.level_off[l - 1] : 0;
    f.next_base = l + 1 < t.nlevels ? (int)t.level_off[l + 1] : 0;
    f.level = l;
    f.has_next = l + 1 < t.nlevels;
    f.pbase = pbase;
    f.parent_local = L.parent_local.p;
    f.slot = L.slot.p;
    f.split = L.split.p;
    f.rank = L.rank.p;
    f.pcount = L.pcount.p;
    f.inst = L.inst.p;
    f.instv = t.inst.p;
    f.pstart = L.pstart.p;
    f.corners = L.corners.p;
    f.side = L.side.p;
    f.parent = t.parent.p;
    f.child = t.child.p;
    f.is_leaf = t.is_leaf.p;
    f.levelv = t.level.p;
    f.slotv = t.slot.p;
    f.pcountv = t.pcount.p;
    f.pstartv = t.pstart.p;
    f.cornersv = t.corners.p;
    f.sidev = t.side.p;
    f.nperm = L.nperm;
    f.perm = L.perm.p;
    f.perm_all = t.perm_all.p;
    f.anc = anc.p;
    const int64_t nthr = std::max<int64_t>((int64_t)kSlLanes * L.count, L.nperm);
    k_flatten<<<ceil_div(nthr, kTB), kTB, 0, s>>>(f, t.sl_nbr.p, t.sl_cnt.p, t.status.p);
    pbase += L.nperm;
  }
  DBuf<int> cpc;
  cpc.alloc(nn);
  t.cp_off.alloc(nn + 1);
  const int ncoarse = (int)ceil_div((int64_t)nn * kMaxDepth, kTB);
  k_coarse<<<ncoarse, kTB, 0, s>>>(nn, anc.p, t.level.p, t.is_leaf.p, t.corners.p, t.sl_nbr.p,
                                   t.sl_cnt.p, nullptr, cpc.p, nullptr, 0);
  exclusive_scan_i32_to_i64(cpc.p, t.cp_off.p, nn, s);
  DBuf<int64_t> cxf;
  cxf.alloc(2);
  k_cp_total<<<1, 64, 0, s>>>(nn, t.cp_off.p, cpc.p, t.status.p, cxf.p);
  // a node's coarse list holds at most kMaxNbrs leaves per coarser level, so
  // the lists are written without waiting for their total
  t.cp_nbr.alloc((size_t)nn * kMaxNbrs * std::max(t.nlevels - 1, 1));
  k_coarse<<<ncoarse, kTB, 0, s>>>(nn, anc.p, t.level.p, t.is_leaf.p, t.corners.p, t.sl_nbr.p,
                                   t.sl_cnt.p, t.cp_off.p, cpc.p, t.cp_nbr.p, 1);

  // ---- host mirrors the factorization needs (the rest on export), one copy
  // each into page-locked memory and one synchronization; the coarse lists'
  // total and the status word come with them
  t.nperm_total = ptot;
  t.mirrors_ready = false;
  // written by a kernel straight into the mapped page-locked block: a few
  // hundred KB, and no copy-engine round trip (one of those, idle before,
  // has been seen to take milliseconds)
  const size_t dbl = 9 * (size_t)nn + 2 * (size_t)ni;  // corners | side | min/max
  const size_t ints = (3 * (size_t)nn + 1) / 2 + 1;      // inst | is_leaf | pcount
  double* hd = pin.get<double>(dbl + ints + 2, s);
  int* hi = reinterpret_cast<int*>(hd + dbl);
  const int64_t* hc = reinterpret_cast<const int64_t*>(hd + dbl + ints);
  {
    double* dd = pin.dget<double>();
    k_mirror_out<<<ceil_div(9 * (int64_t)nn, kTB), kTB, 0, s>>>(
        nn, ni, t.corners.p, t.side.p, mm.p, t.inst.p, t.is_leaf.p, t.pcount.p, dd,
        reinterpret_cast<int*>(dd + dbl));
    KI_HIP(hipGetLastError());
    copy_to_host(pin, s, {{cxf.p, sizeof(double) * (dbl + ints), 2 * sizeof(int64_t)}});
  }
  stream_wait(s);
  ++*syncs;
  if (hc[1]) KI_FAIL(KIGPU_ERR_STATE, "more than 8 same-level neighbors (overlapping boxes)");
  t.ncp = hc[0];
  t.h_corners.assign(hd, hd + 8 * (size_t)nn);
  t.h_side.assign(hd + 8 * (size_t)nn, hd + 9 * (size_t)nn);
  t.inst_min.resize(ni);
  t.inst_max.resize(ni);
  for (int i = 0; i < ni; ++i) {
    t.inst_min[i] = hd[9 * (size_t)nn + 2 * i];
    t.inst_max[i] = hd[9 * (size_t)nn + 2 * i + 1];
  }
  t.min = t.inst_min[0];
  t.max = t.inst_max[0];
  t.h_inst.assign(hi, hi + nn);
  t.h_is_leaf.assign(hi + nn, hi + 2 * (size_t)nn);
  t.h_pcount.assign(hi + 2 * (size_t)nn, hi + 3 * (size_t)nn);
}

void ensure_tree_mirrors(TreeData& t) {
  if (t.mirrors_ready) return;
  KI_HIP(hipDeviceSynchronize());
  const size_t nn = t.nnodes;
  t.h_parent.resize(nn);
  t.h_child.resize(4 * nn);
  t.h_level.resize(nn);
  t.h_pstart.resize(nn);
  t.h_perm.resize(t.nperm_total);
  t.h_sl_nbr.resize(kMaxNbrs * nn);
  t.h_sl_cnt.resize(nn);
  t.h_cp_off.resize(nn + 1);
  t.h_cp_nbr.resize(t.ncp);
  auto get = [](void* dst, const void* src, size_t bytes) {
    if (bytes) KI_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  };
  get(t.h_parent.data(), t.parent.p, sizeof(int) * nn);
  get(t.h_child.data(), t.child.p, sizeof(int) * 4 * nn);
  get(t.h_level.data(), t.level.p, sizeof(int) * nn);
  get(t.h_pstart.data(), t.pstart.p, sizeof(int64_t) * nn);
  get(t.h_perm.data(), t.perm_all.p, sizeof(int) * t.nperm_total);
  get(t.h_sl_nbr.data(), t.sl_nbr.p, sizeof(int) * kMaxNbrs * nn);
  get(t.h_sl_cnt.data(), t.sl_cnt.p, sizeof(int) * nn);
  get(t.h_cp_off.data(), t.cp_off.p, sizeof(int64_t) * (nn + 1));
  get(t.h_cp_nbr.data(), t.cp_nbr.p, sizeof(int) * t.ncp);
  t.mirrors_ready = true;
}

void ensure_factor_mirrors(TreeData& t) {
  if (t.fmirrors_ready) return;
  KI_HIP(hipDeviceSynchronize());
  t.h_sr.resize(t.act_pool_size);
  t.h_act.resize(t.act_pool_size);
  t.h_pos.resize(t.act_pool_size);
  KI_HIP(hipMemcpy(t.h_pos.data(), t.pos_pool.p, sizeof(int) * t.act_pool_size,
                   hipMemcpyDeviceToHost));
  KI_HIP(hipMemcpy(t.h_sr.data(), t.sr_pool.p, sizeof(int) * t.act_pool_size,
                   hipMemcpyDeviceToHost));
  KI_HIP(hipMemcpy(t.h_act.data(), t.act_pool.p, sizeof(int) * t.act_pool_size,
                   hipMemcpyDeviceToHost));
  t.fmirrors_ready = true;
}

}  // namespace kigpu
