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

#include "tree.h"

namespace kigpu {

namespace {

constexpr int kTB = 256;

__global__ void k_minmax(const double* px, const double* py, int n, double* out) {
  __shared__ double smin[kTB], smax[kTB];
  double mn = px[0], mx = px[0];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double a = px[i], b = py[i];
    mn = fmin(mn, fmin(a, b));
    mx = fmax(mx, fmax(a, b));
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int s = kTB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = smin[0];
    out[1] = smax[0];
  }
}

__global__ void k_iota(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
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

// rank of each splitting node among the splitting nodes by (st, path)
__global__ void k_rank(int c, const int* split, const int* st, const uint64_t* path, int* rank,
                       int* nsplit) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c) return;
  if (!split[i]) {
    rank[i] = -1;
    return;
  }
  int si = st[i];
  uint64_t pi = path[i];
  int r = 0;
  for (int j = 0; j < c; ++j) {
    if (!split[j]) continue;
    int sj = st[j];
    if (sj < si || (sj == si && path[j] < pi)) ++r;
  }
  rank[i] = r;
  atomicAdd(nsplit, 1);
}

// records of the 4 children of every splitting node (node_subdivide, quadtree.cpp:203-253)
__global__ void k_children(int c, const int* rank, const double* corners, const double* side,
                           const uint64_t* path, int* n_parent, int* n_slot, uint64_t* n_path,
                           double* n_corners, double* n_side) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c || rank[i] < 0) return;
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
  }
}

// quadrant counts of each splitting node's points -> cnt[4*rank + q]
__global__ void k_quad_count(const int* rank, const int* pcount, const int64_t* pstart,
                             const int* perm, const double* corners, const double* px,
                             const double* py, int* cnt) {
  int i = blockIdx.x;
  if (rank[i] < 0) return;
  __shared__ int sc[4];
  if (threadIdx.x < 4) sc[threadIdx.x] = 0;
  __syncthreads();
  double midx, midy;
  node_mid(corners + 8 * i, &midx, &midy);
  int loc[4] = {0, 0, 0, 0};
  const int* pp = perm + pstart[i];
  for (int j = threadIdx.x; j < pcount[i]; j += blockDim.x) {
    int p = pp[j];
    loc[quadrant(px[p], py[p], midx, midy)]++;
  }
  for (int q = 0; q < 4; ++q)
    if (loc[q]) atomicAdd(&sc[q], loc[q]);
  __syncthreads();
  if (threadIdx.x < 4) cnt[4 * rank[i] + threadIdx.x] = sc[threadIdx.x];
}

// stable 4-way partition of each splitting node's points into its children
__global__ __launch_bounds__(kTB) void k_partition(const int* rank, const int* pcount,
                                                   const int64_t* pstart, const int* perm,
                                                   const double* corners, const double* px,
                                                   const double* py, const int64_t* child_start,
                                                   int* next_perm) {
  int i = blockIdx.x;
  if (rank[i] < 0) return;
  __shared__ int wtot[kTB / kWave][4];
  __shared__ int running[4];
  if (threadIdx.x < 4) running[threadIdx.x] = 0;
  double midx, midy;
  node_mid(corners + 8 * i, &midx, &midy);
  int64_t base[4];
  for (int q = 0; q < 4; ++q) base[q] = child_start[4 * rank[i] + q];
  const int* pp = perm + pstart[i];
  int cnt = pcount[i];
  int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  __syncthreads();
  for (int c0 = 0; c0 < cnt; c0 += kTB) {
    int j = c0 + threadIdx.x;
    int p = -1, q = -1;
    if (j < cnt) {
      p = pp[j];
      q = quadrant(px[p], py[p], midx, midy);
    }
    int pre = 0;
    for (int qq = 0; qq < 4; ++qq) {
      unsigned long long m = __ballot(q == qq);
      if (q == qq) pre = __popcll(m & ((1ull << lane) - 1));
      if (lane == 0) wtot[wv][qq] = __popcll(m);
    }
    __syncthreads();
    if (q >= 0) {
      int off = running[q] + pre;
      for (int w = 0; w < wv; ++w) off += wtot[w][q];
      next_perm[base[q] + off] = p;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      int s = 0;
      for (int w = 0; w < kTB / kWave; ++w) s += wtot[w][threadIdx.x];
      running[threadIdx.x] += s;
    }
    __syncthreads();
  }
}

// same-level neighbor lists of level [a, b): siblings, then cousins
// (compute_neighbor_lists, quadtree.cpp:33-50)
__global__ void k_same_level(int a, int b, const int* parent, const int* slot, const int* child,
                             const int* is_leaf, const int* level, const double* corners,
                             const double* side, int* sl_nbr, int* sl_cnt, int* status) {
#pragma clang fp contract(off)
  int x = a + blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= b) return;
  int p = parent[x];
  int cnt = 0;
  int* out = sl_nbr + kMaxNbrs * x;
  if (p >= 0) {
    for (int s = 0; s < 4; ++s)
      if (s != slot[x]) out[cnt++] = child[4 * p + s];
    double thr = side[x] * sqrt(2.0) + 1e-5;
    double x0 = corners[8 * x], y0 = corners[8 * x + 1];
    for (int j = 0; j < sl_cnt[p]; ++j) {
      int y = sl_nbr[kMaxNbrs * p + j];
      if (is_leaf[y]) continue;
      for (int s = 0; s < 4; ++s) {
        int z = child[4 * y + s];
        if (z < 0 || level[z] != level[x]) continue;
        double dx = x0 - corners[8 * z];
        double dy = y0 - corners[8 * z + 1];
        if (sqrt(dx * dx + dy * dy) < thr) {
          if (cnt >= kMaxNbrs) {
            atomicOr(status, 1);
            continue;
          }
          out[cnt++] = z;
        }
      }
    }
  }
  sl_cnt[x] = cnt;
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
__global__ void k_coarse(int nnodes, const int* parent, const int* level, const int* is_leaf,
                         const double* corners, const int* sl_nbr, const int* sl_cnt,
                         const int64_t* cp_off, int* cp_cnt, int* cp_nbr, int pass) {
  int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nnodes) return;
  int l = level[d];
  int anc[64];
  int cur = d;
  for (int k = l; k >= 0; --k) {
    anc[k] = cur;
    cur = parent[cur];
  }
  const double* dc = corners + 8 * d;
  int cnt = 0;
  int64_t o = pass ? cp_off[d] : 0;
  for (int lb = 1; lb < l; ++lb) {
    int n = anc[lb];
    int cand[kMaxNbrs];
    int nc = 0;
    for (int j = 0; j < sl_cnt[n]; ++j) {
      int bb = sl_nbr[kMaxNbrs * n + j];
      if (is_leaf[bb]) cand[nc++] = bb;
    }
    for (int i = 1; i < nc; ++i) {  // level order = ascending node id
      int v = cand[i], k = i - 1;
      while (k >= 0 && cand[k] > v) {
        cand[k + 1] = cand[k];
        --k;
      }
      cand[k + 1] = v;
    }
    for (int i = 0; i < nc; ++i) {
      if (wall_touch(corners + 8 * cand[i], dc)) {
        if (pass) cp_nbr[o + cnt] = cand[i];
        ++cnt;
      }
    }
  }
  if (!pass) cp_cnt[d] = cnt;
}

template <typename T>
__global__ void k_gather_copy(const T* src, T* dst, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void k_level_fill(int* level, int a, int b, int l) {
  int i = a + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < b) level[i] = l;
}

__global__ void k_offset_i32(int* p, int n, int add) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && p[i] >= 0) p[i] += add;
}

__global__ void k_offset_i64(int64_t* p, int n, int64_t add) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] += add;
}

// children ids of level-l nodes: child[4*x + s] = next_base + 4*rank + s
__global__ void k_child_ids(int c, const int* rank, int next_base, int* child) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c) return;
  for (int s = 0; s < 4; ++s) child[4 * i + s] = rank[i] < 0 ? -1 : next_base + 4 * rank[i] + s;
}

struct LevelBuf {
  int count = 0;
  DBuf<int> parent_local, slot, split, st, rank, pcount, child;
  DBuf<int64_t> pstart;
  DBuf<uint64_t> path;
  DBuf<double> corners, side;
  DBuf<int> perm;
  int nperm = 0;
};

}  // namespace

void build_tree(TreeData& t, hipStream_t s, int* syncs) {
  const int n = t.npts;
  // root box from min/max pooled over x and y (quadtree.cpp:80-105)
  DBuf<double> mm;
  mm.alloc(2);
  k_minmax<<<1, kTB, 0, s>>>(t.px.p, t.py.p, n, mm.p);
  double hmm[2];
  mm.download(hmm, 2, s);
  KI_HIP(hipStreamSynchronize(s));
  ++*syncs;
  t.min = hmm[0];
  t.max = hmm[1];
  const double lo = t.min - 0.01, hi = t.max + 0.01;

  std::vector<std::unique_ptr<LevelBuf>> lv;
  {
    auto r = std::make_unique<LevelBuf>();
    r->count = 1;
    int h_pl = -1, h_slot = 0, h_pc = n;
    int64_t h_ps = 0;
    uint64_t h_path = 0;
    double h_c[8] = {lo, lo, lo, hi, hi, hi, hi, lo};
    double h_side = hi - lo;
    r->parent_local.upload(&h_pl, 1, s);
    r->slot.upload(&h_slot, 1, s);
    r->pcount.upload(&h_pc, 1, s);
    r->pstart.upload(&h_ps, 1, s);
    r->path.upload(&h_path, 1, s);
    r->corners.upload(h_c, 8, s);
    r->side.upload(&h_side, 1, s);
    r->perm.alloc(n);
    r->nperm = n;
    k_iota<<<ceil_div(n, kTB), kTB, 0, s>>>(r->perm.p, n);
    lv.push_back(std::move(r));
  }
  DBuf<int> nsplit;
  nsplit.alloc(1);
  DBuf<int> cnt;
  DBuf<int64_t> cstart, cnt64;
  for (int l = 0;; ++l) {
    LevelBuf& L = *lv[l];
    const int c = L.count;
    L.split.alloc(c);
    L.st.alloc(c);
    L.rank.alloc(c);
    L.child.alloc(4 * (size_t)c);
    const int* pst = l ? lv[l - 1]->st.p : nullptr;
    k_split_info<<<ceil_div(c, kTB), kTB, 0, s>>>(c, L.pcount.p, L.pstart.p, L.perm.p,
                                                  L.parent_local.p, pst, t.sd, t.leaf_dofs,
                                                  L.split.p, L.st.p);
    nsplit.zero(s);
    k_rank<<<ceil_div(c, kTB), kTB, 0, s>>>(c, L.split.p, L.st.p, L.path.p, L.rank.p, nsplit.p);
    int S = 0;
    nsplit.download(&S, 1, s);
    KI_HIP(hipStreamSynchronize(s));
    ++*syncs;
    if (S == 0) break;
    if (l >= 30) KI_FAIL(KIGPU_ERR_NOTSUPP, "quadtree deeper than 31 levels");
    auto nx = std::make_unique<LevelBuf>();
    LevelBuf& N = *nx;
    N.count = 4 * S;
    N.parent_local.alloc(N.count);
    N.slot.alloc(N.count);
    N.path.alloc(N.count);
    N.corners.alloc(8 * (size_t)N.count);
    N.side.alloc(N.count);
    N.pcount.alloc(N.count);
    N.pstart.alloc(N.count);
    k_children<<<ceil_div(c, kTB), kTB, 0, s>>>(c, L.rank.p, L.corners.p, L.side.p, L.path.p,
                                                N.parent_local.p, N.slot.p, N.path.p,
                                                N.corners.p, N.side.p);
    cnt.alloc(N.count);
    k_quad_count<<<c, kTB, 0, s>>>(L.rank.p, L.pcount.p, L.pstart.p, L.perm.p, L.corners.p,
                                   t.px.p, t.py.p, cnt.p);
    exclusive_scan_i32_to_i64(cnt.p, N.pstart.p, N.count, s);
    KI_HIP(hipMemcpyAsync(N.pcount.p, cnt.p, sizeof(int) * N.count, hipMemcpyDeviceToDevice, s));
    // every point of a splitting node goes to the next level
    int64_t last_start = 0;
    int last_cnt = 0;
    KI_HIP(hipMemcpyAsync(&last_start, N.pstart.p + N.count - 1, sizeof(int64_t),
                          hipMemcpyDeviceToHost, s));
    KI_HIP(hipMemcpyAsync(&last_cnt, cnt.p + N.count - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    KI_HIP(hipStreamSynchronize(s));
    ++*syncs;
    N.nperm = (int)(last_start + last_cnt);
    N.perm.alloc(std::max(N.nperm, 1));
    k_partition<<<c, kTB, 0, s>>>(L.rank.p, L.pcount.p, L.pstart.p, L.perm.p, L.corners.p,
                                  t.px.p, t.py.p, N.pstart.p, N.perm.p);
    lv.push_back(std::move(nx));
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
  int64_t ptot = 0;
  for (auto& L : lv) ptot += L->nperm;
  t.perm_all.alloc(std::max<int64_t>(ptot, 1));
  int64_t pbase = 0;
  for (int l = 0; l < t.nlevels; ++l) {
    LevelBuf& L = *lv[l];
    const int c = L.count;
    const int64_t b = t.level_off[l];
    const int g = ceil_div(c, kTB);
    KI_HIP(hipMemcpyAsync(t.parent.p + b, L.parent_local.p, sizeof(int) * c,
                          hipMemcpyDeviceToDevice, s));
    if (l > 0) k_offset_i32<<<g, kTB, 0, s>>>(t.parent.p + b, c, (int)t.level_off[l - 1]);
    if (l + 1 < t.nlevels) {
      k_child_ids<<<g, kTB, 0, s>>>(c, L.rank.p, (int)t.level_off[l + 1], t.child.p + 4 * b);
    } else {
      KI_HIP(hipMemsetAsync(t.child.p + 4 * b, 0xff, sizeof(int) * 4 * c, s));
    }
    KI_HIP(hipMemcpyAsync(t.slot.p + b, L.slot.p, sizeof(int) * c, hipMemcpyDeviceToDevice, s));
    KI_HIP(hipMemcpyAsync(t.corners.p + 8 * b, L.corners.p, sizeof(double) * 8 * c,
                          hipMemcpyDeviceToDevice, s));
    KI_HIP(hipMemcpyAsync(t.side.p + b, L.side.p, sizeof(double) * c, hipMemcpyDeviceToDevice, s));
    KI_HIP(hipMemcpyAsync(t.pcount.p + b, L.pcount.p, sizeof(int) * c, hipMemcpyDeviceToDevice,
                          s));
    KI_HIP(hipMemcpyAsync(t.pstart.p + b, L.pstart.p, sizeof(int64_t) * c,
                          hipMemcpyDeviceToDevice, s));
    k_offset_i64<<<g, kTB, 0, s>>>(t.pstart.p + b, c, pbase);
    if (L.nperm)
      KI_HIP(hipMemcpyAsync(t.perm_all.p + pbase, L.perm.p, sizeof(int) * L.nperm,
                            hipMemcpyDeviceToDevice, s));
    pbase += L.nperm;
    // a node is a leaf iff it did not split (the last level never splits)
    KI_HIP(hipMemcpyAsync(t.is_leaf.p + b, L.split.p ? L.split.p : L.pcount.p, sizeof(int) * c,
                          hipMemcpyDeviceToDevice, s));
    k_level_fill<<<g, kTB, 0, s>>>(t.level.p, (int)b, (int)(b + c), l);
  }
  // is_leaf currently holds the split flag
  {
    std::vector<int> sp(nn);
    t.is_leaf.download(sp.data(), nn, s);
    KI_HIP(hipStreamSynchronize(s));
    ++*syncs;
    for (int i = 0; i < nn; ++i) sp[i] = sp[i] ? 0 : 1;
    t.is_leaf.upload(sp, s);
  }

  // ---- neighbor lists ----
  t.sl_nbr.alloc(kMaxNbrs * (size_t)nn);
  t.sl_cnt.alloc(nn);
  t.status.alloc(1);
  t.status.zero(s);
  for (int l = 0; l < t.nlevels; ++l) {
    int a = (int)t.level_off[l], b = (int)t.level_off[l + 1];
    k_same_level<<<ceil_div(b - a, kTB), kTB, 0, s>>>(a, b, t.parent.p, t.slot.p, t.child.p,
                                                       t.is_leaf.p, t.level.p, t.corners.p,
                                                       t.side.p, t.sl_nbr.p, t.sl_cnt.p,
                                                       t.status.p);
  }
  DBuf<int> cpc;
  cpc.alloc(nn);
  t.cp_off.alloc(nn + 1);
  k_coarse<<<ceil_div(nn, kTB), kTB, 0, s>>>(nn, t.parent.p, t.level.p, t.is_leaf.p,
                                              t.corners.p, t.sl_nbr.p, t.sl_cnt.p, nullptr,
                                              cpc.p, nullptr, 0);
  exclusive_scan_i32_to_i64(cpc.p, t.cp_off.p, nn + 1 > nn ? nn : nn, s);
  // total = last offset + last count
  int64_t lo_ = 0;
  int lc_ = 0;
  KI_HIP(hipMemcpyAsync(&lo_, t.cp_off.p + nn - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  KI_HIP(hipMemcpyAsync(&lc_, cpc.p + nn - 1, sizeof(int), hipMemcpyDeviceToHost, s));
  int hstat = 0;
  t.status.download(&hstat, 1, s);
  KI_HIP(hipStreamSynchronize(s));
  ++*syncs;
  if (hstat) KI_FAIL(KIGPU_ERR_STATE, "more than 8 same-level neighbors (overlapping boxes)");
  const int64_t ncp = lo_ + lc_;
  KI_HIP(hipMemcpyAsync(t.cp_off.p + nn, &ncp, sizeof(int64_t), hipMemcpyHostToDevice, s));
  t.cp_nbr.alloc(std::max<int64_t>(ncp, 1));
  k_coarse<<<ceil_div(nn, kTB), kTB, 0, s>>>(nn, t.parent.p, t.level.p, t.is_leaf.p,
                                              t.corners.p, t.sl_nbr.p, t.sl_cnt.p, t.cp_off.p,
                                              cpc.p, t.cp_nbr.p, 1);

  // ---- host mirrors ----
  t.h_parent.resize(nn);
  t.h_child.resize(4 * (size_t)nn);
  t.h_is_leaf.resize(nn);
  t.h_level.resize(nn);
  t.h_pcount.resize(nn);
  t.h_pstart.resize(nn);
  t.h_corners.resize(8 * (size_t)nn);
  t.h_side.resize(nn);
  t.h_perm.resize(ptot);
  t.h_sl_nbr.resize(kMaxNbrs * (size_t)nn);
  t.h_sl_cnt.resize(nn);
  t.h_cp_off.resize(nn + 1);
  t.h_cp_nbr.resize(ncp);
  t.parent.download(t.h_parent.data(), nn, s);
  t.child.download(t.h_child.data(), 4 * (size_t)nn, s);
  t.is_leaf.download(t.h_is_leaf.data(), nn, s);
  t.level.download(t.h_level.data(), nn, s);
  t.pcount.download(t.h_pcount.data(), nn, s);
  t.pstart.download(t.h_pstart.data(), nn, s);
  t.corners.download(t.h_corners.data(), 8 * (size_t)nn, s);
  t.side.download(t.h_side.data(), nn, s);
  t.perm_all.download(t.h_perm.data(), ptot, s);
  t.sl_nbr.download(t.h_sl_nbr.data(), kMaxNbrs * (size_t)nn, s);
  t.sl_cnt.download(t.h_sl_cnt.data(), nn, s);
  t.cp_off.download(t.h_cp_off.data(), nn + 1, s);
  t.cp_nbr.download(t.h_cp_nbr.data(), ncp, s);
  KI_HIP(hipStreamSynchronize(s));
  ++*syncs;
}

}  // namespace kigpu
