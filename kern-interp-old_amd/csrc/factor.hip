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
#include "elim_dev.h"
#include "id_dev.h"
#include "entries.h"
#include "factor.h"

namespace kigpu {

namespace {

using namespace elim;

constexpr int kTB = 256;
constexpr int kNW = kTB / kWave;


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

// leaf DOF counts, and the per-node factor records and status word cleared
__global__ void k_factor_init(int nnodes, const int* is_leaf, const int* pcount, int sd, int* cnt,
                              uint64_t* fac_ptr, uint64_t* acc_ptr, int* osize, int* status) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nnodes) return;
  cnt[x] = is_leaf[x] ? pcount[x] * sd : 0;
  fac_ptr[x] = 0;
  acc_ptr[x] = 0;
  osize[x] = 0;
  if (x == 0) *status = 0;
}

// |active_box| (remove_inactive_dofs_at_box, quadtree.cpp:694-720)
// size of the list node y hands to its parent once its ID has run: its rank
// if compressed, else its active count (the elimination stores the same
// value in osize; the lists of the parent level are built from the ID's
// output before that elimination runs)
__device__ __forceinline__ int out_size(int y, const int* rank, const int* nact) {
  return rank[y] >= 0 ? rank[y] : nact[y];
}

// active sizes of level [a, b) and their exclusive prefix sums from `add`
// (the level's offsets in act_pool); one workgroup, chunks of kLsTB nodes.
// The sizes also go to host_nact (page-locked, device-visible).
constexpr int kLsTB = 1024;
__global__ __launch_bounds__(kLsTB) void k_active_size(int a, int b, const int* is_leaf,
                                                       const int* child, const int* pcount,
                                                       int sd, int* nact, int* rank,
                                                       int64_t* act_off, int64_t add,
                                                       int* host_nact) {
  __shared__ int64_t wtot[kLsTB / kWave];
  __shared__ int64_t carry;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  if (tid == 0) carry = add;
  __syncthreads();
  for (int base = a; base < b; base += kLsTB) {
    const int x = base + tid;
    int n = 0;
    if (x < b) {
      if (is_leaf[x]) {
        n = pcount[x] * sd;
      } else {
        for (int s = 0; s < 4; ++s) n += out_size(child[4 * x + s], rank, nact);
      }
      nact[x] = n;
      rank[x] = -1;
      if (host_nact) host_nact[x - a] = n;
    }
    int64_t incl = n;
    for (int d = 1; d < kWave; d <<= 1) {
      const int64_t t = __shfl_up(incl, d, kWave);
      if (lane >= d) incl += t;
    }
    if (lane == kWave - 1) wtot[wv] = incl;
    __syncthreads();
    int64_t before = carry;
    for (int w = 0; w < wv; ++w) before += wtot[w];
    if (x < b) act_off[x] = before + incl - n;
    __syncthreads();
    if (tid == kLsTB - 1) carry = before + incl;
    __syncthreads();
  }
}

// the list a node hands to its parent: skel (compressed) or active_box
__device__ __forceinline__ const int* out_list(int y, const int* rank, const int64_t* act_off,
                                               const int* act_pool, const int* sr_pool) {
  return rank[y] >= 0 ? sr_pool + act_off[y] : act_pool + act_off[y];
}

__global__ void k_active_fill(int a, const int* is_leaf, const int* child, const int64_t* obox_off,
                              const int* obox, const int* rank, const int64_t* act_off,
                              const int* nact, int* act_pool, const int* sr_pool) {
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
    int cnt = out_size(y, rank, nact);
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
// one wave per node.  pass 0 counts into nrows, pass 1 writes rows_pool and,
// with fac_ptr, every node's factor slab address (slab + soff * tsz).
__global__ void k_inner_rows(int a, int b, int min_dofs, GeomView g, const int* is_leaf,
                             const int64_t* obox_off, const int* obox, const int* pcount,
                             const int64_t* act_off, const int* nact, const int* act_pool,
                             const int* sl_nbr, const int* sl_cnt, const int64_t* cp_off,
                             const int* cp_nbr, const double* corners, const double* side,
                             int* nrows, const int64_t* rows_off, int* rows_pool, int pass,
                             int* host_rows, uint64_t* fac_ptr = nullptr, uint64_t slab = 0,
                             const int64_t* soff = nullptr, size_t tsz = 0) {
#pragma clang fp contract(off)
  int x = a + blockIdx.x;
  if (x >= b) return;
  const int lane = threadIdx.x;
  if (fac_ptr && lane == 0) fac_ptr[x] = slab + (uint64_t)soff[x - a] * tsz;
  if (nact[x] < min_dofs) {
    if (!pass && lane == 0) {
      nrows[x] = 0;
      if (host_rows) host_rows[x - a] = 0;
    }
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
  if (!pass && lane == 0) {
    nrows[x] = cnt;
    if (host_rows) host_rows[x - a] = cnt;
  }
}

// outside rows for levels 1 and 2 (kernel.cpp:164-207): every other node's
// active_box at this level, then (level 2) the level-1 leaves' original_box.
// With several instances only the box's own instance counts.
__global__ void k_outside_rows(const IdWork* work, int a, int b, int level, int l1a, int l1b,
                               const int* is_leaf, const int64_t* obox_off, const int* obox,
                               const int* pcount, int sd, const int64_t* act_off, const int* nact,
                               const int* act_pool, const int* inst, int* rows_pool) {
  const IdWork w = work[blockIdx.x];
  const int me = inst[w.node];
  int* out = rows_pool + w.rows_off;
  int o = 0;
  for (int y = a; y < b; ++y) {
    if (y == w.node || inst[y] != me) continue;
    const int* src = act_pool + act_off[y];
    for (int j = threadIdx.x; j < nact[y]; j += blockDim.x) out[o + j] = src[j];
    o += nact[y];
  }
  if (level == 2) {
    for (int y = l1a; y < l1b; ++y) {
      if (!is_leaf[y] || inst[y] != me) continue;
      const int* src = obox + obox_off[y];
      int n = pcount[y] * sd;
      for (int j = threadIdx.x; j < n; j += blockDim.x) out[o + j] = src[j];
      o += n;
    }
  }
}

// One launch per level over every box to compress: boxes whose ID matrix
// fits in LDS run there, the others keep it in the global workspace ws (put
// first in the work list, since they take longest), so both kinds share the
// chip instead of running one launch after the other.
template <typename T, typename TS>
__global__ __launch_bounds__(kIdTB) void k_id(const IdWork* work, GeomView g, int pde, double tol,
                                              const int* rows_pool, const int* act_pool,
                                              const int64_t* act_off, const double* ring, T* ws,
                                              int* rank, int* sr_pool, int* pos_pool,
                                              const uint64_t* fac_ptr, int* status) {
  extern __shared__ __align__(16) unsigned char smem[];
  const IdWork w = work[blockIdx.x];
  if (w.in_lds)
    id_box<T, TS, true>(w, smem, g, pde, tol, rows_pool, act_pool, act_off, ring, ws, rank,
                        sr_pool, pos_pool, fac_ptr, status);
  else
    id_box<T, TS, false>(w, smem, g, pde, tol, rows_pool, act_pool, act_off, ring, ws, rank,
                         sr_pool, pos_pool, fac_ptr, status);
}

// The boxes of id_wide_box, one 512-thread block each, on the same level.
template <typename T, typename TS>
__global__ __launch_bounds__(kIdWideTB) void k_id_wide(
    const IdWork* work, GeomView g, int pde, double tol, const int* rows_pool,
    const int* act_pool, const int64_t* act_off, const double* ring, T* ws, int* rank,
    int* sr_pool, int* pos_pool, const uint64_t* fac_ptr, int* status) {
  extern __shared__ __align__(16) unsigned char smem[];
  const IdWork w = work[blockIdx.x];
  if (w.in_lds)
    id_box<T, TS, true, false, kIdWideTB>(w, smem, g, pde, tol, rows_pool, act_pool, act_off,
                                          ring, ws, rank, sr_pool, pos_pool, fac_ptr, status);
  else
    id_box<T, TS, false, false, kIdWideTB>(w, smem, g, pde, tol, rows_pool, act_pool, act_off,
                                           ring, ws, rank, sr_pool, pos_pool, fac_ptr, status);
}

// The same pivoted QR and T on one given m x n matrix (one block): in LDS
// when it fits (w.in_lds), else in place in `given` (global memory).
// Outputs as for a tree box: rank[0] (-1 when not compressed), the pivots in
// pos_pool[0..n), T into *fac_ptr[0].
__global__ __launch_bounds__(kIdTB) void k_id_given(IdWork w, double* given, double tol,
                                                    const int* iota, const int64_t* zero_off,
                                                    int* rank, int* sr_pool, int* pos_pool,
                                                    const uint64_t* fac_ptr, int* status) {
  extern __shared__ __align__(16) unsigned char smem[];
  const GeomView g{};
  if (w.in_lds)
    id_box<double, double, true, true>(w, smem, g, 0, tol, nullptr, iota, zero_off, nullptr,
                                       given, rank, sr_pool, pos_pool, fac_ptr, status, given);
  else
    id_box<double, double, false, true>(w, smem, g, 0, tol, nullptr, iota, zero_off, nullptr,
                                        given, rank, sr_pool, pos_pool, fac_ptr, status, given);
}

template <typename TS, int PDE>
__global__ __launch_bounds__(kGramTB) void k_id_gram(const IdWork* work, GeomView g,
                                                     double tol, const int* rows_pool,
                                                     const int* act_pool, const int64_t* act_off,
                                                     const double* ring, int* rank, int* sr_pool,
                                                     int* pos_pool, const uint64_t* fac_ptr,
                                                     int* status) {
  extern __shared__ __align__(16) unsigned char smem[];
  id_gram_box<TS, PDE>(work[blockIdx.x], smem, g, tol, rows_pool, act_pool, act_off, ring, rank,
                       sr_pool, pos_pool, fac_ptr, status);
}

// elimination with one 1024-thread workgroup per box (decouple_block_box in
// elim_dev.h).  At most 64 VGPRs (8 waves per SIMD) so that two boxes share
// a CU when their LDS fits: the elimination of one box is a chain of
// latency-bound steps, and a second box fills the gaps.
template <typename T, bool LDS>
__global__ __launch_bounds__(kDecTB) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_decouple(const int* nodes, int a, GeomView g, int pde, const int* is_leaf,
                const int* child, const int* osize_in, const uint64_t* acc_ptr_in,
                const int* acc_ld_in, const int* nact, const int* rank, const int64_t* act_off,
                const int* sr_pool, const int* pos_pool, int* piv_pool, const uint64_t* fac_ptr,
                int* osize, uint64_t* acc_ptr, int* acc_ld, T* ws, size_t ws_stride,
                int* status) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_p;
  const int x = nodes ? nodes[blockIdx.x] : a + blockIdx.x;
  unsigned char* base = LDS ? smem : reinterpret_cast<unsigned char*>(ws + ws_stride * blockIdx.x);
  decouple_block_box<T>(x, base, &s_p, g, pde, is_leaf, child, osize_in, acc_ptr_in, acc_ld_in,
                        nact, rank, act_off, sr_pool, pos_pool, piv_pool, fac_ptr, osize,
                        acc_ptr, acc_ld, status);
}

// elimination with one wavefront per box (decouple_wave_box in elim_dev.h)
template <typename T, int RMAX>
__global__ __launch_bounds__(kWave) void k_decouple_wave(
    const int* nodes, GeomView g, int pde, const int* is_leaf, const int* child,
    const int* osize_in, const uint64_t* acc_ptr_in, const int* acc_ld_in, const int* nact,
    const int* rank, const int64_t* act_off, const int* sr_pool, const int* pos_pool,
    int* piv_pool, const uint64_t* fac_ptr, int* osize, uint64_t* acc_ptr, int* acc_ld,
    int* status) {
  extern __shared__ __align__(16) unsigned char smem[];
  decouple_wave_box<T, RMAX>(nodes[blockIdx.x], smem, g, pde, is_leaf, child, osize_in,
                             acc_ptr_in, acc_ld_in, nact, rank, act_off, sr_pool, pos_pool,
                             piv_pool, fac_ptr, osize, acc_ptr, acc_ld, status);
}

// ---------------------------------------------------------------------------
// root: allskel_mat = K(allskel, allskel) - the children's updates
// (skel_factorization.cpp:116-128); its inverse comes from dense_inverse.
// ---------------------------------------------------------------------------
// Grid (column groups, instances); the root of instance b is node b, and
// child s's update covers the diagonal block at the offset of the children
// before it (up to the first missing child).
template <typename T>
__global__ __launch_bounds__(kTB) void k_root_assemble(GeomView g, int pde, const int* child,
                                                       const int* osize, const uint64_t* acc_ptr,
                                                       const int* acc_ld, const int* root_n,
                                                       const int64_t* root_doff,
                                                       const int64_t* root_moff,
                                                       const int* dofs_all, double* Amat_all) {
  const int x = blockIdx.y;
  const int n = root_n[x];
  const int* dofs = dofs_all + root_doff[x];
  double* Amat = Amat_all + root_moff[x];
  int o[5];
  const T* acc[4];
  int ld[4];
  o[0] = 0;
  bool open = true;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int y = child[4 * x + s];
    open = open && y >= 0;
    const int sz = open ? osize[y] : 0;
    acc[s] = open ? reinterpret_cast<const T*>(acc_ptr[y]) : nullptr;
    ld[s] = open ? acc_ld[y] : 0;
    o[s + 1] = o[s] + sz;
  }
  const int lane = threadIdx.x & (kWave - 1);
  for (int j = blockIdx.x * kNW + (threadIdx.x >> 6); j < n; j += gridDim.x * kNW) {
    const int sj = (j >= o[1]) + (j >= o[2]) + (j >= o[3]);
    const bool inb = j < o[4];
    const int b0 = o[sj];
    const T* aj = nullptr;
    int lj = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (s == sj) {
        aj = acc[s];
        lj = ld[s];
      }
    for (int i = lane; i < n; i += kWave) {
      double v = kernel_entry(g, pde, dofs[i], dofs[j]);
      if (inb && aj && i >= b0 && i < o[sj + 1])
        v -= (double)aj[(i - b0) + (size_t)(j - b0) * lj];
      Amat[i + (size_t)j * n] = v;
    }
  }
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
  const size_t lds_cap = c.lds_max;
  // at most 9 markers per level (lists start/end, rows, ID start, ID kernel
  // start/end, ID end, elimination start/end) plus 2 for the root
  c.ensure_events(10 * (size_t)t.nlevels + 8);
  int ev = 0;
  auto mark = [&]() {
    if ((size_t)ev >= c.evpool.size()) KI_FAIL(KIGPU_ERR_STATE, "event pool exhausted");
    KI_HIP(hipEventRecord(c.evpool[ev], s));
    return ev++;
  };

  // leaves' original_box DOF lists
  DBuf<int> cnt;
  cnt.alloc(nn);
  t.obox_off.alloc(nn);
  t.act_off.alloc(nn);
  t.nact.alloc(nn);
  t.nrows.alloc(nn);
  t.rank.alloc(nn);
  t.osize.alloc(nn);
  t.fac_ptr.alloc(nn);
  t.acc_ptr.alloc(nn);
  t.acc_ld.alloc(nn);
  t.status.alloc(1);
  k_factor_init<<<ceil_div(nn, kTB), kTB, 0, s>>>(nn, t.is_leaf.p, t.pcount.p, t.sd, cnt.p,
                                                 t.fac_ptr.p, t.acc_ptr.p, t.osize.p, t.status.p);
  exclusive_scan_i32_to_i64(cnt.p, t.obox_off.p, nn, s);
  t.obox_pool.alloc(t.ndof);
  k_obox<<<nn, 64, 0, s>>>(nn, t.is_leaf.p, t.pcount.p, t.pstart.p, t.perm_all.p, t.obox_off.p,
                           t.sd, t.obox_pool.p);
  const int64_t pool = (int64_t)t.ndof * t.nlevels + 64;
  t.act_pool.alloc(pool);
  t.sr_pool.alloc(pool);
  t.pos_pool.alloc(pool);
  t.piv_pool.alloc(pool);
  t.act_pool_size = pool;
  t.drop_graphs();  // they hold the previous factorization's launches
  t.work.clear();
  t.work.resize(t.nlevels);
  std::vector<int> h_n, h_q;
  int64_t pool_base = 0;
  struct LevelEv {
    int l0, l1, pre, r0, id0, id1, dec0, dec1, idl0, idl1;
  };
  std::vector<LevelEv> lev;

  // ---- active lists of level l (remove_inactive_dofs_at_level,
  // quadtree.cpp:664-691) and, for l >= 3, the inner-row counts, whose sizes
  // are copied to page-locked memory (read on the host once marker l1 is
  // done).  They need only the ID of level l + 1, so they are queued ahead of
  // that level's elimination, which then runs while the host builds level
  // l's work lists.
  auto launch_lists = [&](int l, LevelEv& E) -> int* {
    const int a = (int)t.level_off[l], b = (int)t.level_off[l + 1], cn = b - a;
    E = LevelEv{};
    E.idl0 = E.idl1 = -1;
    E.l0 = mark();
    // sizes and row counts go straight to the page-locked block (the level
    // above's were read before this call)
    int* hp = l > 0 ? c.pin_down.get<int>(2 * (size_t)cn, s) : nullptr;
    int* dp = l > 0 ? c.pin_down.dget<int>() : nullptr;
    k_active_size<<<1, kLsTB, 0, s>>>(a, b, t.is_leaf.p, t.child.p, t.pcount.p, t.sd, t.nact.p,
                                      t.rank.p, t.act_off.p, pool_base, dp);
    k_active_fill<<<cn, 64, 0, s>>>(a, t.is_leaf.p, t.child.p, t.obox_off.p, t.obox_pool.p,
                                    t.rank.p, t.act_off.p, t.nact.p, t.act_pool.p, t.sr_pool.p);
    if (l >= 3)
      k_inner_rows<<<cn, kWave, 0, s>>>(a, b, o.min_dofs, g, t.is_leaf.p, t.obox_off.p,
                                        t.obox_pool.p, t.pcount.p, t.act_off.p, t.nact.p,
                                        t.act_pool.p, t.sl_nbr.p, t.sl_cnt.p, t.cp_off.p,
                                        t.cp_nbr.p, t.corners.p, t.side.p, t.nrows.p, nullptr,
                                        nullptr, 0, dp + cn);
    E.l1 = mark();
    E.pre = E.l1;
    return hp;
  };

  LevelEv E;
  int* hlists = launch_lists(t.nlevels - 1, E);
  for (int l = t.nlevels - 1; l >= 0; --l) {
    const int a = (int)t.level_off[l], b = (int)t.level_off[l + 1], cn = b - a;
    auto lw = std::make_unique<LevelWork>();
    lw->level = l;
    lw->a = a;
    lw->b = b;
    if (l == 0) {  // the roots' lists (queued ahead of level 1's elimination)
      lev.push_back(E);
      t.work[l] = std::move(lw);
      break;
    }
    event_wait(c.evpool[E.l1]);
    ++c.tim.syncs;
    h_n.assign(hlists, hlists + cn);
    if (l >= 3) h_q.assign(hlists + cn, hlists + 2 * cn);
    else h_q.assign(cn, 0);
    int64_t level_dofs = 0;
    for (int i = 0; i < cn; ++i) level_dofs += h_n[i];
    if (l <= 2) {
      // rows of a box: the other boxes of its instance on this level, plus
      // (level 2) its instance's level-1 leaves
      std::vector<int64_t> inst_dofs(t.ninst, 0), extra(t.ninst, 0);
      for (int i = 0; i < cn; ++i) inst_dofs[t.h_inst[a + i]] += h_n[i];
      if (l == 2) {
        for (int y = (int)t.level_off[1]; y < (int)t.level_off[2]; ++y)
          if (t.h_is_leaf[y]) extra[t.h_inst[y]] += (int64_t)t.h_pcount[y] * t.sd;
      }
      for (int i = 0; i < cn; ++i) {
        const int ii = t.h_inst[a + i];
        h_q[i] = (int)(inst_dofs[ii] - h_n[i] + extra[ii]);
      }
    }
    // ---- work list: boxes with >= min_dofs active DOFs and a non-empty ID matrix
    std::vector<IdWork> wl_lds, wl_glb, wide_lds, wide_glb, wl_gram;
    const bool use_gram =
        c.id_method == 2 || (c.id_method == 0 && sizeof(T) == 4 && o.id_tol >= kGramMinTol);
    size_t gram_bytes = 0;
    std::vector<int64_t> rows_off(cn, 0), soff(cn);
    int64_t rtot = 0, atot = 0, stot = 0;
    int nmax = 0;
    size_t lds_bytes = 0, glds = 0, wide_bytes = 0;
    for (int i = 0; i < cn; ++i) {
      nmax = std::max(nmax, h_n[i]);
      soff[i] = stot;
      stot += (int64_t)slab_elems(h_n[i]);
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
      if (use_gram && w.n <= kGramMaxN) {  // ID from the Gram matrix: no stored A
        gram_bytes = std::max(gram_bytes, gram_layout(w.n, w.q).total);
        wl_gram.push_back(w);
        continue;
      }
      const size_t in_l = id_lds_layout(w.m, w.n, w.q, true, sizeof(double)).total;
      w.in_lds = in_l <= lds_cap;
      w.a_off = atot;
      const bool wide = id_wide_box(w.m, w.n);
      if (w.in_lds) {
        (wide ? wide_bytes : lds_bytes) = std::max(wide ? wide_bytes : lds_bytes, in_l);
        (wide ? wide_lds : wl_lds).push_back(w);
      } else {
        atot += (int64_t)w.m * w.n;
        const size_t gl = id_lds_layout(w.m, w.n, w.q, false, sizeof(double)).total;
        (wide ? wide_bytes : glds) = std::max(wide ? wide_bytes : glds, gl);
        (wide ? wide_glb : wl_glb).push_back(w);
      }
    }
    lw->nmax = nmax;
    // one ID work list: the global-workspace boxes first, then the LDS ones;
    // the wide boxes in a list of their own (k_id_wide), after the main one
    // in the same upload and the same row pool
    std::vector<IdWork> wl(wl_glb);
    wl.insert(wl.end(), wl_lds.begin(), wl_lds.end());
    const int nwl = (int)wl.size();
    wl.insert(wl.end(), wide_glb.begin(), wide_glb.end());
    wl.insert(wl.end(), wide_lds.begin(), wide_lds.end());
    const int nwide = (int)wl.size() - nwl;
    wl.insert(wl.end(), wl_gram.begin(), wl_gram.end());
    const int ngram = (int)wl.size() - nwl - nwide;
    // elimination: boxes of at most kWdecMaxN active DOFs one wavefront
    // each, in two size classes (their LDS differs fourfold), largest
    // first; the larger boxes a workgroup each
    std::vector<int> dec_small, dec_mid, dec_large;
    int nmax_small = 0, nmax_mid = 0, nmax_large = 0;
    {
      std::vector<int> ord(cn);
      for (int i = 0; i < cn; ++i) ord[i] = i;
      std::stable_sort(ord.begin(), ord.end(), [&](int p, int q) { return h_n[p] > h_n[q]; });
      for (int i : ord) {
        const int nb = h_n[i];
        const bool wave = nb <= kWdecMaxN && c.tune.dec_wave[std::max(0, (nb - 1) / 16)];
        if (!wave) {
          dec_large.push_back(a + i);
          nmax_large = std::max(nmax_large, nb);
        } else if (nb <= kWdecMaxN / 2) {
          dec_small.push_back(a + i);
          nmax_small = std::max(nmax_small, nb);
        } else if (nb <= kWdecMaxN) {
          dec_mid.push_back(a + i);
          nmax_mid = std::max(nmax_mid, nb);
        } else {
          dec_large.push_back(a + i);
          nmax_large = std::max(nmax_large, nb);
        }
      }
    }
    // ---- the level's lists in one copy
    UploadPack pk;
    const size_t o_soff = pk.add(soff.data(), sizeof(int64_t) * cn);
    const size_t o_roff = pk.add(rows_off.data(), sizeof(int64_t) * cn);
    const size_t o_wl = pk.add(wl.data(), sizeof(IdWork) * wl.size());
    const size_t o_ds = pk.add(dec_small.data(), sizeof(int) * dec_small.size());
    const size_t o_dm = pk.add(dec_mid.data(), sizeof(int) * dec_mid.size());
    const size_t o_dl = pk.add(dec_large.data(), sizeof(int) * dec_large.size());
    unsigned char* dpk = pk.send(c.pin_up, lw->pack, s);
    const int* d_dec_small = reinterpret_cast<const int*>(dpk + o_ds);
    const int* d_dec_mid = reinterpret_cast<const int*>(dpk + o_dm);
    const int* d_dec_large = reinterpret_cast<const int*>(dpk + o_dl);
    const int64_t* d_soff = reinterpret_cast<const int64_t*>(dpk + o_soff);
    const int64_t* d_roff = reinterpret_cast<const int64_t*>(dpk + o_roff);
    const IdWork* d_wl = reinterpret_cast<const IdWork*>(dpk + o_wl);
    // ---- factor slabs for every node of the level
    E.r0 = mark();
    T* slab;
    if (sizeof(T) == 8) {
      lw->slab64.alloc(stot);
      slab = reinterpret_cast<T*>(lw->slab64.p);
    } else {
      lw->slab32.alloc(stot);
      slab = reinterpret_cast<T*>(lw->slab32.p);
    }
    // (levels >= 3: set by the row kernel below)
    if (l <= 2)
      k_set_ptr<<<ceil_div(cn, kTB), kTB, 0, s>>>(t.fac_ptr.p, a, b, t.nact.p,
                                                 reinterpret_cast<uint64_t>(slab), d_soff,
                                                 sizeof(T));
    // ---- ID rows
    lw->rows.alloc(std::max<int64_t>(rtot, 1));
    if (l >= 3) {
      k_inner_rows<<<cn, kWave, 0, s>>>(a, b, o.min_dofs, g, t.is_leaf.p, t.obox_off.p,
                                        t.obox_pool.p, t.pcount.p, t.act_off.p, t.nact.p,
                                        t.act_pool.p, t.sl_nbr.p, t.sl_cnt.p, t.cp_off.p,
                                        t.cp_nbr.p, t.corners.p, t.side.p, t.nrows.p,
                                        d_roff - a, lw->rows.p, 1, nullptr, t.fac_ptr.p,
                                        reinterpret_cast<uint64_t>(slab), d_soff, sizeof(T));
    }
    if (l <= 2 && nwl + nwide + ngram > 0)
      k_outside_rows<<<nwl + nwide + ngram, kTB, 0, s>>>(
          d_wl, a, b, l, (int)t.level_off[1], (int)t.level_off[2], t.is_leaf.p, t.obox_off.p,
          t.obox_pool.p, t.pcount.p, t.sd, t.act_off.p, t.nact.p, t.act_pool.p, t.inst.p,
          lw->rows.p);
    // ---- ID
    E.id0 = mark();
    if (nwl + nwide + ngram > 0) {
      if (!wl_glb.empty() || !wide_glb.empty()) lw->ws.alloc(atot);
      E.idl0 = mark();
      if (ngram > 0) {
        auto kern = pde == KIGPU_STOKES           ? k_id_gram<T, KIGPU_STOKES>
                    : pde == KIGPU_LAPLACE_NEUMANN ? k_id_gram<T, KIGPU_LAPLACE_NEUMANN>
                                                   : k_id_gram<T, KIGPU_LAPLACE>;
        allow_lds(kern, gram_bytes);
        kern<<<ngram, kGramTB, gram_bytes, s>>>(d_wl + nwl + nwide, g, o.id_tol, lw->rows.p,
                                                t.act_pool.p, t.act_off.p, c.ring.p, t.rank.p,
                                                t.sr_pool.p, t.pos_pool.p, t.fac_ptr.p,
                                                t.status.p);
        KI_HIP(hipGetLastError());
      }
      // the wide boxes on the side stream, beside the main launch
      if (nwide > 0) {
        KI_HIP(hipEventRecord(c.ev_fork, s));
        KI_HIP(hipStreamWaitEvent(c.side, c.ev_fork, 0));
        allow_lds(k_id_wide<double, T>, wide_bytes);
        k_id_wide<double, T><<<nwide, kIdWideTB, wide_bytes, c.side>>>(
            d_wl + nwl, g, pde, o.id_tol, lw->rows.p, t.act_pool.p, t.act_off.p, c.ring.p,
            lw->ws.p, t.rank.p, t.sr_pool.p, t.pos_pool.p, t.fac_ptr.p, t.status.p);
        KI_HIP(hipGetLastError());
      }
      if (nwl > 0) {
        const size_t ids = std::max(lds_bytes, glds);
        allow_lds(k_id<double, T>, ids);
        k_id<double, T><<<nwl, kIdTB, ids, s>>>(d_wl, g, pde, o.id_tol, lw->rows.p, t.act_pool.p,
                                                t.act_off.p, c.ring.p, lw->ws.p, t.rank.p,
                                                t.sr_pool.p, t.pos_pool.p, t.fac_ptr.p,
                                                t.status.p);
        KI_HIP(hipGetLastError());
      }
      if (nwide > 0) {
        KI_HIP(hipEventRecord(c.ev_join, c.side));
        KI_HIP(hipStreamWaitEvent(s, c.ev_join, 0));
      }
      E.idl1 = mark();
      lw->id_boxes = wl;
      lw->ngram = ngram;
    }
    E.id1 = mark();
    // the next level's lists, then this level's elimination
    pool_base += level_dofs;  // the next level's lists start after this level's
    LevelEv En;
    hlists = launch_lists(l - 1, En);
    // ---- elimination for every node of the level
    E.dec0 = mark();
    {
      auto wave_launch = [&](const int* nodes, int count, int nmx, hipStream_t st) {
        const size_t lds = wdec_lds_bound(std::max(nmx, 1), sizeof(T));
        auto kern = nmx <= kWdecMaxN / 2 ? k_decouple_wave<T, kWdecMaxN / 2>
                                         : k_decouple_wave<T, kWdecMaxN>;
        allow_lds(kern, lds);
        kern<<<count, kWave, lds, st>>>(
            nodes, g, pde, t.is_leaf.p, t.child.p, t.osize.p, t.acc_ptr.p, t.acc_ld.p,
            t.nact.p, t.rank.p, t.act_off.p, t.sr_pool.p, t.pos_pool.p, t.piv_pool.p,
            t.fac_ptr.p, t.osize.p, t.acc_ptr.p, t.acc_ld.p, t.status.p);
      };
      auto large_launch = [&](hipStream_t st) {
        const int nl = (int)dec_large.size();
        DecLds L = dec_layout(nmax_large, sizeof(T));
        if (L.total <= lds_cap) {
          allow_lds(k_decouple<T, true>, L.total);
          k_decouple<T, true><<<nl, kDecTB, L.total, st>>>(
              d_dec_large, a, g, pde, t.is_leaf.p, t.child.p, t.osize.p, t.acc_ptr.p, t.acc_ld.p,
              t.nact.p, t.rank.p, t.act_off.p, t.sr_pool.p, t.pos_pool.p, t.piv_pool.p,
              t.fac_ptr.p, t.osize.p, t.acc_ptr.p, t.acc_ld.p, (T*)nullptr, 0,
              t.status.p);
          return;
        }
        size_t stride = (L.total + sizeof(T) - 1) / sizeof(T);
        T* ws;
        if (sizeof(T) == 8) {
          lw->dws64.alloc(stride * nl);
          ws = reinterpret_cast<T*>(lw->dws64.p);
        } else {
          lw->dws32.alloc(stride * nl);
          ws = reinterpret_cast<T*>(lw->dws32.p);
        }
        k_decouple<T, false><<<nl, kDecTB, 0, st>>>(
            d_dec_large, a, g, pde, t.is_leaf.p, t.child.p, t.osize.p, t.acc_ptr.p, t.acc_ld.p,
            t.nact.p, t.rank.p, t.act_off.p, t.sr_pool.p, t.pos_pool.p, t.piv_pool.p,
            t.fac_ptr.p, t.osize.p, t.acc_ptr.p, t.acc_ld.p, ws, stride,
            t.status.p);
      };
      // the middle size class on the side stream beside the small boxes and
      // then the large ones on the main stream
      const bool side = c.tune.mid_on_side && !dec_mid.empty() &&
                        (!dec_small.empty() || !dec_large.empty());
      if (side) {
        KI_HIP(hipEventRecord(c.ev_fork, s));
        KI_HIP(hipStreamWaitEvent(c.side, c.ev_fork, 0));
        wave_launch(d_dec_mid, (int)dec_mid.size(), nmax_mid, c.side);
      } else if (!dec_mid.empty()) {
        wave_launch(d_dec_mid, (int)dec_mid.size(), nmax_mid, s);
      }
      if (!dec_small.empty()) wave_launch(d_dec_small, (int)dec_small.size(), nmax_small, s);
      if (!dec_large.empty()) large_launch(s);
      if (side) {
        KI_HIP(hipEventRecord(c.ev_join, c.side));
        KI_HIP(hipStreamWaitEvent(s, c.ev_join, 0));
      }
      KI_HIP(hipGetLastError());
    }
    E.dec1 = mark();
    lev.push_back(E);
    t.work[l] = std::move(lw);
    En.pre = E.dec1;  // the next level's upload follows this elimination
    E = En;
  }

  // ---- roots (level 0), one per instance
  const int e_root0 = mark();
  const int ni = t.ninst;
  int64_t* pin = c.pin_down.get<int64_t>(2 * (size_t)ni, s);
  copy_to_host(c.pin_down, s,
               {{t.act_off.p, 0, sizeof(int64_t) * ni}, {t.nact.p, sizeof(int64_t) * ni,
                                                         sizeof(int) * ni}});
  stream_wait(s);
  ++c.tim.syncs;
  const int64_t r_off0 = pin[0];
  const int* pn = reinterpret_cast<const int*>(pin + ni);
  t.h_root_n.assign(pn, pn + ni);
  t.h_root_doff.assign(ni + 1, 0);
  t.h_root_moff.assign(ni + 1, 0);
  t.nroot_max = 0;
  for (int i = 0; i < ni; ++i) {
    const int64_t n0 = t.h_root_n[i];
    t.h_root_doff[i + 1] = t.h_root_doff[i] + n0;
    t.h_root_moff[i + 1] = t.h_root_moff[i] + n0 * n0;
    t.nroot_max = std::max(t.nroot_max, (int)n0);
  }
  t.nroot = (int)t.h_root_doff[ni];
  if (t.nroot_max > kDenseInvMax) KI_FAIL(KIGPU_ERR_NOTSUPP, "root skeleton larger than 1024 DOFs");
  if (t.nroot > 0) {
    // the roots' active lists are contiguous in act_pool (level 0 is last)
    t.root_dofs.alloc(t.nroot);
    KI_HIP(hipMemcpyAsync(t.root_dofs.p, t.act_pool.p + r_off0, sizeof(int) * t.nroot,
                          hipMemcpyDeviceToDevice, s));
    UploadPack pk;
    const size_t o_n = pk.add(t.h_root_n.data(), sizeof(int) * ni);
    const size_t o_d = pk.add(t.h_root_doff.data(), sizeof(int64_t) * (ni + 1));
    const size_t o_m = pk.add(t.h_root_moff.data(), sizeof(int64_t) * (ni + 1));
    unsigned char* dpk = pk.send(c.pin_up, t.root_pack, s);
    t.root_n_p = reinterpret_cast<const int*>(dpk + o_n);
    t.root_doff_p = reinterpret_cast<const int64_t*>(dpk + o_d);
    t.root_moff_p = reinterpret_cast<const int64_t*>(dpk + o_m);
    const size_t msz = std::max<int64_t>(t.h_root_moff[ni], 1);
    t.root_lu.alloc(msz);   // allskel_mat^{-1}
    t.root_mat.alloc(msz);  // allskel_mat
    t.root_piv.alloc(t.nroot);
    k_root_assemble<T><<<dim3(ceil_div(t.nroot_max, kNW), ni), kTB, 0, s>>>(
        g, pde, t.child.p, t.osize.p, t.acc_ptr.p, t.acc_ld.p, t.root_n_p, t.root_doff_p,
        t.root_moff_p, t.root_dofs.p, t.root_mat.p);
    KI_HIP(hipGetLastError());
    dense_inverse(c, s, ni, t.nroot_max, t.root_n_p, t.root_doff_p, t.root_moff_p, 0,
                  (size_t)t.h_root_moff[ni], (size_t)t.nroot, t.root_mat.p, t.root_lu.p,
                  t.root_piv.p, t.status.p, 8);
    KI_HIP(hipGetLastError());
  }
  const int e_root1 = mark();
  stream_wait(s);
  ++c.tim.syncs;
  auto el = [&](int e0, int e1) {
    float ms = 0;
    KI_HIP(hipEventElapsedTime(&ms, c.evpool[e0], c.evpool[e1]));
    return ms * 1e-3;
  };
  for (const LevelEv& E : lev) {
    // the lists' kernels, and the wait for the host's work list plus the
    // ID rows (from the marker before the level's upload to the ID)
    c.tim.t_lists += el(E.l0, E.l1);
    if (E.id0 > E.r0 && E.r0 > 0) c.tim.t_lists += el(E.pre, E.id0);
    if (E.id1 > E.id0) {
      c.tim.t_id += el(E.id0, E.id1);
      c.tim.t_decouple += el(E.dec0, E.dec1);
    }
    if (E.idl0 >= 0) {
      c.tim.id_kernel_s += el(E.idl0, E.idl1);
      c.tim.id_launches += 1;
    }
    if (E.dec1 > E.dec0) {
      c.tim.dec_kernel_s += el(E.dec0, E.dec1);
      c.tim.dec_launches += 1;
    }
  }
  c.tim.t_root += el(e_root0, e_root1);
}

void interp_decomp(Context& c, const double* h_A, int m, int n, double tol, int* k_out,
                   int* perm, double* h_T) {
  hipStream_t s = c.stream;
  DBuf<double> A, T;
  DBuf<int> iota, rank, srp, posp, status;
  DBuf<int64_t> zero;
  DBuf<uint64_t> fp;
  A.upload(h_A, (size_t)m * n, s);
  std::vector<int> io(n);
  for (int j = 0; j < n; ++j) io[j] = j;
  iota.upload(io, s);
  const int64_t z = 0;
  zero.upload(&z, 1, s);
  T.alloc((size_t)n * n);
  const uint64_t tp = reinterpret_cast<uint64_t>(T.p);
  fp.upload(&tp, 1, s);
  rank.alloc(1);
  srp.alloc(n);
  posp.alloc(n);
  status.alloc(1);
  status.zero(s);
  IdWork w{};
  w.node = 0;
  w.n = n;
  w.q = 0;
  w.m = m;
  const size_t in_l = id_lds_layout(m, n, 0, true, sizeof(double)).total;
  w.in_lds = in_l <= c.lds_max;
  const size_t lds = w.in_lds ? in_l : id_lds_layout(m, n, 0, false, sizeof(double)).total;
  allow_lds(k_id_given, lds);
  k_id_given<<<1, kIdTB, lds, s>>>(w, A.p, tol, iota.p, zero.p, rank.p, srp.p, posp.p, fp.p,
                                   status.p);
  KI_HIP(hipGetLastError());
  int rk = -1, st = 0;
  KI_HIP(hipMemcpyAsync(&rk, rank.p, sizeof(int), hipMemcpyDeviceToHost, s));
  KI_HIP(hipMemcpyAsync(&st, status.p, sizeof(int), hipMemcpyDeviceToHost, s));
  stream_wait(s);
  if (st & 2) KI_FAIL(KIGPU_ERR_RANK, "ID matrix with fewer rows than columns");
  *k_out = rk > 0 ? rk : 0;
  if (*k_out > 0) {
    if (perm) posp.download(perm, n, s);
    if (h_T) T.download(h_T, (size_t)*k_out * (n - *k_out), s);
    stream_wait(s);
  }
}

// Algorithmic flops of the ID kernel for one box, as executed: entry
// generation (c_gen flops per entry), initial column norms, the QR steps up to
// the rank-revealing one (all n steps when the box is not compressed) and the
// T = R11^{-1} R12 solve.
static double id_flops(int m, int n, int k, double c_gen) {
  double f = c_gen * m * n + 2.0 * m * n;
  const double st = k > 0 ? k + 1 : n;  // QR steps
  // sum_{i<st} 2(m - i)  +  sum_{i<st-1} 4(m - i)(n - 1 - i), in closed form
  const double s1 = st - 1;                       // terms of the second sum
  const double j1 = s1 * (s1 - 1) / 2;            // sum of i
  const double j2 = (s1 - 1) * s1 * (2 * s1 - 1) / 6;  // sum of i^2
  f += 2.0 * (m * st - st * (st - 1) / 2);
  if (s1 > 0) f += 4.0 * ((double)m * (n - 1) * s1 - (double)(m + n - 1) * j1 + j2);
  if (k > 0) f += (double)k * k * (n - k);
  return f;
}

// The same for the Gram-matrix ID: generation, one triangle of A^T A
// (m n (n + 1)), the pivoted Cholesky's rows of R up to the rank-revealing
// step (2 i (n - i) at step i) and the T solve.
static double gram_id_flops(int m, int n, int k, double c_gen) {
  double f = c_gen * m * n + (double)m * n * (n + 1);
  const double st = k > 0 ? k + 1 : n;
  f += 2.0 * n * (st * (st - 1) / 2) - 2.0 * ((st - 1) * st * (2 * st - 1) / 6);
  if (k > 0) f += (double)k * k * (n - k);
  return f;
}

// elimination: K_BN generation plus 6k^2r + 8kr^2 + 2/3 r^3
static double dec_flops(int n, int k, double c_gen) {
  if (k <= 0) return 0;
  const double r = n - k;
  return c_gen * n * n + 6.0 * k * k * r + 8.0 * k * r * r + 2.0 / 3.0 * r * r * r;
}

TreeData::~TreeData() { drop_graphs(); }

void factor_tree(Context& c, TreeData& t, const kigpu_factor_opts& o) {
  c.tim.t_lists = c.tim.t_id = c.tim.t_decouple = c.tim.t_root = 0;
  c.tim.id_launches = c.tim.dec_launches = 0;
  c.tim.id_kernel_s = c.tim.id_flops = c.tim.dec_kernel_s = c.tim.dec_flops = 0;
  if (o.precision == KIGPU_FP32_REFINE) factor_levels<float>(c, t, o);
  else factor_levels<double>(c, t, o);
  if (t.nborder > 0) {
    if (o.precision == KIGPU_FP32_REFINE) border_factor<float>(c, t);
    else border_factor<double>(c, t);
  }
  // status word and host copies of the per-node results, one synchronization
  const int nn = t.nnodes;
  hipStream_t s = c.stream;
  int64_t* pin = c.pin_down.get<int64_t>(2 * (size_t)nn + 8, s);
  int* pi = reinterpret_cast<int*>(pin + nn);
  const size_t b8 = sizeof(int64_t) * nn, b4 = sizeof(int) * nn;
  copy_to_host(c.pin_down, s,
               {{t.act_off.p, 0, b8}, {t.rank.p, b8, b4}, {t.nact.p, b8 + b4, b4},
                {t.status.p, b8 + 2 * b4, sizeof(int)}});
  stream_wait(s);
  const int st = pi[2 * nn];
  if (st & 2) KI_FAIL(KIGPU_ERR_RANK, "ID matrix with fewer rows than columns");
  if (st & kStatusSingular) {
    std::string what = "zero or non-finite pivot in";
    if (st & 4) what += " an X_rr inverse";
    if (st & 8) what += " the root inverse";
    if (st & 16) what += " the bordered inverse";
    KI_FAIL(KIGPU_ERR_SINGULAR, what);
  }
  t.h_act_off.assign(pin, pin + nn);
  t.h_rank.assign(pi, pi + nn);
  t.h_nact.assign(pi + nn, pi + 2 * nn);
  t.fmirrors_ready = false;
  const double c_gen = o.pde == KIGPU_STOKES ? 20.0 : 11.0;
  for (auto& lw : t.work) {
    if (!lw) continue;
    const size_t g0 = lw->id_boxes.size() - lw->ngram;
    for (size_t i = 0; i < lw->id_boxes.size(); ++i) {
      const IdWork& w = lw->id_boxes[i];
      c.tim.id_flops += (i >= g0 ? gram_id_flops : id_flops)(w.m, w.n, t.h_rank[w.node], c_gen);
    }
    for (int x = lw->a; x < lw->b; ++x) c.tim.dec_flops += dec_flops(t.h_nact[x], t.h_rank[x], c_gen);
  }
  t.opts = o;
  t.factored = true;
}

}  // namespace kigpu
