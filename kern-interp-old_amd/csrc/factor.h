// Factorization / solve internals shared by factor.hip and capi.cpp.
#pragma once

#include "tree.h"

namespace kigpu {

// Bits of the device status word (TreeData::status) that the factorization
// kernels set: 2 an ID matrix with fewer rows than columns, 4 an X_rr inverse
// with a zero or non-finite pivot, 8 the same for a root inverse, 16 for the
// bordered inverse.  The host turns 2 into KIGPU_ERR_RANK and 4 | 8 | 16
// into KIGPU_ERR_SINGULAR.
constexpr int kStatusSingular = 4 | 8 | 16;

// one box to compress at a level
struct LevelWork {
  int level = 0;
  int a = 0, b = 0;         // node id range
  int nmax = 0;             // largest |active_box| on the level
  std::vector<IdWork> id_boxes;  // the level's ID work list (flop accounting)
  int ngram = 0;                 // its last ngram boxes took the Gram-matrix ID
  DBuf<double> slab64;      // factor slabs of this level's nodes
  DBuf<float> slab32;
  // per-level device temporaries, kept until the tree is freed so that no
  // hipFree stalls the level loop
  DBuf<unsigned char> pack;  // slab offsets, ID row offsets and the two ID work lists
  DBuf<int> rows;
  DBuf<double> ws, dws64;
  DBuf<float> dws32;
};

// Launch choices the autotuner sets (kigpu_tuning in kigpu.h mirrors it).
struct Tuning {
  // elimination: per box-size class n in 1-16, 17-32, 33-48, 49-64 active
  // DOFs, 1 = one wavefront per box, 0 = one 1024-thread workgroup per box
  int dec_wave[4] = {1, 1, 1, 1};
  int mid_on_side = 1;       // the 33-64 wavefront class on the side stream
  int pair_sweeps = 1;       // two tree levels per sweep launch where LDS allows
  int apply_split_cap = 64;  // source splits of the direct K apply (one RHS block)
  int fwd_split_cap = 256;   // source splits of the forward / transposed applies
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // second stream for work that overlaps the main one
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev[16];
  std::vector<hipEvent_t> evpool;  // per-level phase markers of a factorization
  kigpu_timings tim{};
  int refine_apply = 0;     // refinement residual: 0 treecode where it applies, 1 direct sum
  int id_method = 0;        // ID: 0 by precision (FP32: Gram matrix, FP64: pivoted QR), 1 pivoted QR, 2 Gram matrix
  bool graphs = false;      // replay the solve sweeps from captured HIP graphs
  Tuning tune;
  DBuf<double> ring;        // proxy unit vectors cos | sin (host libm)
  size_t lds_max = 65536;
  PinnedHost pin_down;      // per-level downloads
  PinnedHost pin_up;        // packed per-level uploads
  void ensure_events(size_t n) { kigpu::ensure_events(evpool, n); }
  ~Context() {
    for (auto e : evpool) (void)hipEventDestroy(e);
  }
};

void factor_tree(Context& c, TreeData& t, const kigpu_factor_opts& o);
// ID of one host matrix (m x n, column-major) by the factorization's pivoted
// QR: rank k (0 = not compressed), pivots perm[n], T (k x (n - k)).
void interp_decomp(Context& c, const double* h_A, int m, int n, double tol, int* k,
                   int* perm, double* h_T);
// the proxy block of a box against explicit proxy points (blocks.hip)
void proxy_block(Context& c, const GeomView& g, int pde, const double* h_pp, const double* h_pn,
                 int np, double pw, const int* h_box, int nbox, double* h_out);
// the elimination of one box from its K_BN (blocks.hip)
void schur_eliminate(Context& c, const double* h_K, int n, int k, const int* h_perm,
                     const double* h_T, double* X_rr, double* X_lu, int* piv, double* L,
                     double* U, double* schur, int* singular);
// Largest order of a root or bordered matrix that dense_inverse takes.
constexpr int kDenseInvMax = 1024;
// Inverses of ni dense FP64 matrices (dense_inv.hip).  Matrix x has order
// d_n[x] and sits column-major at offset d_moff[x] of A; its inverse goes to
// the same offset of Ainv and the pivot row of each elimination step to
// offset d_doff[x] of piv (if not null).  d_n == nullptr: one matrix of
// order n0 at offset 0.  mtot, dtot: the totals of the orders squared and of
// the orders.  A zero or non-finite pivot sets `bit` in *status.  Queued on
// stream s; no host wait.
void dense_inverse(Context& c, hipStream_t s, int ni, int nmax, const int* d_n,
                   const int64_t* d_doff, const int64_t* d_moff, int n0, size_t mtot, size_t dtot,
                   const double* A, double* Ainv, int* piv, int* status, int bit);
// bordered system (border.hip)
template <typename T>
void border_factor(Context& c, TreeData& t);
template <typename T>
void border_step(Context& c, TreeData& t, T* x, int ldx, int nrhs, const double* g);
// Largest ndof * nrhs of solve_tree_noalloc on an FP32 factor: the
// per-GPU static scratch holds that many floats (16 MiB).
constexpr int kNoAllocMaxEntries = 1 << 22;
// the scratch's device address on `device`, once per device (context creation)
void noalloc_init(int device);
void solve_tree_noalloc(Context& c, TreeData& t, const double* d_b, int nrhs, double* d_x);
void solve_tree(Context& c, TreeData& t, const double* d_b, int nrhs, double* d_x,
                double* d_alpha, const kigpu_refine_opts* ro, kigpu_solve_info* info,
                bool trans = false);
// treecode.hip: K mu for the Laplace double layer by a treecode (series
// truncation below 1e-14 relative); the refinement's apply.  mu, y: ndof x
// nrhs, column-major with leading dimension ndof
bool treecode_usable(const TreeData& t);
// tev: three events recorded around the series and evaluation kernels (or
// nullptr); treecode_timing adds their times to the context's timings once
// they have completed.  No synchronization inside.
void treecode_apply(Context& c, const TreeData& t, const double* d_mu, int nrhs, double* d_y,
                    const hipEvent_t* tev);
void treecode_timing(Context& c, const hipEvent_t* tev);
// the tree's treecode lists and box centres, on stream st (once per tree)
void treecode_prepare(Context& c, const TreeData& t, hipStream_t st);
void kernel_apply(Context& c, const TreeData& t, int pde, const double* d_mu, int nrhs,
                  double* d_y);
void kernel_apply_transpose(Context& c, const TreeData& t, int pde, const double* d_lam, int nrhs,
                            double* d_y);
void forward_apply_transpose(Context& c, const TreeData& t, int pde, const double* d_tgt, int ntgt,
                             const double* d_v, int nrhs, double* d_y);
void split_points(Context& c, const double* d_xy, int n, DBuf<double>* x, DBuf<double>* y);
// shape sensitivities of the Laplace double layer (adjoint.hip)
void kernel_sensitivity(Context& c, const TreeData& t, const double* d_lam, const double* d_mu,
                        double* d_sens);
void forward_sensitivity(Context& c, const TreeData& t, const double* d_tgt, int ntgt,
                         const double* d_v, const double* d_mu, double* d_sens);
void kernel_block(Context& c, const GeomView& g, const double* dpx, const double* dpy, int pde,
                  int sd, const int* h_tgt, int ntgt, const int* h_src, int nsrc, int forward,
                  double* h_out);
// in-domain mask and domain evaluation (domain.hip)
void domain_mask(Context& c, const kigpu_region& r, const double* d_txy, int ntgt,
                 unsigned char* d_inside);
void domain_solution(Context& c, const TreeData& t, int pde, const double* d_txy, int ntgt,
                     const double* d_mu, int nrhs, const double* d_alpha, const double* h_hc,
                     int nholes, const kigpu_region* region, double* d_u);
void forward_apply(Context& c, const TreeData& t, int pde, const double* d_tgt, int ntgt,
                   const double* d_mu, int nrhs, double* d_u);

}  // namespace kigpu
