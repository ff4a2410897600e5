/*
 * kigpu.h -- C ABI of the MI355X recursive-skeletonization factor/solve.
 *
 * Plain C types only: no exceptions, no std:: types and no torch types cross
 * this boundary.  Every entry point returns a kigpu_status (0 = success);
 * kigpu_status_string() names the code and kigpu_last_error() gives the
 * detail of the last failure on the calling thread.
 *
 * Host buffers are caller-owned, FP64, column-major; DOF indices are 32-bit.
 * Device memory is owned by the handles.  Entry points ending in _device
 * take device pointers (memory already resident in HBM) and run on the
 * context's stream.
 *
 * The reference interface each entry point replaces is cited beside it
 * (paths relative to the reference repository root).
 */
#ifndef KIGPU_H_
#define KIGPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int kigpu_status;
enum {
  KIGPU_OK = 0,
  KIGPU_ERR_ARG = 1,          /* invalid argument (null pointer, bad size) */
  KIGPU_ERR_HIP = 2,          /* a HIP runtime call failed */
  KIGPU_ERR_NOMEM = 3,        /* device allocation failed */
  KIGPU_ERR_STATE = 4,        /* wrong call order (e.g. solve before factor) */
  KIGPU_ERR_RANK = 5,         /* ID matrix with fewer rows than columns */
  KIGPU_ERR_NOCOMPRESS = 6,   /* bordered system with nothing compressed */
  KIGPU_ERR_NOTSUPP = 7,      /* combination not supported */
  KIGPU_ERR_NOCONV = 8,       /* iterative refinement did not converge */
  KIGPU_ERR_SINGULAR = 9      /* a zero or non-finite pivot in an X_rr, root or
                                 bordered inverse: the factorization is unusable
                                 (the reference leaves dgetrf's info unchecked,
                                 ki_mat.cpp:431-432, and goes on with inf/NaN) */
};

/* Kernel::Pde (kern_interp/kernel/kernel.h:16-20) */
enum { KIGPU_LAPLACE = 0, KIGPU_LAPLACE_NEUMANN = 1, KIGPU_STOKES = 2 };

/* arithmetic of the factorization */
enum {
  KIGPU_FP64 = 0,         /* everything in FP64 (the reference's arithmetic) */
  KIGPU_FP32_REFINE = 1   /* FP32 factor/solve; FP64 norms, root LU, residuals */
};

typedef struct kigpu_ctx kigpu_ctx;
typedef struct kigpu_tree kigpu_tree;

/* circular hole: kern_interp/boundaries/boundary.h:12-16 */
typedef struct {
  double cx, cy, radius;
  int num_nodes;
} kigpu_hole;

typedef struct {
  int pde;               /* KIGPU_LAPLACE / KIGPU_LAPLACE_NEUMANN / KIGPU_STOKES */
  double id_tol;         /* SkelFactorization::id_tol (skel_factorization.h:21) */
  int precision;         /* KIGPU_FP64 or KIGPU_FP32_REFINE */
  int min_dofs;          /* MIN_DOFS_TO_COMPRESS, 16 (skel_factorization.h:11) */
} kigpu_factor_opts;

typedef struct {
  int max_iters;         /* refinement steps (FP32_REFINE only); 0 = plain solve */
  double rel_residual;   /* stop when ||f - K mu|| / ||f|| <= this */
} kigpu_refine_opts;

typedef struct {
  int iterations;        /* refinement steps taken */
  double rel_residual;   /* last FP64 residual measured (-1 if none) */
} kigpu_solve_info;

typedef struct {
  int num_levels, num_nodes, num_points, solution_dim, leaf_dofs;
  double min, max;       /* QuadTree::min/max (quadtree.h:78) */
} kigpu_tree_info;

typedef struct {
  /* device time per phase of the last tree build / factor / solve (seconds) */
  double t_tree, t_lists, t_id, t_decouple, t_root, t_solve, t_refine_apply;
  int syncs;             /* host synchronizations in the last tree build + factor */
  /* the ID kernel (ID matrix generation + pivoted QR + T), one launch per
   * level over every box compressed there: launches, summed launch
   * durations (s) and algorithmic flops */
  int id_launches;
  double id_kernel_s, id_flops;
  /* the same for the elimination kernel */
  int dec_launches;
  double dec_kernel_s, dec_flops;
  /* FP64 refinement applies (Laplace, one right-hand side): launches, their
   * summed kernel time and the number of (target, source) pairs */
  int apply_launches;
  double apply_kernel_s, apply_pairs;
  /* the treecode apply the FP64 refinement uses by default (Laplace double
   * layer): applies, their summed kernel time (series +
   * evaluation), the evaluation kernel's share (s), the (target, series
   * box) and (target, source) pairs of one apply, and the series terms
   * summed (each pair sums the terms its box needs for its leaf, <= 48) */
  int tc_launches;
  double tc_kernel_s, tc_eval_s, tc_series_pairs, tc_direct_pairs, tc_series_terms;
  /* sweeps of the last solve replayed from a captured graph, and captured */
  int graph_replays, graph_captures;
} kigpu_timings;

/* ---- context ---------------------------------------------------------- */
kigpu_status kigpu_context_create(int device, kigpu_ctx** ctx);
kigpu_status kigpu_context_destroy(kigpu_ctx* ctx);
const char* kigpu_status_string(kigpu_status s);
const char* kigpu_last_error(void);
/* device-side phase timings of the last factor/solve on this context */
/* ---- launch tuning ------------------------------------------------------
 * No counterpart in the reference.  Launch choices of this library that do
 * not change what is computed, only how it is cut into launches:
 *   dec_wave[b]      elimination of boxes with 16 b + 1 .. 16 b + 16 active
 *                    DOFs: 1 one wavefront per box (matrix-core products,
 *                    register Gauss-Jordan; the default), 0 one 1024-thread
 *                    workgroup per box (other values are rejected);
 *   mid_on_side      the 33-64 wavefront class on the context's side stream,
 *                    beside the other classes;
 *   pair_sweeps      two tree levels per solve-sweep launch where LDS allows;
 *   apply_split_cap  most source splits (blocks per target tile) of the
 *                    direct K apply; fwd_split_cap the same for the forward
 *                    and transposed applies.
 * Results may differ in the last bits between settings (other summation
 * orders), never more. */
typedef struct {
  int version; /* KIGPU_TUNING_VERSION */
  int dec_wave[4];
  int mid_on_side, pair_sweeps, apply_split_cap, fwd_split_cap;
} kigpu_tuning;
enum { KIGPU_TUNING_VERSION = 1 };
kigpu_status kigpu_get_tuning(kigpu_ctx* ctx, kigpu_tuning* out);
kigpu_status kigpu_set_tuning(kigpu_ctx* ctx, const kigpu_tuning* in);
/* Measures the choices on this tree: factors it (opts) reps times per
 * elimination choice and keeps the fastest per box-size class, then times
 * solves and direct applies for the sweep and split choices; the context
 * keeps the winners, *out gets them (the tree is left factored with them).
 * Time per choice: the fastest of reps runs, from HIP events. */
kigpu_status kigpu_autotune(kigpu_ctx* ctx, kigpu_tree* tree, const kigpu_factor_opts* opts,
                            int reps, kigpu_tuning* out);

/* Solve sweeps from HIP graphs (no counterpart in the reference): with on = 1,
 * each tree captures the launches of its up/down sweeps and root (or
 * bordered) step on the second solve with the same vector, shape and border,
 * and replays the graph from then on (a new factorization drops them).  For
 * callers that solve many times with one factorization; the results are those
 * of the direct launches, bit for bit.  Default 0. */
kigpu_status kigpu_set_graphs(kigpu_ctx* ctx, int on);
kigpu_status kigpu_last_timings(kigpu_ctx* ctx, kigpu_timings* out);

/* ---- geometry + tree -------------------------------------------------- */
/* QuadTree::initialize_tree (kern_interp/quadtree/quadtree.cpp:72-114) plus
 * the Kernel's copies of the boundary arrays (kern_interp/kernel/kernel.h:22-32).
 * points/normals: 2*npts interleaved; weights/curvatures: npts.
 * leaf_dofs replaces MAX_LEAF_DOFS (quadtree.h:10).  Built on the GPU. */
kigpu_status kigpu_tree_build(kigpu_ctx* ctx, const double* points,
                              const double* normals, const double* weights,
                              const double* curvatures, int npts, int sd,
                              int leaf_dofs, kigpu_tree** tree);
/* the same from device arrays (SoA, already resident in HBM): x, y, normal
 * x, normal y, weights, curvatures, each npts doubles */
kigpu_status kigpu_tree_build_device(kigpu_ctx* ctx, const double* d_px, const double* d_py,
                                     const double* d_nx, const double* d_ny, const double* d_w,
                                     const double* d_kappa, int npts, int sd, int leaf_dofs,
                                     kigpu_tree** tree);
/* A tree described on the host instead of built from the points: QuadTree's
 * node structs after QuadTree::perturb or QuadTree::copy_into
 * (quadtree.cpp:359-641).  Nodes numbered level by level: level_off[nlevels
 * + 1] (level 0 = the root alone); parent[nn] (-1 for the root); child[4 nn]
 * in the slots {BL, TL, TR, BR} (all four or none, -1); corners[8 nn]
 * ({BL, TL, TR, BR} x, y) and side[nn]; each leaf's original_box DOF list
 * box_dofs[box_off[x] .. box_off[x + 1]) (interior entries ignored: their
 * points are the union of their leaves'); each node's neighbor list
 * nbrs[nbr_off[x] .. nbr_off[x + 1]) in QuadTree order (coarser leaves, then
 * same-level boxes, then finer boxes, which are not read); the root box's
 * min / max (QuadTree::min/max).  Every point in exactly one leaf.  The
 * factorization runs on the result as on a built tree. */
kigpu_status kigpu_tree_import(kigpu_ctx* ctx, const double* points, const double* normals,
                               const double* weights, const double* curvatures, int npts, int sd,
                               int leaf_dofs, int nlevels, const int64_t* level_off,
                               const int* parent, const int* child, const double* corners,
                               const double* side, const int64_t* box_off, const int* box_dofs,
                               const int64_t* nbr_off, const int* nbrs, double tree_min,
                               double tree_max, kigpu_tree** tree);
kigpu_status kigpu_tree_destroy(kigpu_tree* tree);
/* A batch of independent boundaries (instance i: npts[i] points, stored one
 * after another in the arrays, interleaved like kigpu_tree_build) built as
 * one forest.  No counterpart in the reference, which would loop over
 * QuadTree::initialize_tree (quadtree.cpp:72-114); each instance's quadtree
 * is the one kigpu_tree_build gives it.  kigpu_factor then factors every
 * instance with one launch per tree level over the boxes of all instances,
 * kigpu_solve / kigpu_solve_transpose take the instances' right-hand sides
 * stacked (sum of sd*npts rows), kigpu_kernel_apply is block diagonal and
 * FP32 refinement stops when every instance meets rel_residual.  Holes,
 * domain targets and sensitivities need a single instance
 * (KIGPU_ERR_NOTSUPP). */
kigpu_status kigpu_tree_build_batch(kigpu_ctx* ctx, int ninst, const int* npts,
                                    const double* points, const double* normals,
                                    const double* weights, const double* curvatures, int sd,
                                    int leaf_dofs, kigpu_tree** tree);
/* the same from device arrays (SoA, resident in HBM, instances one after
 * another); npts is a host array */
kigpu_status kigpu_tree_build_batch_device(kigpu_ctx* ctx, int ninst, const int* npts,
                                           const double* d_px, const double* d_py,
                                           const double* d_nx, const double* d_ny,
                                           const double* d_w, const double* d_kappa, int sd,
                                           int leaf_dofs, kigpu_tree** tree);
/* ninst; point_off[ninst + 1]; bounds[2 * ninst] = root-box min/max
 * (QuadTree::min/max per instance); root_skel[ninst] = |allskel| per
 * instance after kigpu_factor (-1 before).  Output pointers may be NULL. */
kigpu_status kigpu_tree_instances(const kigpu_tree* tree, int* ninst, int64_t* point_off,
                                  double* bounds, int* root_skel);
kigpu_status kigpu_tree_get_info(const kigpu_tree* tree, kigpu_tree_info* info);

/* Node arrays in level order (node id = position in levels[0], levels[1], ...):
 * level_offsets[num_levels+1]; parent[num_nodes]; children[4*num_nodes] (-1 =
 * none); is_leaf[num_nodes]; corners[8*num_nodes] ({BL,TL,TR,BR}, quadtree.h:55);
 * side[num_nodes].  Any pointer may be NULL to skip that array. */
kigpu_status kigpu_tree_export_nodes(const kigpu_tree* tree, int64_t* level_offsets,
                                     int* parent, int* children, int* is_leaf,
                                     double* corners, double* side);
/* InteractionLists::original_box of every node (quadtree.h:15) as CSR:
 * offsets[num_nodes+1], then dofs[offsets[num_nodes]].  Call with dofs=NULL
 * to get the offsets (and so the size) first. */
kigpu_status kigpu_tree_export_boxes(const kigpu_tree* tree, int64_t* offsets, int* dofs);
/* QuadTreeNode::neighbors (quadtree.h:50) as CSR of node ids, in the order
 * compute_neighbor_lists builds them (quadtree.cpp:27-69, 119-161). */
kigpu_status kigpu_tree_export_neighbors(const kigpu_tree* tree, int64_t* offsets, int* nbrs);

/* ---- factorization ---------------------------------------------------- */
/* Border matrices for holes (QuadTree::U / QuadTree::Psi, quadtree.h:81),
 * column-major: U is ndof x nborder, Psi is nborder x ndof.  Set before
 * kigpu_factor; nborder = 0 clears them.  (linear_solve.cpp:154-163) */
kigpu_status kigpu_set_border(kigpu_tree* tree, const double* U, const double* Psi,
                              int nborder);
/* SkelFactorization::skeletonize (kern_interp/skel_factorization/
 * skel_factorization.cpp:94-243): one batched launch per tree level. */
kigpu_status kigpu_factor(kigpu_ctx* ctx, kigpu_tree* tree, const kigpu_factor_opts* opts);
/* Per-node results: rank[num_nodes] = #skeleton DOFs (-1: not compressed),
 * active[num_nodes] = |active_box| when the node was processed. */
kigpu_status kigpu_factor_ranks(const kigpu_tree* tree, int* rank, int* active);
/* skel / redundant DOF lists of every node as CSR (InteractionLists::skel and
 * ::redundant, quadtree.h:17-18; filled by set_rs_ranges, quadtree.h:23-35). */
kigpu_status kigpu_factor_export_lists(const kigpu_tree* tree, int64_t* skel_off, int* skel,
                                       int64_t* red_off, int* red);
/* InteractionLists::active_box of every node as used by the factorization
 * (CSR like above), and ::permutation (the ID pivots, positions in the
 * active list; empty for nodes that were not compressed).  perm_off may be
 * NULL when perm is NULL. */
kigpu_status kigpu_factor_export_active(const kigpu_tree* tree, int64_t* act_off, int* active,
                                        int64_t* perm_off, int* perm);
/* the root skeleton QuadTree::root->dof_lists.active_box after factoring
 * (for a batch: every instance's, one after another) */
kigpu_status kigpu_factor_allskel(const kigpu_tree* tree, int* n, int* allskel);

/* ---- solve ------------------------------------------------------------ */
/* SkelFactorization::solve (skel_factorization.cpp:341-403) for nrhs columns,
 * or multiply_connected_solve (:406-551) when a border is set (then alpha,
 * nborder x nrhs, receives the hole unknowns).  With precision FP32_REFINE and
 * refine->max_iters > 0, FP64 iterative refinement against the matrix-free
 * FP64 kernel apply follows.  Host buffers. */
kigpu_status kigpu_solve(kigpu_ctx* ctx, kigpu_tree* tree, const double* b, int ldb,
                         int nrhs, double* x, int ldx, double* alpha,
                         const kigpu_refine_opts* refine, kigpu_solve_info* info);
/* the same with device pointers (b, x: ndof x nrhs, leading dimension ndof) */
kigpu_status kigpu_solve_device(kigpu_ctx* ctx, kigpu_tree* tree, const double* d_b,
                                int nrhs, double* d_x, double* d_alpha,
                                const kigpu_refine_opts* refine, kigpu_solve_info* info);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same x = F^{-1} b as kigpu_solve_device with
 * refine->max_iters = 0, through the same kernels, allocating nothing.  An
 * FP64 factor solves in place in d_x; an FP32_REFINE factor stages x in
 * FP32 in a per-GPU static device buffer of 4 194 304 floats (16 MiB, set up
 * by kigpu_context_create), so ndof * nrhs above that returns
 * KIGPU_ERR_ARG.  Trees with a border return KIGPU_ERR_NOTSUPP.  Every call
 * on one GPU shares that buffer: calls that overlap on one GPU race.
 * Waits for the result like kigpu_solve_device. */
kigpu_status kigpu_solve_device_noalloc(kigpu_ctx* ctx, kigpu_tree* tree, const double* d_b,
                                        int nrhs, double* d_x);

/* ---- dense kernel blocks ------------------------------------------------ */
typedef struct kigpu_geom kigpu_geom;
/* Device copies of a Kernel's arrays (kern_interp/kernel/kernel.h:22-37):
 * boundary points/normals (2*npts interleaved), weights/curvatures (npts),
 * domain points (2*ndomain interleaved, may be 0). */
kigpu_status kigpu_geometry_create(kigpu_ctx* ctx, const double* points, const double* normals,
                                   const double* weights, const double* curvatures, int npts,
                                   const double* domain_points, int ndomain, kigpu_geom** geom);
kigpu_status kigpu_geometry_destroy(kigpu_geom* geom);
/* out (ntgt x nsrc, column-major) = Kernel::operator()(tgt, src, forward)
 * (kernel.cpp:99-150 Laplace/Neumann, :47-96 Stokes); sd = 2 for Stokes.
 * forward = 1 takes targets from the domain points (Kernel::forward, :337-357). */
kigpu_status kigpu_kernel_block(kigpu_ctx* ctx, const kigpu_geom* geom, int pde, int sd,
                                const int* tgt, int ntgt, const int* src, int nsrc, int forward,
                                double* out);

/* Kernel::laplace_proxy_get / stokes_proxy_get (kernel.cpp:268-334): the
 * rows [proxy <- box; box <- proxy] of the box DOFs against np explicit proxy
 * points (pxy_p, pxy_n: 2*np interleaved) of quadrature weight pxy_w.
 * out: (2*np for Laplace, 4*np for Stokes) x nbox, column-major. */
kigpu_status kigpu_proxy_block(kigpu_ctx* ctx, const kigpu_geom* geom, int pde,
                               const double* pxy_p, const double* pxy_n, int np, double pxy_w,
                               const int* box, int nbox, double* out);
/* ki_Mat::id (ki_mat.cpp:505-541) of a host matrix A (m x n, column-major) by
 * the pivoted QR of the factorization: k = the first i >= 1 with
 * |R_ii| < tol |R_00| (0: not compressible), perm[n] the column pivots
 * (0-based; skeleton first), T (k x (n - k)) = R11^{-1} R12.  perm and T are
 * written only when k > 0.  m < n returns KIGPU_ERR_RANK. */
kigpu_status kigpu_interp_decomp(kigpu_ctx* ctx, const double* A, int m, int n, double tol,
                                 int* k, int* perm, double* T);
/* The arithmetic of SkelFactorization::decouple (skel_factorization.cpp:55-91)
 * for one box: K_BN (n x n, active order, the descendants' updates already
 * subtracted), the ID's perm (skeleton positions first) and T (k x (n - k)).
 * Out: X_rr, its dgetrf factors X_rr_lu and X_rr_piv (1-based), L (k x r) =
 * K_sn_r X_rr^{-1}, U (r x k) = X_rr^{-1} K_r_sn, schur_update (k x k) =
 * L K_r_sn, r = n - k.  Any output may be NULL.  A zero or non-finite pivot
 * fills the outputs and returns KIGPU_ERR_SINGULAR. */
kigpu_status kigpu_schur_eliminate(kigpu_ctx* ctx, const double* K_BN, int n, int k,
                                   const int* perm, const double* T, double* X_rr,
                                   double* X_rr_lu, int* X_rr_piv, double* L, double* U,
                                   double* schur_update);

/* ---- matrix-free kernel apply ----------------------------------------- */
/* y = K(all, all) mu in FP64 on the fly, Kernel::operator() over all DOFs
 * (kernel.cpp:99-150, 47-96).  mu, y: ndof x nrhs. */
kigpu_status kigpu_kernel_apply(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                const double* mu, int nrhs, double* y);
kigpu_status kigpu_kernel_apply_device(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                       const double* d_mu, int nrhs, double* d_y);
/* The same product for the Laplace double layer by a treecode: box series of
 * 48 terms used where the box radius is at most half the distance from its
 * centre to the target leaf, the rest summed directly; the truncation stays
 * below 1e-14 relative to the direct sum.  mu, y: num_points x nrhs,
 * column-major.  This is the FP64 refinement's residual apply (below); it has
 * no counterpart in the reference, whose path has no refinement. */
kigpu_status kigpu_kernel_apply_treecode(kigpu_ctx* ctx, const kigpu_tree* tree,
                                         const double* mu, int nrhs, double* y);
/* Residual apply of the FP64 refinement on this context: 0 (default) the
 * treecode where it applies (Laplace double layer, no coincident nodes, not
 * the adjoint), 1 always the direct sum of kigpu_kernel_apply.  No
 * counterpart in the reference. */
kigpu_status kigpu_set_refine_apply(kigpu_ctx* ctx, int mode);
/* How the factorizations on this context compute each box's ID (replaces
 * the dgeqp3 + dtrtrs of ki_Mat::id, kern_interp/ki_mat.cpp:505-541):
 * 0 (default) by precision -- FP32_REFINE with id_tol >= 1e-7 from the
 * box's Gram matrix A^T A (FP64, matrix cores) and a pivoted Cholesky of it,
 * otherwise by Householder QR with column pivoting; 1 always the pivoted QR;
 * 2 always the Gram matrix (whose rank rule cannot resolve tolerances much
 * below 1e-7: remaining norms squared below ~eps |A|^2 are noise).
 * Both give the QR's pivots, rank and T in exact arithmetic; the Gram route
 * can differ at pivots or ranks that FP64 cannot separate (near ties). */
kigpu_status kigpu_set_id_method(kigpu_ctx* ctx, int method);
/* u = Kernel::forward() mu (kernel.cpp:337-357): boundary -> domain targets,
 * targets: 2*ntgt interleaved.  u: (sd*ntgt) x nrhs. */
kigpu_status kigpu_forward_apply(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                 const double* targets, int ntgt, const double* mu,
                                 int nrhs, double* u);

/* ---- in-domain test and domain evaluation ------------------------------ */
/* Boundary::is_in_domain as data (one rule per geometry family):
 *  KIGPU_REGION_DISC    Circle, Annulus, Donut (circle.cpp:34-42,
 *                       annulus.cpp:66-78, donut.cpp:67-79): outside when
 *                       |t - c| + eps > radius, or |t - h_k| - eps < r_k for
 *                       a hole k (holes: x, y, r per hole);
 *  KIGPU_REGION_WINDING CubicBoundary (cubic_boundary.cpp:182-230): crossing
 *                       counts of the node polygon, nodes = the boundary
 *                       points (2 per node), num_outer outer nodes first, then
 *                       hole_nodes[k] nodes per hole; the outer count must be
 *                       odd, a hole's count of 1 excludes, a node closer than
 *                       eps excludes;
 *  KIGPU_REGION_STAR    a Fourier star about the origin (no counterpart in
 *                       the reference): inside when
 *                       |t| + eps < 1 + sum a_k cos(f_k atan2(t) + phi_k). */
enum { KIGPU_REGION_DISC = 0, KIGPU_REGION_WINDING = 1, KIGPU_REGION_STAR = 2 };
typedef struct {
  int kind;
  double eps;                 /* 1e-2 in every reference geometry */
  double cx, cy, radius;      /* DISC */
  int nholes;
  const double* holes;        /* DISC: 3 * nholes */
  const double* nodes;        /* WINDING */
  int num_outer;              /* WINDING */
  const int* hole_nodes;      /* WINDING: nholes */
  int nterms;                 /* STAR */
  const double* amps;
  const double* freqs;
  const double* phases;
} kigpu_region;
/* inside[i] = 1 when target i (2*ntgt interleaved) is in the region. */
kigpu_status kigpu_domain_mask(kigpu_ctx* ctx, const kigpu_region* region, const double* targets,
                               int ntgt, unsigned char* inside);
/* The domain solution of boundary_integral_solve (linear_solve.cpp:143-180):
 * u = Kernel::forward() mu + U_forward alpha at the targets, zero at targets
 * outside region (NULL: no mask).  U_forward is initialize_U_mat at the
 * targets (linear_solve.cpp:18-64) for holes centred at hole_centers (2 per
 * hole); alpha ((3 for Stokes, else 1) * nholes) x nrhs, or NULL for no hole
 * terms (simply connected, or LAPLACE_NEUMANN).  mu: ndof x nrhs;
 * u: (sd*ntgt) x nrhs. */
kigpu_status kigpu_domain_solution(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                   const double* targets, int ntgt, const double* mu,
                                   const double* alpha, const double* hole_centers, int nholes,
                                   int nrhs, const kigpu_region* region, double* u);

/* ---- adjoint ------------------------------------------------------------ */
/* No counterpart in the reference, whose experiments differentiate by finite
 * differences of whole solves (experiments/experiment_three.cpp:22-56,
 * 160-171).  x = F^{-T} b for the factorization F of kigpu_factor: the
 * transposes of the sweeps of SkelFactorization::solve
 * (skel_factorization.cpp:341-403) in reverse order.  With FP32_REFINE and
 * refine->max_iters > 0 it is refined against the FP64 K^T apply.  Trees
 * with a border (holes) return KIGPU_ERR_NOTSUPP. */
kigpu_status kigpu_solve_transpose(kigpu_ctx* ctx, kigpu_tree* tree, const double* b, int ldb,
                                   int nrhs, double* x, int ldx,
                                   const kigpu_refine_opts* refine, kigpu_solve_info* info);
kigpu_status kigpu_solve_transpose_device(kigpu_ctx* ctx, kigpu_tree* tree, const double* d_b,
                                          int nrhs, double* d_x,
                                          const kigpu_refine_opts* refine,
                                          kigpu_solve_info* info);
/* y = K(all, all)^T lam in FP64 on the fly (entries of kernel.cpp:11-96). */
kigpu_status kigpu_kernel_apply_transpose(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                          const double* lam, int nrhs, double* y);
/* y = Kernel::forward()^T v (kernel.cpp:337-357): v is (sd*ntgt) x nrhs,
 * y is ndof x nrhs; targets: 2*ntgt interleaved. */
kigpu_status kigpu_forward_apply_transpose(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                          const double* targets, int ntgt, const double* v,
                                          int nrhs, double* y);
/* Shape sensitivities (Laplace double layer, KIGPU_LAPLACE only): for every
 * boundary point p, sens[6p .. 6p+5] = d/d(x_p, y_p, nx_p, ny_p, w_p, kappa_p)
 * of lam^T K mu (kernel_sensitivity) or of v^T F mu with F the forward
 * matrix at fixed targets (forward_sensitivity).  lam, mu: npts; v: ntgt.
 * Other PDEs return KIGPU_ERR_NOTSUPP. */
kigpu_status kigpu_kernel_sensitivity(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                      const double* lam, const double* mu, double* sens);
kigpu_status kigpu_forward_sensitivity(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                       const double* targets, int ntgt, const double* v,
                                       const double* mu, double* sens);

#ifdef __cplusplus
}
#endif

#endif /* KIGPU_H_ */
