// Drop-in for kern_interp/skel_factorization/skel_factorization.h
// (reference skel_factorization.h:19-62).  skeletonize/solve run on the GPU
// through kigpu_factor / kigpu_solve (the factorization stays on the device).
// The per-box building blocks work on the host node structs like the
// reference's: id_compress (ID on the GPU, kigpu_interp_decomp), decouple
// (kigpu_schur_eliminate) and the get_*_updates bookkeeping.
#ifndef KERN_INTERP_SKEL_FACTORIZATION_SKEL_FACTORIZATION_H_
#define KERN_INTERP_SKEL_FACTORIZATION_SKEL_FACTORIZATION_H_

#include <vector>

#include "kern_interp/kernel/kernel.h"
#include "kern_interp/ki_mat.h"
#include "kern_interp/quadtree/quadtree.h"

#define MIN_DOFS_TO_COMPRESS 16

namespace kern_interp {

class SkelFactorization {
 public:
  double id_tol;
  int fact_threads;  // accepted for compatibility; the GPU batches every box of a level
  // arithmetic of the factorization: false = FP64 (the reference's), true =
  // FP32 elimination/solve with FP64 ID and FP64 iterative refinement
  bool fp32_refine = false;
  int refine_iters = 10;
  double refine_tol = 1e-10;

  SkelFactorization() {}
  SkelFactorization(double id_tol, int fact_threads);
  ~SkelFactorization() {}

  void decouple(const Kernel& K, QuadTreeNode* node);
  int id_compress(const Kernel& K, const QuadTree* tree, QuadTreeNode* node);

  void get_all_schur_updates(ki_Mat* updates, const std::vector<int>& BN,
                             const QuadTreeNode* node) const;
  void get_descendents_updates(ki_Mat* updates, const std::vector<int>& BN,
                               const QuadTreeNode* node) const;
  void get_update(ki_Mat* updates, const std::vector<int>& BN, const QuadTreeNode* node) const;

  void skeletonize(const Kernel& K, QuadTree* tree);

  void apply_sweep_matrix(const ki_Mat& mat, ki_Mat* vec, const std::vector<int>& a,
                          const std::vector<int>& b, bool transpose) const;
  void apply_diag_matrix(const ki_Mat& mat, ki_Mat* vec, const std::vector<int>& range) const;
  void apply_diag_inv_matrix(const ki_Mat& mat, const std::vector<lapack_int>& piv, ki_Mat* vec,
                             const std::vector<int>& range) const;

  void solve(const QuadTree& quadtree, ki_Mat* x, const ki_Mat& b) const;
  void multiply_connected_solve(const QuadTree& quadtree, ki_Mat* x, ki_Mat* alpha,
                                const ki_Mat& b) const;
};

}  // namespace kern_interp

#endif  // KERN_INTERP_SKEL_FACTORIZATION_SKEL_FACTORIZATION_H_
