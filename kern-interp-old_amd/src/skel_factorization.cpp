// SkelFactorization of the drop-in API (interface of the reference's
// skel_factorization.h:19-62): skeletonize and the solves run on the GPU;
// per-node results are copied back into the QuadTree's node lists.
#include "kern_interp/skel_factorization/skel_factorization.h"

#include <cassert>

#include "runtime.h"

namespace kern_interp {

SkelFactorization::SkelFactorization(double tol, int threads) {
  assert(tol > 0 && "id_tol must be greater than zero");
  id_tol = tol;
  fact_threads = threads;
}

void SkelFactorization::skeletonize(const Kernel& K, QuadTree* tree) {
  kigpu_tree* g = tree->gpu();
  assert(g && "initialize_tree before skeletonize");
  const int h = tree->U.width();
  detail::check(kigpu_set_border(g, h ? tree->U.mat : nullptr, h ? tree->Psi.mat : nullptr, h),
                "kigpu_set_border");
  kigpu_factor_opts o;
  o.pde = (int)K.pde;
  o.id_tol = id_tol;
  o.precision = fp32_refine ? KIGPU_FP32_REFINE : KIGPU_FP64;
  o.min_dofs = MIN_DOFS_TO_COMPRESS;
  detail::check(kigpu_factor(detail::context(), g, &o), "kigpu_factor");
  const auto& nodes = tree->nodes_by_id();
  const int nn = nodes.size();
  std::vector<int> rank(nn), nact(nn);
  detail::check(kigpu_factor_ranks(g, rank.data(), nact.data()), "kigpu_factor_ranks");
  std::vector<int64_t> aoff(nn + 1), poff(nn + 1);
  detail::check(kigpu_factor_export_active(g, aoff.data(), nullptr, poff.data(), nullptr),
                "export_active");
  std::vector<int> act(aoff[nn]), perm(poff[nn]);
  detail::check(kigpu_factor_export_active(g, aoff.data(), act.data(), poff.data(), perm.data()),
                "export_active");
  for (int x = 0; x < nn; ++x) {
    QuadTreeNode* nd = nodes[x];
    InteractionLists& dl = nd->dof_lists;
    dl.active_box.assign(act.begin() + aoff[x], act.begin() + aoff[x + 1]);
    dl.skel.clear();
    dl.redundant.clear();
    dl.skelnear.clear();
    dl.permutation.clear();
    nd->compressed = rank[x] >= 0;
    nd->X_rr_is_LU_factored = nd->compressed;
    if (nd->compressed) {
      std::vector<int> p(perm.begin() + poff[x], perm.begin() + poff[x + 1]);
      dl.set_rs_ranges(p, rank[x], nact[x] - rank[x]);
      dl.set_skelnear_range();
    }
  }
}

// vec(b) += op(mat) vec(a) (skel_factorization.cpp:301-318)
void SkelFactorization::apply_sweep_matrix(const ki_Mat& mat, ki_Mat* vec,
                                           const std::vector<int>& a, const std::vector<int>& b,
                                           bool transpose) const {
  if (a.size() * b.size() == 0) return;
  ki_Mat prod = (transpose ? mat.transpose() : mat) * (*vec)(a, 0, vec->width());
  vec->set_submatrix(b, 0, vec->width(), prod + (*vec)(b, 0, vec->width()));
}

void SkelFactorization::apply_diag_matrix(const ki_Mat& mat, ki_Mat* vec,
                                          const std::vector<int>& range) const {
  if (range.empty()) return;
  vec->set_submatrix(range, 0, vec->width(), mat * (*vec)(range, 0, vec->width()));
}

void SkelFactorization::apply_diag_inv_matrix(const ki_Mat& mat, const std::vector<lapack_int>& piv,
                                              ki_Mat* vec, const std::vector<int>& range) const {
  if (range.empty()) return;
  ki_Mat prod;
  mat.left_multiply_inverse((*vec)(range, 0, vec->width()), piv, &prod);
  vec->set_submatrix(range, 0, vec->width(), prod);
}

static void gpu_solve(const SkelFactorization& f, const QuadTree& tree, ki_Mat* x, ki_Mat* alpha,
                      const ki_Mat& b) {
  kigpu_tree* g = tree.gpu();
  assert(g && x->height() == b.height());
  *x = ki_Mat(b.height(), b.width());
  const int h = tree.U.width();
  if (alpha) *alpha = ki_Mat(h, b.width());
  kigpu_refine_opts ro{f.fp32_refine ? f.refine_iters : 0, f.refine_tol};
  kigpu_solve_info info;
  detail::check(kigpu_solve(detail::context(), g, b.mat, b.height(), b.width(), x->mat,
                            x->height(), alpha && h ? alpha->mat : nullptr, &ro, &info),
                "kigpu_solve");
}

void SkelFactorization::solve(const QuadTree& quadtree, ki_Mat* x, const ki_Mat& b) const {
  gpu_solve(*this, quadtree, x, nullptr, b);
}

void SkelFactorization::multiply_connected_solve(const QuadTree& quadtree, ki_Mat* mu,
                                                 ki_Mat* alpha, const ki_Mat& b) const {
  gpu_solve(*this, quadtree, mu, alpha, b);
}

}  // namespace kern_interp
