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

// skel_factorization.cpp:32-52: ID of the node's ID matrix; returns the
// number of redundant DOFs (0: not compressed)
int SkelFactorization::id_compress(const Kernel& kernel, const QuadTree* tree,
                                   QuadTreeNode* node) {
  assert(node != nullptr && "InterpolativeDecomposition fails on null node.");
  assert(node->dof_lists.active_box.size() > 0 &&
         "Num of DOFs must be positive in InterpolativeDecomposition.");
  ki_Mat pxy = kernel.get_id_mat(tree, node);
  if (pxy.height() == 0) return 0;
  std::vector<int> p;
  const int numskel = pxy.id(&p, &node->T, id_tol);
  if (numskel == 0) return 0;
  node->dof_lists.set_rs_ranges(p, node->T.height(), node->T.width());
  node->dof_lists.set_skelnear_range();
  return node->T.width();
}

// skel_factorization.cpp:55-91: K_BN minus the descendants' updates, then
// X_rr, its LU, L, U and the Schur update (kigpu_schur_eliminate)
void SkelFactorization::decouple(const Kernel& kernel, QuadTreeNode* node) {
  const int nr = node->T.width(), ns = node->T.height();
  const std::vector<int>& BN = node->dof_lists.active_box;
  const int n = BN.size();
  assert(ns + nr == n && (int)node->dof_lists.permutation.size() == n);
  ki_Mat update(n, n);
  get_all_schur_updates(&update, BN, node);
  ki_Mat K_BN = kernel(BN, BN) - update;
  node->X_rr = ki_Mat(nr, nr);
  node->X_rr_lu = ki_Mat(nr, nr);
  node->X_rr_piv.assign(nr, 0);
  node->L = ki_Mat(ns, nr);
  node->U = ki_Mat(nr, ns);
  node->schur_update = ki_Mat(ns, ns);
  std::vector<int> piv(nr);
  const kigpu_status st = kigpu_schur_eliminate(
      detail::context(), K_BN.mat, n, ns, node->dof_lists.permutation.data(), node->T.mat,
      node->X_rr.mat, node->X_rr_lu.mat, piv.data(), node->L.mat, node->U.mat,
      node->schur_update.mat);
  // a singular X_rr leaves its factors as computed (the reference does not
  // check dgetrf's info, ki_mat.cpp:431-432)
  if (st != KIGPU_ERR_SINGULAR) detail::check(st, "kigpu_schur_eliminate");
  for (int i = 0; i < nr; ++i) node->X_rr_piv[i] = piv[i];
  node->X_rr_is_LU_factored = true;
  node->compressed = true;
}

// skel_factorization.cpp:246-294: the Schur updates of every compressed
// descendant, entered at the positions of its skelnear DOFs in BN
void SkelFactorization::get_all_schur_updates(ki_Mat* updates, const std::vector<int>& BN,
                                              const QuadTreeNode* node) const {
  assert(node != nullptr && "get_all_schur_updates fails on null node.");
  assert(BN.size() > 0 && "get_all_schur_updates needs positive num of DOFs");
  if (!node->is_leaf) get_descendents_updates(updates, BN, node);
}

void SkelFactorization::get_descendents_updates(ki_Mat* updates, const std::vector<int>& BN,
                                                const QuadTreeNode* node) const {
  assert(node != nullptr && !node->is_leaf &&
         "get_descendents_updates must be called on non-leaf.");
  for (QuadTreeNode* child : node->children) {
    if (child->compressed) get_update(updates, BN, child);
    if (!child->is_leaf) get_descendents_updates(updates, BN, child);
  }
}

void SkelFactorization::get_update(ki_Mat* update, const std::vector<int>& BN,
                                   const QuadTreeNode* node) const {
  // positions of the node's skelnear DOFs that BN holds, in skelnear order
  std::vector<int> at_bn, at_sn;
  const std::vector<int>& sn = node->dof_lists.skelnear;
  for (size_t a = 0; a < sn.size(); ++a)
    for (size_t b = 0; b < BN.size(); ++b)
      if (BN[b] == sn[a]) {
        at_sn.push_back(a);
        at_bn.push_back(b);
      }
  for (size_t i = 0; i < at_bn.size(); ++i)
    for (size_t j = 0; j < at_bn.size(); ++j)
      update->addset(at_bn[i], at_bn[j], node->schur_update.get(at_sn[i], at_sn[j]));
}

void SkelFactorization::skeletonize(const Kernel& K, QuadTree* tree) {
  const detail::CtxRef lock = detail::context();  // the whole factor and its exports
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
  tree->fac_pde_ = o.pde;
  tree->fac_tol_ = o.id_tol;
  tree->fac_precision_ = o.precision;
  tree->fac_min_dofs_ = o.min_dofs;
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
  const detail::CtxRef lock = detail::context();
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
