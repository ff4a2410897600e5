// QuadTree of the drop-in API: built on the GPU (kigpu_tree_build, the
// reference's QuadTree::initialize_tree, quadtree.cpp:72-114) and mirrored
// into the reference's node structs (quadtree.h:14-102).
#include "kern_interp/quadtree/quadtree.h"

#include <cassert>

#include "runtime.h"

namespace kern_interp {

void InteractionLists::set_rs_ranges(const std::vector<int>& prm, int sk, int rd) {
  assert((int)prm.size() == sk + rd);
  for (int i = 0; i < sk + rd; ++i) {
    (i < sk ? skel : redundant).push_back(active_box[prm[i]]);
    permutation.push_back(prm[i]);
  }
}

void InteractionLists::set_skelnear_range() { skelnear = skel; }

QuadTree::~QuadTree() { clear(); }

void QuadTree::clear() {
  for (QuadTreeLevel* l : levels) delete l;
  levels.clear();
  by_id_.clear();
  root = nullptr;
  if (gpu_) kigpu_tree_destroy(gpu_);
  gpu_ = nullptr;
}

void QuadTree::initialize_tree(Boundary* boundary, int sd, int dd) {
  assert(!boundary->points.empty() && "number of boundary_points to init tree cannot be 0.");
  clear();
  boundary_points = boundary->points;
  solution_dimension = sd;
  domain_dimension = dd;
  const int npts = boundary->weights.size();
  detail::check(kigpu_tree_build(detail::context(), boundary->points.data(),
                                 boundary->normals.data(), boundary->weights.data(),
                                 boundary->curvatures.data(), npts, sd, leaf_dofs, &gpu_),
                "kigpu_tree_build");
  kigpu_tree_info info;
  detail::check(kigpu_tree_get_info(gpu_, &info), "kigpu_tree_get_info");
  min = info.min;
  max = info.max;
  const int nn = info.num_nodes, nl = info.num_levels;
  std::vector<int64_t> loff(nl + 1);
  std::vector<int> parent(nn), child(4 * nn), leaf(nn);
  std::vector<double> corners(8 * nn), side(nn);
  detail::check(kigpu_tree_export_nodes(gpu_, loff.data(), parent.data(), child.data(),
                                        leaf.data(), corners.data(), side.data()),
                "kigpu_tree_export_nodes");
  std::vector<int64_t> boff(nn + 1), noff(nn + 1);
  detail::check(kigpu_tree_export_boxes(gpu_, boff.data(), nullptr), "export_boxes");
  std::vector<int> dofs(boff[nn]);
  detail::check(kigpu_tree_export_boxes(gpu_, boff.data(), dofs.data()), "export_boxes");
  detail::check(kigpu_tree_export_neighbors(gpu_, noff.data(), nullptr), "export_neighbors");
  std::vector<int> nbrs(noff[nn]);
  detail::check(kigpu_tree_export_neighbors(gpu_, noff.data(), nbrs.data()), "export_neighbors");
  by_id_.resize(nn);
  for (int l = 0; l < nl; ++l) {
    QuadTreeLevel* lev = new QuadTreeLevel();
    for (int64_t x = loff[l]; x < loff[l + 1]; ++x) {
      QuadTreeNode* nd = new QuadTreeNode();
      nd->level = l;
      nd->is_leaf = leaf[x];
      nd->side_length = side[x];
      for (int j = 0; j < 8; ++j) nd->corners[j] = corners[8 * x + j];
      nd->dof_lists.original_box.assign(dofs.begin() + boff[x], dofs.begin() + boff[x + 1]);
      nd->dofs_below = (int)(boff[x + 1] - boff[x]);
      lev->nodes.push_back(nd);
      by_id_[x] = nd;
    }
    levels.push_back(lev);
  }
  for (int x = 0; x < nn; ++x) {
    QuadTreeNode* nd = by_id_[x];
    nd->parent = parent[x] >= 0 ? by_id_[parent[x]] : nullptr;
    for (int s = 0; s < 4; ++s) nd->children[s] = child[4 * x + s] >= 0 ? by_id_[child[4 * x + s]] : nullptr;
    for (int64_t e = noff[x]; e < noff[x + 1]; ++e) nd->neighbors.push_back(by_id_[nbrs[e]]);
  }
  root = by_id_[0];
}

void QuadTree::reset() { clear(); }

void QuadTree::reset(Boundary* boundary_) {
  initialize_tree(boundary_, solution_dimension, domain_dimension);
}

// quadtree.cpp:694-720: active list from the children (skel if compressed)
void QuadTree::remove_inactive_dofs_at_box(QuadTreeNode* node) {
  node->dof_lists.skel.clear();
  node->dof_lists.skelnear.clear();
  node->dof_lists.redundant.clear();
  node->dof_lists.active_box.clear();
  if (node->is_leaf) {
    node->dof_lists.active_box = node->dof_lists.original_box;
    return;
  }
  for (QuadTreeNode* c : node->children) {
    const std::vector<int>& src = c->compressed ? c->dof_lists.skel : c->dof_lists.active_box;
    node->dof_lists.active_box.insert(node->dof_lists.active_box.end(), src.begin(), src.end());
  }
}

// quadtree.cpp:664-691
void QuadTree::remove_inactive_dofs_at_level(int level) {
  for (QuadTreeNode* node : levels[level]->nodes)
    if (!node->compressed) remove_inactive_dofs_at_box(node);
  for (QuadTreeNode* a : levels[level]->nodes) {
    a->dof_lists.near.clear();
    for (QuadTreeNode* nb : a->neighbors) {
      if (nb->level > a->level) continue;
      const std::vector<int>& src = nb->is_leaf ? nb->dof_lists.original_box : nb->dof_lists.active_box;
      a->dof_lists.near.insert(a->dof_lists.near.end(), src.begin(), src.end());
    }
  }
}

void QuadTree::remove_inactive_dofs_at_all_boxes() {
  for (int l = (int)levels.size() - 1; l >= 0; --l) remove_inactive_dofs_at_level(l);
}

}  // namespace kern_interp
