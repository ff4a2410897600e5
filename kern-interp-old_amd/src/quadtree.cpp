// QuadTree of the drop-in API: built on the GPU (kigpu_tree_build, the
// reference's QuadTree::initialize_tree, quadtree.cpp:72-114) and mirrored
// into the reference's node structs (quadtree.h:14-102).
#include "kern_interp/quadtree/quadtree.h"

#include <cassert>
#include <cmath>

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

// quadtree.cpp:27-68.  Levels are walked in order, so a parent's list is
// complete before its children read it: siblings first, then the same-level
// children of the parent's neighbors whose bottom-left corner lies within
// side*sqrt(2) (+1e-5); a leaf then collects the smaller boxes touching it
// from its same-level neighbors' subtrees.  The neighbor vector grows while
// it is scanned; the entries added there are deeper and are skipped.
void QuadTree::compute_neighbor_lists() {
  const double reach = std::sqrt(2.0);
  for (QuadTreeLevel* lev : levels) {
    for (QuadTreeNode* a : lev->nodes) {
      if (a->parent) {
        for (QuadTreeNode* sib : a->parent->children)
          if (sib != a) a->neighbors.push_back(sib);
        for (QuadTreeNode* pn : a->parent->neighbors)
          for (QuadTreeNode* c : pn->children) {
            if (!c || c->level != a->level) continue;
            const double dx = a->corners[0] - c->corners[0], dy = a->corners[1] - c->corners[1];
            if (std::sqrt(dx * dx + dy * dy) < a->side_length * reach + 1e-5)
              a->neighbors.push_back(c);
          }
      }
      if (!a->is_leaf) continue;
      for (size_t k = 0; k < a->neighbors.size(); ++k) {
        QuadTreeNode* nb = a->neighbors[k];
        if (nb->level != a->level) continue;
        for (QuadTreeNode* c : nb->children)
          if (c) get_descendent_neighbors(a, c);
      }
    }
  }
}

// quadtree.cpp:118-158: `small` (a descendant of a same-level neighbor of the
// leaf `big`) becomes a mutual neighbor when one of its corners lies on big's
// top or bottom edge strictly between big's left and right, or on big's left
// or right edge with y below both big's top and big's bottom (the reference's
// test, kept as written).  The subtree of `small` is searched either way.
void QuadTree::get_descendent_neighbors(QuadTreeNode* big, QuadTreeNode* small) {
  assert(big->level < small->level);
  const double top = big->corners[3], bottom = big->corners[1];
  const double left = big->corners[0], right = big->corners[4];
  for (int c = 0; c < 4; ++c) {
    const double x = small->corners[2 * c], y = small->corners[2 * c + 1];
    const bool on_tb = x > left && x < right &&
                       (std::fabs(y - top) < 1e-14 || std::fabs(y - bottom) < 1e-14);
    const bool on_lr = y < top && y < bottom &&
                       (std::fabs(x - left) < 1e-14 || std::fabs(x - right) < 1e-14);
    if (on_tb || on_lr) {
      big->neighbors.push_back(small);
      small->neighbors.push_back(big);
      break;
    }
  }
  for (QuadTreeNode* c : small->children)
    if (c) get_descendent_neighbors(big, c);
}

namespace {
// child slot of (x, y) in a box with midpoints (mx, my): {BL, TL, TR, BR}
int quadrant(double x, double y, double mx, double my) {
  if (x < mx) return y < my ? 0 : 1;
  return y < my ? 3 : 2;
}
double mid_x(const QuadTreeNode* n) { return (n->corners[6] - n->corners[0]) / 2.0 + n->corners[0]; }
double mid_y(const QuadTreeNode* n) { return (n->corners[3] - n->corners[1]) / 2.0 + n->corners[1]; }
}  // namespace

// quadtree.cpp:161-192: the point's DOFs join every box on its path; the
// leaf that overflows leaf_dofs without a child for the point subdivides.
void QuadTree::recursive_add(QuadTreeNode* node, double x, double y, int point_ind) {
  assert(node != nullptr && "recursive_add fails on null node.");
  for (int i = 0; i < solution_dimension; ++i)
    node->dof_lists.original_box.push_back(solution_dimension * point_ind + i);
  QuadTreeNode* child = node->children[quadrant(x, y, mid_x(node), mid_y(node))];
  if (child)
    recursive_add(child, x, y, point_ind);
  else if (node->is_leaf && (int)node->dof_lists.original_box.size() > leaf_dofs)
    node_subdivide(node);
}

// quadtree.cpp:198-285: four children with the parent's corners cut at the
// midpoints, appended to the next level, the parent's DOFs dealt out by
// point, and any child over capacity subdivided in slot order.
void QuadTree::node_subdivide(QuadTreeNode* node) {
  assert(node != nullptr && "node_subdivide fails on null node.");
  node->is_leaf = false;
  const double mx = mid_x(node), my = mid_y(node);
  // corner entries each child moves to a midpoint: {index, 0 = x / 1 = y}
  static const int cut[4][4] = {{3, 4, 5, 6}, {1, 4, 6, 7}, {0, 1, 2, 7}, {0, 2, 3, 5}};
  for (int s = 0; s < 4; ++s) {
    QuadTreeNode* c = new QuadTreeNode();
    for (int j = 0; j < 8; ++j) c->corners[j] = node->corners[j];
    for (int j : cut[s]) c->corners[j] = (j % 2 == 0) ? mx : my;
    c->level = node->level + 1;
    c->side_length = node->side_length / 2.0;
    c->parent = node;
    node->children[s] = c;
  }
  if ((int)levels.size() < node->level + 2) levels.push_back(new QuadTreeLevel());
  for (QuadTreeNode* c : node->children) levels[node->level + 1]->nodes.push_back(c);
  const std::vector<int>& ob = node->dof_lists.original_box;
  for (size_t k = 0; k < ob.size(); k += solution_dimension) {
    const int pt = ob[k] / solution_dimension;
    const double x = boundary_points[pt * domain_dimension];
    const double y = boundary_points[pt * domain_dimension + 1];
    std::vector<int>& dst = node->children[quadrant(x, y, mx, my)]->dof_lists.original_box;
    for (int i = 0; i < solution_dimension; ++i) dst.push_back(ob[k] + i);
  }
  for (QuadTreeNode* c : node->children)
    if ((int)c->dof_lists.original_box.size() > leaf_dofs) node_subdivide(c);
}

void QuadTree::remove_inactive_dofs_at_all_boxes() {
  for (int l = (int)levels.size() - 1; l >= 0; --l) remove_inactive_dofs_at_level(l);
}

}  // namespace kern_interp
