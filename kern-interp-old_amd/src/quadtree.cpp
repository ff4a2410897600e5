// QuadTree of the drop-in API: built on the GPU (kigpu_tree_build, the
// reference's QuadTree::initialize_tree, quadtree.cpp:72-114) and mirrored
// into the reference's node structs (quadtree.h:14-102).
#include "kern_interp/quadtree/quadtree.h"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <unordered_map>
#include <unordered_set>
#include <utility>

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
  fac_pde_ = -1;
}

void QuadTree::initialize_tree(Boundary* boundary, int sd, int dd) {
  const detail::CtxRef lock = detail::context();  // the build and its exports
  assert(!boundary->points.empty() && "number of boundary_points to init tree cannot be 0.");
  clear();
  boundary_points = boundary->points;
  normals_ = boundary->normals;
  weights_ = boundary->weights;
  curvatures_ = boundary->curvatures;
  fac_pde_ = -1;
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

// ---- incremental edits ----------------------------------------------------

// quadtree.cpp:288-313: the box, its neighbors and its ancestors (with theirs)
// lose their compression and factor blocks
void QuadTree::mark_neighbors_and_parents(QuadTreeNode* node) {
  auto reset = [](QuadTreeNode* n) {
    n->compressed = false;
    n->X_rr_is_LU_factored = false;
    InteractionLists& d = n->dof_lists;
    d.active_box.clear();
    d.skel.clear();
    d.skelnear.clear();
    d.redundant.clear();
    d.permutation.clear();
    n->T = ki_Mat(0, 0);
    n->U = ki_Mat(0, 0);
    n->L = ki_Mat(0, 0);
  };
  for (; node; node = node->parent) {
    reset(node);
    for (QuadTreeNode* nb : node->neighbors) reset(nb);
  }
}

// quadtree.cpp:316-356: the box becomes a leaf holding its leaves' DOFs (in
// breadth-first order); its descendants leave their levels and are deleted
void QuadTree::consolidate_node(QuadTreeNode* node) {
  node->dof_lists.original_box.clear();
  std::vector<QuadTreeNode*> queue{node}, gone;
  for (size_t i = 0; i < queue.size(); ++i) {
    QuadTreeNode* cur = queue[i];
    if (cur->is_leaf) {
      const std::vector<int>& ob = cur->dof_lists.original_box;
      node->dof_lists.original_box.insert(node->dof_lists.original_box.end(), ob.begin(), ob.end());
    } else {
      for (QuadTreeNode* c : cur->children) queue.push_back(c);
    }
    if (cur != node) gone.push_back(cur);
  }
  for (QuadTreeNode* g : gone) {
    std::vector<QuadTreeNode*>& lv = levels[g->level]->nodes;
    lv.erase(std::find(lv.begin(), lv.end(), g));
    delete g;
  }
  for (int i = 0; i < 4; ++i) node->children[i] = nullptr;
  node->is_leaf = true;
}

namespace {
struct PairHash {
  size_t operator()(const std::pair<double, double>& p) const {
    const size_t a = std::hash<double>()(p.first), b = std::hash<double>()(p.second);
    return a ^ (b + 0x9e3779b97f4a7c15ull + (a << 6) + (a >> 2));
  }
};

// the leaf holding (x, y), by the midpoint rule of recursive_add
QuadTreeNode* leaf_of(QuadTreeNode* cur, double x, double y, QuadTreeNode** first_sparse = nullptr,
                      int cap = 0) {
  while (!cur->is_leaf) {
    if (first_sparse && !*first_sparse && cur->dofs_below < cap) *first_sparse = cur;
    cur = cur->children[quadrant(x, y, mid_x(cur), mid_y(cur))];
  }
  return cur;
}
}  // namespace

// quadtree.cpp:359-560.  Points are matched by exact coordinates (the last
// new index of a repeated point wins, as with the reference's hash map).
void QuadTree::perturb(const Boundary& nb) {
  const int sd = solution_dimension;
  const std::vector<double> old_pts = boundary_points;
  const std::vector<double>& new_pts = nb.points;
  std::unordered_map<std::pair<double, double>, int, PairHash> new_index;
  for (size_t i = 0; i < new_pts.size(); i += 2) new_index[{new_pts[i], new_pts[i + 1]}] = i / 2;
  std::vector<char> found(new_pts.size() / 2, 0);
  std::vector<int> old_to_new(old_pts.size() / 2, -1), additions, deletions;
  for (size_t i = 0; i < old_pts.size(); i += 2) {
    auto it = new_index.find({old_pts[i], old_pts[i + 1]});
    if (it != new_index.end()) {
      old_to_new[i / 2] = it->second;
      found[it->second] = 1;
    } else {
      deletions.push_back(i / 2);
    }
  }
  for (size_t i = 0; i < found.size(); ++i)
    if (!found[i]) additions.push_back(i);
  // renumber the DOF lists (leaves' original boxes; every box's active,
  // skeleton, redundant and skelnear lists), dropping deleted points
  auto remap = [&](std::vector<int>& v) {
    std::vector<int> out;
    for (int d : v) {
      const int q = old_to_new[d / sd];
      if (q >= 0) out.push_back(sd * q + d % sd);
    }
    v.swap(out);
  };
  for (QuadTreeLevel* lv : levels)
    for (QuadTreeNode* n : lv->nodes) {
      InteractionLists& d = n->dof_lists;
      if (n->is_leaf) remap(d.original_box);
      remap(d.active_box);
      remap(d.skel);
      remap(d.redundant);
      remap(d.skelnear);
    }
  // additions go to the leaves holding them
  std::vector<QuadTreeNode*> maybe_bursting;
  for (int a : additions) {
    QuadTreeNode* leaf = leaf_of(root, new_pts[2 * a], new_pts[2 * a + 1]);
    for (int j = 0; j < sd; ++j) leaf->dof_lists.original_box.push_back(sd * a + j);
    maybe_bursting.push_back(leaf);
    mark_neighbors_and_parents(leaf);
  }
  for (QuadTreeLevel* lv : levels)
    for (QuadTreeNode* n : lv->nodes)
      n->dofs_below = n->is_leaf ? (int)n->dof_lists.original_box.size() : 0;
  for (int l = (int)levels.size() - 1; l >= 1; --l)
    for (QuadTreeNode* n : levels[l]->nodes) n->parent->dofs_below += n->dofs_below;
  // deletions: mark their leaves; the first box on the way down holding
  // fewer than leaf_dofs DOFs is merged below
  std::unordered_set<QuadTreeNode*> sparse;
  for (int d : deletions) {
    QuadTreeNode* first = nullptr;
    QuadTreeNode* leaf = leaf_of(root, old_pts[2 * d], old_pts[2 * d + 1], &first, leaf_dofs);
    if (first) sparse.insert(first);
    mark_neighbors_and_parents(leaf);
  }
  boundary_points = new_pts;
  normals_ = nb.normals;
  weights_ = nb.weights;
  curvatures_ = nb.curvatures;
  for (QuadTreeNode* n : maybe_bursting)
    if (n->is_leaf && (int)n->dof_lists.original_box.size() > leaf_dofs) node_subdivide(n);
  for (QuadTreeNode* n : sparse) consolidate_node(n);
  for (QuadTreeLevel* lv : levels)
    for (QuadTreeNode* n : lv->nodes) n->neighbors.clear();
  compute_neighbor_lists();
  rebuild_device();
}

// quadtree.cpp:563-641
void QuadTree::copy_into(QuadTree* dst) const {
  const detail::CtxRef lock = detail::context();
  dst->clear();
  std::unordered_map<const QuadTreeNode*, QuadTreeNode*> to_new;
  to_new[nullptr] = nullptr;
  for (QuadTreeLevel* lv : levels) {
    QuadTreeLevel* nl = new QuadTreeLevel();
    for (QuadTreeNode* o : lv->nodes) {
      QuadTreeNode* n = new QuadTreeNode();
      n->level = o->level;
      n->dofs_below = o->dofs_below;
      n->is_leaf = o->is_leaf;
      n->X_rr_is_LU_factored = o->X_rr_is_LU_factored;
      n->compressed = o->compressed;
      n->side_length = o->side_length;
      n->dof_lists = o->dof_lists;
      n->T = o->T;
      n->L = o->L;
      n->U = o->U;
      n->X_rr = o->X_rr;
      n->schur_update = o->schur_update;
      n->X_rr_lu = o->X_rr_lu;
      n->X_rr_piv = o->X_rr_piv;
      for (int i = 0; i < 8; ++i) n->corners[i] = o->corners[i];
      to_new[o] = n;
      nl->nodes.push_back(n);
    }
    dst->levels.push_back(nl);
  }
  for (QuadTreeLevel* lv : levels)
    for (QuadTreeNode* o : lv->nodes) {
      QuadTreeNode* n = to_new[o];
      n->parent = to_new[o->parent];
      for (int c = 0; c < 4; ++c) n->children[c] = to_new[o->children[c]];
      for (QuadTreeNode* nb : o->neighbors) n->neighbors.push_back(to_new[nb]);
    }
  dst->root = to_new[root];
  dst->solution_dimension = solution_dimension;
  dst->domain_dimension = domain_dimension;
  dst->no_proxy_level = no_proxy_level;
  dst->leaf_dofs = leaf_dofs;
  dst->min = min;
  dst->max = max;
  dst->boundary_points = boundary_points;
  dst->normals_ = normals_;
  dst->weights_ = weights_;
  dst->curvatures_ = curvatures_;
  dst->allskel_mat = allskel_mat;
  dst->allskel_mat_lu = allskel_mat_lu;
  dst->U = U;
  dst->Psi = Psi;
  dst->S_LU = S_LU;
  dst->allskel_mat_piv = allskel_mat_piv;
  dst->S_piv = S_piv;
  if (!gpu_) return;
  dst->rebuild_device();
  if (fac_pde_ < 0) return;
  // the same factorization on the copy's device tree
  const int h = U.width();
  detail::check(kigpu_set_border(dst->gpu_, h ? dst->U.mat : nullptr, h ? dst->Psi.mat : nullptr, h),
                "kigpu_set_border");
  kigpu_factor_opts o;
  o.pde = fac_pde_;
  o.id_tol = fac_tol_;
  o.precision = fac_precision_;
  o.min_dofs = fac_min_dofs_;
  detail::check(kigpu_factor(detail::context(), dst->gpu_, &o), "kigpu_factor");
  dst->fac_pde_ = fac_pde_;
  dst->fac_tol_ = fac_tol_;
  dst->fac_precision_ = fac_precision_;
  dst->fac_min_dofs_ = fac_min_dofs_;
}

// the device tree from the host structs: nodes numbered level by level,
// empty trailing levels (left by consolidate_node) dropped
void QuadTree::rebuild_device() {
  const detail::CtxRef lock = detail::context();
  while (!levels.empty() && levels.back()->nodes.empty()) {
    delete levels.back();
    levels.pop_back();
  }
  by_id_.clear();
  std::vector<int64_t> loff{0};
  for (QuadTreeLevel* lv : levels) {
    by_id_.insert(by_id_.end(), lv->nodes.begin(), lv->nodes.end());
    loff.push_back(by_id_.size());
  }
  const int nn = by_id_.size();
  std::unordered_map<const QuadTreeNode*, int> id;
  for (int x = 0; x < nn; ++x) id[by_id_[x]] = x;
  std::vector<int> parent(nn), child(4 * (size_t)nn, -1), dofs, nbrs;
  std::vector<double> corners(8 * (size_t)nn), side(nn);
  std::vector<int64_t> boff{0}, noff{0};
  for (int x = 0; x < nn; ++x) {
    const QuadTreeNode* n = by_id_[x];
    parent[x] = n->parent ? id.at(n->parent) : -1;
    if (!n->is_leaf)
      for (int c = 0; c < 4; ++c) child[4 * (size_t)x + c] = id.at(n->children[c]);
    for (int j = 0; j < 8; ++j) corners[8 * (size_t)x + j] = n->corners[j];
    side[x] = n->side_length;
    if (n->is_leaf)
      dofs.insert(dofs.end(), n->dof_lists.original_box.begin(), n->dof_lists.original_box.end());
    boff.push_back(dofs.size());
    for (const QuadTreeNode* nb : n->neighbors) nbrs.push_back(id.at(nb));
    noff.push_back(nbrs.size());
  }
  kigpu_tree* g = nullptr;
  detail::check(kigpu_tree_import(detail::context(), boundary_points.data(), normals_.data(),
                                  weights_.data(), curvatures_.data(), (int)weights_.size(),
                                  solution_dimension, leaf_dofs, (int)levels.size(), loff.data(),
                                  parent.data(), child.data(), corners.data(), side.data(),
                                  boff.data(), dofs.data(), noff.data(), nbrs.data(), min, max, &g),
                "kigpu_tree_import");
  if (gpu_) kigpu_tree_destroy(gpu_);
  gpu_ = g;
  fac_pde_ = -1;
}

void QuadTree::remove_inactive_dofs_at_all_boxes() {
  for (int l = (int)levels.size() - 1; l >= 0; --l) remove_inactive_dofs_at_level(l);
}

}  // namespace kern_interp
