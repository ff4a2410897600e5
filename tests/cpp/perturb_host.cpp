// CPU driver for the host-side incremental edits of the drop-in QuadTree
// (QuadTree::perturb, copy_into, mark_neighbors_and_parents,
// consolidate_node in kern-interp-old_amd/src/quadtree.cpp): builds a tree
// on the host with the construction members (the steps of
// initialize_tree, quadtree.cpp:72-114), optionally marks every node as
// factored, copies it, perturbs the copy and prints both trees as JSON.
//
// The device calls these sources make are replaced by test doubles: the
// context is a null handle, and kigpu_tree_import (the device rebuild at
// the end of perturb) records the arrays it is given, which are printed
// too.  Every other device entry point fails loudly if reached.
//
// stdin: sd leaf_dofs factored nold x y ... nnew x y ...
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <unordered_map>
#include <vector>

#include "kern_interp/quadtree/quadtree.h"
#include "runtime.h"

namespace kern_interp {
namespace detail {
CtxRef context() {
  static std::recursive_mutex mu;
  return CtxRef{std::unique_lock<std::recursive_mutex>(mu), nullptr};
}
}  // namespace detail
}  // namespace kern_interp

namespace {
struct Imported {
  int nlevels = -1;
  std::vector<int64_t> loff, boff, noff;
  std::vector<int> parent, child, dofs, nbrs;
} g_import;
int g_fake_tree;
[[noreturn]] void unexpected(const char* f) {
  std::fprintf(stderr, "unexpected device call %s\n", f);
  std::abort();
}
}  // namespace

extern "C" {
const char* kigpu_status_string(kigpu_status) { return "stub"; }
const char* kigpu_last_error(void) { return "stub"; }
kigpu_status kigpu_tree_destroy(kigpu_tree*) { return KIGPU_OK; }
kigpu_status kigpu_tree_import(kigpu_ctx*, const double*, const double*, const double*,
                               const double*, int, int, int, int num_levels,
                               const int64_t* level_offsets, const int* parent,
                               const int* children, const double*, const double*,
                               const int64_t* box_offsets, const int* box_dofs,
                               const int64_t* nbr_offsets, const int* nbrs, double, double,
                               kigpu_tree** out) {
  Imported& m = g_import;
  m.nlevels = num_levels;
  m.loff.assign(level_offsets, level_offsets + num_levels + 1);
  const int64_t nn = m.loff.back();
  m.parent.assign(parent, parent + nn);
  m.child.assign(children, children + 4 * nn);
  m.boff.assign(box_offsets, box_offsets + nn + 1);
  m.dofs.assign(box_dofs, box_dofs + m.boff.back());
  m.noff.assign(nbr_offsets, nbr_offsets + nn + 1);
  m.nbrs.assign(nbrs, nbrs + m.noff.back());
  *out = reinterpret_cast<kigpu_tree*>(&g_fake_tree);
  return KIGPU_OK;
}
kigpu_status kigpu_tree_build(kigpu_ctx*, const double*, const double*, const double*,
                              const double*, int, int, int, kigpu_tree**) {
  unexpected("kigpu_tree_build");
}
kigpu_status kigpu_tree_get_info(const kigpu_tree*, kigpu_tree_info*) {
  unexpected("kigpu_tree_get_info");
}
kigpu_status kigpu_tree_export_nodes(const kigpu_tree*, int64_t*, int*, int*, int*, double*,
                                     double*) {
  unexpected("kigpu_tree_export_nodes");
}
kigpu_status kigpu_tree_export_boxes(const kigpu_tree*, int64_t*, int*) {
  unexpected("kigpu_tree_export_boxes");
}
kigpu_status kigpu_tree_export_neighbors(const kigpu_tree*, int64_t*, int*) {
  unexpected("kigpu_tree_export_neighbors");
}
kigpu_status kigpu_set_border(kigpu_tree*, const double*, const double*, int) {
  unexpected("kigpu_set_border");
}
kigpu_status kigpu_factor(kigpu_ctx*, kigpu_tree*, const kigpu_factor_opts*) {
  unexpected("kigpu_factor");
}
kigpu_status kigpu_interp_decomp(kigpu_ctx*, const double*, int, int, double, int*, int*,
                                 double*) {
  unexpected("kigpu_interp_decomp");
}
}

namespace {
using kern_interp::QuadTree;
using kern_interp::QuadTreeNode;

struct Pts : kern_interp::Boundary {
  std::unique_ptr<kern_interp::Boundary> clone() const override {
    return std::make_unique<Pts>(*this);
  }
  void initialize(int, kern_interp::BoundaryCondition) override {}
  bool is_in_domain(const kern_interp::Vec2&) const override { return false; }
};

void print_list(const char* key, const std::vector<int>& v) {
  std::printf("\"%s\": [", key);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%d", i ? ", " : "", v[i]);
  std::printf("]");
}
template <typename T>
void print_arr(const char* key, const std::vector<T>& v) {
  std::printf("\"%s\": [", key);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%lld", i ? ", " : "", (long long)v[i]);
  std::printf("]");
}

// levels, node by node: position (level, index) of parent, children and
// neighbors; corners; flags; DOF lists
void print_tree(const QuadTree& t) {
  std::unordered_map<const QuadTreeNode*, std::pair<int, int>> at;
  for (size_t l = 0; l < t.levels.size(); ++l)
    for (size_t i = 0; i < t.levels[l]->nodes.size(); ++i)
      at[t.levels[l]->nodes[i]] = {(int)l, (int)i};
  auto pos = [&](const QuadTreeNode* n) {
    if (!n) return std::string("null");
    auto it = at.find(n);
    if (it == at.end()) throw std::runtime_error("pointer to a node outside the levels");
    return "[" + std::to_string(it->second.first) + ", " + std::to_string(it->second.second) + "]";
  };
  std::printf("[");
  for (size_t l = 0; l < t.levels.size(); ++l) {
    std::printf("%s[", l ? ",\n " : "");
    const auto& nodes = t.levels[l]->nodes;
    for (size_t i = 0; i < nodes.size(); ++i) {
      const QuadTreeNode* n = nodes[i];
      std::printf("%s{\"level\": %d, \"leaf\": %d, \"compressed\": %d, \"dofs_below\": %d, ",
                  i ? ",\n  " : "", n->level, (int)n->is_leaf, (int)n->compressed, n->dofs_below);
      std::printf("\"parent\": %s, \"children\": [", pos(n->parent).c_str());
      for (int c = 0; c < 4; ++c) std::printf("%s%s", c ? ", " : "", pos(n->children[c]).c_str());
      std::printf("], \"neighbors\": [");
      for (size_t k = 0; k < n->neighbors.size(); ++k)
        std::printf("%s%s", k ? ", " : "", pos(n->neighbors[k]).c_str());
      std::printf("], \"corners\": [");
      for (int c = 0; c < 8; ++c) std::printf("%s%.17g", c ? ", " : "", n->corners[c]);
      std::printf("], \"side\": %.17g, ", n->side_length);
      print_list("original_box", n->dof_lists.original_box);
      std::printf(", ");
      print_list("active_box", n->dof_lists.active_box);
      std::printf(", ");
      print_list("skel", n->dof_lists.skel);
      std::printf(", ");
      print_list("redundant", n->dof_lists.redundant);
      std::printf(", ");
      print_list("skelnear", n->dof_lists.skelnear);
      std::printf("}");
    }
    std::printf("]");
  }
  std::printf("]");
}

std::vector<double> read_points() {
  int n = 0;
  if (std::scanf("%d", &n) != 1) throw std::runtime_error("bad input");
  std::vector<double> p(2 * (size_t)n);
  for (double& v : p)
    if (std::scanf("%lf", &v) != 1) throw std::runtime_error("bad input");
  return p;
}
}  // namespace

int main() {
  int sd = 0, leaf = 0, factored = 0;
  if (std::scanf("%d %d %d", &sd, &leaf, &factored) != 3) return 2;
  const std::vector<double> oldp = read_points(), newp = read_points();
  QuadTree t;
  t.solution_dimension = sd;
  t.domain_dimension = 2;
  t.leaf_dofs = leaf;
  t.boundary_points = oldp;
  // root box: min/max pooled over x and y, widened by 0.01 (quadtree.cpp:80-110)
  double lo = oldp[0], hi = oldp[0];
  for (double v : oldp) {
    lo = std::min(lo, v);
    hi = std::max(hi, v);
  }
  t.min = lo;
  t.max = hi;
  QuadTreeNode* root = new QuadTreeNode();
  const double a = lo - 0.01, b = hi + 0.01;
  const double c[8] = {a, a, a, b, b, b, b, a};
  for (int i = 0; i < 8; ++i) root->corners[i] = c[i];
  root->side_length = b - a;
  root->parent = nullptr;
  t.root = root;
  t.levels.push_back(new kern_interp::QuadTreeLevel());
  t.levels[0]->nodes.push_back(root);
  for (size_t i = 0; i < oldp.size() / 2; ++i)
    t.recursive_add(root, oldp[2 * i], oldp[2 * i + 1], (int)i);
  t.compute_neighbor_lists();
  if (factored) {  // every node compressed, with lists from its own box
    for (auto* lv : t.levels)
      for (QuadTreeNode* n : lv->nodes) {
        auto& d = n->dof_lists;
        n->compressed = true;
        d.active_box = d.original_box;
        d.skel.assign(d.active_box.begin(), d.active_box.begin() + d.active_box.size() / 2);
        d.redundant.assign(d.active_box.begin() + d.active_box.size() / 2, d.active_box.end());
        d.skelnear = d.skel;
      }
  }
  QuadTree cp;
  t.copy_into(&cp);
  std::printf("{\"before\": ");
  print_tree(t);
  std::printf(",\n\"copy\": ");
  print_tree(cp);
  // the copy shares no node with the original
  bool shared = false;
  for (auto* lv : t.levels)
    for (QuadTreeNode* n : lv->nodes)
      for (auto* lc : cp.levels)
        for (QuadTreeNode* m : lc->nodes) shared |= n == m;
  std::printf(",\n\"copy_shares_nodes\": %d", (int)shared);
  Pts nb;
  nb.points = newp;
  const size_t nn = newp.size() / 2;
  nb.normals.assign(2 * nn, 0.0);
  nb.weights.assign(nn, 1.0);
  nb.curvatures.assign(nn, 0.0);
  cp.perturb(nb);
  std::printf(",\n\"after\": ");
  print_tree(cp);
  const Imported& m = g_import;
  std::printf(",\n\"import\": {\"nlevels\": %d, ", m.nlevels);
  print_arr("loff", m.loff);
  std::printf(", ");
  print_arr("parent", m.parent);
  std::printf(", ");
  print_arr("child", m.child);
  std::printf(", ");
  print_arr("boff", m.boff);
  std::printf(", ");
  print_arr("dofs", m.dofs);
  std::printf(", ");
  print_arr("noff", m.noff);
  std::printf(", ");
  print_arr("nbrs", m.nbrs);
  std::printf("}}\n");
  return 0;
}
