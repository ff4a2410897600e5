// Drop-in for kern_interp/quadtree/quadtree.h (reference quadtree.h:14-102).
// initialize_tree builds the tree on the GPU (kigpu_tree_build) and mirrors
// nodes, DOF lists and neighbor lists into these host structs; the factor
// data stays device-resident in the handle (gpu()).  After host edits
// (perturb, copy_into) the device tree is rebuilt from the structs.
#ifndef KERN_INTERP_QUADTREE_QUADTREE_H_
#define KERN_INTERP_QUADTREE_QUADTREE_H_

#include <cassert>
#include <vector>

#include "kern_interp/boundaries/boundary.h"
#include "kern_interp/ki_mat.h"

#ifndef MAX_LEAF_DOFS
#define MAX_LEAF_DOFS 32
#endif

struct kigpu_tree;

namespace kern_interp {

struct InteractionLists {
  std::vector<int> original_box, active_box, redundant, skel, near, skelnear, permutation;

  void set_rs_ranges(const std::vector<int>& prm, int sk, int rd);
  void set_skelnear_range();
};

struct QuadTreeNode {
  int level = 0, dofs_below = 0;
  bool is_leaf, X_rr_is_LU_factored = false, compressed = false;
  double side_length;
  QuadTreeNode* parent;
  QuadTreeNode* children[4];
  std::vector<QuadTreeNode*> neighbors;
  InteractionLists dof_lists;
  // factor blocks live on the device; these stay empty in this build
  ki_Mat T, L, U, X_rr, schur_update, X_rr_lu;
  std::vector<lapack_int> X_rr_piv;
  // format is {BL, TL, TR, BR}
  double corners[8];
  QuadTreeNode() {
    is_leaf = true;
    for (int i = 0; i < 4; i++) children[i] = nullptr;
  }
};

struct QuadTreeLevel {
  std::vector<QuadTreeNode*> nodes;
  ~QuadTreeLevel() {
    for (QuadTreeNode* node : nodes) delete node;
  }
};

class QuadTree {
 public:
  int solution_dimension, domain_dimension;
  int no_proxy_level = 0;
  // leaf capacity in DOFs (the reference fixes it with MAX_LEAF_DOFS)
  int leaf_dofs = MAX_LEAF_DOFS;
  double min, max;
  std::vector<double> boundary_points;
  QuadTreeNode* root = nullptr;
  ki_Mat allskel_mat, allskel_mat_lu, U, Psi, S_LU;
  std::vector<lapack_int> allskel_mat_piv, S_piv;
  std::vector<QuadTreeLevel*> levels;

  QuadTree() = default;
  QuadTree(const QuadTree&) = delete;
  QuadTree& operator=(const QuadTree&) = delete;
  ~QuadTree();
  void initialize_tree(Boundary* boundary, int solution_dimension_, int domain_dimension_);
  void reset();
  void reset(Boundary* boundary_);

  void remove_inactive_dofs_at_level(int level);
  void remove_inactive_dofs_at_all_boxes();
  void remove_inactive_dofs_at_box(QuadTreeNode* node);

  // Host construction steps on these node structs (quadtree.cpp:27-285),
  // with leaf_dofs as the leaf capacity.  initialize_tree does not use them:
  // it builds the same tree on the GPU.
  void compute_neighbor_lists();
  void recursive_add(QuadTreeNode* node, double x, double y, int point_ind);
  void get_descendent_neighbors(QuadTreeNode* big, QuadTreeNode* small);
  void node_subdivide(QuadTreeNode* node);
  void consolidate_node(QuadTreeNode* node);

  // Incremental edits (quadtree.cpp:288-641) on the host structs; the device
  // tree is then rebuilt from them (kigpu_tree_import).  perturb maps the DOF
  // lists to the new boundary's numbering, adds and deletes points, marks
  // the touched boxes, their neighbors and ancestors uncompressed, splits
  // overfull leaves, merges sparse subtrees and recomputes the neighbor
  // lists.  The next skeletonize factors the new device tree (every box: one
  // batched launch per level costs about what the marked boxes alone would).
  // copy_into deep-copies the structs and, when this tree is factored, gives
  // the copy the same factorization.
  void copy_into(QuadTree* new_tree) const;
  void mark_neighbors_and_parents(QuadTreeNode* node);
  void perturb(const Boundary& new_boundary);

  // device handle (geometry, tree, factorization)
  kigpu_tree* gpu() const { return gpu_; }
  // node ids in level order, as the C ABI numbers them
  const std::vector<QuadTreeNode*>& nodes_by_id() const { return by_id_; }

 private:
  friend class SkelFactorization;
  void clear();
  void rebuild_device();  // device tree from the host structs
  kigpu_tree* gpu_ = nullptr;
  std::vector<QuadTreeNode*> by_id_;
  std::vector<double> normals_, weights_, curvatures_;  // the tree's boundary arrays
  // the factorization on the device tree (pde < 0: none)
  int fac_pde_ = -1, fac_precision_ = 0, fac_min_dofs_ = 0;
  double fac_tol_ = 0;
};

}  // namespace kern_interp

#endif  // KERN_INTERP_QUADTREE_QUADTREE_H_
