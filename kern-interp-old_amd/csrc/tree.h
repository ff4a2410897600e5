// Device-resident quadtree + factorization state owned by a kigpu_tree handle.
#pragma once

#include <memory>

#include "common.h"

namespace kigpu {

struct LevelWork;  // per-level factor work lists (factor.cpp)

struct TreeData {
  // ---- geometry (Kernel's copies of the boundary arrays) ----
  int npts = 0, sd = 1, leaf_dofs = 32, ndof = 0;
  double min = 0, max = 0;  // root box bounds of instance 0
  DBuf<double> px, py, nx, ny, w, kappa;

  // ---- instances: a batch of independent boundaries factored together ----
  // Instance i owns points [inst_pt[i], inst_pt[i+1]) and its own quadtree;
  // every tree level holds the nodes of all instances, instance by instance,
  // so each level is one launch over all of them.  A single tree is ninst = 1.
  int ninst = 1;
  std::vector<int64_t> inst_pt;        // ninst + 1 point offsets
  std::vector<double> inst_min, inst_max;
  DBuf<int64_t> d_inst_pt;
  DBuf<int> inst;                      // node -> instance
  std::vector<int> h_inst;

  // ---- tree, node ids in level order ----
  int nlevels = 0, nnodes = 0;
  std::vector<int64_t> level_off;  // nlevels + 1
  DBuf<int> parent, child, is_leaf, level, slot;
  DBuf<double> corners, side;
  DBuf<int64_t> pstart;            // node's points: perm_all[pstart .. pstart + pcount)
  DBuf<int> pcount, perm_all;
  DBuf<int> sl_nbr, sl_cnt;        // same-level neighbors, kMaxNbrs per node, in list order
  DBuf<int64_t> cp_off;            // coarse leaves pushed into a node (CSR)
  DBuf<int> cp_nbr;

  // host mirrors (the export-only ones are filled on demand)
  bool mirrors_ready = false, fmirrors_ready = false;
  int64_t nperm_total = 0, ncp = 0;
  std::vector<int> h_parent, h_child, h_is_leaf, h_level, h_pcount, h_perm, h_sl_nbr, h_sl_cnt,
      h_cp_nbr;
  std::vector<int64_t> h_pstart, h_cp_off;
  std::vector<double> h_corners, h_side;

  // coincident nodes: -1 not checked yet, 0 none, 1 some (the fast apply
  // assumes none)
  mutable int dup_state = -1;
  // block-pair table of the symmetric double-layer apply, built on first use:
  // 6 ints per pair (target block start and length, source block start and
  // length, target slot, source slot or -1 on a diagonal block) and 3 ints
  // per block (start, length, slots of its instance)
  mutable DBuf<int> sym_pairs, sym_blocks;
  mutable DBuf<int64_t> d_dof_off;     // instance DOF offsets (residual norms)
  mutable int sym_npairs = 0, sym_nblocks = 0, sym_slots = 0;
  // treecode apply of the refinement residual (treecode.hip): box centres and
  // series (per apply), interaction lists of every leaf (once per tree),
  // pair counts of one apply
  mutable bool tc_ready = false;
  mutable DBuf<double4> tc_ctr;
  mutable DBuf<double2> tc_coef;
  mutable DBuf<int> tc_far, tc_near, tc_nfar, tc_nnear;
  mutable DBuf<unsigned long long> tc_work;
  mutable DBuf<int4> tc_wlist;           // series work list (4 boxes or 1 large box per block)
  mutable int tc_nwork = 0;

  // ---- factorization ----
  bool factored = false;
  kigpu_factor_opts opts{};
  // solve sweeps captured as HIP graphs (contexts with graphs on): one per
  // (element size, vector, leading dimension, columns, border vector,
  // transpose); exec == nullptr marks a sweep seen once and not captured yet
  struct SweepGraph {
    int tsz;
    const void* x;
    int ldx, nrhs;
    const void* g;
    bool trans;
    hipGraphExec_t exec;
    // the bordered step's scratch (bgram, bM, alpha_dev) the capture used: a
    // later solve with more right-hand sides can move it
    const void* scr[3];
  };
  std::vector<SweepGraph> sweep_graphs;
  DBuf<double> gx;  // FP64 solves with graphs: the sweeps' vector, kept at one address
  void drop_graphs() {
    for (auto& e : sweep_graphs)
      if (e.exec) (void)hipGraphExecDestroy(e.exec);
    sweep_graphs.clear();
  }
  ~TreeData();  // factor.hip, where LevelWork is complete
  DBuf<int64_t> obox_off;  // original_box DOF list of each leaf in obox_pool
  DBuf<int> obox_pool;
  // per-node factor bookkeeping (device), node ids as above
  DBuf<int64_t> act_off;   // offset of the node's active list in act_pool / sr_pool
  DBuf<int> nact;          // |active_box|
  DBuf<int> nrows;         // ID rows (inner-circle or outside DOFs) of the node
  DBuf<int> rank;          // #skeleton DOFs, -1 = not compressed
  DBuf<int> osize;         // size of the list handed to the parent
  DBuf<uint64_t> fac_ptr;  // device address of the node's factor slab
  DBuf<uint64_t> acc_ptr;  // accumulated Schur update on the parent-facing list (0 = none)
  DBuf<int> acc_ld;
  DBuf<int> act_pool;      // active DOF lists
  DBuf<int> sr_pool;       // active lists permuted to [skel | redundant]
  DBuf<int> pos_pool;      // the ID pivots (positions in the active list)
  DBuf<int> piv_pool;      // LU pivots of X_rr
  DBuf<int> rows_pool;     // ID row lists of the current level
  DBuf<double> ws64;       // ID / elimination workspace when a box exceeds LDS
  DBuf<float> ws32;
  DBuf<int> status;        // device status word
  int64_t act_pool_size = 0;
  std::vector<int> h_rank, h_nact;
  std::vector<int64_t> h_act_off;
  std::vector<int> h_sr;   // host copy of sr_pool after factoring
  std::vector<int> h_act;  // host copy of act_pool after factoring
  std::vector<int> h_pos;  // host copy of pos_pool after factoring

  // roots, one per instance: instance i's allskel is root_dofs[root_doff[i] ..
  // + root_n[i]), its matrices start at root_moff[i] (n^2 entries each)
  int nroot = 0;                       // total root skeleton DOFs
  int nroot_max = 0;
  std::vector<int> h_root_n;
  std::vector<int64_t> h_root_doff, h_root_moff;
  DBuf<unsigned char> root_pack;       // root_n | root_doff | root_moff on the device
  const int* root_n_p = nullptr;
  const int64_t* root_doff_p = nullptr;
  const int64_t* root_moff_p = nullptr;
  DBuf<int> root_dofs;
  DBuf<double> root_lu;
  DBuf<int> root_piv;

  // per-level compressed-node lists for the solve sweeps (device), by level
  std::vector<std::unique_ptr<LevelWork>> work;

  // border (holes): U (ndof x h), Psi (h x ndof) as given; their swept forms
  int nborder = 0, nS = 0;
  std::vector<double> h_U, h_Psi;
  DBuf<double> bU64in, bP64in;         // U and Psi^T in FP64 (residuals)
  DBuf<double> bU64, bD64, bP64;       // swept U, D^{-1} C, swept Psi^T (FP64 factor)
  DBuf<float> bU32, bD32, bP32;        // the same for the FP32 factor
  DBuf<int> is_skel;                   // 1 on the root skeleton DOFs
  DBuf<double> root_mat;               // allskel_mat before inversion
  DBuf<double> Sinv;                   // inverse of the bordered matrix S
  DBuf<double> bgram, bM;              // solve scratch

  // solve scratch
  DBuf<double> x64, r64, f64, rb64, alpha64;
  DBuf<float> x32;
  DBuf<double> alpha_dev;
};

// tree.hip
// evs: an event pool (grown to the tree's depth bound) for the pipelined levels
void build_tree(TreeData& t, hipStream_t s, int* syncs, PinnedHost& pin,
                std::vector<hipEvent_t>& evs);
// A quadtree described on the host, nodes numbered level by level
// (tree_import.cpp): level_off[nlevels + 1]; parent[nn] (-1: root);
// child[4 nn] (all -1 or all set); corners[8 nn]; side[nn]; every leaf's
// DOF list box_dofs[box_off[x] .. box_off[x + 1]) (interior entries are not
// read); neighbor lists nbrs[nbr_off[x] .. nbr_off[x + 1]) in QuadTree order:
// coarser leaves, then same-level boxes, then finer boxes (not read).
struct TreeImport {
  int nlevels = 0;
  const int64_t* level_off = nullptr;
  const int* parent = nullptr;
  const int* child = nullptr;
  const double* corners = nullptr;
  const double* side = nullptr;
  const int64_t* box_off = nullptr;
  const int* box_dofs = nullptr;
  const int64_t* nbr_off = nullptr;
  const int* nbrs = nullptr;
  double min = 0, max = 0;
};
void import_tree(TreeData& t, const TreeImport& in, hipStream_t s);
void ensure_tree_mirrors(TreeData& t);    // host copies used by the export calls
void ensure_factor_mirrors(TreeData& t);

}  // namespace kigpu
