// A device tree from a host description (QuadTree's node structs after
// perturb or copy_into, quadtree.cpp:359-641): the same arrays build_tree
// derives on the GPU, validated and filled on the host, so that the
// factorization and solves run on it unchanged.
#include <algorithm>
#include <string>

#include "tree.h"

namespace kigpu {

namespace {
[[noreturn]] void bad(const std::string& what) { KI_FAIL(KIGPU_ERR_ARG, "tree import: " + what); }
}  // namespace

void import_tree(TreeData& t, const TreeImport& in, hipStream_t s) {
  const int nl = in.nlevels;
  if (nl < 1 || nl > 31) bad("level count must be 1 .. 31");
  if (in.level_off[0] != 0 || in.level_off[1] != 1) bad("level 0 must hold the root alone");
  for (int l = 0; l < nl; ++l)
    if (in.level_off[l + 1] <= in.level_off[l]) bad("empty level");
  const int64_t nn64 = in.level_off[nl];
  if (nn64 >= (1ll << 30)) bad("too many nodes");
  const int nn = (int)nn64, sd = t.sd;
  std::vector<int> level(nn), slot(nn, 0), is_leaf(nn);
  for (int l = 0; l < nl; ++l)
    for (int64_t x = in.level_off[l]; x < in.level_off[l + 1]; ++x) level[x] = l;
  // parent / child links
  for (int x = 0; x < nn; ++x) {
    const int* c = in.child + 4 * (size_t)x;
    const int set = (c[0] >= 0) + (c[1] >= 0) + (c[2] >= 0) + (c[3] >= 0);
    if (set != 0 && set != 4) bad("a box must have four children or none");
    is_leaf[x] = set == 0;
    for (int q = 0; q < 4 && set; ++q) {
      if (c[q] >= nn || level[c[q]] != level[x] + 1 || in.parent[c[q]] != x)
        bad("child link inconsistent");
      slot[c[q]] = q;
    }
    if (x == 0 ? in.parent[0] != -1 : (in.parent[x] < 0 || in.parent[x] >= nn))
      bad("parent link inconsistent");
  }
  // points of every box: a leaf's from its DOF list (sd consecutive DOFs per
  // point), an interior box's the ascending union of its children's
  std::vector<std::vector<int>> pts(nn);
  std::vector<char> seen(t.npts, 0);
  int64_t covered = 0;
  for (int x = 0; x < nn; ++x) {
    if (!is_leaf[x]) continue;
    const int64_t a = in.box_off[x], b = in.box_off[x + 1];
    if (b < a || (b - a) % sd) bad("leaf DOF list size");
    for (int64_t e = a; e < b; e += sd) {
      const int d = in.box_dofs[e];
      if (d < 0 || d % sd || d / sd >= t.npts) bad("DOF out of range");
      for (int q = 1; q < sd; ++q)
        if (in.box_dofs[e + q] != d + q) bad("a point's DOFs must be consecutive");
      if (seen[d / sd]++) bad("a point in two leaves");
      pts[x].push_back(d / sd);
    }
    covered += pts[x].size();
  }
  if (covered != t.npts) bad("the leaves do not hold every point");
  for (int l = nl - 1; l >= 0; --l)
    for (int64_t x = in.level_off[l]; x < in.level_off[l + 1]; ++x) {
      if (is_leaf[x]) continue;
      for (int q = 0; q < 4; ++q) {
        const std::vector<int>& c = pts[in.child[4 * x + q]];
        pts[x].insert(pts[x].end(), c.begin(), c.end());
      }
      std::sort(pts[x].begin(), pts[x].end());
    }
  std::vector<int64_t> pstart(nn);
  std::vector<int> pcount(nn), perm;
  for (int x = 0; x < nn; ++x) {
    pstart[x] = perm.size();
    pcount[x] = pts[x].size();
    perm.insert(perm.end(), pts[x].begin(), pts[x].end());
  }
  // neighbor lists: coarser leaves (cp) before same-level boxes (sl)
  std::vector<int> sl(kMaxNbrs * (size_t)nn, -1), slc(nn, 0), cp;
  std::vector<int64_t> cpo(nn + 1, 0);
  for (int x = 0; x < nn; ++x) {
    int phase = 0;
    for (int64_t e = in.nbr_off[x]; e < in.nbr_off[x + 1]; ++e) {
      const int y = in.nbrs[e];
      if (y < 0 || y >= nn || y == x) bad("neighbor out of range");
      const int ph = level[y] < level[x] ? 0 : level[y] == level[x] ? 1 : 2;
      if (ph < phase) bad("neighbor lists must list coarser, same-level, finer boxes in that order");
      phase = ph;
      if (ph == 0) {
        if (!is_leaf[y]) bad("a coarser neighbor must be a leaf");
        cp.push_back(y);
      } else if (ph == 1) {
        if (slc[x] == kMaxNbrs) bad("more than 8 same-level neighbors");
        sl[kMaxNbrs * (size_t)x + slc[x]++] = y;
      }
    }
    cpo[x + 1] = cp.size();
  }
  std::vector<int> parent(in.parent, in.parent + nn), child(in.child, in.child + 4 * (size_t)nn);
  std::vector<double> corners(in.corners, in.corners + 8 * (size_t)nn), side(in.side, in.side + nn);

  t.inst_pt = {0, (int64_t)t.npts};
  t.ninst = 1;
  t.d_inst_pt.upload(t.inst_pt, s);
  t.nlevels = nl;
  t.level_off.assign(in.level_off, in.level_off + nl + 1);
  t.nnodes = nn;
  t.parent.upload(parent, s);
  t.child.upload(child, s);
  t.is_leaf.upload(is_leaf, s);
  t.level.upload(level, s);
  t.slot.upload(slot, s);
  t.corners.upload(corners, s);
  t.side.upload(side, s);
  t.pstart.upload(pstart, s);
  t.pcount.upload(pcount, s);
  t.perm_all.upload(perm, s);
  t.sl_nbr.upload(sl, s);
  t.sl_cnt.upload(slc, s);
  t.cp_off.upload(cpo, s);
  if (cp.empty()) cp.push_back(-1);  // a one-entry buffer; ncp stays 0
  const int64_t ncp = cpo[nn];
  t.cp_nbr.upload(cp, s);
  t.inst.alloc(nn);
  KI_HIP(hipMemsetAsync(t.inst.p, 0, sizeof(int) * nn, s));
  t.status.alloc(1);
  t.status.zero(s);
  t.nperm_total = perm.size();
  t.ncp = ncp;
  t.min = in.min;
  t.max = in.max;
  t.inst_min = {in.min};
  t.inst_max = {in.max};
  // host mirrors, complete: the export calls need no download
  t.h_corners = corners;
  t.h_side = side;
  t.h_inst.assign(nn, 0);
  t.h_is_leaf = is_leaf;
  t.h_pcount = pcount;
  t.h_parent = parent;
  t.h_child = child;
  t.h_level = level;
  t.h_pstart = pstart;
  t.h_perm = perm;
  t.h_sl_nbr = sl;
  t.h_sl_cnt = slc;
  t.h_cp_off = cpo;
  t.h_cp_nbr.assign(cp.begin(), cp.begin() + ncp);
  t.mirrors_ready = true;
  stream_wait(s);
}

}  // namespace kigpu
