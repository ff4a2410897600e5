// Kernel of the drop-in API (interface of the reference's kernel.h:15-69).
// Blocks K(tgt, src) and proxy rows come from the GPU (kigpu_kernel_block,
// kigpu_proxy_block); the ID-matrix assembly follows kernel.cpp:153-235.
#include "kern_interp/kernel/kernel.h"

#include <cmath>

#include "runtime.h"

namespace kern_interp {

kigpu_geom* Kernel::geom() const {
  if (!geom_) {
    kigpu_geom* g = nullptr;
    detail::check(kigpu_geometry_create(detail::context(), boundary_points.data(),
                                        boundary_normals.data(), boundary_weights.data(),
                                        boundary_curvatures.data(),
                                        (int)boundary_weights.size(), domain_points.data(),
                                        (int)domain_points.size() / 2, &g),
                  "kigpu_geometry_create");
    geom_ = std::shared_ptr<kigpu_geom>(g, [](kigpu_geom* p) { kigpu_geometry_destroy(p); });
  }
  return geom_.get();
}

void Kernel::update_boundary(Boundary* boundary) {
  boundary_points = boundary->points;
  boundary_normals = boundary->normals;
  boundary_weights = boundary->weights;
  boundary_curvatures = boundary->curvatures;
  geom_.reset();
}

ki_Mat Kernel::operator()(const std::vector<int>& tgt, const std::vector<int>& src,
                          bool forward) const {
  ki_Mat out(tgt.size(), src.size());
  if (tgt.empty() || src.empty()) return out;
  detail::check(kigpu_kernel_block(detail::context(), geom(), (int)pde, solution_dimension,
                                   tgt.data(), (int)tgt.size(), src.data(), (int)src.size(),
                                   forward ? 1 : 0, out.mat),
                "kigpu_kernel_block");
  return out;
}

ki_Mat Kernel::laplace_get(const std::vector<int>& t, const std::vector<int>& s, bool f) const {
  return (*this)(t, s, f);
}

ki_Mat Kernel::stokes_get(const std::vector<int>& t, const std::vector<int>& s, bool f) const {
  return (*this)(t, s, f);
}

ki_Mat Kernel::forward() const {
  const int sd = solution_dimension;
  std::vector<int> tgt((domain_points.size() / 2) * (sd == 1 ? 1 : 2));
  std::vector<int> src(boundary_weights.size() * sd);
  for (size_t i = 0; i < tgt.size(); ++i) tgt[i] = i;
  for (size_t j = 0; j < src.size(); ++j) src[j] = j;
  return (*this)(tgt, src, true);
}

// kernel.cpp:11-36, one entry: the double-layer entry for LAPLACE, the
// Neumann entry (or its single-layer log form when forward) for
// LAPLACE_NEUMANN; r = target - source, diagonal when r is exactly zero.
void Kernel::laplace(int mat_idx, ki_Mat* ret, double r1, double r2, double sw, double sc,
                     double sn1, double sn2, double tn1, double tn2, bool forward) const {
  const double inv2pi = 1.0 / (2 * M_PI);
  double v = 0;
  if (pde == LAPLACE) {
    if (r1 == 0. && r2 == 0.) v = 0.5 + 0.5 * sc * sw * inv2pi;
    else v = -sw * inv2pi * (r1 * sn1 + r2 * sn2) / (r1 * r1 + r2 * r2);
  } else if (pde == LAPLACE_NEUMANN) {
    if (forward) v = sw * inv2pi * std::log(std::sqrt(r1 * r1 + r2 * r2));
    else if (r1 == 0. && r2 == 0.) v = sw - 0.5 + 0.5 * sc * sw * inv2pi;
    else v = sw + sw * inv2pi * (r1 * tn1 + r2 * tn2) / (r1 * r1 + r2 * r2);
  } else {
    return;
  }
  ret->mat[mat_idx] = v;
}

static ki_Mat proxy_rows(kigpu_geom* g, int pde, const std::vector<double>& pp,
                         const std::vector<double>& pn, double pw, const std::vector<int>& box) {
  const int np = pp.size() / 2;
  ki_Mat ret((pde == Kernel::STOKES ? 4 : 2) * np, box.size());
  if (np == 0 || box.empty()) return ret;
  detail::check(kigpu_proxy_block(detail::context(), g, pde, pp.data(), pn.data(), np, pw,
                                  box.data(), (int)box.size(), ret.mat),
                "kigpu_proxy_block");
  return ret;
}

ki_Mat Kernel::laplace_proxy_get(const std::vector<double>& pxy_p,
                                 const std::vector<double>& pxy_n, double pxy_w,
                                 const std::vector<int>& box_inds) const {
  return proxy_rows(geom(), pde == LAPLACE_NEUMANN ? LAPLACE_NEUMANN : LAPLACE, pxy_p, pxy_n,
                    pxy_w, box_inds);
}

ki_Mat Kernel::stokes_proxy_get(const std::vector<double>& pxy_p,
                                const std::vector<double>& pxy_n, double pxy_w,
                                const std::vector<int>& box_inds) const {
  return proxy_rows(geom(), STOKES, pxy_p, pxy_n, pxy_w, box_inds);
}

// kernel.cpp:236-265: NUM_PROXY_POINTS points on a circle of radius r, unit
// outward normals, weight 2 pi r / NUM_PROXY_POINTS
ki_Mat Kernel::get_proxy_mat(double cx, double cy, double r, const QuadTree*,
                             const std::vector<int>& box) const {
  const double pw = 2.0 * M_PI * r / NUM_PROXY_POINTS;
  std::vector<double> pp, pn;
  for (int i = 0; i < NUM_PROXY_POINTS; ++i) {
    const double ang = 2 * M_PI * i * (1.0 / NUM_PROXY_POINTS);
    pp.push_back(cx + r * std::cos(ang));
    pp.push_back(cy + r * std::sin(ang));
    pn.push_back(std::cos(ang));
    pn.push_back(std::sin(ang));
  }
  return pde == STOKES ? stokes_proxy_get(pp, pn, pw, box) : laplace_proxy_get(pp, pn, pw, box);
}

// kernel.cpp:153-235: [K(rows, B); K(B, rows)^T (; proxy)] from the node's lists
ki_Mat Kernel::get_id_mat(const QuadTree* tree, const QuadTreeNode* node) const {
  const std::vector<int>& box = node->dof_lists.active_box;
  std::vector<int> rows;
  bool proxy = false;
  if (node->level == 1 || node->level == 2) {
    for (QuadTreeNode* o : tree->levels[node->level]->nodes)
      if (o != node) rows.insert(rows.end(), o->dof_lists.active_box.begin(), o->dof_lists.active_box.end());
    if (node->level == 2)
      for (QuadTreeNode* o : tree->levels[1]->nodes)
        if (o->is_leaf)
          rows.insert(rows.end(), o->dof_lists.original_box.begin(), o->dof_lists.original_box.end());
  } else {
    proxy = true;
    const double cx = node->corners[0] + node->side_length / 2.0;
    const double cy = node->corners[1] + node->side_length / 2.0;
    for (int d : node->dof_lists.near) {
      const int pt = d / solution_dimension;
      const double dx = cx - boundary_points[2 * pt], dy = cy - boundary_points[2 * pt + 1];
      if (std::sqrt(dx * dx + dy * dy) < RADIUS_RATIO * node->side_length) rows.push_back(d);
    }
  }
  ki_Mat pxy;
  if (proxy) {
    const double cx = node->corners[0] + node->side_length / 2.0;
    const double cy = node->corners[1] + node->side_length / 2.0;
    pxy = get_proxy_mat(cx, cy, node->side_length * RADIUS_RATIO, tree, box);
  }
  const int q = rows.size();
  ki_Mat mat(2 * q + pxy.height(), box.size());
  if (q) {
    mat.set_submatrix(0, q, 0, box.size(), (*this)(rows, box));
    mat.set_submatrix(q, 2 * q, 0, box.size(), (*this)(box, rows), true);
  }
  if (proxy) mat.set_submatrix(2 * q, 2 * q + pxy.height(), 0, box.size(), pxy);
  return mat;
}

}  // namespace kern_interp
