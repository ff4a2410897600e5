// Kernel of the drop-in API (interface of the reference's kernel.h:15-69).
// Blocks K(tgt, src) come from the GPU (kigpu_kernel_block); the proxy rows
// and the ID-matrix assembly follow kernel.cpp:153-334.
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

// proxy rows against box DOFs (kernel.cpp:238-334): 128 points on a circle of
// radius r, rows [proxy <- box; box <- proxy]
ki_Mat Kernel::get_proxy_mat(double cx, double cy, double r, const QuadTree*,
                             const std::vector<int>& box) const {
  const int P = NUM_PROXY_POINTS;
  const double pw = 2.0 * M_PI * r / P;
  const double inv2pi = 1.0 / (2 * M_PI), invpi = 1.0 / M_PI;
  const bool stokes = pde == STOKES;
  ki_Mat ret((stokes ? 4 : 2) * P, box.size());
  for (size_t j = 0; j < box.size(); ++j) {
    const int d = box[j], pt = stokes ? d / 2 : d, comp = d % 2;
    const double bx = boundary_points[2 * pt], by = boundary_points[2 * pt + 1];
    const double bnx = boundary_normals[2 * pt], bny = boundary_normals[2 * pt + 1];
    const double bw = boundary_weights[pt];
    for (int i = 0; i < P; ++i) {
      const double ang = 2 * M_PI * i * (1.0 / P);
      const double pc = std::cos(ang), ps = std::sin(ang);
      const double r1 = (cx + r * pc) - bx, r2 = (cy + r * ps) - by;
      const double rr = r1 * r1 + r2 * r2;
      if (!stokes) {
        if (pde == LAPLACE) {
          ret.set(i, j, -bw * inv2pi * (r1 * bnx + r2 * bny) / rr);
          ret.set(P + i, j, -pw * inv2pi * ((-r1) * pc + (-r2) * ps) / rr);
        } else {
          ret.set(i, j, bw + bw * inv2pi * (r1 * pc + r2 * ps) / rr);
          ret.set(P + i, j, pw + pw * inv2pi * ((-r1) * bnx + (-r2) * bny) / rr);
        }
      } else {
        const double rv[2] = {r1, r2}, sn[2] = {bnx, bny}, tn[2] = {pc, ps};
        const double pot = bw * invpi * (r1 * bnx + r2 * bny) / (rr * rr);
        const double pot2 = -pw * invpi * (r1 * pc + r2 * ps) / (rr * rr);
        for (int a = 0; a < 2; ++a) {
          ret.set(2 * i + a, j, pot * rv[a] * rv[comp] + bw * tn[a] * sn[comp]);
          ret.set(2 * P + 2 * i + a, j, pot2 * rv[a] * rv[comp] + pw * sn[comp] * tn[a]);
        }
      }
    }
  }
  return ret;
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
