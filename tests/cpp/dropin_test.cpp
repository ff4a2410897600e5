// The drop-in C++ API on the GPU, exercised the way the reference's own test
// suite uses it (kern_interp_test.cpp:26-95 backward error, :297-316 and
// :392-410 analytic agreement): same headers, classes and calls.
#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "kern_interp/boundaries/annulus.h"
#include "kern_interp/boundaries/circle.h"
#include "kern_interp/boundaries/star.h"
#include "kern_interp/kernel/kernel.h"
#include "kern_interp/ki_mat.h"
#include "kern_interp/linear_solve.h"
#include "kern_interp/quadtree/quadtree.h"
#include "kern_interp/skel_factorization/skel_factorization.h"

using namespace kern_interp;

static int g_fail = 0;

static void report(const std::string& name, double err, double bound) {
  const bool ok = err <= bound;
  std::printf("%s %s err=%.3e bound=%.1e\n", ok ? "PASS" : "FAIL", name.c_str(), err, bound);
  std::fflush(stdout);
  if (!ok) ++g_fail;
}

// |K mu - f| / |f| with a dense K from Kernel::operator() (plus the border)
static double backward_error(const Kernel& kernel, Boundary* b, int leaf, double tol) {
  QuadTree tree;
  tree.leaf_dofs = leaf;
  tree.initialize_tree(b, kernel.solution_dimension, kernel.domain_dimension);
  SkelFactorization fac(tol, 4);
  std::vector<int> all(kernel.solution_dimension * b->weights.size());
  for (size_t i = 0; i < all.size(); ++i) all[i] = i;
  const bool holes = !b->holes.empty() && kernel.pde != Kernel::LAPLACE_NEUMANN;
  ki_Mat mu, alpha;
  if (holes) {
    tree.U = initialize_U_mat(kernel.pde, b->holes, b->points);
    tree.Psi = initialize_Psi_mat(kernel.pde, b->holes, *b);
    fac.skeletonize(kernel, &tree);
    linear_solve(fac, tree, b->boundary_values, &mu, &alpha);
  } else {
    fac.skeletonize(kernel, &tree);
    linear_solve(fac, tree, b->boundary_values, &mu);
  }
  ki_Mat r = kernel(all, all) * mu - b->boundary_values;
  double e2 = r.vec_two_norm() * r.vec_two_norm();
  if (holes) {
    r += tree.U * alpha;
    e2 = r.vec_two_norm() * r.vec_two_norm();
    ki_Mat r2 = tree.Psi * mu - alpha;
    e2 += r2.vec_two_norm() * r2.vec_two_norm();
  }
  return std::sqrt(e2) / b->boundary_values.vec_two_norm();
}

int main() {
  {
    Circle c;
    c.initialize(1024, SINGLE_ELECTRON);
    Kernel k(1, 2, Kernel::LAPLACE, &c, {});
    report("laplace_circle_backward_error", backward_error(k, &c, 32, 1e-6), 1e-5);
  }
  {
    StarCurve s;
    s.initialize(4096, DEFAULT);
    Kernel k(1, 2, Kernel::LAPLACE, &s, {});
    report("laplace_star_backward_error", backward_error(k, &s, 64, 1e-6), 1e-5);
  }
  {
    Circle c;
    c.initialize(1024, ALL_ONES);
    Kernel k(1, 2, Kernel::LAPLACE_NEUMANN, &c, {});
    report("neumann_circle_backward_error", backward_error(k, &c, 32, 1e-6), 1e-5);
  }
  {
    Circle c;
    c.initialize(1024, TANGENT_VEC);
    Kernel k(2, 2, Kernel::STOKES, &c, {});
    report("stokes_circle_backward_error", backward_error(k, &c, 32, 1e-6), 1e-5);
  }
  {
    Annulus a;
    a.initialize(1024, SINGLE_ELECTRON);
    Kernel k(1, 2, Kernel::LAPLACE, &a, {});
    report("laplace_annulus_backward_error", backward_error(k, &a, 32, 1e-6), 1e-5);
  }
  {
    // analytic agreement: N = 2^12, id_tol 1e-13, 20 x 20 grid over the tree box
    Circle c;
    c.initialize(4096, SINGLE_ELECTRON);
    QuadTree tree;
    tree.initialize_tree(&c, 1, 2);
    std::vector<double> pts;
    get_domain_points(20, &pts, tree.min, tree.max, tree.min, tree.max);
    Kernel k(1, 2, Kernel::LAPLACE, &c, pts);
    ki_Mat sol = boundary_integral_solve(k, c, &tree, 1e-13, 4, pts);
    double dn = 0, tn = 0;
    for (size_t i = 0; i < pts.size(); i += 2) {
      if (!c.is_in_domain(Vec2(pts[i], pts[i + 1]))) continue;
      const double u = std::log(std::sqrt(std::pow(pts[i] + 3, 2) + std::pow(pts[i + 1] + 2, 2))) /
                       (2 * M_PI);
      dn += std::pow(u - sol.get(i / 2, 0), 2);
      tn += u * u;
    }
    report("laplace_circle_analytic", std::sqrt(dn / tn), 1e-12);
  }
  {
    // FP32 elimination + FP64 refinement through the same API
    StarCurve s;
    s.initialize(4096, DEFAULT);
    Kernel k(1, 2, Kernel::LAPLACE, &s, {});
    QuadTree tree;
    tree.leaf_dofs = 64;
    tree.initialize_tree(&s, 1, 2);
    SkelFactorization fac(1e-6, 4);
    fac.fp32_refine = true;
    fac.skeletonize(k, &tree);
    ki_Mat mu;
    linear_solve(fac, tree, s.boundary_values, &mu);
    std::vector<int> all(4096);
    for (int i = 0; i < 4096; ++i) all[i] = i;
    ki_Mat r = k(all, all) * mu - s.boundary_values;
    report("laplace_star_fp32_refine_residual",
           r.vec_two_norm() / s.boundary_values.vec_two_norm(), 1e-10);
  }
  std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "OK", g_fail);
  return g_fail ? 1 : 0;
}
