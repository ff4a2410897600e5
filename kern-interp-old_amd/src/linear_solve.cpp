// linear_solve of the drop-in API (interface of the reference's
// linear_solve.h:17-35): hole border matrices and the end-to-end driver.
// The domain evaluation is one device call (kigpu_domain_solution): the
// matrix-free forward apply instead of the reference's dense forward matrix,
// the hole terms and the in-domain mask (linear_solve.cpp:143-180).
#include "kern_interp/linear_solve.h"

#include <cmath>

#include "runtime.h"

namespace kern_interp {

// linear_solve.cpp:18-64
ki_Mat initialize_U_mat(const Kernel::Pde pde, const std::vector<Hole>& holes,
                        const std::vector<double>& tgt) {
  if (holes.empty() || pde == Kernel::LAPLACE_NEUMANN) return ki_Mat(0, 0);
  const int np = tgt.size() / 2, h = holes.size();
  if (pde == Kernel::LAPLACE) {
    ki_Mat U(np, h);
    for (int i = 0; i < np; ++i)
      for (int k = 0; k < h; ++k)
        U.set(i, k, std::log((Vec2(tgt[2 * i], tgt[2 * i + 1]) - holes[k].center).norm()));
    return U;
  }
  ki_Mat U(2 * np, 3 * h);
  const double sc = 1.0 / (4 * M_PI);
  for (int i = 0; i < np; ++i)
    for (int k = 0; k < h; ++k) {
      const Vec2 r = Vec2(tgt[2 * i], tgt[2 * i + 1]) - holes[k].center;
      const double n2 = r.norm() * r.norm(), lg = std::log(1 / r.norm());
      U.set(2 * i, 3 * k, sc * (lg + r.a[0] * r.a[0] / n2));
      U.set(2 * i + 1, 3 * k, sc * (r.a[1] * r.a[0] / n2));
      U.set(2 * i, 3 * k + 1, sc * (r.a[0] * r.a[1] / n2));
      U.set(2 * i + 1, 3 * k + 1, sc * (lg + r.a[1] * r.a[1] / n2));
      U.set(2 * i, 3 * k + 2, r.a[1] * (sc / n2));
      U.set(2 * i + 1, 3 * k + 2, -r.a[0] * (sc / n2));
    }
  return U;
}

// linear_solve.cpp:67-110: a node belongs to the first hole within radius + 1e-8
ki_Mat initialize_Psi_mat(const Kernel::Pde pde, const std::vector<Hole>& holes,
                          const Boundary& b) {
  if (holes.empty() || pde == Kernel::LAPLACE_NEUMANN) return ki_Mat(0, 0);
  const int np = b.points.size() / 2, h = holes.size();
  const bool stokes = pde == Kernel::STOKES;
  ki_Mat Psi(stokes ? 3 * h : h, stokes ? 2 * np : np);
  for (int i = 0; i < np; ++i) {
    const Vec2 x(b.points[2 * i], b.points[2 * i + 1]);
    for (int k = 0; k < h; ++k) {
      if ((x - holes[k].center).norm() >= holes[k].radius + 1e-8) continue;
      const double w = b.weights[i];
      if (!stokes) {
        Psi.set(k, i, w);
      } else {
        Psi.set(3 * k, 2 * i, w);
        Psi.set(3 * k + 1, 2 * i + 1, w);
        Psi.set(3 * k + 2, 2 * i, w * x.a[1]);
        Psi.set(3 * k + 2, 2 * i + 1, -w * x.a[0]);
      }
      break;
    }
  }
  return Psi;
}

void linear_solve(const SkelFactorization& f, const QuadTree& tree, const ki_Mat& rhs, ki_Mat* mu,
                  ki_Mat* alpha) {
  *mu = ki_Mat(rhs.height(), 1);
  if (!alpha) {
    f.solve(tree, mu, rhs);
  } else {
    f.multiply_connected_solve(tree, mu, alpha, rhs);
  }
}

void schur_solve(const SkelFactorization& f, const QuadTree& tree, const ki_Mat& U, const ki_Mat&,
                 const ki_Mat& rhs, const ki_Mat& K_domain, const ki_Mat& U_forward,
                 ki_Mat* solution) {
  ki_Mat mu;
  if (U.width() == 0) {
    linear_solve(f, tree, rhs, &mu);
    *solution = K_domain * mu;
  } else {
    ki_Mat alpha;
    linear_solve(f, tree, rhs, &mu, &alpha);
    *solution = (K_domain * mu) + (U_forward * alpha);
  }
}

ki_Mat boundary_integral_solve(const Kernel& kernel, const Boundary& boundary, QuadTree* tree,
                               double id_tol, int fact_threads,
                               const std::vector<double>& domain_points) {
  SkelFactorization fac(id_tol, fact_threads);
  ki_Mat U = initialize_U_mat(kernel.pde, boundary.holes, boundary.points);
  ki_Mat Psi = initialize_Psi_mat(kernel.pde, boundary.holes, boundary);
  tree->U = U;
  tree->Psi = Psi;
  fac.skeletonize(kernel, tree);
  ki_Mat mu, alpha;
  if (U.width() == 0) linear_solve(fac, *tree, boundary.boundary_values, &mu);
  else linear_solve(fac, *tree, boundary.boundary_values, &mu, &alpha);
  const int ntgt = domain_points.size() / 2, sd = kernel.solution_dimension;
  ki_Mat sol(ntgt * sd, 1);
  if (!ntgt) return sol;
  // forward apply, hole terms and the in-domain mask in one device call
  const DomainRule rule = boundary.domain_rule();
  kigpu_region reg{};
  reg.kind = rule.kind;
  reg.eps = rule.eps;
  reg.cx = rule.cx;
  reg.cy = rule.cy;
  reg.radius = rule.radius;
  if (rule.kind == DomainRule::kDisc) {
    reg.nholes = rule.holes.size() / 3;
    reg.holes = rule.holes.data();
  } else if (rule.kind == DomainRule::kWinding) {
    reg.nodes = boundary.points.data();
    reg.num_outer = boundary.num_outer_nodes;
    reg.nholes = rule.hole_nodes.size();
    reg.hole_nodes = rule.hole_nodes.data();
  } else if (rule.kind == DomainRule::kStar) {
    reg.nterms = rule.amps.size();
    reg.amps = rule.amps.data();
    reg.freqs = rule.freqs.data();
    reg.phases = rule.phases.data();
  }
  std::vector<double> hc;
  for (const Hole& h : boundary.holes) hc.insert(hc.end(), {h.center.a[0], h.center.a[1]});
  const bool with_holes = U.width() > 0;
  detail::check(kigpu_domain_solution(detail::context(), tree->gpu(), (int)kernel.pde,
                                      domain_points.data(), ntgt, mu.mat,
                                      with_holes ? alpha.mat : nullptr, hc.data(),
                                      with_holes ? (int)boundary.holes.size() : 0, 1,
                                      rule.kind == DomainRule::kNone ? nullptr : &reg, sol.mat),
                "kigpu_domain_solution");
  // a Boundary subclass without a device rule: its own host is_in_domain
  if (rule.kind == DomainRule::kNone)
    for (int i = 0; i < ntgt; ++i) {
      if (boundary.is_in_domain(Vec2(domain_points[2 * i], domain_points[2 * i + 1]))) continue;
      for (int c = 0; c < sd; ++c) sol.set(sd * i + c, 0, 0.);
    }
  return sol;
}

void get_domain_points(int n, std::vector<double>* points, double x0, double x1, double y0,
                       double y1) {
  for (int i = 0; i < n; ++i) {
    const double x = x0 + ((i + 0.0) / (n - 1)) * (x1 - x0);
    for (int j = 0; j < n; ++j) {
      points->push_back(x);
      points->push_back(y0 + ((j + 0.0) / (n - 1)) * (y1 - y0));
    }
  }
}

}  // namespace kern_interp
