// Experiment three of the reference (experiments/experiment_three.cpp:15-227)
// through the drop-in API: place two star-shaped holes in the starfish so as
// to maximise the flow -u_x at a point (Stokes, EX3B data).  Each step takes
// the fourth-order finite-difference gradient in both hole angles from four
// perturbed copies of the tree per angle (QuadTree::copy_into + perturb,
// solved from four host threads), then a backtracking line search that
// keeps the holes pi/4 apart.
//   experiment_three [steps=30] [N=4096]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "kern_interp/boundaries/ex3boundary.h"
#include "kern_interp/kernel/kernel.h"
#include "kern_interp/linear_solve.h"
#include "kern_interp/quadtree/quadtree.h"

using namespace kern_interp;

namespace {
constexpr double kDelta = 1e-4;  // DELTA_X: offset of the observation point
constexpr double kIdTol = 1e-6;

double objective(const ki_Mat& sol) { return -sol.get(0, 0); }

// objective at four values of one hole angle, each on its own perturbed copy
void fin_diff_vals(const double* samples, Boundary* boundary, int param, int n,
                   const QuadTree& quadtree, const Kernel& kernel,
                   const std::vector<double>& pts, double* out) {
  QuadTree trees[4];
  std::unique_ptr<Boundary> bs[4];
  Kernel ks[4];
  for (int i = 0; i < 4; ++i) quadtree.copy_into(&trees[i]);
  for (int i = 0; i < 4; ++i) {
    boundary->perturbation_parameters[param] = samples[i];
    boundary->initialize(n, EX3B);
    bs[i] = boundary->clone();
    ks[i] = Kernel(kernel.solution_dimension, kernel.domain_dimension, kernel.pde, bs[i].get(), pts);
    trees[i].perturb(*bs[i]);
  }
#pragma omp parallel for num_threads(4)
  for (int i = 0; i < 4; ++i)
    out[i] = objective(boundary_integral_solve(ks[i], *bs[i], &trees[i], kIdTol, 1, pts));
}

// experiment_three.cpp:64-95: wrap both angles to [0, 2 pi] and push them
// apart symmetrically when closer than pi / 4 (on the circle)
void enforce_separation(double* a1, double* a2) {
  double c1 = *a1, c2 = *a2;
  while (c1 > 2 * M_PI) c1 -= 2 * M_PI;
  while (c1 < 0) c1 += 2 * M_PI;
  while (c2 > 2 * M_PI) c2 -= 2 * M_PI;
  while (c2 < 0) c2 += 2 * M_PI;
  double* lo = c1 < c2 ? &c1 : &c2;
  double* hi = c1 < c2 ? &c2 : &c1;
  const double inner = *hi - *lo, outer = *lo + 2 * M_PI - *hi;
  const double dist = std::min(inner, outer);
  if (dist < M_PI / 4.) {
    const double gap = (M_PI / 4.) - dist;
    if (inner < outer) {
      *hi += gap;
      *lo -= gap;
    } else {
      *hi -= gap;
      *lo += gap;
    }
  }
  *a1 = c1;
  *a2 = c2;
}
}  // namespace

int main(int argc, char** argv) {
  const int steps = argc > 1 ? std::atoi(argv[1]) : 30;
  const int n = argc > 2 ? std::atoi(argv[2]) : 1 << 12;
  const double h = 1e-4, alpha0 = 1, decay = 0.8;
  std::unique_ptr<Boundary> boundary(new Ex3Boundary());
  boundary->initialize(n, EX3B);
  double ang1 = 1.4, ang2 = 3;
  boundary->perturbation_parameters[0] = ang1;
  boundary->perturbation_parameters[1] = ang2;
  boundary->initialize(n, EX3B);
  QuadTree quadtree;
  quadtree.initialize_tree(boundary.get(), 2, 2);
  const std::vector<double> pts{0.5 - kDelta, 0.5 + kDelta};
  Kernel kernel(2, 2, Kernel::STOKES, boundary.get(), pts);
  double best = objective(boundary_integral_solve(kernel, *boundary, &quadtree, kIdTol, 1, pts));
  std::printf("start: theta1 %.6f theta2 %.6f objective %.9f\n", ang1, ang2, best);
  for (int step = 0; step < steps; ++step) {
    double f1[4], f2[4];
    const double s1[4] = {ang1 - 2 * h, ang1 - h, ang1 + h, ang1 + 2 * h};
    const double s2[4] = {ang2 - 2 * h, ang2 - h, ang2 + h, ang2 + 2 * h};
    fin_diff_vals(s1, boundary.get(), 0, n, quadtree, kernel, pts, f1);
    boundary->perturbation_parameters[0] = ang1;
    fin_diff_vals(s2, boundary.get(), 1, n, quadtree, kernel, pts, f2);
    const double g1 = (f1[0] - 8 * f1[1] + 8 * f1[2] - f1[3]) / (12 * h);
    const double g2 = (f2[0] - 8 * f2[1] + 8 * f2[2] - f2[3]) / (12 * h);
    double alpha = alpha0;
    for (; alpha > 0.01; alpha *= decay) {
      double t1 = ang1 + alpha * g1, t2 = ang2 + alpha * g2;
      enforce_separation(&t1, &t2);
      boundary->perturbation_parameters[0] = t1;
      boundary->perturbation_parameters[1] = t2;
      boundary->initialize(n, EX3B);
      quadtree.perturb(*boundary);
      kernel.update_boundary(boundary.get());
      const double f =
          objective(boundary_integral_solve(kernel, *boundary, &quadtree, kIdTol, 1, pts));
      if (best < f) {
        std::printf("theta1: %.6f theta2: %.6f objective: %.9f\n", t1, t2, f);
        std::fflush(stdout);
        best = f;
        ang1 = t1;
        ang2 = t2;
        break;
      }
    }
    if (alpha <= 0.01) {
      std::printf("Line search did not terminate\n");
      return 0;
    }
  }
  return 0;
}
