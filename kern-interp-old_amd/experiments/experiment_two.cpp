// Experiment two of the reference (experiments/experiment_two.cpp:18-81)
// through the drop-in API: Stokes channel flow past eight circular holes;
// each frame moves one hole to a random height, updates the tree with
// QuadTree::perturb and solves again on a domain_size^2 grid over the nodes'
// bounding box.  Frames are written for output/animation.py.
//   experiment_two [bake_dir=output/bake] [frames=100] [N=4096] [domain_size=350]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "kern_interp/boundaries/ex2boundary.h"
#include "kern_interp/kernel/kernel.h"
#include "kern_interp/linear_solve.h"
#include "kern_interp/output_dump.h"
#include "kern_interp/quadtree/quadtree.h"

using namespace kern_interp;

int main(int argc, char** argv) {
  const std::string bake = argc > 1 ? argv[1] : "output/bake";
  const int frames = argc > 2 ? std::atoi(argv[2]) : 100;
  const int n = argc > 3 ? std::atoi(argv[3]) : 1 << 12;
  const int domain_size = argc > 4 ? std::atoi(argv[4]) : 350;
  const double id_tol = 1e-6;
  srand(0);
  std::unique_ptr<Boundary> boundary(new Ex2Boundary());
  boundary->initialize(n, DEFAULT);
  QuadTree quadtree;
  quadtree.initialize_tree(boundary.get(), 2, 2);
  double x0 = boundary->points[0], x1 = x0, y0 = boundary->points[1], y1 = y0;
  for (size_t i = 0; i < boundary->points.size(); i += 2) {
    x0 = std::min(x0, boundary->points[i]);
    x1 = std::max(x1, boundary->points[i]);
    y0 = std::min(y0, boundary->points[i + 1]);
    y1 = std::max(y1, boundary->points[i + 1]);
  }
  std::vector<double> domain_points;
  get_domain_points(domain_size, &domain_points, x0, x1, y0, y1);
  Kernel kernel(2, 2, Kernel::STOKES, boundary.get(), domain_points);
  ki_Mat solution =
      boundary_integral_solve(kernel, *boundary, &quadtree, id_tol, 4, domain_points);
  if (!write_frame(bake, 0, domain_points, solution, 2, *boundary)) {
    std::fprintf(stderr, "cannot write to %s/{sol,boundary}\n", bake.c_str());
    return 1;
  }
  for (int frame = 0; frame < frames; ++frame) {
    const int k = (int)std::floor(8 * (rand() / (0. + RAND_MAX)));
    boundary->perturbation_parameters[k] = 0.35 + 0.3 * (rand() / (0. + RAND_MAX));
    boundary->initialize(n, DEFAULT);
    quadtree.perturb(*boundary);
    kernel.update_boundary(boundary.get());
    solution = boundary_integral_solve(kernel, *boundary, &quadtree, id_tol, 4, domain_points);
    if (!write_frame(bake, frame + 1, domain_points, solution, 2, *boundary)) return 1;
    std::printf("frame %d: hole %d at y = %.4f\n", frame + 1, k,
                boundary->perturbation_parameters[k]);
    std::fflush(stdout);
  }
  return 0;
}
