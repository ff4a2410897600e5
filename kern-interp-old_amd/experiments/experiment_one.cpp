// Experiment one of the reference (experiments/experiment_one.cpp:19-61)
// through the drop-in API: Stokes flow in the Ex1 domain (a wavy square with
// two star and two circular holes), N = 2^12 boundary nodes, solved on a
// domain_size^2 grid over the tree's box, written for output/domain_plot.py.
//   experiment_one [out_dir=output/data] [N=4096] [domain_size=200]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "kern_interp/boundaries/ex1boundary.h"
#include "kern_interp/kernel/kernel.h"
#include "kern_interp/linear_solve.h"
#include "kern_interp/output_dump.h"
#include "kern_interp/quadtree/quadtree.h"

using namespace kern_interp;

int main(int argc, char** argv) {
  const std::string out = argc > 1 ? argv[1] : "output/data";
  const int n = argc > 2 ? std::atoi(argv[2]) : 1 << 12;
  const int domain_size = argc > 3 ? std::atoi(argv[3]) : 200;
  const double id_tol = 1e-6;
  std::unique_ptr<Boundary> boundary(new Ex1Boundary());
  boundary->initialize(n, DEFAULT);
  QuadTree quadtree;
  quadtree.initialize_tree(boundary.get(), 2, 2);
  std::vector<double> domain_points;
  get_domain_points(domain_size, &domain_points, quadtree.min, quadtree.max, quadtree.min,
                    quadtree.max);
  Kernel kernel(2, 2, Kernel::STOKES, boundary.get(), domain_points);
  ki_Mat solution =
      boundary_integral_solve(kernel, *boundary, &quadtree, id_tol, 4, domain_points);
  if (!write_solution_file(out + "/ie_solver_solution.txt", domain_points, solution, 2) ||
      !write_boundary_file(out + "/ie_solver_boundary.txt", *boundary)) {
    std::fprintf(stderr, "cannot write to %s\n", out.c_str());
    return 1;
  }
  std::printf("experiment one: %zu boundary nodes, %zu domain points -> %s\n",
              boundary->weights.size(), domain_points.size() / 2, out.c_str());
  return 0;
}
