// Host side of the drop-in boundaries, for tests/test_cpp_boundaries.py (no
// GPU call): writes a geometry's arrays, answers is_in_domain for targets, or
// the real eigenvalues of a matrix, as raw binary on stdout.
//   boundary_dump geom <name> <N> <bc> [perturbation parameters...]
//   boundary_dump domain <name> <N> <bc>   (stdin: int n, 2n doubles)
//   boundary_dump eig                       (stdin: int n, n*n doubles column-major)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "kern_interp/boundaries/annulus.h"
#include "kern_interp/boundaries/circle.h"
#include "kern_interp/boundaries/cubic_spline.h"
#include "kern_interp/boundaries/donut.h"
#include "kern_interp/boundaries/ex1boundary.h"
#include "kern_interp/boundaries/ex2boundary.h"
#include "kern_interp/boundaries/ex3boundary.h"

using namespace kern_interp;

static std::unique_ptr<Boundary> make(const std::string& name) {
  if (name == "circle") return std::unique_ptr<Boundary>(new Circle());
  if (name == "annulus") return std::unique_ptr<Boundary>(new Annulus());
  if (name == "donut") return std::unique_ptr<Boundary>(new Donut());
  if (name == "starfish") return std::unique_ptr<Boundary>(new CubicSpline());
  if (name == "ex1") return std::unique_ptr<Boundary>(new Ex1Boundary());
  if (name == "ex2") return std::unique_ptr<Boundary>(new Ex2Boundary());
  if (name == "ex3") return std::unique_ptr<Boundary>(new Ex3Boundary());
  std::fprintf(stderr, "unknown geometry %s\n", name.c_str());
  std::exit(2);
}

static BoundaryCondition bc_of(const std::string& s) {
  static const char* names[] = {"SINGLE_ELECTRON", "ALL_ONES", "ALL_NEG_ONES", "ALL_ZEROS",
                                "TANGENT_VEC", "REVERSE_TANGENT_VEC", "NORMAL_VEC",
                                "REVERSE_NORMAL_VEC", "LEFT_TO_RIGHT_FLOW", "NO_SLIP",
                                "HORIZONTAL_VEC", "EX3A", "EX3B", "DEFAULT"};
  for (int i = 0; i < 14; ++i)
    if (s == names[i]) return (BoundaryCondition)i;
  std::fprintf(stderr, "unknown boundary condition %s\n", s.c_str());
  std::exit(2);
}

template <class T>
static void put(const T* p, size_t n) {
  std::fwrite(p, sizeof(T), n, stdout);
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string mode = argv[1];
  if (mode == "eig") {
    int n;
    if (std::fread(&n, sizeof n, 1, stdin) != 1) return 3;
    ki_Mat A(n, n);
    if (std::fread(A.mat, sizeof(double), (size_t)n * n, stdin) != (size_t)n * n) return 3;
    std::vector<double> e = A.real_eigenvalues();
    const int m = e.size();
    put(&m, 1);
    put(e.data(), m);
    return 0;
  }
  if (argc < 5) return 2;
  std::unique_ptr<Boundary> b = make(argv[2]);
  for (int i = 5; i < argc; ++i) b->perturbation_parameters.push_back(std::atof(argv[i]));
  b->initialize(std::atoi(argv[3]), bc_of(argv[4]));
  if (mode == "geom") {
    const int np = b->weights.size(), nh = b->holes.size(), nb = b->boundary_values.height();
    const int head[4] = {np, b->num_outer_nodes, nh, nb};
    put(head, 4);
    for (const Hole& h : b->holes) {
      const double d[3] = {h.center.a[0], h.center.a[1], h.radius};
      put(d, 3);
      put(&h.num_nodes, 1);
    }
    put(b->points.data(), 2 * np);
    put(b->normals.data(), 2 * np);
    put(b->weights.data(), np);
    put(b->curvatures.data(), np);
    put(b->boundary_values.mat, nb);
    return 0;
  }
  if (mode == "domain") {
    int n;
    if (std::fread(&n, sizeof n, 1, stdin) != 1) return 3;
    std::vector<double> t(2 * (size_t)n);
    if (std::fread(t.data(), sizeof(double), t.size(), stdin) != t.size()) return 3;
    std::vector<unsigned char> in(n);
    for (int i = 0; i < n; ++i) in[i] = b->is_in_domain(Vec2(t[2 * i], t[2 * i + 1]));
    put(in.data(), n);
    return 0;
  }
  return 2;
}
