// Drop-in for kern_interp/boundaries/boundary.h (reference boundary.h:12-70):
// the input contract of Kernel and QuadTree, and the cubic-spline curves.
#ifndef KERN_INTERP_BOUNDARIES_BOUNDARY_H_
#define KERN_INTERP_BOUNDARIES_BOUNDARY_H_

#include <memory>
#include <vector>

#include "kern_interp/ki_mat.h"
#include "kern_interp/vec2.h"

namespace kern_interp {

struct Hole {
  Vec2 center;
  double radius;
  int num_nodes;
};

enum BoundaryCondition {
  SINGLE_ELECTRON,
  ALL_ONES,
  ALL_NEG_ONES,
  ALL_ZEROS,
  TANGENT_VEC,
  REVERSE_TANGENT_VEC,
  NORMAL_VEC,
  REVERSE_NORMAL_VEC,
  LEFT_TO_RIGHT_FLOW,
  NO_SLIP,
  HORIZONTAL_VEC,
  EX3A,
  EX3B,
  DEFAULT
};

// is_in_domain as data, for the GPU's in-domain kernel (no counterpart in the
// reference).  kind: kDisc (outside when |t - c| + eps > radius or
// |t - h| - eps < r_h for a hole), kWinding (the crossing count of the node
// polygon, outer nodes first, then hole_nodes[k] nodes per hole), kStar (inside
// when |t| + eps < 1 + sum a_k cos(f_k atan2(t) + phi_k)), kNone (no device
// form: the masking calls is_in_domain on the host).
struct DomainRule {
  enum Kind { kNone = -1, kDisc = 0, kWinding = 1, kStar = 2 };
  int kind = kNone;
  double eps = 1e-2, cx = 0, cy = 0, radius = 0;
  std::vector<double> holes;                // kDisc: x, y, r per hole
  std::vector<int> hole_nodes;              // kWinding: nodes per hole
  std::vector<double> amps, freqs, phases;  // kStar
};

class Boundary {
 public:
  int num_outer_nodes = -1;
  std::vector<double> perturbation_parameters;
  std::vector<double> points, normals, curvatures, weights;
  std::vector<Hole> holes;
  ki_Mat boundary_values;

  virtual std::unique_ptr<Boundary> clone() const = 0;
  virtual ~Boundary() {}
  void set_boundary_values_size(BoundaryCondition bc);
  void apply_boundary_condition(int start_idx, int end_idx, BoundaryCondition bc);
  virtual void initialize(int n, BoundaryCondition bc) = 0;
  virtual bool is_in_domain(const Vec2& a) const = 0;
  virtual DomainRule domain_rule() const { return DomainRule(); }
};

// Closed cubic-spline curves (reference boundary.h:54-70,
// cubic_boundary.cpp:10-230).  The outer nodes come first in the arrays.
class CubicBoundary : public Boundary {
 public:
  virtual void get_spline_points(std::vector<double>* outer_x0_points,
                                 std::vector<double>* outer_x1_points) = 0;
  void get_cubics(const std::vector<double>& x0_points, const std::vector<double>& x1_points,
                  std::vector<std::vector<double>>* x0_cubics,
                  std::vector<std::vector<double>>* x1_cubics);
  void interpolate(bool is_interior, int nodes_per_spline,
                   const std::vector<std::vector<double>>& x0_cubics,
                   const std::vector<std::vector<double>>& x1_cubics);
  void find_real_roots_of_cubic(const std::vector<double>& y_cubic, std::vector<double>* t_vals);
  bool is_in_domain(const Vec2& a) const override;
  DomainRule domain_rule() const override;
};

}  // namespace kern_interp

#endif  // KERN_INTERP_BOUNDARIES_BOUNDARY_H_
