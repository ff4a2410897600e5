// Drop-in for kern_interp/boundaries/boundary.h (reference boundary.h:12-51):
// the input contract of Kernel and QuadTree.
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
};

}  // namespace kern_interp

#endif  // KERN_INTERP_BOUNDARIES_BOUNDARY_H_
