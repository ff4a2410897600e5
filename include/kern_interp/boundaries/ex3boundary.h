// Drop-in for kern_interp/boundaries/ex3boundary.h: experiment three's domain (reference ex3boundary.h:10-26).
#ifndef KERN_INTERP_BOUNDARIES_EX3BOUNDARY_H_
#define KERN_INTERP_BOUNDARIES_EX3BOUNDARY_H_

#include <memory>
#include <vector>

#include "kern_interp/boundaries/boundary.h"

namespace kern_interp {

class Ex3Boundary : public CubicBoundary {
 public:
  void initialize(int N, BoundaryCondition bc) override;
  void get_spline_points(std::vector<double>* x0_points, std::vector<double>* x1_points) override;
  void get_star_spline_points(double x, double y, std::vector<double>* star_x0_spline_points,
                              std::vector<double>* star_x1_spline_points);
  std::unique_ptr<Boundary> clone() const override {
    return std::unique_ptr<Boundary>(new Ex3Boundary(*this));
  }
};

}  // namespace kern_interp

#endif  // KERN_INTERP_BOUNDARIES_EX3BOUNDARY_H_
