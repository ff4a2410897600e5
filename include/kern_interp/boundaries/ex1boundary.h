// Drop-in for kern_interp/boundaries/ex1boundary.h: experiment one's domain (reference ex1boundary.h:10-25).
#ifndef KERN_INTERP_BOUNDARIES_EX1BOUNDARY_H_
#define KERN_INTERP_BOUNDARIES_EX1BOUNDARY_H_

#include <memory>
#include <vector>

#include "kern_interp/boundaries/boundary.h"

namespace kern_interp {

class Ex1Boundary : public CubicBoundary {
 public:
  void initialize(int N, BoundaryCondition bc) override;
  void get_spline_points(std::vector<double>* x0_points, std::vector<double>* x1_points) override;
  void get_star_spline_points(double x, double y, std::vector<double>* star_x0_spline_points,
                              std::vector<double>* star_x1_spline_points);
  std::unique_ptr<Boundary> clone() const override {
    return std::unique_ptr<Boundary>(new Ex1Boundary(*this));
  }
};

}  // namespace kern_interp

#endif  // KERN_INTERP_BOUNDARIES_EX1BOUNDARY_H_
