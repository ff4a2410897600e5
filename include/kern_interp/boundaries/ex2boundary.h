// Drop-in for kern_interp/boundaries/ex2boundary.h: experiment two's channel (reference ex2boundary.h:10-22).
#ifndef KERN_INTERP_BOUNDARIES_EX2BOUNDARY_H_
#define KERN_INTERP_BOUNDARIES_EX2BOUNDARY_H_

#include <memory>
#include <vector>

#include "kern_interp/boundaries/boundary.h"

namespace kern_interp {

class Ex2Boundary : public CubicBoundary {
 public:
  void initialize(int N, BoundaryCondition bc) override;
  void get_spline_points(std::vector<double>* x0_points, std::vector<double>* x1_points) override;
  std::unique_ptr<Boundary> clone() const override {
    return std::unique_ptr<Boundary>(new Ex2Boundary(*this));
  }
};

}  // namespace kern_interp

#endif  // KERN_INTERP_BOUNDARIES_EX2BOUNDARY_H_
