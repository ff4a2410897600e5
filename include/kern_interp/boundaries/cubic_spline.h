// Drop-in for kern_interp/boundaries/cubic_spline.h: the starfish of 20 spline knots (reference cubic_spline.h:10-22).
#ifndef KERN_INTERP_BOUNDARIES_CUBIC_SPLINE_H_
#define KERN_INTERP_BOUNDARIES_CUBIC_SPLINE_H_

#include <memory>
#include <vector>

#include "kern_interp/boundaries/boundary.h"

namespace kern_interp {

class CubicSpline : public CubicBoundary {
 public:
  void initialize(int N, BoundaryCondition bc) override;
  void get_spline_points(std::vector<double>* x0_points, std::vector<double>* x1_points) override;
  std::unique_ptr<Boundary> clone() const override {
    return std::unique_ptr<Boundary>(new CubicSpline(*this));
  }
};

}  // namespace kern_interp

#endif  // KERN_INTERP_BOUNDARIES_CUBIC_SPLINE_H_
