// Workload geometries without a reference counterpart (new Boundary
// subclasses): the trapezoidal star r(t) = 1 + sum_k a_k cos(f_k t + phi_k)
// and the unit disc with circular holes.
#ifndef KERN_INTERP_BOUNDARIES_STAR_H_
#define KERN_INTERP_BOUNDARIES_STAR_H_

#include <memory>
#include <vector>

#include "kern_interp/boundaries/boundary.h"

namespace kern_interp {

class StarCurve : public Boundary {
 public:
  // default: r = 1 + 0.3 cos(5 t)
  std::vector<double> amps{0.3}, freqs{5.0}, phases{0.0};
  // point-source location of SINGLE_ELECTRON-style data on this curve
  Vec2 source = Vec2(3.0, 3.0);
  void initialize(int n, BoundaryCondition bc) override;
  bool is_in_domain(const Vec2& a) const override;
  DomainRule domain_rule() const override;
  double radius(double t) const;
  std::unique_ptr<Boundary> clone() const override {
    return std::unique_ptr<Boundary>(new StarCurve(*this));
  }
};

class HoledDisc : public Boundary {
 public:
  // outer unit circle at the origin; holes must be set before initialize()
  int hole_nodes = 256;
  void initialize(int n, BoundaryCondition bc) override;
  bool is_in_domain(const Vec2& a) const override;
  DomainRule domain_rule() const override;
  std::unique_ptr<Boundary> clone() const override {
    return std::unique_ptr<Boundary>(new HoledDisc(*this));
  }
};

}  // namespace kern_interp

#endif  // KERN_INTERP_BOUNDARIES_STAR_H_
