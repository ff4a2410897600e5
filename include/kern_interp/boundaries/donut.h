// Drop-in for kern_interp/boundaries/donut.h: the reference's donut test geometry.
#ifndef KERN_INTERP_BOUNDARIES_DONUT_H_
#define KERN_INTERP_BOUNDARIES_DONUT_H_

#include <memory>

#include "kern_interp/boundaries/boundary.h"

namespace kern_interp {

class Donut : public Boundary {
 public:
  void initialize(int n, BoundaryCondition bc) override;
  bool is_in_domain(const Vec2& a) const override;
  DomainRule domain_rule() const override;
  std::unique_ptr<Boundary> clone() const override {
    return std::unique_ptr<Boundary>(new Donut(*this));
  }
};

}  // namespace kern_interp

#endif  // KERN_INTERP_BOUNDARIES_DONUT_H_
