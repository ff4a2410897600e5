// Drop-in for kern_interp/boundaries/annulus.h: the reference's annulus test geometry.
#ifndef KERN_INTERP_BOUNDARIES_ANNULUS_H_
#define KERN_INTERP_BOUNDARIES_ANNULUS_H_

#include <memory>

#include "kern_interp/boundaries/boundary.h"

namespace kern_interp {

class Annulus : public Boundary {
 public:
  void initialize(int n, BoundaryCondition bc) override;
  bool is_in_domain(const Vec2& a) const override;
  DomainRule domain_rule() const override;
  std::unique_ptr<Boundary> clone() const override {
    return std::unique_ptr<Boundary>(new Annulus(*this));
  }
};

}  // namespace kern_interp

#endif  // KERN_INTERP_BOUNDARIES_ANNULUS_H_
