// Vec2, Boundary and the test/workload geometries of the drop-in API
// (interfaces of kern_interp/vec2.h, boundaries/boundary.h:12-51 and the
// geometries of boundaries/{circle,annulus,donut}.cpp).
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <stdexcept>

#include "kern_interp/boundaries/annulus.h"
#include "kern_interp/boundaries/boundary.h"
#include "kern_interp/boundaries/circle.h"
#include "kern_interp/boundaries/donut.h"
#include "kern_interp/boundaries/star.h"
#include "kern_interp/vec2.h"

namespace kern_interp {

Vec2::Vec2() { a[0] = a[1] = 0; }
Vec2::Vec2(double m, double n) {
  a[0] = m;
  a[1] = n;
}
double Vec2::norm() const { return std::sqrt(a[0] * a[0] + a[1] * a[1]); }
double Vec2::dot(const Vec2& o) const { return a[0] * o.a[0] + a[1] * o.a[1]; }
Vec2 Vec2::operator-(const Vec2& o) const { return Vec2(a[0] - o.a[0], a[1] - o.a[1]); }
Vec2 Vec2::operator*(const double d) const { return Vec2(a[0] * d, a[1] * d); }

void Boundary::set_boundary_values_size(BoundaryCondition bc) {
  const int np = weights.size();
  switch (bc) {
    case SINGLE_ELECTRON: case EX3A: case ALL_ONES: case ALL_NEG_ONES: case ALL_ZEROS:
      boundary_values = ki_Mat(np, 1);
      break;
    case DEFAULT:
      throw std::invalid_argument("DEFAULT boundary condition has no size of its own");
    default:
      boundary_values = ki_Mat(2 * np, 1);
  }
}

// boundary.cpp:43-120 of the reference: the same data per condition
void Boundary::apply_boundary_condition(int s, int e, BoundaryCondition bc) {
  for (int p = s; p < e; ++p) {
    const double x = points[2 * p], y = points[2 * p + 1];
    const double nx = normals[2 * p], ny = normals[2 * p + 1];
    switch (bc) {
      case SINGLE_ELECTRON:
        boundary_values.set(p, 0, std::log(std::sqrt((x + 3) * (x + 3) + (y + 2) * (y + 2))) /
                                      (2 * M_PI));
        break;
      case ALL_ONES: boundary_values.set(p, 0, 1.); break;
      case ALL_NEG_ONES: boundary_values.set(p, 0, -1.); break;
      case ALL_ZEROS: boundary_values.set(p, 0, 0.); break;
      case TANGENT_VEC:
        boundary_values.set(2 * p, 0, -ny);
        boundary_values.set(2 * p + 1, 0, nx);
        break;
      case REVERSE_TANGENT_VEC:
        boundary_values.set(2 * p, 0, ny);
        boundary_values.set(2 * p + 1, 0, -nx);
        break;
      case NORMAL_VEC:
        boundary_values.set(2 * p, 0, nx);
        boundary_values.set(2 * p + 1, 0, ny);
        break;
      case REVERSE_NORMAL_VEC:
        boundary_values.set(2 * p, 0, -nx);
        boundary_values.set(2 * p + 1, 0, -ny);
        break;
      case HORIZONTAL_VEC:
        boundary_values.set(2 * p, 0, 1.);
        boundary_values.set(2 * p + 1, 0, 0.);
        break;
      case LEFT_TO_RIGHT_FLOW: {
        const bool side = x < -0.99 || x > 1.99;
        boundary_values.set(2 * p, 0, side ? 1. : 0.);
        boundary_values.set(2 * p + 1, 0, 0.);
        break;
      }
      case NO_SLIP:
        boundary_values.set(2 * p, 0, 0.);
        boundary_values.set(2 * p + 1, 0, 0.);
        break;
      default:
        throw std::invalid_argument("boundary condition not applicable node by node");
    }
  }
}

// the disc rules of circle.cpp:34-42, annulus.cpp:66-78, donut.cpp:67-79:
// outside when |a - c| + 1e-2 > R or |a - h| - 1e-2 < r_h for a hole
static double dist_to(const Vec2& a, double cx, double cy) {
  const double x = a.a[0] - cx, y = a.a[1] - cy;
  return std::sqrt(x * x + y * y);
}

static bool in_disc(const Vec2& a, double cx, double cy, double R, const std::vector<Hole>& holes) {
  if (dist_to(a, cx, cy) + 1e-2 > R) return false;
  for (const Hole& h : holes)
    if ((a - h.center).norm() - 1e-2 < h.radius) return false;
  return true;
}

static DomainRule disc_rule(double cx, double cy, double R, const std::vector<Hole>& holes) {
  DomainRule r;
  r.kind = DomainRule::kDisc;
  r.cx = cx;
  r.cy = cy;
  r.radius = R;
  for (const Hole& h : holes) r.holes.insert(r.holes.end(), {h.center.a[0], h.center.a[1], h.radius});
  return r;
}

static void push_ring(Boundary* b, double cx, double cy, double rad, int n, bool inward) {
  for (int i = 0; i < n; ++i) {
    const double ang = i * 2.0 * M_PI / n;
    const double c = std::cos(ang), s = std::sin(ang);
    b->points.push_back(cx + rad * c);
    b->points.push_back(cy + rad * s);
    b->normals.push_back(inward ? -c : c);
    b->normals.push_back(inward ? -s : s);
    b->curvatures.push_back(inward ? -1.0 / rad : 1.0 / rad);
    b->weights.push_back(2 * rad * M_PI / n);
  }
}

static void clear_arrays(Boundary* b) {
  b->points.clear();
  b->normals.clear();
  b->weights.clear();
  b->curvatures.clear();
}

void Circle::initialize(int n, BoundaryCondition bc) {
  clear_arrays(this);
  push_ring(this, 0.5, 0.5, 0.25, n, false);
  if (bc == DEFAULT) {
    boundary_values = ki_Mat(n, 1);
    apply_boundary_condition(0, n, SINGLE_ELECTRON);
  } else {
    set_boundary_values_size(bc);
    apply_boundary_condition(0, n, bc);
  }
}

// circle.cpp:34-42: outside when |a - (0.5, 0.5)| + 1e-2 > 0.25
bool Circle::is_in_domain(const Vec2& a) const {
  return !(dist_to(a, 0.5, 0.5) + 1e-2 > 0.25);
}

DomainRule Circle::domain_rule() const { return disc_rule(0.5, 0.5, 0.25, {}); }

static void ring_with_holes(Boundary* b, int n, double cx, double cy, double rad) {
  clear_arrays(b);
  push_ring(b, cx, cy, rad, n, false);
  for (auto& h : b->holes) push_ring(b, h.center.a[0], h.center.a[1], h.radius, h.num_nodes, true);
}

void Annulus::initialize(int n, BoundaryCondition bc) {
  if (holes.empty()) {
    holes.push_back(Hole{Vec2(0.9, 0.5), 0.2, n / 5});
    holes.push_back(Hole{Vec2(0.1, 0.5), 0.2, n / 5});
  }
  for (auto& h : holes) h.num_nodes = n / 5;
  ring_with_holes(this, n, 0.5, 0.5, 1.0);
  const int hn = n / 5, np = weights.size();
  if (bc == DEFAULT) {
    boundary_values = ki_Mat(np, 1);
    apply_boundary_condition(0, n, ALL_ZEROS);
    apply_boundary_condition(n, n + hn, ALL_ONES);
    apply_boundary_condition(n + hn, n + 2 * hn, ALL_NEG_ONES);
  } else {
    set_boundary_values_size(bc);
    apply_boundary_condition(0, np, bc);
  }
}

// annulus.cpp:66-78
bool Annulus::is_in_domain(const Vec2& a) const { return in_disc(a, 0.5, 0.5, 1.0, holes); }

DomainRule Annulus::domain_rule() const { return disc_rule(0.5, 0.5, 1.0, holes); }

void Donut::initialize(int n, BoundaryCondition bc) {
  if (holes.empty()) holes.push_back(Hole{Vec2(0.5, 0.5), 0.5, n / 5});
  for (auto& h : holes) h.num_nodes = n / 5;
  ring_with_holes(this, n, 0.5, 0.5, 1.0);
  const int np = weights.size();
  if (bc == DEFAULT) {
    boundary_values = ki_Mat(np, 1);
    for (int i = 0; i < n; ++i) boundary_values.set(i, 0, 1);
    for (int i = n; i < np; ++i) boundary_values.set(i, 0, -1. / holes[0].radius);
  } else {
    set_boundary_values_size(bc);
    apply_boundary_condition(0, np, bc);
  }
}

// donut.cpp:67-79
bool Donut::is_in_domain(const Vec2& a) const { return in_disc(a, 0.5, 0.5, 1.0, holes); }

DomainRule Donut::domain_rule() const { return disc_rule(0.5, 0.5, 1.0, holes); }

double StarCurve::radius(double t) const {
  double r = 1.0;
  for (size_t k = 0; k < amps.size(); ++k) r += amps[k] * std::cos(freqs[k] * t + phases[k]);
  return r;
}

void StarCurve::initialize(int n, BoundaryCondition bc) {
  clear_arrays(this);
  for (int j = 0; j < n; ++j) {
    const double t = 2 * M_PI * j / n;
    double r = 1.0, dr = 0.0, ddr = 0.0;
    for (size_t k = 0; k < amps.size(); ++k) {
      const double arg = freqs[k] * t + phases[k];
      r += amps[k] * std::cos(arg);
      dr -= amps[k] * freqs[k] * std::sin(arg);
      ddr -= amps[k] * freqs[k] * freqs[k] * std::cos(arg);
    }
    const double c = std::cos(t), s = std::sin(t);
    const double dx = dr * c - r * s, dy = dr * s + r * c;
    const double ddx = ddr * c - 2 * dr * s - r * c, ddy = ddr * s + 2 * dr * c - r * s;
    const double sp = std::sqrt(dx * dx + dy * dy);
    points.push_back(r * c);
    points.push_back(r * s);
    normals.push_back(dy / sp);
    normals.push_back(-dx / sp);
    weights.push_back(sp * 2 * M_PI / n);
    curvatures.push_back((dx * ddy - ddx * dy) / (sp * sp * sp));
  }
  if (bc == DEFAULT) {
    boundary_values = ki_Mat(n, 1);
    for (int p = 0; p < n; ++p) {
      const double ex = points[2 * p] - source.a[0], ey = points[2 * p + 1] - source.a[1];
      boundary_values.set(p, 0, std::log(std::sqrt(ex * ex + ey * ey)) / (2 * M_PI));
    }
  } else {
    set_boundary_values_size(bc);
    apply_boundary_condition(0, n, bc);
  }
}

bool StarCurve::is_in_domain(const Vec2& a) const {
  const double t = std::atan2(a.a[1], a.a[0]);
  return a.norm() + 1e-2 < radius(t);
}

DomainRule StarCurve::domain_rule() const {
  DomainRule r;
  r.kind = DomainRule::kStar;
  r.amps = amps;
  r.freqs = freqs;
  r.phases = phases;
  return r;
}

void HoledDisc::initialize(int n, BoundaryCondition bc) {
  for (auto& h : holes) h.num_nodes = hole_nodes;
  ring_with_holes(this, n, 0.0, 0.0, 1.0);
  if (bc == DEFAULT) throw std::invalid_argument("HoledDisc has no default boundary data");
  set_boundary_values_size(bc);
  apply_boundary_condition(0, weights.size(), bc);
}

bool HoledDisc::is_in_domain(const Vec2& a) const { return in_disc(a, 0.0, 0.0, 1.0, holes); }

DomainRule HoledDisc::domain_rule() const { return disc_rule(0.0, 0.0, 1.0, holes); }

}  // namespace kern_interp
