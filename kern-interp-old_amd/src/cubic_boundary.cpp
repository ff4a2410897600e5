// Cubic-spline boundaries of the drop-in API: CubicBoundary
// (cubic_boundary.cpp:10-230), the starfish CubicSpline (cubic_spline.cpp:13-50)
// and the experiment domains Ex1Boundary / Ex2Boundary / Ex3Boundary
// (ex1boundary.cpp, ex2boundary.cpp, ex3boundary.cpp).  Host generators of the
// input arrays; the in-domain rule also has a device form (domain_rule).
#include <cmath>
#include <stdexcept>

#include "kern_interp/boundaries/cubic_spline.h"
#include "kern_interp/boundaries/ex1boundary.h"
#include "kern_interp/boundaries/ex2boundary.h"
#include "kern_interp/boundaries/ex3boundary.h"

namespace kern_interp {

namespace {
constexpr int kQuadSize = 10;       // polyline samples per node for the weights
constexpr int kStarfishKnots = 20;  // NUM_SPLINE_POINTS of cubic_spline.cpp

double cubic_at(const std::vector<double>& c, double t) {
  return c[0] + t * c[1] + std::pow(t, 2) * c[2] + std::pow(t, 3) * c[3];
}

// 20 knots of c + scale (cos a, sin a)(sin 5a + 4)
void star_knots(double cx, double cy, double scale, std::vector<double>* xs,
                std::vector<double>* ys) {
  for (int i = 0; i < kStarfishKnots; ++i) {
    const double ang = 2 * M_PI * (i / (kStarfishKnots + 0.));
    xs->push_back(cx + scale * std::cos(ang) * (std::sin(5 * ang) + 4));
    ys->push_back(cy + scale * std::sin(ang) * (std::sin(5 * ang) + 4));
  }
}

void clear_arrays(Boundary* b) {
  b->points.clear();
  b->normals.clear();
  b->weights.clear();
  b->curvatures.clear();
  b->holes.clear();
}

// the nodes of a circular hole: inward normals, curvature -1/r (ex1boundary.cpp:128-147)
void circle_hole(Boundary* b, const Hole& h) {
  const int n = h.num_nodes;
  for (int i = 0; i < n; ++i) {
    const double ang = (2.0 * M_PI * i) / n;
    b->points.push_back(h.center.a[0] + h.radius * std::cos(ang));
    b->points.push_back(h.center.a[1] + h.radius * std::sin(ang));
    b->normals.push_back(-std::cos(ang));
    b->normals.push_back(-std::sin(ang));
    b->curvatures.push_back(-1.0 / h.radius);
    b->weights.push_back(2.0 * M_PI * h.radius / n);
  }
}

void spline_curve(CubicBoundary* b, bool interior, int nps, const std::vector<double>& xs,
                  const std::vector<double>& ys) {
  std::vector<std::vector<double>> xc, yc;
  b->get_cubics(xs, ys, &xc, &yc);
  b->interpolate(interior, nps, xc, yc);
}
}  // namespace

// cubic_boundary.cpp:10-77: periodic C2 cubics through the knots.  The knot
// derivatives d solve the cyclic system 4 d_i + d_{i-1} + d_{i+1} =
// 3 (p_{i+1} - p_{i-1}); interval i is p_i + d_i t + c2 t^2 + c3 t^3.
void CubicBoundary::get_cubics(const std::vector<double>& x0, const std::vector<double>& x1,
                               std::vector<std::vector<double>>* x0_cubics,
                               std::vector<std::vector<double>>* x1_cubics) {
  const int n = x0.size();
  ki_Mat A(n, n);
  for (int i = 0; i < n; ++i) {
    A.set(i, i, 4.0);
    A.set(i, (i + 1) % n, 1.0);
    A.set(i, (i + n - 1) % n, 1.0);
  }
  const std::vector<double>* pts[2] = {&x0, &x1};
  std::vector<std::vector<double>>* out[2] = {x0_cubics, x1_cubics};
  for (int c = 0; c < 2; ++c) {
    const std::vector<double>& p = *pts[c];
    ki_Mat rhs(n, 1), d;
    for (int i = 0; i < n; ++i) rhs.set(i, 0, 3 * (p[(i + 1) % n] - p[(i + n - 1) % n]));
    A.left_multiply_inverse(rhs, &d);
    for (int i = 0; i < n; ++i) {
      const int j = (i + 1) % n;
      const double di = d.get(i, 0), dj = d.get(j, 0);
      out[c]->push_back({p[i], di, 3 * (p[j] - p[i]) - 2 * di - dj, 2 * (p[i] - p[j]) + di + dj});
    }
  }
}

// cubic_boundary.cpp:80-89: the real roots of c0 + c1 t + c2 t^2 + c3 t^3 as
// the real eigenvalues of its companion matrix
void CubicBoundary::find_real_roots_of_cubic(const std::vector<double>& c,
                                             std::vector<double>* t_vals) {
  ki_Mat comp(3, 3);
  comp.set(1, 0, 1.0);
  comp.set(2, 1, 1.0);
  comp.set(0, 2, -c[0] / c[3]);
  comp.set(1, 2, -c[1] / c[3]);
  comp.set(2, 2, -c[2] / c[3]);
  *t_vals = comp.real_eigenvalues();
}

// cubic_boundary.cpp:90-178: nodes at t = j / nps of every interval; unit
// normal (y', -x') / |p'| and curvature (x'y'' - x''y') / |p'|^3, both negated
// for a hole; the weight of node i is the mean of the polyline lengths (10
// samples each) of the node's own piece and the previous one, cyclically.
void CubicBoundary::interpolate(bool interior, int nps,
                                const std::vector<std::vector<double>>& xc,
                                const std::vector<std::vector<double>>& yc) {
  const int nsp = xc.size(), nn = nsp * nps;
  std::vector<double> quad;
  quad.reserve(2 * (size_t)kQuadSize * nn);
  for (int i = 0; i < nsp; ++i) {
    const std::vector<double>& cx = xc[i];
    const std::vector<double>& cy = yc[i];
    for (int j = 0; j < nps; ++j) {
      const double t = j / (nps + 0.0);
      points.push_back(cubic_at(cx, t));
      points.push_back(cubic_at(cy, t));
      for (int q = 0; q < kQuadSize; ++q) {
        const double tq = ((j * kQuadSize) + q) / (0.0 + kQuadSize * nps);
        quad.push_back(cubic_at(cx, tq));
        quad.push_back(cubic_at(cy, tq));
      }
      double xp = cx[1] + 2 * t * cx[2] + 3 * std::pow(t, 2) * cx[3];
      double yp = cy[1] + 2 * t * cy[2] + 3 * std::pow(t, 2) * cy[3];
      const double xpp = 2 * cx[2] + 6 * t * cx[3];
      const double ypp = 2 * cy[2] + 6 * t * cy[3];
      const double sp = std::sqrt(std::pow(xp, 2) + std::pow(yp, 2));
      const double kap = (xp * ypp - xpp * yp) / std::pow(sp, 3);
      xp /= sp;
      yp /= sp;
      curvatures.push_back(interior ? -kap : kap);
      normals.push_back(interior ? -yp : yp);
      normals.push_back(interior ? xp : -xp);
    }
  }
  const size_t L = quad.size();
  std::vector<double> piece(nn);
  for (int i = 0; i < nn; ++i) {
    double s = 0;
    for (int j = 0; j < kQuadSize; ++j) {
      const size_t idx = 2 * ((size_t)kQuadSize * i + j);
      s += std::sqrt(std::pow(quad[idx] - quad[(idx + 2) % L], 2) +
                     std::pow(quad[idx + 1] - quad[(idx + 3) % L], 2));
    }
    piece[i] = s;
  }
  weights.push_back((piece[nn - 1] + piece[0]) / 2.);
  for (int i = 1; i < nn; ++i) weights.push_back((piece[i - 1] + piece[i]) / 2.);
}

// cubic_boundary.cpp:182-230: per node section (outer, then each hole) the
// count of edges from a node right of the target that cross its horizontal,
// +1 upward, -1 downward; a node closer than eps rejects.  The outer count
// must be odd; a hole whose count is 1 (C remainder) rejects.
bool CubicBoundary::is_in_domain(const Vec2& a) const {
  const double eps = 1e-2, vx = a.a[0], vy = a.a[1];
  int start = 0;
  for (int sec = -1; sec < (int)holes.size(); ++sec) {
    const int n = sec < 0 ? num_outer_nodes : holes[sec].num_nodes;
    int wind = 0;
    for (int k = 0; k < n; ++k) {
      const int i = 2 * (start + k), nx = 2 * (start + (k + 1) % n);
      const double dx = vx - points[i], dy = vy - points[i + 1];
      if (std::sqrt(dx * dx + dy * dy) < eps) return false;
      if (points[i] > vx) {
        if (points[i + 1] < vy && points[nx + 1] > vy) ++wind;
        else if (points[i + 1] > vy && points[nx + 1] < vy) --wind;
      }
    }
    if (sec < 0 ? wind % 2 == 0 : wind % 2 == 1) return false;
    start += n;
  }
  return true;
}

DomainRule CubicBoundary::domain_rule() const {
  DomainRule r;
  r.kind = DomainRule::kWinding;
  for (const Hole& h : holes) r.hole_nodes.push_back(h.num_nodes);
  return r;
}

// ---- CubicSpline: the starfish (cubic_spline.cpp:13-50) -------------------
void CubicSpline::get_spline_points(std::vector<double>* x0, std::vector<double>* x1) {
  star_knots(0.5, 0.5, 0.375, x0, x1);
}

// With DEFAULT the reference applies SINGLE_ELECTRON to N nodes of the
// 20 (N / 20) it made, past its arrays; here it covers the nodes made.
void CubicSpline::initialize(int N, BoundaryCondition bc) {
  clear_arrays(this);
  const int nps = N / kStarfishKnots;
  std::vector<double> xs, ys;
  get_spline_points(&xs, &ys);
  spline_curve(this, false, nps, xs, ys);
  num_outer_nodes = nps * kStarfishKnots;
  const int np = weights.size();
  if (bc == DEFAULT) {
    boundary_values = ki_Mat(np, 1);
    apply_boundary_condition(0, np, SINGLE_ELECTRON);
  } else {
    set_boundary_values_size(bc);
    apply_boundary_condition(0, np, bc);
  }
}

// ---- Ex1Boundary (ex1boundary.cpp:10-173) ---------------------------------
// a unit square with wavy walls: 24 knots alternating between the wall and
// 0.05 inside it
void Ex1Boundary::get_spline_points(std::vector<double>* x0, std::vector<double>* x1) {
  for (int side = 0; side < 4; ++side)
    for (int i = 0; i < 3; ++i) {
      const double s = i / 3.0, h = (i / 3.0) + (1.0 / 6.0);
      const double hx = 1.0 - (1.0 / 6.0) - (i / 3.0);
      switch (side) {
        case 0: x0->insert(x0->end(), {s, h}); x1->insert(x1->end(), {0, 0.05}); break;
        case 1: x0->insert(x0->end(), {1.0, 0.95}); x1->insert(x1->end(), {s, h}); break;
        case 2: x0->insert(x0->end(), {1.0 - s, hx}); x1->insert(x1->end(), {1.0, 0.95}); break;
        default: x0->insert(x0->end(), {0, 0.05}); x1->insert(x1->end(), {1.0 - s, hx});
      }
    }
}

void Ex1Boundary::get_star_spline_points(double x, double y, std::vector<double>* x0,
                                         std::vector<double>* x1) {
  star_knots(x, y, 0.015, x0, x1);
}

void Ex1Boundary::initialize(int N, BoundaryCondition bc) {
  clear_arrays(this);
  if (perturbation_parameters.empty()) perturbation_parameters.push_back(M_PI / 2.);
  const int ospl = 24, sspl = kStarfishKnots;
  const int onps = (2 * N / 3) / ospl, snps = (N / 12) / sspl, ncirc = N / 12;
  num_outer_nodes = ospl * onps;
  const double th = perturbation_parameters[0];
  holes.push_back(Hole{Vec2(0.5 + 0.3 * std::cos(M_PI + th), 0.5 + 0.3 * std::sin(M_PI + th)),
                       0.05, sspl * snps});
  holes.push_back(Hole{Vec2(0.4, 0.5), 0.05, sspl * snps});
  holes.push_back(Hole{Vec2(0.6, 0.5), 0.05, ncirc});
  holes.push_back(Hole{Vec2(0.5 + 0.3 * std::cos(th), 0.5 + 0.3 * std::sin(th)), 0.05, ncirc});
  std::vector<double> xs, ys;
  get_spline_points(&xs, &ys);
  spline_curve(this, false, onps, xs, ys);
  for (int k = 0; k < 2; ++k) {
    std::vector<double> sx, sy;
    get_star_spline_points(holes[k].center.a[0], holes[k].center.a[1], &sx, &sy);
    spline_curve(this, true, snps, sx, sy);
  }
  circle_hole(this, holes[2]);
  circle_hole(this, holes[3]);
  if (bc != DEFAULT) {
    set_boundary_values_size(bc);
    apply_boundary_condition(0, weights.size(), bc);
    return;
  }
  boundary_values = ki_Mat(weights.size() * 2, 1);
  const int b1 = ospl * onps, b2 = b1 + sspl * snps, b3 = b2 + sspl * snps;
  const int b4 = b3 + ncirc, b5 = b4 + ncirc;
  apply_boundary_condition(0, b1, TANGENT_VEC);
  apply_boundary_condition(b1, b2, REVERSE_NORMAL_VEC);
  apply_boundary_condition(b2, b3, NORMAL_VEC);
  apply_boundary_condition(b3, b4, REVERSE_NORMAL_VEC);
  apply_boundary_condition(b4, b5, NORMAL_VEC);
}

// ---- Ex2Boundary (ex2boundary.cpp:10-94) ----------------------------------
// the channel [-1, 2] x [0.25, 0.75]: 12 knots per long side, 2 per end
void Ex2Boundary::get_spline_points(std::vector<double>* x0, std::vector<double>* x1) {
  for (int i = 0; i < 12; ++i) {
    x0->push_back(-1 + 3 * (i / 12.0));
    x1->push_back(0.25);
  }
  for (int i = 0; i < 2; ++i) {
    x0->push_back(2);
    x1->push_back(0.25 + 0.5 * i / 2.0);
  }
  for (int i = 0; i < 12; ++i) {
    x0->push_back(2 - 3 * (i / 12.0));
    x1->push_back(0.75);
  }
  for (int i = 0; i < 2; ++i) {
    x0->push_back(-1);
    x1->push_back(0.75 - 0.5 * i / 2.0);
  }
}

void Ex2Boundary::initialize(int N, BoundaryCondition bc) {
  clear_arrays(this);
  if (perturbation_parameters.empty())
    for (int i = 0; i <= 7; ++i) perturbation_parameters.push_back(0.5);
  const int ospl = 28, onps = (3 * N / 4) / ospl, ncirc = (N / 4) / 8;
  num_outer_nodes = onps * ospl;
  for (int i = 0; i <= 7; ++i)
    holes.push_back(Hole{Vec2(-0.7 + (i / 3.), perturbation_parameters[i]), 0.05, ncirc});
  std::vector<double> xs, ys;
  get_spline_points(&xs, &ys);
  spline_curve(this, false, onps, xs, ys);
  for (const Hole& h : holes) circle_hole(this, h);
  if (bc == DEFAULT) {
    boundary_values = ki_Mat(weights.size() * 2, 1);
    apply_boundary_condition(0, weights.size(), LEFT_TO_RIGHT_FLOW);
  } else {
    set_boundary_values_size(bc);
    apply_boundary_condition(0, weights.size(), bc);
  }
}

// ---- Ex3Boundary (ex3boundary.cpp:10-132) ---------------------------------
void Ex3Boundary::get_spline_points(std::vector<double>* x0, std::vector<double>* x1) {
  star_knots(0.5, 0.5, 0.375, x0, x1);
}

void Ex3Boundary::get_star_spline_points(double x, double y, std::vector<double>* x0,
                                         std::vector<double>* x1) {
  star_knots(x, y, 0.06, x0, x1);
}

// the starfish with two star holes centred on 0.5 + 0.2 (cos a, sin a)(sin 5a + 4)
// at a = perturbation_parameters[0], [1]; the reference exits on DEFAULT,
// this throws
void Ex3Boundary::initialize(int N, BoundaryCondition bc) {
  clear_arrays(this);
  const int ospl = kStarfishKnots, sspl = kStarfishKnots;
  const int onps = (2 * N / 3) / ospl, snps = (N / 6) / sspl;
  num_outer_nodes = ospl * onps;
  if (perturbation_parameters.empty()) {
    perturbation_parameters.push_back(0);
    perturbation_parameters.push_back(M_PI);
  }
  for (int k = 0; k < 2; ++k) {
    const double a = perturbation_parameters[k];
    const double x = 0.2 * std::cos(a) * (std::sin(5 * a) + 4);
    const double y = 0.2 * std::sin(a) * (std::sin(5 * a) + 4);
    holes.push_back(Hole{Vec2(0.5 + x, 0.5 + y), 0.3, sspl * snps});
  }
  std::vector<double> xs, ys;
  get_spline_points(&xs, &ys);
  spline_curve(this, false, onps, xs, ys);
  for (int k = 0; k < 2; ++k) {
    std::vector<double> sx, sy;
    get_star_spline_points(holes[k].center.a[0], holes[k].center.a[1], &sx, &sy);
    spline_curve(this, true, snps, sx, sy);
  }
  const int b1 = ospl * onps, b2 = b1 + sspl * snps, b3 = b2 + sspl * snps;
  if (bc == EX3A) {
    boundary_values = ki_Mat(weights.size(), 1);
    apply_boundary_condition(0, b1, ALL_ZEROS);
    apply_boundary_condition(b1, b2, ALL_ONES);
    apply_boundary_condition(b2, b3, ALL_NEG_ONES);
  } else if (bc == EX3B) {
    boundary_values = ki_Mat(2 * weights.size(), 1);
    apply_boundary_condition(0, b1, HORIZONTAL_VEC);
    apply_boundary_condition(b1, b2, REVERSE_NORMAL_VEC);
    apply_boundary_condition(b2, b3, NORMAL_VEC);
  } else if (bc == DEFAULT) {
    throw std::invalid_argument("Ex3Boundary has no default boundary condition");
  } else {
    set_boundary_values_size(bc);
    apply_boundary_condition(0, weights.size(), bc);
  }
}

}  // namespace kern_interp
