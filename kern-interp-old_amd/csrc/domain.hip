// In-domain tests and the domain evaluation of boundary_integral_solve
// (kern_interp/linear_solve.cpp:143-180) on the GPU: Boundary::is_in_domain
// of the reference's geometries for every target at once, the hole terms
// U_forward alpha (initialize_U_mat at the targets, linear_solve.cpp:18-64)
// added to the forward apply, and the targets outside the domain zeroed.
//
// One thread per target.  The crossing-count rule streams the boundary nodes
// through LDS in tiles of 1024 (x, y, successor's y, section-end flag), so a
// block reads every node once from HBM and each thread does a few flops per
// node; 10^5 targets against 6.5 * 10^4 nodes is ~7 * 10^9 node tests.
#include <algorithm>
#include <cmath>

#include "factor.h"

namespace kigpu {

namespace {

constexpr int kMaskTB = 256;
constexpr int kNodeTile = 1024;

// Circle / Annulus / Donut (circle.cpp:34-42, annulus.cpp:66-78,
// donut.cpp:67-79): outside when |t - c| + eps > R or |t - h| - eps < r_h
__global__ void k_mask_disc(const double* __restrict__ txy, int ntgt, double cx, double cy,
                            double R, double eps, const double* __restrict__ holes, int nholes,
                            unsigned char* __restrict__ inside) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntgt) return;
  const double x = txy[2 * i], y = txy[2 * i + 1];
  const double dx = x - cx, dy = y - cy;
  bool in = !(sqrt(dx * dx + dy * dy) + eps > R);
  for (int k = 0; in && k < nholes; ++k) {
    const double hx = x - holes[3 * k], hy = y - holes[3 * k + 1];
    if (sqrt(hx * hx + hy * hy) - eps < holes[3 * k + 2]) in = false;
  }
  inside[i] = in;
}

// StarCurve: inside when |t| + eps < 1 + sum_k a_k cos(f_k atan2(t) + phi_k)
__global__ void k_mask_star(const double* __restrict__ txy, int ntgt, double eps,
                            const double* __restrict__ terms, int nterms,
                            unsigned char* __restrict__ inside) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntgt) return;
  const double x = txy[2 * i], y = txy[2 * i + 1];
  const double th = atan2(y, x);
  double r = 1.0;
  for (int k = 0; k < nterms; ++k) r += terms[3 * k] * cos(terms[3 * k + 1] * th + terms[3 * k + 2]);
  inside[i] = sqrt(x * x + y * y) + eps < r;
}

// CubicBoundary::is_in_domain (cubic_boundary.cpp:182-230).  node = (x, y,
// y of the successor within its section, 1 on the last node of a section).
// Per section: edges from a node right of the target crossing its horizontal
// count +1 upward and -1 downward; the outer count must be odd (C remainder
// != 0), a hole excludes at remainder 1; a node closer than eps excludes.
__global__ __launch_bounds__(kMaskTB) void k_mask_winding(const double* __restrict__ txy,
                                                          int ntgt,
                                                          const double4* __restrict__ nodes,
                                                          int nnodes, double eps,
                                                          unsigned char* __restrict__ inside) {
#pragma clang fp contract(off)
  __shared__ double4 tile[kNodeTile];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < ntgt;
  const double x = valid ? txy[2 * i] : 0.0, y = valid ? txy[2 * i + 1] : 0.0;
  bool out = !valid;
  int wind = 0, sec = 0;
  for (int base = 0; base < nnodes; base += kNodeTile) {
    const int m = min(kNodeTile, nnodes - base);
    __syncthreads();
    for (int k = threadIdx.x; k < m; k += blockDim.x) tile[k] = nodes[base + k];
    __syncthreads();
    for (int k = 0; k < m && !out; ++k) {
      const double4 q = tile[k];
      const double dx = x - q.x, dy = y - q.y;
      if (sqrt(dx * dx + dy * dy) < eps) out = true;
      if (q.x > x) {
        if (q.y < y && q.z > y) ++wind;
        else if (q.y > y && q.z < y) --wind;
      }
      if (q.w != 0.0) {
        const int r = wind % 2;
        if (sec == 0 ? r == 0 : r == 1) out = true;
        wind = 0;
        ++sec;
      }
    }
  }
  if (valid) inside[i] = !out;
}

// u(:, c) += U_forward(t) alpha(:, c) (linear_solve.cpp:18-64 at the
// targets), then zero where the target is outside (inside may be null)
__global__ void k_holes_and_mask(int stokes, const double* __restrict__ txy, int ntgt,
                                 const double* __restrict__ hc, int nholes,
                                 const double* __restrict__ alpha, int lda, int nrhs,
                                 const unsigned char* __restrict__ inside, double* __restrict__ u,
                                 int ldu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntgt) return;
  const double x = txy[2 * i], y = txy[2 * i + 1];
  const bool zero = inside && !inside[i];
  const double sc = 1.0 / (4 * M_PI);
  for (int c = 0; c < nrhs; ++c) {
    const double* a = alpha ? alpha + (size_t)c * lda : nullptr;
    double* uc = u + (size_t)c * ldu;
    if (!stokes) {
      double acc = 0;
      for (int k = 0; a && k < nholes; ++k) {
        const double rx = x - hc[2 * k], ry = y - hc[2 * k + 1];
        acc += log(sqrt(rx * rx + ry * ry)) * a[k];
      }
      uc[i] = zero ? 0.0 : uc[i] + acc;
    } else {
      double a0 = 0, a1 = 0;
      for (int k = 0; a && k < nholes; ++k) {
        const double rx = x - hc[2 * k], ry = y - hc[2 * k + 1];
        const double nr = sqrt(rx * rx + ry * ry), n2 = nr * nr, lg = log(1 / nr);
        a0 += sc * (lg + rx * rx / n2) * a[3 * k] + sc * (rx * ry / n2) * a[3 * k + 1] +
              ry * (sc / n2) * a[3 * k + 2];
        a1 += sc * (ry * rx / n2) * a[3 * k] + sc * (lg + ry * ry / n2) * a[3 * k + 1] -
              rx * (sc / n2) * a[3 * k + 2];
      }
      uc[2 * i] = zero ? 0.0 : uc[2 * i] + a0;
      uc[2 * i + 1] = zero ? 0.0 : uc[2 * i + 1] + a1;
    }
  }
}

}  // namespace

void domain_mask(Context& c, const kigpu_region& r, const double* d_txy, int ntgt,
                 unsigned char* d_inside) {
  hipStream_t s = c.stream;
  const int nb = ceil_div(ntgt, kMaskTB);
  if (ntgt == 0) return;
  if (r.kind == KIGPU_REGION_DISC) {
    DBuf<double> holes;
    holes.upload(r.holes, 3 * (size_t)r.nholes, s);
    k_mask_disc<<<nb, kMaskTB, 0, s>>>(d_txy, ntgt, r.cx, r.cy, r.radius, r.eps, holes.p,
                                       r.nholes, d_inside);
    KI_HIP(hipGetLastError());
    stream_wait(s);
  } else if (r.kind == KIGPU_REGION_STAR) {
    std::vector<double> terms;
    for (int k = 0; k < r.nterms; ++k)
      terms.insert(terms.end(), {r.amps[k], r.freqs[k], r.phases[k]});
    DBuf<double> dterms;
    dterms.upload(terms, s);
    k_mask_star<<<nb, kMaskTB, 0, s>>>(d_txy, ntgt, r.eps, dterms.p, r.nterms, d_inside);
    KI_HIP(hipGetLastError());
    stream_wait(s);
  } else if (r.kind == KIGPU_REGION_WINDING) {
    // sections: the outer nodes, then hole_nodes[k] per hole
    std::vector<int> sec{r.num_outer};
    for (int k = 0; k < r.nholes; ++k) sec.push_back(r.hole_nodes[k]);
    std::vector<double4> nd;
    int o = 0;
    for (int n : sec) {
      for (int k = 0; k < n; ++k) {
        const int a = o + k, b = o + (k + 1) % n;
        nd.push_back(make_double4(r.nodes[2 * a], r.nodes[2 * a + 1], r.nodes[2 * b + 1],
                                  k == n - 1 ? 1.0 : 0.0));
      }
      o += n;
    }
    DBuf<double4> dn;
    dn.upload(nd, s);
    k_mask_winding<<<nb, kMaskTB, 0, s>>>(d_txy, ntgt, dn.p, (int)nd.size(), r.eps, d_inside);
    KI_HIP(hipGetLastError());
    stream_wait(s);
  } else {
    KI_FAIL(KIGPU_ERR_ARG, "unknown region kind");
  }
}

void domain_solution(Context& c, const TreeData& t, int pde, const double* d_txy, int ntgt,
                     const double* d_mu, int nrhs, const double* d_alpha, const double* h_hc,
                     int nholes, const kigpu_region* region, double* d_u) {
  hipStream_t s = c.stream;
  forward_apply(c, t, pde, d_txy, ntgt, d_mu, nrhs, d_u);
  DBuf<unsigned char> in;
  if (region) {
    in.alloc(ntgt);
    domain_mask(c, *region, d_txy, ntgt, in.p);
  }
  DBuf<double> hc;
  const bool holes = d_alpha && nholes > 0;
  if (holes) hc.upload(h_hc, 2 * (size_t)nholes, s);
  const int stokes = pde == KIGPU_STOKES, sd = stokes ? 2 : 1;
  k_holes_and_mask<<<ceil_div(ntgt, kMaskTB), kMaskTB, 0, s>>>(
      stokes, d_txy, ntgt, hc.p, holes ? nholes : 0, holes ? d_alpha : nullptr,
      (stokes ? 3 : 1) * nholes, nrhs, region ? in.p : nullptr, d_u, sd * ntgt);
  KI_HIP(hipGetLastError());
  stream_wait(s);
}

}  // namespace kigpu
