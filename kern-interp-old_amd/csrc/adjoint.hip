// Shape sensitivities of the Laplace double layer for adjoint gradients.
//
// The reference has no adjoint; its experiments differentiate by finite
// differences of whole solves (experiments/experiment_three.cpp:22-56,
// 160-171).  With mu = K^{-1} f, u = F mu at fixed domain targets and a loss
// L(u), the gradient with respect to the boundary geometry is
//   dL = v^T dF mu + lam^T df - lam^T dK mu,   v = dL/du,  lam = K^{-T} F^T v,
// so the device work is one transposed solve and the two bilinear forms
// lam^T dK mu and v^T dF mu, taken here per boundary point p with respect to
// its own data (x_p, y_p, nx_p, ny_p, w_p, kappa_p).  The caller chains them
// through its parameterisation of the curve.
//
// Entry (kernel.cpp:11-36), target t, source s, r = x_t - x_s, q = r . n_s:
//   K(t, s) = -w_s / (2 pi) q / |r|^2,    K(s, s) = 1/2 + kappa_s w_s / (4 pi)
// with the diagonal taken where r is exactly zero, as the entries do.
#include "device_util.h"
#include "entries.h"
#include "factor.h"

namespace kigpu {

namespace {

constexpr int kST = 256;

// sens[6 p + (0..5)] = d/d(x_p, y_p, nx_p, ny_p, w_p, kappa_p) of
//   KSENS:  sum_{t,s} lam_t K(t, s) mu_s       (rows = the boundary points)
//   !KSENS: sum_{t,s} v_t F(t, s) mu_s         (rows = domain targets)
// One thread per boundary point p; the rows are staged through LDS.
template <bool KSENS>
__global__ __launch_bounds__(kST) void k_sens(int nrows, const double* rx, const double* ry,
                                              const double* rnx, const double* rny,
                                              const double* rw, const double* rcoef,
                                              GeomView g, const double* lam, const double* mu,
                                              double* part, int rows_per_split) {
  __shared__ double sx[kST], sy[kST], snx[kST], sny[kST], sw[kST], sl[kST], sm[kST];
  const int p = blockIdx.x * kST + threadIdx.x;
  const bool valid = p < g.npts;
  const double px = valid ? g.px[p] : 0, py = valid ? g.py[p] : 0;
  const double pnx = valid ? g.nx[p] : 0, pny = valid ? g.ny[p] : 0;
  const double pw = valid ? g.w[p] : 0, pk = valid ? g.kappa[p] : 0;
  const double mp = valid ? mu[p] : 0;
  const double lp = (valid && KSENS) ? lam[p] : 0;
  const double c = kInv2Pi;
  double gx = 0, gy = 0, gnx = 0, gny = 0, gw = 0, gk = 0;
  const int j0 = blockIdx.y * rows_per_split;
  const int j1 = min(nrows, j0 + rows_per_split);
  for (int b0 = j0; b0 < j1; b0 += kST) {
    const int nb = min(kST, j1 - b0);
    __syncthreads();
    if (threadIdx.x < nb) {
      const int j = b0 + threadIdx.x;
      sx[threadIdx.x] = rx[j];
      sy[threadIdx.x] = ry[j];
      sl[threadIdx.x] = rcoef[j];  // lam_j (KSENS) or v_j
      if (KSENS) {
        snx[threadIdx.x] = rnx[j];
        sny[threadIdx.x] = rny[j];
        sw[threadIdx.x] = rw[j];
        sm[threadIdx.x] = mu[j];
      }
    }
    __syncthreads();
    if (!valid) continue;
    for (int j = 0; j < nb; ++j) {
      // row point j as target, p as source: r' = x_j - x_p
      const double a1 = sx[j] - px, a2 = sy[j] - py;
      const double cb = sl[j] * mp;
      if (a1 == 0. && a2 == 0.) {  // diagonal entry 1/2 + kappa_p w_p c / 2
        gw += cb * 0.5 * pk * c;
        gk += cb * 0.5 * pw * c;
        continue;
      }
      const double inv = 1.0 / (a1 * a1 + a2 * a2);
      const double qp = a1 * pnx + a2 * pny;
      // d/d x_p of -w_p c q'/rho with q' = r'.n_p, r' = x_j - x_p
      const double fb = pw * c * inv;
      gx += cb * fb * (pnx - 2.0 * qp * a1 * inv);
      gy += cb * fb * (pny - 2.0 * qp * a2 * inv);
      gnx -= cb * fb * a1;
      gny -= cb * fb * a2;
      gw -= cb * c * qp * inv;
      if (KSENS) {
        // p as target, row point j as source: r = x_p - x_j = -r'
        const double ca = lp * sm[j];
        const double qj = -(a1 * snx[j] + a2 * sny[j]);
        const double fa = sw[j] * c * inv;
        // d/d x_p of -w_j c q/rho, q = r.n_j
        gx -= ca * fa * (snx[j] + 2.0 * qj * a1 * inv);
        gy -= ca * fa * (sny[j] + 2.0 * qj * a2 * inv);
      }
    }
  }
  if (valid) {
    double* o = part + (size_t)blockIdx.y * 6 * g.npts + 6 * (size_t)p;
    o[0] = gx;
    o[1] = gy;
    o[2] = gnx;
    o[3] = gny;
    o[4] = gw;
    o[5] = gk;
  }
}

__global__ void k_sum6(const double* part, int nsplit, size_t len, double* out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double s = 0;
  for (int q = 0; q < nsplit; ++q) s += part[(size_t)q * len + i];
  out[i] = s;
}

template <bool KSENS>
void run_sens(Context& c, const TreeData& t, int nrows, const double* rx, const double* ry,
              const double* rnx, const double* rny, const double* rw, const double* rcoef,
              const double* d_lam, const double* d_mu, double* d_sens) {
  hipStream_t s = c.stream;
  const GeomView g{t.px.p, t.py.p, t.nx.p, t.ny.p, t.w.p, t.kappa.p, t.npts, t.sd};
  const int tb = ceil_div(t.npts, kST);
  int split = std::max(1, std::min(256, ceil_div(2048, tb)));
  int per = ceil_div(std::max(nrows, 1), split);
  per = ceil_div(per, kST) * kST;
  split = std::max(1, ceil_div(nrows, per));
  const size_t len = 6 * (size_t)t.npts;
  DBuf<double> part;
  part.alloc(len * split);
  k_sens<KSENS><<<dim3(tb, split), kST, 0, s>>>(nrows, rx, ry, rnx, rny, rw, rcoef, g, d_lam,
                                                d_mu, part.p, per);
  k_sum6<<<ceil_div((long)len, 256), 256, 0, s>>>(part.p, split, len, d_sens);
  KI_HIP(hipGetLastError());
  stream_wait(s);
}

}  // namespace

void kernel_sensitivity(Context& c, const TreeData& t, const double* d_lam, const double* d_mu,
                        double* d_sens) {
  run_sens<true>(c, t, t.npts, t.px.p, t.py.p, t.nx.p, t.ny.p, t.w.p, d_lam, d_lam, d_mu, d_sens);
}

void forward_sensitivity(Context& c, const TreeData& t, const double* d_tgt, int ntgt,
                         const double* d_v, const double* d_mu, double* d_sens) {
  DBuf<double> tx, ty;
  split_points(c, d_tgt, ntgt, &tx, &ty);
  run_sens<false>(c, t, ntgt, tx.p, ty.p, nullptr, nullptr, nullptr, d_v, nullptr, d_mu, d_sens);
}

}  // namespace kigpu
