// Dense blocks on the GPU for the C++ drop-in classes: kernel blocks
// K(tgt, src) (Kernel::operator(), kern_interp/kernel/kernel.cpp:99-150,
// :47-96), proxy blocks against explicit proxy points (laplace_proxy_get
// :310-334, stokes_proxy_get :268-307), and the elimination of one box from
// its K_BN (SkelFactorization::decouple, skel_factorization.cpp:55-91).
#include <algorithm>

#include "device_util.h"
#include "entries.h"
#include "factor.h"

namespace kigpu {

namespace {

__global__ void k_block(GeomView g, const double* dpx, const double* dpy, int pde, int sd,
                        const int* tgt, int ntgt, const int* src, int nsrc, int forward,
                        double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntgt) return;
  for (int j = blockIdx.y; j < nsrc; j += gridDim.y) {
  const int td = tgt[i], sdf = src[j];
  double v;
  if (!forward) {
    v = kernel_entry(g, pde, td, sdf);
  } else {
    const int tp = td / sd, sp = sdf / sd;
    const double tx = dpx[tp], ty = dpy[tp];
    if (pde == KIGPU_STOKES) {
      v = stokes_entry(td, sdf, tx, ty, 0.0, 0.0, g.px[sp], g.py[sp], g.nx[sp], g.ny[sp], g.w[sp],
                       g.kappa[sp]);
    } else {
#pragma clang fp contract(off)
      v = laplace_entry(pde, tx - g.px[sp], ty - g.py[sp], g.w[sp], g.kappa[sp], g.nx[sp],
                        g.ny[sp], 0.0, 0.0, true);
    }
  }
  out[i + (size_t)j * ntgt] = v;
  }
}

// Rows [proxy <- box; box <- proxy] against box DOF j: Laplace 2 np rows
// (row i < np: proxy point i as target of the box DOF; row np + i: the box
// DOF as target of proxy point i, weight pw); Stokes 4 np rows, two
// components per proxy point in each half.
__global__ void k_proxy(GeomView g, int pde, const double* pp, const double* pn, int np,
                        double pw, const int* box, int nbox, double* out) {
#pragma clang fp contract(off)
  const int stokes = pde == KIGPU_STOKES;
  const int rows = (stokes ? 4 : 2) * np;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= rows || j >= nbox) return;
  const int d = box[j];
  const int bp = stokes ? d >> 1 : d;
  const double bx = g.px[bp], by = g.py[bp], bnx = g.nx[bp], bny = g.ny[bp], bw = g.w[bp];
  double v;
  if (!stokes) {
    const int half = i >= np, p = half ? i - np : i;
    const double r1 = pp[2 * p] - bx, r2 = pp[2 * p + 1] - by;
    const double tn1 = pn[2 * p], tn2 = pn[2 * p + 1];
    v = !half ? laplace_entry(pde, r1, r2, bw, 0.0, bnx, bny, tn1, tn2, false)
              : laplace_entry(pde, -r1, -r2, pw, 0.0, tn1, tn2, bnx, bny, false);
  } else {
    const int half = i >= 2 * np, ii = half ? i - 2 * np : i;
    const int p = ii >> 1, a = ii & 1, comp = d & 1;
    const double r[2] = {pp[2 * p] - bx, pp[2 * p + 1] - by};
    const double tn[2] = {pn[2 * p], pn[2 * p + 1]};
    const double sn[2] = {bnx, bny};
    const double rr = r[0] * r[0] + r[1] * r[1];
    if (!half) {
      const double pot = bw * kInvPi * (r[0] * sn[0] + r[1] * sn[1]) / (rr * rr);
      v = pot * r[a] * r[comp] + bw * tn[a] * sn[comp];
    } else {
      const double pot = -pw * kInvPi * (r[0] * tn[0] + r[1] * tn[1]) / (rr * rr);
      v = pot * r[a] * r[comp] + pw * sn[comp] * tn[a];
    }
  }
  out[i + (size_t)j * rows] = v;
}

// The elimination of one box (decouple, skel_factorization.cpp:55-91) from
// its K_BN (n x n, active order, the descendants' updates already taken
// off), with the reference's LU semantics: dgetrf of X_rr (first-maximum
// row pivots, piv 1-based), L = K_sn_r X_rr^{-1} and U = X_rr^{-1} K_r_sn by
// LU solves, schur_update = L K_r_sn.  One workgroup, FP64, global memory;
// ws holds K_r_sn (r x k) and K_sn_r (k x r).
constexpr int kElimTB = 1024;
constexpr int kElimNW = kElimTB / kWave;
__global__ __launch_bounds__(kElimTB) void k_schur_one(const double* K, int n, int k,
                                                       const int* perm, const double* T,
                                                       double* X, double* Xlu, int* piv,
                                                       double* L, double* U, double* S,
                                                       double* ws, int* status) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int r = n - k;
  double* Krsn = ws;                    // r x k
  double* Ksnr = ws + (size_t)r * k;    // k x r
  auto KB = [&](int i, int j) { return K[perm[i] + (size_t)perm[j] * n]; };  // [s r] order
  for (int e = tid; e < r * k; e += kElimTB) {
    const int i = e % r, j = e / r;
    double v = KB(k + i, j);
    for (int t = 0; t < k; ++t) v -= T[t + (size_t)i * k] * KB(t, j);
    Krsn[e] = v;
  }
  for (int e = tid; e < k * r; e += kElimTB) {
    const int i = e % k, j = e / k;
    double v = KB(i, k + j);
    for (int t = 0; t < k; ++t) v -= KB(i, t) * T[t + (size_t)j * k];
    Ksnr[e] = v;
  }
  __syncthreads();
  for (int e = tid; e < r * r; e += kElimTB) {
    const int i = e % r, j = e / r;
    double v = KB(k + i, k + j);
    for (int t = 0; t < k; ++t) v -= T[t + (size_t)i * k] * KB(t, k + j);
    for (int t = 0; t < k; ++t) v -= Krsn[i + (size_t)t * r] * T[t + (size_t)j * k];
    X[e] = v;
    Xlu[e] = v;
  }
  __syncthreads();
  // dgetrf: right-looking LU with partial pivoting (first largest |a|)
  __shared__ int s_p;
  for (int c = 0; c < r; ++c) {
    if (wv == 0) {
      double best = -1.0;
      int bi = c;
      for (int i = c + lane; i < r; i += kWave) {
        const double a = fabs(Xlu[i + (size_t)c * r]);
        if (a > best) {
          best = a;
          bi = i;
        }
      }
      wave_argmax(best, bi);
      if (lane == 0) {
        s_p = bi;
        piv[c] = bi + 1;
      }
    }
    __syncthreads();
    const int p = s_p;
    if (p != c)
      for (int j = tid; j < r; j += kElimTB) {
        const double t0 = Xlu[c + (size_t)j * r];
        Xlu[c + (size_t)j * r] = Xlu[p + (size_t)j * r];
        Xlu[p + (size_t)j * r] = t0;
      }
    __syncthreads();
    const double d = Xlu[c + (size_t)c * r];
    if (!pivot_usable(d) && tid == 0) atomicOr(status, 1);
    for (int i = c + 1 + tid; i < r; i += kElimTB) Xlu[i + (size_t)c * r] /= d;
    __syncthreads();
    const int m = r - c - 1;
    for (int e = tid; e < m * m; e += kElimTB) {
      const int i = c + 1 + e % m, j = c + 1 + e / m;
      Xlu[i + (size_t)j * r] -= Xlu[i + (size_t)c * r] * Xlu[c + (size_t)j * r];
    }
    __syncthreads();
  }
  // the solves take 0-based pivots
  __shared__ int zpiv[1024];
  for (int c = tid; c < r; c += kElimTB) zpiv[c] = piv[c] - 1;
  __syncthreads();
  // L^T = X^{-T} K_sn_r^T: every row of K_sn_r (stride k) solved with X^T
  block_lu_solve<double, kElimNW>(Xlu, r, zpiv, r, Ksnr, 1, k, k, true);
  // U = X^{-1} K_r_sn: every column
  for (int e = tid; e < r * k; e += kElimTB) U[e] = Krsn[e];
  __syncthreads();
  block_lu_solve<double, kElimNW>(Xlu, r, zpiv, r, U, r, 1, k, false);
  __syncthreads();
  for (int e = tid; e < k * r; e += kElimTB) L[e] = Ksnr[e];
  __syncthreads();
  for (int e = tid; e < k * k; e += kElimTB) {
    const int i = e % k, j = e / k;
    double v = 0;
    for (int t = 0; t < r; ++t) v += L[i + (size_t)t * k] * Krsn[t + (size_t)j * r];
    S[e] = v;
  }
}

}  // namespace

void proxy_block(Context& c, const GeomView& g, int pde, const double* h_pp, const double* h_pn,
                 int np, double pw, const int* h_box, int nbox, double* h_out) {
  hipStream_t s = c.stream;
  const int rows = (pde == KIGPU_STOKES ? 4 : 2) * np;
  DBuf<double> pp, pn, out;
  DBuf<int> box;
  pp.upload(h_pp, 2 * (size_t)np, s);
  pn.upload(h_pn, 2 * (size_t)np, s);
  box.upload(h_box, nbox, s);
  out.alloc((size_t)rows * nbox);
  if (rows > 0 && nbox > 0) {
    k_proxy<<<dim3(ceil_div(rows, 256), nbox), 256, 0, s>>>(g, pde, pp.p, pn.p, np, pw, box.p,
                                                           nbox, out.p);
    KI_HIP(hipGetLastError());
  }
  out.download(h_out, (size_t)rows * nbox, s);
  stream_wait(s);
}

void schur_eliminate(Context& c, const double* h_K, int n, int k, const int* h_perm,
                     const double* h_T, double* X_rr, double* X_lu, int* piv, double* L,
                     double* U, double* schur, int* singular) {
  hipStream_t s = c.stream;
  const int r = n - k;
  if (k < 1 || r < 1 || r > 1024) KI_FAIL(KIGPU_ERR_ARG, "need 1 <= k < n and n - k <= 1024");
  DBuf<double> K, T, X, Xlu, Ld, Ud, S, ws;
  DBuf<int> pm, pv, st;
  K.upload(h_K, (size_t)n * n, s);
  T.upload(h_T, (size_t)k * r, s);
  pm.upload(h_perm, n, s);
  X.alloc((size_t)r * r);
  Xlu.alloc((size_t)r * r);
  Ld.alloc((size_t)k * r);
  Ud.alloc((size_t)r * k);
  S.alloc((size_t)k * k);
  ws.alloc(2 * (size_t)k * r);
  pv.alloc(r);
  st.alloc(1);
  st.zero(s);
  k_schur_one<<<1, kElimTB, 0, s>>>(K.p, n, k, pm.p, T.p, X.p, Xlu.p, pv.p, Ld.p, Ud.p, S.p,
                                    ws.p, st.p);
  KI_HIP(hipGetLastError());
  if (X_rr) X.download(X_rr, (size_t)r * r, s);
  if (X_lu) Xlu.download(X_lu, (size_t)r * r, s);
  if (piv) pv.download(piv, r, s);
  if (L) Ld.download(L, (size_t)k * r, s);
  if (U) Ud.download(U, (size_t)r * k, s);
  if (schur) S.download(schur, (size_t)k * k, s);
  int sg = 0;
  KI_HIP(hipMemcpyAsync(&sg, st.p, sizeof(int), hipMemcpyDeviceToHost, s));
  stream_wait(s);
  if (singular) *singular = sg;
}

void kernel_block(Context& c, const GeomView& g, const double* dpx, const double* dpy, int pde,
                  int sd, const int* h_tgt, int ntgt, const int* h_src, int nsrc, int forward,
                  double* h_out) {
  hipStream_t s = c.stream;
  DBuf<int> dt, ds;
  DBuf<double> dout;
  dt.upload(h_tgt, ntgt, s);
  ds.upload(h_src, nsrc, s);
  dout.alloc((size_t)ntgt * nsrc);
  if (ntgt > 0 && nsrc > 0) {
    GeomView gv = g;
    gv.sd = sd;
    k_block<<<dim3(ceil_div(ntgt, 256), std::min(nsrc, 65535)), 256, 0, s>>>(gv, dpx, dpy, pde, sd, dt.p, ntgt,
                                                           ds.p, nsrc, forward, dout.p);
    KI_HIP(hipGetLastError());
  }
  dout.download(h_out, (size_t)ntgt * nsrc, s);
  stream_wait(s);
}

}  // namespace kigpu
