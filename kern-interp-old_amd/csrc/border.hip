// Bordered system for multiply connected domains: the hole unknowns alpha of
//   [K U; Psi -I] [mu; alpha] = [f; g]
// (skel_factorization.cpp:130-242 for the factor, :406-551 for the solve,
// where g = 0).
//
// Factor: U and Psi^T go through the same upward sweeps as a right-hand side
// (U: x_r -= T^T x_s, x_s -= L x_r; Psi^T: x_r -= T^T x_s, x_s -= U^T x_r),
// D^{-1} C = X_rr^{-1} applied to the swept U on the redundant DOFs, and
//   S = [allskel_mat, U'(s,:); Psi'(:,s), -I - Psi'(:,red) D^{-1} C]
// is assembled and inverted in FP64.
// Solve: after the upward pass with D^{-1} fused in, x(red) already holds
// D^{-1} w and x(s) holds z, so
//   y = S^{-1} [z; g - Psi'(:,red) D^{-1} w],  alpha = y(ns:),
//   x(s) = y(:ns),  x(red) = D^{-1} w - (D^{-1} C) alpha,
// then the downward pass.
#include <algorithm>
#include <cmath>

#include "device_util.h"
#include "factor.h"

namespace kigpu {

template <typename T>
void up_sweeps(Context& c, TreeData& t, T* x, int ldx, int nrhs, int mode);

namespace {

constexpr int kTB = 256;
constexpr int kNW = kTB / kWave;

// out(i, j) = sum over non-skeleton DOFs d of A(d, i) B(d, j)
template <typename TA, typename TB>
__global__ __launch_bounds__(kTB) void k_masked_gram(const TA* A, int lda, const TB* B, int ldb,
                                                     const int* is_skel, int ndof, int h,
                                                     double* out) {
  __shared__ double red[kNW];
  const int i = blockIdx.x, j = blockIdx.y;
  const TA* a = A + (size_t)i * lda;
  const TB* b = B + (size_t)j * ldb;
  double s = 0;
  for (int d = threadIdx.x; d < ndof; d += kTB)
    if (!is_skel[d]) s += (double)a[d] * (double)b[d];
  s = block_sum<kNW>(s, red);
  if (threadIdx.x == 0) out[i + (size_t)j * h] = s;
}

__global__ void k_mark(const int* dofs, int n, int* flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[dofs[i]] = 1;
}

template <typename A, typename B>
__global__ void k_cast(const A* in, B* out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (B)in[i];
}

// one block per column of S
template <typename T>
__global__ void k_assemble_s(int ns, int h, int ndof, const int* allskel, const double* root_mat,
                             const T* sU, const T* sP, const double* G, double* S) {
  const int nS = ns + h;
  const int j = blockIdx.x;
  for (int i = threadIdx.x; i < nS; i += blockDim.x) {
    double v;
    if (j < ns) {
      v = i < ns ? root_mat[i + (size_t)j * ns] : (double)sP[allskel[j] + (size_t)(i - ns) * ndof];
    } else {
      const int jj = j - ns;
      if (i < ns) v = (double)sU[allskel[i] + (size_t)jj * ndof];
      else v = (i - ns == jj ? -1.0 : 0.0) - G[(i - ns) + (size_t)jj * h];
    }
    S[i + (size_t)j * nS] = v;
  }
}

// M = [x(allskel, :); g - Psi'(:, red) x(red, :)] with gram = Psi'(:, red) x(red, :)
template <typename T>
__global__ void k_border_rhs(int ns, int h, const int* allskel, const T* x, int ldx,
                             const double* gram, const double* g, double* M) {
  const int c = blockIdx.x, nS = ns + h;
  for (int i = threadIdx.x; i < nS; i += blockDim.x) {
    double v;
    if (i < ns) v = (double)x[allskel[i] + (size_t)c * ldx];
    else v = (g ? g[(i - ns) + (size_t)c * h] : 0.0) - gram[(i - ns) + (size_t)c * h];
    M[i + (size_t)c * nS] = v;
  }
}

// y = S^{-1} M; x(allskel) = y(:ns); alpha = y(ns:)
template <typename T>
__global__ void k_border_apply(int ns, int h, const int* allskel, const double* Sinv,
                               const double* M, T* x, int ldx, double* alpha) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* m = reinterpret_cast<double*>(smem);
  const int c = blockIdx.x, nS = ns + h;
  for (int i = threadIdx.x; i < nS; i += blockDim.x) m[i] = M[i + (size_t)c * nS];
  __syncthreads();
  for (int i = threadIdx.x; i < nS; i += blockDim.x) {
    double s = 0;
    for (int j = 0; j < nS; ++j) s += Sinv[i + (size_t)j * nS] * m[j];
    if (i < ns) x[allskel[i] + (size_t)c * ldx] = (T)s;
    else alpha[(i - ns) + (size_t)c * h] = s;
  }
}

// x(red, :) -= (D^{-1} C)(red, :) alpha
template <typename T>
__global__ void k_border_fix(int ndof, int h, const int* is_skel, const T* sD, const double* alpha,
                             T* x, int ldx) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (d >= ndof || is_skel[d]) return;
  double s = 0;
  for (int j = 0; j < h; ++j) s += (double)sD[d + (size_t)j * ndof] * alpha[j + (size_t)c * h];
  x[d + (size_t)c * ldx] = (T)((double)x[d + (size_t)c * ldx] - s);
}

}  // namespace

template <typename T>
static T* bptr(DBuf<double>& d64, DBuf<float>& d32, size_t n) {
  if (sizeof(T) == 8) {
    d64.alloc(n);
    return reinterpret_cast<T*>(d64.p);
  }
  d32.alloc(n);
  return reinterpret_cast<T*>(d32.p);
}

template <typename T>
void border_factor(Context& c, TreeData& t) {
  hipStream_t s = c.stream;
  const int h = t.nborder, nd = t.ndof, ns = t.nroot;
  const size_t len = (size_t)nd * h;
  if (ns >= nd) KI_FAIL(KIGPU_ERR_NOCOMPRESS, "No compression possible");
  std::vector<double> psit(len);
  for (int i = 0; i < h; ++i)
    for (int d = 0; d < nd; ++d) psit[d + (size_t)i * nd] = t.h_Psi[i + (size_t)d * h];
  t.bU64in.upload(t.h_U, s);
  t.bP64in.upload(psit, s);
  T* sU = bptr<T>(t.bU64, t.bU32, len);
  T* sD = bptr<T>(t.bD64, t.bD32, len);
  T* sP = bptr<T>(t.bP64, t.bP32, len);
  const int g = ceil_div((long)len, 256);
  k_cast<double, T><<<g, 256, 0, s>>>(t.bU64in.p, sU, len);
  k_cast<double, T><<<g, 256, 0, s>>>(t.bU64in.p, sD, len);
  k_cast<double, T><<<g, 256, 0, s>>>(t.bP64in.p, sP, len);
  up_sweeps<T>(c, t, sU, nd, h, 0);
  up_sweeps<T>(c, t, sD, nd, h, 1);
  up_sweeps<T>(c, t, sP, nd, h, 2);
  t.is_skel.alloc(nd);
  t.is_skel.zero(s);
  k_mark<<<ceil_div(ns, 256), 256, 0, s>>>(t.root_dofs.p, ns, t.is_skel.p);
  DBuf<double> G, S;
  G.alloc((size_t)h * h);
  k_masked_gram<T, T><<<dim3(h, h), kTB, 0, s>>>(sP, nd, sD, nd, t.is_skel.p, nd, h, G.p);
  const int nS = ns + h;
  if (nS > kDenseInvMax) KI_FAIL(KIGPU_ERR_NOTSUPP, "bordered system larger than 1024");
  S.alloc((size_t)nS * nS);
  k_assemble_s<T><<<nS, 256, 0, s>>>(ns, h, nd, t.root_dofs.p, t.root_mat.p, sU, sP, G.p, S.p);
  KI_HIP(hipGetLastError());
  t.Sinv.alloc((size_t)nS * nS);
  dense_inverse(c, s, 1, nS, nullptr, nullptr, nullptr, nS, (size_t)nS * nS, (size_t)nS, S.p,
                t.Sinv.p, nullptr, t.status.p, 16);
  stream_wait(s);
  t.nS = nS;
}
template void border_factor<double>(Context&, TreeData&);
template void border_factor<float>(Context&, TreeData&);

template <typename T>
void border_step(Context& c, TreeData& t, T* x, int ldx, int nrhs, const double* g) {
  hipStream_t s = c.stream;
  const int h = t.nborder, nd = t.ndof, ns = t.nroot, nS = t.nS;
  const T* sD = sizeof(T) == 8 ? reinterpret_cast<const T*>(t.bD64.p)
                               : reinterpret_cast<const T*>(t.bD32.p);
  const T* sP = sizeof(T) == 8 ? reinterpret_cast<const T*>(t.bP64.p)
                               : reinterpret_cast<const T*>(t.bP32.p);
  t.bgram.alloc((size_t)h * nrhs);
  t.bM.alloc((size_t)nS * nrhs);
  t.alpha_dev.alloc((size_t)h * nrhs);
  k_masked_gram<T, T><<<dim3(h, nrhs), kTB, 0, s>>>(sP, nd, x, ldx, t.is_skel.p, nd, h,
                                                    t.bgram.p);
  k_border_rhs<T><<<nrhs, 256, 0, s>>>(ns, h, t.root_dofs.p, x, ldx, t.bgram.p, g, t.bM.p);
  k_border_apply<T><<<nrhs, 256, sizeof(double) * nS, s>>>(ns, h, t.root_dofs.p, t.Sinv.p,
                                                           t.bM.p, x, ldx, t.alpha_dev.p);
  k_border_fix<T><<<dim3(ceil_div(nd, 256), nrhs), 256, 0, s>>>(nd, h, t.is_skel.p, sD,
                                                                t.alpha_dev.p, x, ldx);
  KI_HIP(hipGetLastError());
}
template void border_step<double>(Context&, TreeData&, double*, int, int, const double*);
template void border_step<float>(Context&, TreeData&, float*, int, int, const double*);

}  // namespace kigpu
