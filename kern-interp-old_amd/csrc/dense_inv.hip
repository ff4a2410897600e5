// Dense FP64 inverses of the root skeleton matrix and of the bordered system
// (the reference factors them with dgetrf and solves with dgetrs:
// skel_factorization.cpp:116-128, :236-242, ki_mat.cpp:427-493).
//
// Blocked Gauss-Jordan by exchange steps with partial pivoting, one
// 256-thread workgroup (one wave per SIMD) per matrix:
//  * panel of 16 columns: every row's 16 panel entries sit in the registers
//    of one lane; a step takes the largest |B(i, c)| over the rows not
//    pivoted yet (near-ties to the first row, NaN below every number),
//    i.e. dgetrf's pivot, by one 64-bit key maximum over the wavefront.
//    Up to 256 rows the panel is held by one wave (rows lane + 64 u) and a
//    step has no workgroup barrier; above, by the four waves (rows
//    tid + 256 u), which meet once per step to pick the winner of their
//    four candidates;
//  * rows are never interchanged: the pivot of step c stays in its row
//    p(c), so the final array holds inv(A)(c, q) at (p(c), c') with
//    p(c') = q, gathered on the way out;
//  * after a panel, the other columns take the panel's bw steps at once:
//      B(i, J) = [i not a pivot row of the panel] B(i, J) + W(i, :) R
//    with W the transformed panel and R the panel's pivot rows of B(:, J)
//    before the update, one 16-deep product per 16 x 16 tile on the FP64
//    matrix cores (v_mfma_f64_16x16x4f64).
// The matrix lives in LDS when it fits, else in a global workspace (L2).
#include <algorithm>
#include <cmath>

#include "dense_inv.h"
#include "factor.h"

namespace kigpu {

namespace {

using namespace gj;

template <int U, bool ONE, bool LDS>
__global__ __launch_bounds__(kGjTB) void k_dense_gj(const int* n_arr, const int64_t* doff,
                                                    const int64_t* moff, int n0, const double* A_all,
                                                    double* Ainv_all, double* ws, double* rws,
                                                    int* piv_all, int* status, int bit) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ GjCand cand[2];
  dense_gj_block<U, ONE, LDS>(blockIdx.x, smem, cand, n_arr, doff, moff, n0, A_all, Ainv_all, ws,
                            rws, piv_all, status, bit);
}

template <int U, bool ONE>
void launch_gj(Context& c, hipStream_t s, int ni, int nmax, const int* d_n, const int64_t* d_doff,
               const int64_t* d_moff, int n0, size_t mtot, size_t dtot, const double* A,
               double* Ainv, int* piv, int* status, int bit) {
  const size_t lds_static = sizeof(GjCand) * 2;
  const size_t lb = gj_lds_bytes(nmax, true);
  if (lb + lds_static <= c.lds_max) {
    allow_lds(k_dense_gj<U, ONE, true>, lb);
    k_dense_gj<U, ONE, true><<<ni, kGjTB, lb, s>>>(d_n, d_doff, d_moff, n0, A, Ainv, nullptr,
                                                nullptr, piv, status, bit);
  } else {
    DBuf<double> ws, rws;
    ws.alloc(mtot);
    rws.alloc((size_t)kPB * dtot);
    const size_t lg = gj_lds_bytes(nmax, false);
    allow_lds(k_dense_gj<U, ONE, false>, lg);
    k_dense_gj<U, ONE, false><<<ni, kGjTB, lg, s>>>(d_n, d_doff, d_moff, n0, A, Ainv, ws.p, rws.p,
                                                 piv, status, bit);
  }
  KI_HIP(hipGetLastError());
}

}  // namespace

void dense_inverse(Context& c, hipStream_t s, int ni, int nmax, const int* d_n,
                   const int64_t* d_doff, const int64_t* d_moff, int n0, size_t mtot, size_t dtot,
                   const double* A, double* Ainv, int* piv, int* status, int bit) {
  if (ni <= 0 || nmax <= 0) return;
  if (nmax > kDenseInvMax) KI_FAIL(KIGPU_ERR_NOTSUPP, "dense inverse larger than 1024");
  auto go = [&](auto launch) {
    launch(c, s, ni, nmax, d_n, d_doff, d_moff, n0, mtot, dtot, A, Ainv, piv, status, bit);
  };
  if (nmax <= kGjWave) go(launch_gj<1, true>);
  else if (nmax <= 2 * kGjWave) go(launch_gj<2, true>);
  else if (nmax <= kGjTB) go(launch_gj<4, true>);
  else if (nmax <= 2 * kGjTB) go(launch_gj<2, false>);
  else go(launch_gj<4, false>);
}

}  // namespace kigpu
