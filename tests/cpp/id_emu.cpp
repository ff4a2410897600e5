// CPU driver for the interpolative decomposition of one box as the ID
// kernels run it (kern-interp-old_amd/csrc/id_dev.h): id_box with the
// 1024-thread workgroup of k_id (register QR, or the trailing update in LDS
// or in the global workspace), with the 512-thread workgroup of k_id_wide,
// and id_gram_box with the 256 threads of k_id_gram, on emulated GPU threads
// (tests/cpp/simt), built with AddressSanitizer by tests/test_id_emulated.py.
// Every array the routines touch -- geometry, row and DOF lists, the proxy
// ring, the LDS, the workspace, the factor slab -- is a heap block of
// exactly the size the box needs.
//
// stdin:  int32 kind (0: k_id, 1: k_id_wide, 2: k_id_gram), tsize (4 or 8:
//         the slab's type), pde, sd, npts, n, q, proxy, in_lds; doubles tol,
//         cx, cy, rad, pw; npts doubles each of x, y, nx, ny, w, kappa; n
//         int32 box DOFs; q int32 row DOFs
// stdout: int32 status, int32 rank, n int32 pivots (position -> column), then
//         k (n - k) doubles of T (column-major)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <hip/hip_runtime.h>
#include "simt_mfma.h"
#include "id_dev.h"

using namespace kigpu;
using namespace kigpu::elim;

static void read_all(void* p, size_t bytes) {
  if (bytes && std::fread(p, 1, bytes, stdin) != bytes) {
    std::fprintf(stderr, "short input\n");
    std::exit(2);
  }
}

template <typename TS>
static int run(int kind, int pde, int sd, int npts, int n, int q, int proxy, int in_lds) {
  double par[5];
  read_all(par, sizeof(par));
  const double tol = par[0];
  std::vector<double> geo(6 * (size_t)npts);
  read_all(geo.data(), sizeof(double) * geo.size());
  double* gx[6];
  for (int c = 0; c < 6; ++c) {
    gx[c] = new double[npts];
    for (int i = 0; i < npts; ++i) gx[c][i] = geo[(size_t)c * npts + i];
  }
  const GeomView g{gx[0], gx[1], gx[2], gx[3], gx[4], gx[5], npts, sd};
  int* act_pool = new int[n];
  int* rows_pool = new int[q > 0 ? q : 1];
  read_all(act_pool, sizeof(int) * n);
  read_all(rows_pool, sizeof(int) * q);
  // proxy ring: unit vectors of the 128 proxy points, host libm
  double* ring = new double[2 * kProxyPoints];
  for (int i = 0; i < kProxyPoints; ++i) {
    const double a = 2.0 * M_PI * i * (1.0 / kProxyPoints);
    ring[i] = std::cos(a);
    ring[kProxyPoints + i] = std::sin(a);
  }
  IdWork w{};
  w.node = 0;
  w.n = n;
  w.q = q;
  w.proxy = proxy;
  const int prow = pde == KIGPU_STOKES ? 4 * kProxyPoints : 2 * kProxyPoints;
  w.m = 2 * q + (proxy ? prow : 0);
  w.in_lds = in_lds;
  w.rows_off = 0;
  w.a_off = 0;
  w.cx = par[1];
  w.cy = par[2];
  w.rad = par[3];
  w.pw = par[4];
  const int m = w.m;
  int64_t act_off[1] = {0};
  int rank[1] = {-7}, status = 0;
  int* sr_pool = new int[n];
  int* pos_pool = new int[n];
  for (int j = 0; j < n; ++j) pos_pool[j] = sr_pool[j] = -1;
  const size_t tmax = (size_t)n * n / 4 + 1;  // k (n - k) <= n^2 / 4
  TS* slab = new TS[tmax];
  uint64_t fac_ptr[1] = {reinterpret_cast<uint64_t>(slab)};
  double* ws = in_lds ? nullptr : new double[(size_t)m * n];
  size_t lds;
  if (kind == 2) lds = gram_layout(n, q).total;
  else lds = id_lds_layout(m, n, q, in_lds != 0, sizeof(double)).total;
  unsigned char* smem = new unsigned char[lds];
  auto go = [&](auto tb) {
    constexpr int TB = decltype(tb)::value;
    simt::launch(TB, [&] {
      if (in_lds)
        id_box<double, TS, true, false, TB>(w, smem, g, pde, tol, rows_pool, act_pool, act_off,
                                            ring, ws, rank, sr_pool, pos_pool, fac_ptr, &status);
      else
        id_box<double, TS, false, false, TB>(w, smem, g, pde, tol, rows_pool, act_pool, act_off,
                                             ring, ws, rank, sr_pool, pos_pool, fac_ptr, &status);
    });
  };
  if (kind == 0) {
    go(std::integral_constant<int, kIdTB>());
  } else if (kind == 1) {
    go(std::integral_constant<int, kIdWideTB>());
  } else {
    simt::launch(kGramTB, [&] {
      if (pde == KIGPU_STOKES)
        id_gram_box<TS, KIGPU_STOKES>(w, smem, g, tol, rows_pool, act_pool, act_off, ring, rank,
                                      sr_pool, pos_pool, fac_ptr, &status);
      else if (pde == KIGPU_LAPLACE_NEUMANN)
        id_gram_box<TS, KIGPU_LAPLACE_NEUMANN>(w, smem, g, tol, rows_pool, act_pool, act_off,
                                               ring, rank, sr_pool, pos_pool, fac_ptr, &status);
      else
        id_gram_box<TS, KIGPU_LAPLACE>(w, smem, g, tol, rows_pool, act_pool, act_off, ring, rank,
                                       sr_pool, pos_pool, fac_ptr, &status);
    });
  }
  std::fwrite(&status, sizeof(int), 1, stdout);
  std::fwrite(&rank[0], sizeof(int), 1, stdout);
  std::fwrite(pos_pool, sizeof(int), n, stdout);
  const int k = rank[0];
  if (k > 0)
    for (size_t e = 0; e < (size_t)k * (n - k); ++e) {
      const double v = (double)slab[e];
      std::fwrite(&v, sizeof(double), 1, stdout);
    }
  delete[] smem;
  delete[] ws;
  delete[] slab;
  delete[] sr_pool;
  delete[] pos_pool;
  delete[] ring;
  delete[] rows_pool;
  delete[] act_pool;
  for (int c = 0; c < 6; ++c) delete[] gx[c];
  return 0;
}

int main() {
  int h[9];
  read_all(h, sizeof(h));
  const int kind = h[0], tsize = h[1], pde = h[2], sd = h[3], npts = h[4], n = h[5], q = h[6],
            proxy = h[7], in_lds = h[8];
  const int prow = pde == KIGPU_STOKES ? 4 * kProxyPoints : 2 * kProxyPoints;
  const int m = 2 * q + (proxy ? prow : 0);
  if (n < 1 || q < 0 || m < 1) return 2;
  if (kind == 1 && !id_wide_box(m, n)) return 3;  // k_id_wide takes wide boxes only
  if (kind == 2 && n > kGramMaxN) return 3;
  if (kind == 0 && id_wide_box(m, n)) return 3;    // the host sends these to k_id_wide
  return tsize == 4 ? run<float>(kind, pde, sd, npts, n, q, proxy, in_lds)
                    : run<double>(kind, pde, sd, npts, n, q, proxy, in_lds);
}
