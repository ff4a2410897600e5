// CPU driver for the elimination of one box, by one wavefront
// (decouple_wave_box, the body of k_decouple_wave) or by one 1024-thread
// workgroup (decouple_block_box, the body of k_decouple), both in
// kern-interp-old_amd/csrc/elim_dev.h, run on emulated GPU threads
// (tests/cpp/simt) and built with AddressSanitizer by
// tests/test_elim_emulated.py.  The box is node 0 of
// a small made-up tree whose children are nodes 1..4; every array the
// routine touches -- geometry, pools, the factor slab, the children's
// updates, the dynamic LDS -- is a heap block of exactly the size the box
// needs, so an index past any of them is reported.
//
// stdin:  int32 tsize (4 or 8), rmax (32 or 64: the wavefront kernel with that
//         register width; 0: the workgroup kernel), pde, sd, npts, n, k, leaf,
//         then 4 int32 child update sizes (0 = no child), then npts doubles
//         each of x, y, nx, ny, w, kappa, then n int32 skeleton-then-
//         redundant DOFs, n int32 positions in the active list, k * (n - k)
//         doubles of T, and each child's update (size^2 doubles, column-major)
// stdout: int32 status, int32 osize, then the slab as doubles
//         (T | L | U | X_rr^{-1} | Acc, or for k = n the passed-up update)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <hip/hip_runtime.h>
#include "simt_mfma.h"
#include "elim_dev.h"

using namespace kigpu;
using namespace kigpu::elim;

static void read_all(void* p, size_t bytes) {
  if (bytes && std::fread(p, 1, bytes, stdin) != bytes) {
    std::fprintf(stderr, "short input\n");
    std::exit(2);
  }
}

template <typename T, int RMAX>
static int run(int pde, int sd, int npts, int n, int k, int leaf, const int* csz) {
  std::vector<double> geo(6 * (size_t)npts);
  read_all(geo.data(), sizeof(double) * geo.size());
  double* gx[6];
  for (int q = 0; q < 6; ++q) {
    gx[q] = new double[npts];
    for (int i = 0; i < npts; ++i) gx[q][i] = geo[(size_t)q * npts + i];
  }
  const GeomView g{gx[0], gx[1], gx[2], gx[3], gx[4], gx[5], npts, sd};
  int* sr_pool = new int[n];
  int* pos_pool = new int[n];
  read_all(sr_pool, sizeof(int) * n);
  read_all(pos_pool, sizeof(int) * n);
  const int kk = k < n ? k : 0, r = n - k;
  std::vector<double> tin((size_t)kk * r);
  read_all(tin.data(), sizeof(double) * tin.size());
  // slab: exactly what the box writes (n^2 + k r entries; n^2 when passed up)
  const size_t slab_n = (size_t)n * n + (size_t)kk * r;
  T* slab = new T[slab_n];
  for (size_t e = 0; e < slab_n; ++e) slab[e] = T(0);
  for (size_t e = 0; e < tin.size(); ++e) slab[e] = (T)tin[e];
  // tree: node 0 and, unless it is a leaf, its four children 1..4 (an
  // empty child hands up nothing, as an empty box does)
  int child[20];
  for (int e = 0; e < 20; ++e) child[e] = -1;
  if (!leaf)
    for (int c = 0; c < 4; ++c) child[c] = c + 1;
  int is_leaf[5] = {leaf, 1, 1, 1, 1};
  int osize_in[5] = {0, 0, 0, 0, 0}, acc_ld_in[5] = {0, 0, 0, 0, 0};
  uint64_t acc_in[5] = {0, 0, 0, 0, 0};
  T* cacc[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int c = 0; c < 4; ++c) {
    if (csz[c] <= 0) continue;
    osize_in[c + 1] = csz[c];
    acc_ld_in[c + 1] = csz[c];
    std::vector<double> a((size_t)csz[c] * csz[c]);
    read_all(a.data(), sizeof(double) * a.size());
    cacc[c] = new T[a.size()];
    for (size_t e = 0; e < a.size(); ++e) cacc[c][e] = (T)a[e];
    acc_in[c + 1] = reinterpret_cast<uint64_t>(cacc[c]);
  }
  int nact[1] = {n}, rank[1] = {k < n ? k : -1};
  int64_t act_off[1] = {0};
  int* piv_pool = new int[n];
  uint64_t fac_ptr[1] = {reinterpret_cast<uint64_t>(slab)};
  int osize[5] = {-7, -7, -7, -7, -7};
  uint64_t acc_ptr[5] = {0, 0, 0, 0, 0};
  int acc_ld[5] = {0, 0, 0, 0, 0};
  int status = 0;
  unsigned char* smem;
  if (RMAX > 0) {  // the wavefront kernel: its LDS layout for this box
    smem = new unsigned char[k < n ? wdec_layout(n, k, sizeof(T)).total : 16];
    simt::launch(kWave, [&] {
      decouple_wave_box<T, (RMAX > 0 ? RMAX : 32)>(0, smem, g, pde, is_leaf, child, osize_in,
                                                   acc_in, acc_ld_in, nact, rank, act_off,
                                                   sr_pool, pos_pool, piv_pool, fac_ptr, osize,
                                                   acc_ptr, acc_ld, &status);
    });
  } else {  // the workgroup kernel: its working arrays for this box, one LDS int
    smem = new unsigned char[dec_layout(n, sizeof(T)).total];
    int* s_p = new int[1];
    simt::launch(kDecTB, [&] {
      decouple_block_box<T>(0, smem, s_p, g, pde, is_leaf, child, osize_in, acc_in, acc_ld_in,
                            nact, rank, act_off, sr_pool, pos_pool, piv_pool, fac_ptr, osize,
                            acc_ptr, acc_ld, &status);
    });
    delete[] s_p;
  }
  std::fwrite(&status, sizeof(int), 1, stdout);
  std::fwrite(&osize[0], sizeof(int), 1, stdout);
  for (size_t e = 0; e < slab_n; ++e) {
    const double v = (double)slab[e];
    std::fwrite(&v, sizeof(double), 1, stdout);
  }
  delete[] smem;
  delete[] slab;
  delete[] piv_pool;
  delete[] sr_pool;
  delete[] pos_pool;
  for (int c = 0; c < 4; ++c) delete[] cacc[c];
  for (int q = 0; q < 6; ++q) delete[] gx[q];
  return 0;
}

int main() {
  int h[12];
  read_all(h, sizeof(h));
  const int tsize = h[0], rmax = h[1], pde = h[2], sd = h[3], npts = h[4], n = h[5], k = h[6],
            leaf = h[7];
  const int* csz = h + 8;
  if (n < 1 || k < 1 || k > n || (rmax != 0 && rmax != 32 && rmax != 64) ||
      (rmax == 32 && n > 32) || (rmax == 64 && n > 64) || n > 512)
    return 2;
  if (tsize == 4) return rmax == 0    ? run<float, 0>(pde, sd, npts, n, k, leaf, csz)
                         : rmax == 32 ? run<float, 32>(pde, sd, npts, n, k, leaf, csz)
                                      : run<float, 64>(pde, sd, npts, n, k, leaf, csz);
  return rmax == 0    ? run<double, 0>(pde, sd, npts, n, k, leaf, csz)
         : rmax == 32 ? run<double, 32>(pde, sd, npts, n, k, leaf, csz)
                      : run<double, 64>(pde, sd, npts, n, k, leaf, csz);
}
