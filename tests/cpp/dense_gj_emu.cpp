// CPU driver for the blocked Gauss-Jordan inverse of
// kern-interp-old_amd/csrc/dense_inv.h (dense_gj_block, the body of
// k_dense_gj), run on emulated GPU threads (tests/cpp/simt) and built with
// AddressSanitizer by tests/test_gj_emulated.py.  Each workgroup's dynamic
// LDS, the candidate buffers, the matrices, the global workspaces and the
// pivot array are separate heap blocks of exactly the size the launcher
// gives them, so an index past any of them is reported.
//
// stdin:  int32 layout (1 matrix in LDS, 0 in the global workspace),
//         int32 ni, ni int32 orders, then the ni matrices column-major,
//         back to back (the root launch's offsets)
// stdout: int32 status word, sum(n) int32 pivot rows, then the inverses
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <hip/hip_runtime.h>
#include "simt_mfma.h"
#include "dense_inv.h"

using namespace kigpu;
using namespace kigpu::gj;

static void read_all(void* p, size_t bytes) {
  if (std::fread(p, 1, bytes, stdin) != bytes) {
    std::fprintf(stderr, "short input\n");
    std::exit(2);
  }
}

template <int U, bool ONE, bool LDS>
static void run(int ni, const int* n_arr, const int64_t* doff, const int64_t* moff,
                const double* A, double* Ainv, double* ws, double* rws, int* piv, int* status) {
  for (int x = 0; x < ni; ++x) {
    unsigned char* smem = new unsigned char[gj_lds_bytes(n_arr[x], LDS)];
    GjCand* cand = new GjCand[2];
    simt::launch(kGjTB, [&] {
      dense_gj_block<U, ONE, LDS>(x, smem, cand, n_arr, doff, moff, 0, A, Ainv, ws, rws, piv,
                                status, 8);
    });
    delete[] cand;
    delete[] smem;
  }
}

int main() {
  int hdr[2];
  read_all(hdr, sizeof(hdr));
  const int lds = hdr[0], ni = hdr[1];
  if (ni <= 0) return 2;
  int* n_arr = new int[ni];
  read_all(n_arr, sizeof(int) * ni);
  int64_t* doff = new int64_t[ni + 1];
  int64_t* moff = new int64_t[ni + 1];
  doff[0] = moff[0] = 0;
  int nmax = 0;
  for (int x = 0; x < ni; ++x) {
    if (n_arr[x] < 0 || n_arr[x] > 1024) return 2;
    doff[x + 1] = doff[x] + n_arr[x];
    moff[x + 1] = moff[x] + (int64_t)n_arr[x] * n_arr[x];
    nmax = n_arr[x] > nmax ? n_arr[x] : nmax;
  }
  const size_t mtot = moff[ni] ? moff[ni] : 1, dtot = doff[ni] ? doff[ni] : 1;
  double* A = new double[mtot];
  read_all(A, sizeof(double) * moff[ni]);
  double* Ainv = new double[mtot];
  int* piv = new int[dtot];
  double* ws = lds ? nullptr : new double[mtot];
  double* rws = lds ? nullptr : new double[(size_t)kPB * dtot];
  int status = 0;
  // the launcher's choice of panel layout (dense_inv.hip)
  if (nmax <= kGjWave) {
    if (lds) run<1, true, true>(ni, n_arr, doff, moff, A, Ainv, ws, rws, piv, &status);
    else run<1, true, false>(ni, n_arr, doff, moff, A, Ainv, ws, rws, piv, &status);
  } else if (nmax <= 2 * kGjWave) {
    if (lds) run<2, true, true>(ni, n_arr, doff, moff, A, Ainv, ws, rws, piv, &status);
    else run<2, true, false>(ni, n_arr, doff, moff, A, Ainv, ws, rws, piv, &status);
  } else if (nmax <= kGjTB) {
    if (lds) run<4, true, true>(ni, n_arr, doff, moff, A, Ainv, ws, rws, piv, &status);
    else run<4, true, false>(ni, n_arr, doff, moff, A, Ainv, ws, rws, piv, &status);
  } else if (nmax <= 2 * kGjTB) {
    if (lds) run<2, false, true>(ni, n_arr, doff, moff, A, Ainv, ws, rws, piv, &status);
    else run<2, false, false>(ni, n_arr, doff, moff, A, Ainv, ws, rws, piv, &status);
  } else {
    if (lds) run<4, false, true>(ni, n_arr, doff, moff, A, Ainv, ws, rws, piv, &status);
    else run<4, false, false>(ni, n_arr, doff, moff, A, Ainv, ws, rws, piv, &status);
  }
  std::fwrite(&status, sizeof(int), 1, stdout);
  std::fwrite(piv, sizeof(int), doff[ni], stdout);
  std::fwrite(Ainv, sizeof(double), moff[ni], stdout);
  delete[] A;
  delete[] Ainv;
  delete[] piv;
  delete[] ws;
  delete[] rws;
  delete[] n_arr;
  delete[] doff;
  delete[] moff;
  return 0;
}
