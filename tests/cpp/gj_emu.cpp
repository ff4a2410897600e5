// CPU driver for the Gauss-Jordan inverses of
// kern-interp-old_amd/csrc/device_util.h (block_gj_inverse_perm of the
// workgroup elimination, wave_gj_reg of the one-wavefront elimination), run
// on emulated GPU threads (tests/cpp/simt) and
// built with AddressSanitizer by tests/test_gj_emulated.py.  Every array the
// routines touch -- the matrix and the LDS scratch vectors -- is a separate
// heap block of exactly the size the kernels give it (n elements, n * n for
// the matrix), so an index past any of them is reported.
//
// stdin:  int32 variant (0 perm, 3 register wave of wave_gj_reg, 4 the
//         same in FP32), int32 n, int32 threads,
//         then n * n doubles, column-major
// stdout: int32 ok, n int32 pivots, n * n doubles of the inverse as the
//         kernels write it out (for the perm variant with the column
//         permutation applied, as k_decouple does:
//         inv(i, j) = A[i + sigma[j] * n]; the register variants write the
//         inverse to a separate r x r array Z, as k_decouple_wave does)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "device_util.h"

using namespace kigpu;

static void read_all(void* p, size_t bytes) {
  if (std::fread(p, 1, bytes, stdin) != bytes) {
    std::fprintf(stderr, "short input\n");
    std::exit(2);
  }
}

int main() {
  int hdr[3];
  read_all(hdr, sizeof(hdr));
  const int variant = hdr[0], n = hdr[1], nthreads = hdr[2];
  if (n <= 0 || nthreads <= 0 || nthreads % 64 || (variant != 0 && variant < 3)) return 2;
  double* A = new double[(size_t)n * n];
  read_all(A, sizeof(double) * (size_t)n * n);
  int* piv = new int[n];
  int* sig = new int[n];
  double* prow = new double[n];
  double* orow = new double[n];
  double* fcol = new double[n];
  int* s_p = new int[1];
  int ok_all = 1;
  if (variant >= 3) {  // one wavefront, rows in registers, n <= 64
    if (n > 64 || nthreads != 64) return 2;
    double* Z = new double[(size_t)n * n];
    float* Af = new float[(size_t)n * n];
    float* Zf = new float[(size_t)n * n];
    double* rprow = new double[64];
    float* prowf = new float[64];
    int* pv = new int[64];
    for (size_t e = 0; e < (size_t)n * n; ++e) Af[e] = (float)A[e];
    simt::launch(64, [&] {
      bool ok;
      if (variant == 3) {
        if (n <= 16) ok = wave_gj_reg<double, 16>(A, n, Z, pv, rprow);
        else if (n <= 32) ok = wave_gj_reg<double, 32>(A, n, Z, pv, rprow);
        else ok = wave_gj_reg<double, 64>(A, n, Z, pv, rprow);
      } else {
        if (n <= 16) ok = wave_gj_reg<float, 16>(Af, n, Zf, pv, prowf);
        else if (n <= 32) ok = wave_gj_reg<float, 32>(Af, n, Zf, pv, prowf);
        else ok = wave_gj_reg<float, 64>(Af, n, Zf, pv, prowf);
      }
      if (threadIdx.x == 0) ok_all = ok;
    });
    if (variant == 4)
      for (size_t e = 0; e < (size_t)n * n; ++e) Z[e] = Zf[e];
    std::fwrite(&ok_all, sizeof(int), 1, stdout);
    std::fwrite(pv, sizeof(int), n, stdout);
    std::fwrite(Z, sizeof(double), (size_t)n * n, stdout);
    delete[] Z;
    delete[] Af;
    delete[] Zf;
    delete[] rprow;
    delete[] prowf;
    delete[] prow;
    delete[] pv;
    delete[] A;
    delete[] piv;
    delete[] sig;
    delete[] orow;
    delete[] fcol;
    delete[] s_p;
    return 0;
  }
  simt::launch(nthreads, [&] {
    bool ok;
    ok = block_gj_inverse_perm(A, n, piv, sig, prow, orow, fcol, s_p);
    if (threadIdx.x == 0) ok_all = ok;
  });
  double* out = new double[(size_t)n * n];
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      out[i + (size_t)j * n] = A[i + (size_t)sig[j] * n];
  std::fwrite(&ok_all, sizeof(int), 1, stdout);
  std::fwrite(piv, sizeof(int), n, stdout);
  std::fwrite(out, sizeof(double), (size_t)n * n, stdout);
  delete[] A;
  delete[] out;
  delete[] piv;
  delete[] sig;
  delete[] prow;
  delete[] orow;
  delete[] fcol;
  delete[] s_p;
  return 0;
}
