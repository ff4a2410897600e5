// Exclusive prefix sums over node-sized arrays (one workgroup, chunked).
#include "common.h"

namespace kigpu {

namespace {

constexpr int kScanTB = 1024;

template <typename TIn>
__global__ __launch_bounds__(kScanTB) void k_scan(const TIn* in, int64_t* out, int n) {
  __shared__ int64_t buf[kScanTB];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += kScanTB) {
    int i = base + threadIdx.x;
    int64_t v = i < n ? (int64_t)in[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanTB; off <<= 1) {
      int64_t t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < n) out[i] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == kScanTB - 1) carry += buf[kScanTB - 1];
    __syncthreads();
  }
}

}  // namespace

void exclusive_scan_i64(const int64_t* in, int64_t* out, int n, hipStream_t s) {
  if (n <= 0) return;
  k_scan<int64_t><<<1, kScanTB, 0, s>>>(in, out, n);
}

void exclusive_scan_i32_to_i64(const int* in, int64_t* out, int n, hipStream_t s) {
  if (n <= 0) return;
  k_scan<int><<<1, kScanTB, 0, s>>>(in, out, n);
}

}  // namespace kigpu
