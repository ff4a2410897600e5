// Exclusive prefix sums over node-sized arrays (one workgroup, chunked).
#include <algorithm>

#include "common.h"

namespace kigpu {

namespace {

constexpr int kScanTB = 1024;

template <typename TIn>
__global__ __launch_bounds__(kScanTB) void k_scan(const TIn* in, int64_t* out, int n,
                                                  int64_t add, const int64_t* dn) {
  __shared__ int64_t buf[kScanTB];
  __shared__ int64_t carry;
  if (dn) n = (int)*dn;  // the length, when only the device has it
  if (threadIdx.x == 0) carry = add;
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
  k_scan<int64_t><<<1, kScanTB, 0, s>>>(in, out, n, 0, nullptr);
}

void exclusive_scan_i32_to_i64(const int* in, int64_t* out, int n, hipStream_t s, int64_t add) {
  if (n <= 0) return;
  k_scan<int><<<1, kScanTB, 0, s>>>(in, out, n, add, nullptr);
}

void exclusive_scan_i32_to_i64_dev(const int* in, int64_t* out, const int64_t* dn, hipStream_t s) {
  k_scan<int><<<1, kScanTB, 0, s>>>(in, out, 0, 0, dn);
}

namespace {
struct HostParts {
  const unsigned* src[4];
  unsigned* dst[4];
  int64_t words[4];
};

__global__ void k_to_host(HostParts p, int n) {
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (int q = 0; q < n; ++q)
    for (int64_t i = i0; i < p.words[q]; i += st) p.dst[q][i] = p.src[q][i];
}
}  // namespace

void copy_to_host(PinnedHost& pin, hipStream_t s, const ToHost* parts, int nparts) {
  HostParts hp{};
  int64_t most = 0;
  int n = 0;
  for (int q = 0; q < nparts; ++q) {
    if (parts[q].bytes == 0) continue;
    if (n == 4) KI_FAIL(KIGPU_ERR_STATE, "copy_to_host: more than four parts");
    if (parts[q].bytes % 4 || parts[q].off % 4 || parts[q].off + parts[q].bytes > pin.cap)
      KI_FAIL(KIGPU_ERR_STATE, "copy_to_host: part outside the page-locked block");
    hp.src[n] = static_cast<const unsigned*>(parts[q].src);
    hp.dst[n] = reinterpret_cast<unsigned*>(static_cast<unsigned char*>(pin.dev) + parts[q].off);
    hp.words[n] = (int64_t)(parts[q].bytes / 4);
    most = std::max(most, hp.words[n]);
    ++n;
  }
  if (n == 0) return;
  const int blocks = (int)std::min<int64_t>(1024, (most + 255) / 256);
  k_to_host<<<blocks, 256, 0, s>>>(hp, n);
  KI_HIP(hipGetLastError());
}

}  // namespace kigpu
