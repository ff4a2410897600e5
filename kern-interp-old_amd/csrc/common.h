// Shared definitions for the HIP sources of libkigpu (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "geom.h"
#include "kigpu.h"
#include "lds_grants.h"

namespace kigpu {


// thread-local error detail returned by kigpu_last_error()
void set_error(const std::string& msg);

struct HipError {
  kigpu_status code;
};

#define KI_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      ::kigpu::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
      throw ::kigpu::HipError{e_ == hipErrorOutOfMemory ? KIGPU_ERR_NOMEM   \
                                                        : KIGPU_ERR_HIP};   \
    }                                                                       \
  } while (0)

#define KI_FAIL(code, msg)          \
  do {                              \
    ::kigpu::set_error(msg);        \
    throw ::kigpu::HipError{code};  \
  } while (0)

// Caching device allocator: freed blocks go to a free list keyed by their
// rounded size and are reused, so the per-level buffers of tree builds,
// factorizations and solves never call hipMalloc/hipFree in steady state.
// A freed block is reused only after the freeing thread calls
// dev_release_pending(), which it does once it has waited for every stream
// that may still touch the block (the end of each synchronous entry point).
void* dev_alloc(size_t bytes);
void dev_free(void* p, size_t bytes);
void dev_release_pending();
size_t dev_round(size_t bytes);

// owning device buffer
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  size_t cap = 0;  // bytes held
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) dev_free(p, cap);
    p = nullptr;
    n = 0;
    cap = 0;
  }
  void alloc(size_t count) {
    size_t bytes = count ? count * sizeof(T) : sizeof(T);
    if (p && bytes <= cap) {
      n = count;
      return;
    }
    release();
    cap = dev_round(bytes);
    p = static_cast<T*>(dev_alloc(cap));
    n = count;
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    alloc(count);
    if (count) KI_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
  void download(T* h, size_t count, hipStream_t s) const {
    if (count) KI_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s) {
    if (n) KI_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
};

// Wait for everything queued on stream s by polling hipStreamQuery.  The
// per-level waits of the tree build, factorization and solve are short; a
// blocking wait can put the host thread to sleep and wake it late.
inline void stream_wait(hipStream_t s) {
  hipError_t e;
  while ((e = hipStreamQuery(s)) == hipErrorNotReady) {
  }
  KI_HIP(e);
}

// Grow an event pool to n events.
inline void ensure_events(std::vector<hipEvent_t>& pool, size_t n) {
  while (pool.size() < n) {
    hipEvent_t e;
    KI_HIP(hipEventCreate(&e));
    pool.push_back(e);
  }
}

// Wait, the same way, for the work queued before event ev only.
inline void event_wait(hipEvent_t ev) {
  hipError_t e;
  while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
  }
  KI_HIP(e);
}

// Page-locked host memory (hipHostMalloc) for the small per-level transfers:
// copies from and to it are plain DMA, where pageable memory goes through a
// staging copy on every call.  Grow-only; growing waits for the stream, so a
// copy still in flight never loses its buffer.
struct PinnedHost {
  void* p = nullptr;
  void* dev = nullptr;  // the same memory seen from the device (kernels may write it)
  size_t cap = 0;
  PinnedHost() = default;
  PinnedHost(const PinnedHost&) = delete;
  PinnedHost& operator=(const PinnedHost&) = delete;
  ~PinnedHost() {
    if (p) (void)hipHostFree(p);
  }
  // device view of the block returned by the last get()
  template <typename T>
  T* dget() const {
    return static_cast<T*>(dev);
  }
  template <typename T>
  T* get(size_t count, hipStream_t s) {
    const size_t bytes = count * sizeof(T) + 64;
    if (bytes > cap) {
      if (p) {
        stream_wait(s);
        KI_HIP(hipHostFree(p));
        p = nullptr;
        cap = 0;
      }
      size_t nb = 4096;
      while (nb < bytes) nb *= 2;
      KI_HIP(hipHostMalloc(&p, nb, hipHostMallocMapped | hipHostMallocCoherent));
      KI_HIP(hipHostGetDevicePointer(&dev, p, 0));
      cap = nb;
    }
    return static_cast<T*>(p);
  }
};

// Device arrays copied into a mapped page-locked block by one kernel launch
// (byte offsets into the block; sizes multiples of 4): the per-level results
// the host needs, without the copy engine.  Read them after stream_wait.
struct ToHost {
  const void* src;
  size_t off;
  size_t bytes;
};
void copy_to_host(PinnedHost& pin, hipStream_t s, const ToHost* parts, int nparts);
inline void copy_to_host(PinnedHost& pin, hipStream_t s, std::initializer_list<ToHost> parts) {
  copy_to_host(pin, s, parts.begin(), (int)parts.size());
}

// Several host arrays sent to the device with one copy through page-locked
// memory: add() reserves a 256-byte aligned slot and returns its offset,
// send() packs the block, copies it into dst and returns the device base
// (slot o is at base + o).  The page-locked block is reused by the next
// send(), so the caller synchronizes the stream between two sends.
struct UploadPack {
  std::vector<std::pair<const void*, size_t>> parts;
  std::vector<size_t> offs;
  size_t total = 0;
  size_t add(const void* h, size_t bytes) {
    const size_t o = total;
    parts.push_back({h, bytes});
    offs.push_back(o);
    total = (o + bytes + 255) / 256 * 256;
    return o;
  }
  unsigned char* send(PinnedHost& pin, DBuf<unsigned char>& dst, hipStream_t s) {
    unsigned char* h = pin.get<unsigned char>(total, s);
    for (size_t i = 0; i < parts.size(); ++i)
      if (parts[i].second) std::memcpy(h + offs[i], parts[i].first, parts[i].second);
    dst.alloc(total);
    if (total) KI_HIP(hipMemcpyAsync(dst.p, h, total, hipMemcpyHostToDevice, s));
    return dst.p;
  }
};


// device-side scan helpers (scan.hip)
void exclusive_scan_i64(const int64_t* in, int64_t* out, int n, hipStream_t s);
// out[i] = add + in[0] + ... + in[i - 1]
void exclusive_scan_i32_to_i64(const int* in, int64_t* out, int n, hipStream_t s,
                               int64_t add = 0);
// the same with the length *dn read on the device
void exclusive_scan_i32_to_i64_dev(const int* in, int64_t* out, const int64_t* dn, hipStream_t s);

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

inline LdsGrants& lds_grants() {
  static LdsGrants g;
  return g;
}

// allow a kernel to use `bytes` of dynamic LDS (above the default limit).
// The attribute is set only when it grows for that kernel on the current
// device: the per-level launches ask again and again, and each set costs a
// runtime call on the host between two launches.
template <typename K>
inline void allow_lds(K kern, size_t bytes) {
  if (bytes == 0) return;
  int dev = 0;
  KI_HIP(hipGetDevice(&dev));
  lds_grants().allow(reinterpret_cast<const void*>(kern), dev, bytes,
                     [](const void* f, size_t b) {
                       KI_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  (int)b));
                     });
}

}  // namespace kigpu
