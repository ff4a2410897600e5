// Device contexts of the C++ drop-in API.
#pragma once

#include <mutex>
#include <stdexcept>
#include <string>

#include "kigpu.h"

namespace kern_interp {
namespace detail {

// The process's kigpu context on the device named by $KIGPU_DEVICE (default
// 0), created on first use, destroyed at exit.  The returned reference holds
// the API's lock until it is destroyed, i.e. to the end of the full
// expression that makes the device call: device calls of several threads are
// serialized.
struct CtxRef {
  std::unique_lock<std::recursive_mutex> lk;
  kigpu_ctx* c;
  operator kigpu_ctx*() const { return c; }
};
CtxRef context();

inline void check(kigpu_status st, const char* what) {
  if (st != KIGPU_OK)
    throw std::runtime_error(std::string(what) + ": " + kigpu_status_string(st) + " (" +
                             kigpu_last_error() + ")");
}

}  // namespace detail
}  // namespace kern_interp
