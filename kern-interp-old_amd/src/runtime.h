// Process-wide device context of the C++ drop-in API.
#pragma once

#include <stdexcept>
#include <string>

#include "kigpu.h"

namespace kern_interp {
namespace detail {

// kigpu context on the device named by $KIGPU_DEVICE (default 0), created on
// first use and destroyed at exit
kigpu_ctx* context();

inline void check(kigpu_status st, const char* what) {
  if (st != KIGPU_OK)
    throw std::runtime_error(std::string(what) + ": " + kigpu_status_string(st) + " (" +
                             kigpu_last_error() + ")");
}

}  // namespace detail
}  // namespace kern_interp
