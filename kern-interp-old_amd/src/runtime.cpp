#include "runtime.h"

#include <cstdlib>
#include <mutex>

namespace kern_interp {
namespace detail {

namespace {
struct Holder {
  kigpu_ctx* ctx = nullptr;
  std::once_flag once;
  ~Holder() {
    if (ctx) kigpu_context_destroy(ctx);
  }
};
Holder g_holder;
}  // namespace

kigpu_ctx* context() {
  std::call_once(g_holder.once, [] {
    const char* env = std::getenv("KIGPU_DEVICE");
    const int dev = env ? std::atoi(env) : 0;
    check(kigpu_context_create(dev, &g_holder.ctx), "kigpu_context_create");
  });
  return g_holder.ctx;
}

}  // namespace detail
}  // namespace kern_interp
