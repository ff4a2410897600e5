#include "runtime.h"

#include <cstdlib>

namespace kern_interp {
namespace detail {

namespace {
// One context (device, two streams, staging buffers) for the process, and
// one lock that every device call of this API holds: calls from several
// threads of a caller (the reference's experiment three factors four trees
// from OpenMP threads) run one after another on the same two streams.  Each
// call fills the device on its own, so little is lost.
struct Holder {
  std::recursive_mutex mu;
  kigpu_ctx* ctx = nullptr;
  ~Holder() {
    if (ctx) kigpu_context_destroy(ctx);
  }
};
Holder g_holder;
}  // namespace

CtxRef context() {
  CtxRef r{std::unique_lock<std::recursive_mutex>(g_holder.mu), nullptr};
  if (!g_holder.ctx) {
    const char* env = std::getenv("KIGPU_DEVICE");
    const int dev = env ? std::atoi(env) : 0;
    check(kigpu_context_create(dev, &g_holder.ctx), "kigpu_context_create");
  }
  r.c = g_holder.ctx;
  return r;
}

}  // namespace detail
}  // namespace kern_interp
