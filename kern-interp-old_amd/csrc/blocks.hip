// Dense kernel blocks K(tgt, src) on the GPU for the C++ Kernel drop-in
// (Kernel::operator(), kern_interp/kernel/kernel.cpp:99-150, :47-96).
#include <algorithm>

#include "entries.h"
#include "factor.h"

namespace kigpu {

namespace {

__global__ void k_block(GeomView g, const double* dpx, const double* dpy, int pde, int sd,
                        const int* tgt, int ntgt, const int* src, int nsrc, int forward,
                        double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntgt) return;
  for (int j = blockIdx.y; j < nsrc; j += gridDim.y) {
  const int td = tgt[i], sdf = src[j];
  double v;
  if (!forward) {
    v = kernel_entry(g, pde, td, sdf);
  } else {
    const int tp = td / sd, sp = sdf / sd;
    const double tx = dpx[tp], ty = dpy[tp];
    if (pde == KIGPU_STOKES) {
      v = stokes_entry(td, sdf, tx, ty, 0.0, 0.0, g.px[sp], g.py[sp], g.nx[sp], g.ny[sp], g.w[sp],
                       g.kappa[sp]);
    } else {
#pragma clang fp contract(off)
      v = laplace_entry(pde, tx - g.px[sp], ty - g.py[sp], g.w[sp], g.kappa[sp], g.nx[sp],
                        g.ny[sp], 0.0, 0.0, true);
    }
  }
  out[i + (size_t)j * ntgt] = v;
  }
}

}  // namespace

void kernel_block(Context& c, const GeomView& g, const double* dpx, const double* dpy, int pde,
                  int sd, const int* h_tgt, int ntgt, const int* h_src, int nsrc, int forward,
                  double* h_out) {
  hipStream_t s = c.stream;
  DBuf<int> dt, ds;
  DBuf<double> dout;
  dt.upload(h_tgt, ntgt, s);
  ds.upload(h_src, nsrc, s);
  dout.alloc((size_t)ntgt * nsrc);
  if (ntgt > 0 && nsrc > 0) {
    GeomView gv = g;
    gv.sd = sd;
    k_block<<<dim3(ceil_div(ntgt, 256), std::min(nsrc, 65535)), 256, 0, s>>>(gv, dpx, dpy, pde, sd, dt.p, ntgt,
                                                           ds.p, nsrc, forward, dout.p);
    KI_HIP(hipGetLastError());
  }
  dout.download(h_out, (size_t)ntgt * nsrc, s);
  stream_wait(s);
}

}  // namespace kigpu
