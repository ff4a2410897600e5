// Nystrom matrix entries of the double-layer kernels, evaluated on the device.
//
// Each formula follows kern_interp/kernel/kernel.cpp with the same sequence of
// IEEE double operations (fused multiply-add contraction is switched off in
// these functions), so a generated entry equals the CPU evaluation bit for
// bit.  The DOF -> point map is point = dof / sd, component = dof % sd
// (quadtree.cpp:167-170).
#pragma once

#include <hip/hip_runtime.h>

#include "geom.h"
#include "kigpu.h"

namespace kigpu {

constexpr double kInv2Pi = 1.0 / (2 * 3.14159265358979323846);  // (1.0 / (2 * M_PI))
constexpr double kInvPi = 1.0 / 3.14159265358979323846;         // 1.0 / (M_PI)

// Kernel::laplace (kernel.cpp:11-36) for r = t - s; `forward` picks the
// single-layer log form of the Neumann kernel at domain targets.
__device__ __forceinline__ double laplace_entry(int pde, double r1, double r2, double sw,
                                                double sc, double sn1, double sn2,
                                                double tn1, double tn2, bool forward) {
#pragma clang fp contract(off)
  if (pde == KIGPU_LAPLACE) {
    if (r1 == 0. && r2 == 0.) return 0.5 + 0.5 * sc * sw * kInv2Pi;
    return -sw * kInv2Pi * (r1 * sn1 + r2 * sn2) / (r1 * r1 + r2 * r2);
  }
  if (forward) return sw * kInv2Pi * log(sqrt(r1 * r1 + r2 * r2));
  if (r1 == 0. && r2 == 0.) return sw - 0.5 + 0.5 * sc * sw * kInv2Pi;
  return sw + sw * kInv2Pi * (r1 * tn1 + r2 * tn2) / (r1 * r1 + r2 * r2);
}

// Kernel::stokes_get entry (kernel.cpp:47-96) for target DOF td at (tx,ty)
// with normal (tnx,tny) and source DOF sd_ at point s.
__device__ __forceinline__ double stokes_entry(int td, int sdof, double tx, double ty,
                                               double tnx, double tny, double sx, double sy,
                                               double snx, double sny, double sw,
                                               double sc) {
#pragma clang fp contract(off)
  const double tn[2] = {tnx, tny};
  const double sn[2] = {snx, sny};
  if (tx == sx && ty == sy) {
    int parity = (td + sdof + 1) % 2;
    return -0.5 * parity + (1 - 2 * parity) * 0.5 * sc * sw * kInvPi * sn[(td + 1) % 2] *
                               sn[(sdof + 1) % 2] +
           sw * tn[td % 2] * sn[sdof % 2];
  }
  const double r[2] = {tx - sx, ty - sy};
  double rr = r[0] * r[0] + r[1] * r[1];
  return sw * kInvPi * (r[0] * sn[0] + r[1] * sn[1]) / (rr * rr) * r[td % 2] * r[sdof % 2] +
         sw * tn[td % 2] * sn[sdof % 2];
}

// K(tgt, src) on boundary DOFs (Kernel::operator(), kernel.cpp:99-111).
__device__ __forceinline__ double kernel_entry(const GeomView& g, int pde, int tdof, int sdof) {
#pragma clang fp contract(off)
  if (pde == KIGPU_STOKES) {
    int tp = tdof >> 1, sp = sdof >> 1;
    return stokes_entry(tdof, sdof, g.px[tp], g.py[tp], g.nx[tp], g.ny[tp], g.px[sp],
                        g.py[sp], g.nx[sp], g.ny[sp], g.w[sp], g.kappa[sp]);
  }
  double r1 = g.px[tdof] - g.px[sdof];
  double r2 = g.py[tdof] - g.py[sdof];
  return laplace_entry(pde, r1, r2, g.w[sdof], g.kappa[sdof], g.nx[sdof], g.ny[sdof],
                       g.nx[tdof], g.ny[tdof], false);
}

// Proxy-ring rows of the ID matrix (laplace_proxy_get kernel.cpp:310-334,
// stokes_proxy_get :268-307).  Row `row` of the 2P (Laplace) or 4P (Stokes)
// proxy block against box DOF `bdof`.  (cx, cy, rad) is the ring; (pc, ps) the
// unit vector of proxy point i, taken from a host libm table.
__device__ __forceinline__ double proxy_entry(const GeomView& g, int pde, int row, int bdof,
                                              double cx, double cy, double rad, double pw,
                                              const double* ring_c, const double* ring_s) {
#pragma clang fp contract(off)
  if (pde != KIGPU_STOKES) {
    int half = row >= kProxyPoints;
    int i = half ? row - kProxyPoints : row;
    double pc = ring_c[i], ps = ring_s[i];
    double ppx = cx + rad * pc;
    double ppy = cy + rad * ps;
    double r1 = ppx - g.px[bdof];
    double r2 = ppy - g.py[bdof];
    if (!half)
      return laplace_entry(pde, r1, r2, g.w[bdof], 0.0, g.nx[bdof], g.ny[bdof], pc, ps, false);
    return laplace_entry(pde, -r1, -r2, pw, 0.0, pc, ps, g.nx[bdof], g.ny[bdof], false);
  }
  // Stokes: rows [0, 2P) proxy <- box, rows [2P, 4P) box <- proxy; within a
  // half, row 2i+c is component c of proxy point i.
  int half = row >= 2 * kProxyPoints;
  int rr_ = half ? row - 2 * kProxyPoints : row;
  int i = rr_ >> 1, comp = rr_ & 1;
  double pc = ring_c[i], ps = ring_s[i];
  double ppx = cx + rad * pc;
  double ppy = cy + rad * ps;
  int bp = bdof >> 1;
  double sx = g.px[bp], sy = g.py[bp];
  double snx = g.nx[bp], sny = g.ny[bp], sw = g.w[bp];
  const double r[2] = {ppx - sx, ppy - sy};
  const double sn[2] = {snx, sny};
  const double tn[2] = {pc, ps};
  double rr = r[0] * r[0] + r[1] * r[1];
  int sc = bdof % 2;
  if (!half) {
    double potential = sw * kInvPi * (r[0] * sn[0] + r[1] * sn[1]) / (rr * rr);
    return potential * r[comp] * r[sc] + sw * tn[comp] * sn[sc];
  }
  double potential = -pw * kInvPi * (r[0] * tn[0] + r[1] * tn[1]) / (rr * rr);
  return potential * r[comp] * r[sc] + pw * sn[sc] * tn[comp];
}

}  // namespace kigpu
