// Device-side basics shared by the kernels: the reference's kernel constants,
// the wavefront width and the boundary arrays as the kernels see them.
// Needs nothing but the HIP device built-ins, so that CPU tests can include
// the device routines built on it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kigpu {

constexpr int kProxyPoints = 128;     // NUM_PROXY_POINTS, kernel.h:10
constexpr double kRadiusRatio = 1.5;  // RADIUS_RATIO, kernel.h:11
constexpr int kMaxNbrs = 8;           // same-level neighbors of a box (8-connectivity)
constexpr int kWave = 64;

// Boundary arrays on the device, SoA for coalesced gathers.
struct GeomView {
  const double* px;
  const double* py;
  const double* nx;
  const double* ny;
  const double* w;
  const double* kappa;
  int npts;
  int sd;
};

// One box's ID work item (host list, read by the ID kernels).
struct IdWork {
  int node;
  int n;          // |active_box|
  int q;          // #near (inner-circle) or outside rows
  int m;          // ID matrix rows: 2q (+ proxy rows)
  int proxy;      // proxy block appended
  int in_lds;     // ID matrix kept in LDS
  int64_t rows_off;
  int64_t a_off;  // global workspace offset when !in_lds
  double cx, cy, rad, pw;
};

}  // namespace kigpu
