// extern "C" entry points of libkigpu (declared in include/kigpu.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <string>

#include "factor.h"

using namespace kigpu;

struct kigpu_ctx {
  Context c;
};

struct kigpu_tree {
  TreeData t;
};

struct kigpu_geom {
  int npts = 0, ndomain = 0;
  DBuf<double> px, py, nx, ny, w, kappa, dx, dy;
};

namespace kigpu {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace kigpu

namespace {

// Runs an entry point's body and maps exceptions to status codes.  With a
// context (entry points that wait for their work before returning), both of
// its streams are waited for once the body's temporaries are gone, and the
// device blocks freed on this thread become reusable.
template <typename F>
kigpu_status guard(F&& f, Context* settle = nullptr) {
  kigpu_status st = KIGPU_OK;
  try {
    f();
  } catch (const HipError& e) {
    st = e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    st = KIGPU_ERR_NOMEM;
  } catch (...) {
    set_error("unexpected exception");
    st = KIGPU_ERR_STATE;
  }
  if (settle) {
    if (hipStreamSynchronize(settle->stream) == hipSuccess &&
        hipStreamSynchronize(settle->side) == hipSuccess)
      dev_release_pending();
  }
  return st;
}

#define KI_REQUIRE(cond, msg)              \
  do {                                     \
    if (!(cond)) {                         \
      set_error(msg);                      \
      return KIGPU_ERR_ARG;                \
    }                                      \
  } while (0)

// get_descendent_neighbors wall test (quadtree.cpp:121-155), host side
bool wall_touch_host(const double* bc, const double* dc) {
  double top = bc[3], bottom = bc[1], left = bc[0], right = bc[4];
  for (int i = 0; i < 8; i += 2) {
    double x = dc[i], y = dc[i + 1];
    if (x > left && x < right) {
      if (std::fabs(y - top) < 1e-14) return true;
      if (std::fabs(y - bottom) < 1e-14) return true;
    }
    if (y < top && y < bottom) {
      if (std::fabs(x - left) < 1e-14) return true;
      if (std::fabs(x - right) < 1e-14) return true;
    }
  }
  return false;
}

void finer_pushes(const TreeData& t, int big, int small, std::vector<int>* out) {
  if (wall_touch_host(&t.h_corners[8 * (size_t)big], &t.h_corners[8 * (size_t)small]))
    out->push_back(small);
  for (int s = 0; s < 4; ++s) {
    int c = t.h_child[4 * (size_t)small + s];
    if (c >= 0) finer_pushes(t, big, c, out);
  }
}

std::vector<int> full_neighbors(const TreeData& t, int x) {
  std::vector<int> v;
  for (int64_t e = t.h_cp_off[x]; e < t.h_cp_off[x + 1]; ++e) v.push_back(t.h_cp_nbr[e]);
  for (int j = 0; j < t.h_sl_cnt[x]; ++j) v.push_back(t.h_sl_nbr[kMaxNbrs * (size_t)x + j]);
  if (t.h_is_leaf[x]) {
    for (int j = 0; j < t.h_sl_cnt[x]; ++j) {
      int nb = t.h_sl_nbr[kMaxNbrs * (size_t)x + j];
      for (int s = 0; s < 4; ++s) {
        int c = t.h_child[4 * (size_t)nb + s];
        if (c >= 0) finer_pushes(t, x, c, &v);
      }
    }
  }
  return v;
}

}  // namespace

static bool pde_matches(const kigpu_tree* tree, int pde) {
  if (pde != KIGPU_LAPLACE && pde != KIGPU_LAPLACE_NEUMANN && pde != KIGPU_STOKES) return false;
  return (pde == KIGPU_STOKES) == (tree->t.sd == 2);
}

extern "C" {

const char* kigpu_status_string(kigpu_status s) {
  switch (s) {
    case KIGPU_OK: return "ok";
    case KIGPU_ERR_ARG: return "invalid argument";
    case KIGPU_ERR_HIP: return "HIP runtime error";
    case KIGPU_ERR_NOMEM: return "out of device memory";
    case KIGPU_ERR_STATE: return "invalid state";
    case KIGPU_ERR_RANK: return "ID matrix has fewer rows than columns";
    case KIGPU_ERR_NOCOMPRESS: return "no compression possible";
    case KIGPU_ERR_NOTSUPP: return "not supported";
    case KIGPU_ERR_NOCONV: return "refinement did not converge";
    case KIGPU_ERR_SINGULAR: return "singular factorization (zero or non-finite pivot)";
    default: return "unknown status";
  }
}

const char* kigpu_last_error(void) { return g_err.c_str(); }

kigpu_status kigpu_context_create(int device, kigpu_ctx** out) {
  KI_REQUIRE(out, "null output handle");
  *out = nullptr;
  auto ctx = std::make_unique<kigpu_ctx>();
  kigpu_status st = guard([&] {
    Context& c = ctx->c;
    c.device = device;
    KI_HIP(hipSetDevice(device));
    KI_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    KI_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
    KI_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
    KI_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
    for (auto& e : c.ev) KI_HIP(hipEventCreate(&e));
    hipDeviceProp_t prop;
    KI_HIP(hipGetDeviceProperties(&prop, device));
    int lds_cu = 0;
    KI_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                 device));
    // a workgroup may use the whole LDS of its CU (160 KiB on gfx950)
    c.lds_max = std::max<size_t>(prop.sharedMemPerBlock, (size_t)lds_cu);
    // proxy ring unit vectors with the host libm (kernel.cpp:242-252)
    std::vector<double> ring(2 * kProxyPoints);
    for (int i = 0; i < kProxyPoints; ++i) {
      double ang = 2 * M_PI * i * (1.0 / kProxyPoints);
      ring[i] = std::cos(ang);
      ring[kProxyPoints + i] = std::sin(ang);
    }
    c.ring.upload(ring, c.stream);
    noalloc_init(device);  // the no-allocation solve's scratch address on this device
    stream_wait(c.stream);
  }, &ctx->c);
  if (st == KIGPU_OK) *out = ctx.release();
  return st;
}

kigpu_status kigpu_context_destroy(kigpu_ctx* ctx) {
  if (!ctx) return KIGPU_OK;
  (void)hipStreamSynchronize(ctx->c.stream);
  ctx->c.ring.release();
  for (auto& e : ctx->c.ev) (void)hipEventDestroy(e);
  (void)hipStreamSynchronize(ctx->c.side);
  (void)hipEventDestroy(ctx->c.ev_fork);
  (void)hipEventDestroy(ctx->c.ev_join);
  (void)hipStreamDestroy(ctx->c.side);
  (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
  return KIGPU_OK;
}

static void to_abi(const Tuning& t, kigpu_tuning* o) {
  o->version = KIGPU_TUNING_VERSION;
  for (int b = 0; b < 4; ++b) o->dec_wave[b] = t.dec_wave[b];
  o->mid_on_side = t.mid_on_side;
  o->pair_sweeps = t.pair_sweeps;
  o->apply_split_cap = t.apply_split_cap;
  o->fwd_split_cap = t.fwd_split_cap;
}

kigpu_status kigpu_get_tuning(kigpu_ctx* ctx, kigpu_tuning* out) {
  KI_REQUIRE(ctx && out, "null argument");
  to_abi(ctx->c.tune, out);
  return KIGPU_OK;
}

kigpu_status kigpu_set_tuning(kigpu_ctx* ctx, const kigpu_tuning* in) {
  KI_REQUIRE(ctx && in, "null argument");
  KI_REQUIRE(in->version == KIGPU_TUNING_VERSION, "tuning version mismatch");
  KI_REQUIRE(in->apply_split_cap >= 1 && in->apply_split_cap <= 4096 && in->fwd_split_cap >= 1 &&
                 in->fwd_split_cap <= 4096,
             "split caps must be 1 .. 4096");
  for (int b = 0; b < 4; ++b)
    KI_REQUIRE(in->dec_wave[b] == 0 || in->dec_wave[b] == 1, "dec_wave entries must be 0 or 1");
  Tuning& t = ctx->c.tune;
  for (int b = 0; b < 4; ++b) t.dec_wave[b] = in->dec_wave[b];
  t.mid_on_side = in->mid_on_side != 0;
  t.pair_sweeps = in->pair_sweeps != 0;
  t.apply_split_cap = in->apply_split_cap;
  t.fwd_split_cap = in->fwd_split_cap;
  return KIGPU_OK;
}

kigpu_status kigpu_autotune(kigpu_ctx* ctx, kigpu_tree* tree, const kigpu_factor_opts* opts,
                            int reps, kigpu_tuning* out) {
  KI_REQUIRE(ctx && tree && opts, "null argument");
  KI_REQUIRE(reps >= 1 && reps <= 100, "reps must be 1 .. 100");
  return guard([&] {
    Context& c = ctx->c;
    TreeData& t = tree->t;
    Tuning& tu = c.tune;
    auto best_of = [&](auto&& run) {
      double b = 1e30;
      for (int r = 0; r < reps; ++r) b = std::min(b, run());
      return b;
    };
    auto factor_time = [&] {
      factor_tree(c, t, *opts);
      return c.tim.t_decouple;
    };
    // elimination kernel per box-size class, then the side stream
    for (int b = 0; b < 4; ++b) {
      double tb[2];
      int bv = 0;
      for (int v = 0; v < 2; ++v) {
        tu.dec_wave[b] = v;
        tb[v] = best_of(factor_time);
        if (tb[v] < tb[bv]) bv = v;
      }
      tu.dec_wave[b] = bv;
    }
    {
      double ts[2];
      for (int v = 0; v < 2; ++v) {
        tu.mid_on_side = v;
        ts[v] = best_of(factor_time);
      }
      tu.mid_on_side = ts[1] <= ts[0];
    }
    factor_tree(c, t, *opts);
    // sweeps: one right-hand side of ones
    const size_t nd = t.ndof;
    std::vector<double> ones(nd, 1.0);
    DBuf<double> db, dx, dy;
    db.upload(ones, c.stream);
    dx.alloc(nd);
    dy.alloc(nd);
    {
      double tp[2];
      for (int v = 0; v < 2; ++v) {
        tu.pair_sweeps = v;
        tp[v] = best_of([&] {
          solve_tree(c, t, db.p, 1, dx.p, nullptr, nullptr, nullptr, false);
          return c.tim.t_solve;
        });
      }
      tu.pair_sweeps = tp[1] <= tp[0];
    }
    // direct K apply with one right-hand side: source splits (the Laplace
    // double layer takes the symmetric apply, which has none)
    if (t.opts.pde != KIGPU_LAPLACE) {
      const int caps[] = {16, 32, 64, 128, 256};
      double bt = 1e30;
      int bc = tu.apply_split_cap;
      for (int cap : caps) {
        tu.apply_split_cap = cap;
        const double tt = best_of([&] {
          KI_HIP(hipEventRecord(c.ev[8], c.stream));
          kernel_apply(c, t, t.opts.pde, db.p, 1, dy.p);
          KI_HIP(hipEventRecord(c.ev[9], c.stream));
          stream_wait(c.stream);
          float ms = 0;
          KI_HIP(hipEventElapsedTime(&ms, c.ev[8], c.ev[9]));
          return (double)ms;
        });
        if (tt < bt) {
          bt = tt;
          bc = cap;
        }
      }
      tu.apply_split_cap = bc;
    }
    if (out) to_abi(tu, out);
  }, &ctx->c);
}

kigpu_status kigpu_set_graphs(kigpu_ctx* ctx, int on) {
  KI_REQUIRE(ctx, "null context");
  ctx->c.graphs = on != 0;
  return KIGPU_OK;
}

kigpu_status kigpu_last_timings(kigpu_ctx* ctx, kigpu_timings* out) {
  KI_REQUIRE(ctx && out, "null argument");
  *out = ctx->c.tim;
  return KIGPU_OK;
}

static kigpu_status tree_build_common(kigpu_ctx* ctx, int npts, int sd, int leaf_dofs,
                                      kigpu_tree** out,
                                      const std::function<void(TreeData&, hipStream_t)>& load,
                                      const std::vector<int64_t>* inst_pt = nullptr,
                                      const TreeImport* imported = nullptr) {
  KI_REQUIRE(ctx && out, "null argument");
  KI_REQUIRE(npts > 0, "number of boundary points to init tree cannot be 0");
  KI_REQUIRE(sd == 1 || sd == 2, "solution dimension must be 1 or 2");
  KI_REQUIRE(leaf_dofs >= sd, "leaf_dofs must hold at least one point");
  KI_REQUIRE((int64_t)npts * sd < (1ll << 31), "too many DOFs for 32-bit indices");
  *out = nullptr;
  auto tr = std::make_unique<kigpu_tree>();
  Context& c = ctx->c;
  kigpu_status st = guard([&] {
    TreeData& t = tr->t;
    t.npts = npts;
    t.sd = sd;
    t.leaf_dofs = leaf_dofs;
    t.ndof = npts * sd;
    if (inst_pt) {
      t.inst_pt = *inst_pt;
      t.ninst = (int)inst_pt->size() - 1;
    } else {
      t.inst_pt = {0, (int64_t)npts};
      t.ninst = 1;
    }
    KI_HIP(hipEventRecord(c.ev[10], c.stream));
    load(t, c.stream);
    c.tim.syncs = 0;
    if (imported) import_tree(t, *imported, c.stream);
    else build_tree(t, c.stream, &c.tim.syncs, c.pin_down, c.evpool);
    KI_HIP(hipEventRecord(c.ev[11], c.stream));
    stream_wait(c.stream);
    float ms = 0;
    KI_HIP(hipEventElapsedTime(&ms, c.ev[10], c.ev[11]));
    c.tim.t_tree = ms * 1e-3;
  }, &ctx->c);
  if (st == KIGPU_OK) *out = tr.release();
  return st;
}

kigpu_status kigpu_tree_build(kigpu_ctx* ctx, const double* points, const double* normals,
                              const double* weights, const double* curvatures, int npts, int sd,
                              int leaf_dofs, kigpu_tree** out) {
  KI_REQUIRE(points && normals && weights && curvatures, "null geometry array");
  return tree_build_common(ctx, npts, sd, leaf_dofs, out, [&](TreeData& t, hipStream_t s) {
    std::vector<double> a(npts), b(npts), na(npts), nb(npts);
    for (int i = 0; i < npts; ++i) {
      a[i] = points[2 * i];
      b[i] = points[2 * i + 1];
      na[i] = normals[2 * i];
      nb[i] = normals[2 * i + 1];
    }
    t.px.upload(a, s);
    t.py.upload(b, s);
    t.nx.upload(na, s);
    t.ny.upload(nb, s);
    t.w.upload(weights, npts, s);
    t.kappa.upload(curvatures, npts, s);
    stream_wait(s);
  });
}

kigpu_status kigpu_tree_import(kigpu_ctx* ctx, const double* points, const double* normals,
                               const double* weights, const double* curvatures, int npts, int sd,
                               int leaf_dofs, int nlevels, const int64_t* level_off,
                               const int* parent, const int* child, const double* corners,
                               const double* side, const int64_t* box_off, const int* box_dofs,
                               const int64_t* nbr_off, const int* nbrs, double tree_min,
                               double tree_max, kigpu_tree** out) {
  KI_REQUIRE(points && normals && weights && curvatures, "null geometry array");
  KI_REQUIRE(nlevels >= 1 && level_off && parent && child && corners && side && box_off &&
                 box_dofs && nbr_off,
             "null tree array");
  KI_REQUIRE(nbrs || nbr_off[level_off[nlevels]] == 0, "null neighbor array");
  TreeImport in;
  in.nlevels = nlevels;
  in.level_off = level_off;
  in.parent = parent;
  in.child = child;
  in.corners = corners;
  in.side = side;
  in.box_off = box_off;
  in.box_dofs = box_dofs;
  in.nbr_off = nbr_off;
  in.nbrs = nbrs;
  in.min = tree_min;
  in.max = tree_max;
  return tree_build_common(
      ctx, npts, sd, leaf_dofs, out,
      [&](TreeData& t, hipStream_t s) {
        std::vector<double> a(npts), b(npts), na(npts), nb(npts);
        for (int i = 0; i < npts; ++i) {
          a[i] = points[2 * i];
          b[i] = points[2 * i + 1];
          na[i] = normals[2 * i];
          nb[i] = normals[2 * i + 1];
        }
        t.px.upload(a, s);
        t.py.upload(b, s);
        t.nx.upload(na, s);
        t.ny.upload(nb, s);
        t.w.upload(weights, npts, s);
        t.kappa.upload(curvatures, npts, s);
        stream_wait(s);
      },
      nullptr, &in);
}

kigpu_status kigpu_tree_build_device(kigpu_ctx* ctx, const double* d_px, const double* d_py,
                                     const double* d_nx, const double* d_ny, const double* d_w,
                                     const double* d_kappa, int npts, int sd, int leaf_dofs,
                                     kigpu_tree** out) {
  KI_REQUIRE(d_px && d_py && d_nx && d_ny && d_w && d_kappa, "null geometry array");
  return tree_build_common(ctx, npts, sd, leaf_dofs, out, [&](TreeData& t, hipStream_t s) {
    const size_t b = sizeof(double) * npts;
    DBuf<double>* dst[6] = {&t.px, &t.py, &t.nx, &t.ny, &t.w, &t.kappa};
    const double* src[6] = {d_px, d_py, d_nx, d_ny, d_w, d_kappa};
    for (int i = 0; i < 6; ++i) {
      dst[i]->alloc(npts);
      KI_HIP(hipMemcpyAsync(dst[i]->p, src[i], b, hipMemcpyDeviceToDevice, s));
    }
  });
}

kigpu_status kigpu_tree_build_batch(kigpu_ctx* ctx, int ninst, const int* npts,
                                    const double* points, const double* normals,
                                    const double* weights, const double* curvatures, int sd,
                                    int leaf_dofs, kigpu_tree** out) {
  KI_REQUIRE(ninst > 0 && npts, "invalid instance count");
  KI_REQUIRE(points && normals && weights && curvatures, "null geometry array");
  std::vector<int64_t> off(ninst + 1, 0);
  for (int i = 0; i < ninst; ++i) {
    KI_REQUIRE(npts[i] > 0, "number of boundary points to init tree cannot be 0");
    off[i + 1] = off[i] + npts[i];
  }
  KI_REQUIRE(off[ninst] < (1ll << 30), "too many points in the batch");
  const int total = (int)off[ninst];
  return tree_build_common(
      ctx, total, sd, leaf_dofs, out,
      [&](TreeData& t, hipStream_t s) {
        std::vector<double> a(total), b(total), na(total), nb(total);
        for (int i = 0; i < total; ++i) {
          a[i] = points[2 * i];
          b[i] = points[2 * i + 1];
          na[i] = normals[2 * i];
          nb[i] = normals[2 * i + 1];
        }
        t.px.upload(a, s);
        t.py.upload(b, s);
        t.nx.upload(na, s);
        t.ny.upload(nb, s);
        t.w.upload(weights, total, s);
        t.kappa.upload(curvatures, total, s);
        stream_wait(s);
      },
      &off);
}

kigpu_status kigpu_tree_build_batch_device(kigpu_ctx* ctx, int ninst, const int* npts,
                                           const double* d_px, const double* d_py,
                                           const double* d_nx, const double* d_ny,
                                           const double* d_w, const double* d_kappa, int sd,
                                           int leaf_dofs, kigpu_tree** out) {
  KI_REQUIRE(ninst > 0 && npts, "invalid instance count");
  KI_REQUIRE(d_px && d_py && d_nx && d_ny && d_w && d_kappa, "null geometry array");
  std::vector<int64_t> off(ninst + 1, 0);
  for (int i = 0; i < ninst; ++i) {
    KI_REQUIRE(npts[i] > 0, "number of boundary points to init tree cannot be 0");
    off[i + 1] = off[i] + npts[i];
  }
  KI_REQUIRE(off[ninst] < (1ll << 30), "too many points in the batch");
  const int total = (int)off[ninst];
  return tree_build_common(
      ctx, total, sd, leaf_dofs, out,
      [&](TreeData& t, hipStream_t s) {
        const size_t b = sizeof(double) * total;
        DBuf<double>* dst[6] = {&t.px, &t.py, &t.nx, &t.ny, &t.w, &t.kappa};
        const double* src[6] = {d_px, d_py, d_nx, d_ny, d_w, d_kappa};
        for (int i = 0; i < 6; ++i) {
          dst[i]->alloc(total);
          KI_HIP(hipMemcpyAsync(dst[i]->p, src[i], b, hipMemcpyDeviceToDevice, s));
        }
      },
      &off);
}

kigpu_status kigpu_tree_instances(const kigpu_tree* tree, int* ninst, int64_t* point_off,
                                  double* bounds, int* root_skel) {
  KI_REQUIRE(tree && ninst, "null argument");
  const TreeData& t = tree->t;
  *ninst = t.ninst;
  for (int i = 0; i < t.ninst; ++i) {
    if (point_off) point_off[i] = t.inst_pt[i];
    if (bounds) {
      bounds[2 * i] = t.inst_min[i];
      bounds[2 * i + 1] = t.inst_max[i];
    }
    if (root_skel) root_skel[i] = t.factored ? t.h_root_n[i] : -1;
  }
  if (point_off) point_off[t.ninst] = t.inst_pt[t.ninst];
  return KIGPU_OK;
}

kigpu_status kigpu_tree_destroy(kigpu_tree* tree) {
  // device blocks go back to the allocator cache: no work may still use them
  (void)hipDeviceSynchronize();
  delete tree;
  return KIGPU_OK;
}

kigpu_status kigpu_tree_get_info(const kigpu_tree* tree, kigpu_tree_info* info) {
  KI_REQUIRE(tree && info, "null argument");
  const TreeData& t = tree->t;
  info->num_levels = t.nlevels;
  info->num_nodes = t.nnodes;
  info->num_points = t.npts;
  info->solution_dim = t.sd;
  info->leaf_dofs = t.leaf_dofs;
  info->min = t.min;
  info->max = t.max;
  return KIGPU_OK;
}

kigpu_status kigpu_tree_export_nodes(const kigpu_tree* tree, int64_t* level_offsets, int* parent,
                                     int* children, int* is_leaf, double* corners, double* side) {
  KI_REQUIRE(tree, "null tree");
  kigpu_status st = guard([&] { ensure_tree_mirrors(const_cast<kigpu_tree*>(tree)->t); });
  if (st != KIGPU_OK) return st;
  const TreeData& t = tree->t;
  const size_t nn = t.nnodes;
  if (level_offsets) std::memcpy(level_offsets, t.level_off.data(), sizeof(int64_t) * (t.nlevels + 1));
  if (parent) std::memcpy(parent, t.h_parent.data(), sizeof(int) * nn);
  if (children) std::memcpy(children, t.h_child.data(), sizeof(int) * 4 * nn);
  if (is_leaf) std::memcpy(is_leaf, t.h_is_leaf.data(), sizeof(int) * nn);
  if (corners) std::memcpy(corners, t.h_corners.data(), sizeof(double) * 8 * nn);
  if (side) std::memcpy(side, t.h_side.data(), sizeof(double) * nn);
  return KIGPU_OK;
}

kigpu_status kigpu_tree_export_boxes(const kigpu_tree* tree, int64_t* offsets, int* dofs) {
  KI_REQUIRE(tree && offsets, "null argument");
  kigpu_status st = guard([&] { ensure_tree_mirrors(const_cast<kigpu_tree*>(tree)->t); });
  if (st != KIGPU_OK) return st;
  const TreeData& t = tree->t;
  int64_t o = 0;
  for (int x = 0; x < t.nnodes; ++x) {
    offsets[x] = o;
    const int cnt = t.h_pcount[x];
    if (dofs) {
      const int* pp = &t.h_perm[t.h_pstart[x]];
      for (int j = 0; j < cnt; ++j)
        for (int c = 0; c < t.sd; ++c) dofs[o + (int64_t)t.sd * j + c] = t.sd * pp[j] + c;
    }
    o += (int64_t)cnt * t.sd;
  }
  offsets[t.nnodes] = o;
  return KIGPU_OK;
}

kigpu_status kigpu_tree_export_neighbors(const kigpu_tree* tree, int64_t* offsets, int* nbrs) {
  KI_REQUIRE(tree && offsets, "null argument");
  kigpu_status st = guard([&] { ensure_tree_mirrors(const_cast<kigpu_tree*>(tree)->t); });
  if (st != KIGPU_OK) return st;
  const TreeData& t = tree->t;
  int64_t o = 0;
  for (int x = 0; x < t.nnodes; ++x) {
    offsets[x] = o;
    std::vector<int> v = full_neighbors(t, x);
    if (nbrs) std::memcpy(nbrs + o, v.data(), sizeof(int) * v.size());
    o += (int64_t)v.size();
  }
  offsets[t.nnodes] = o;
  return KIGPU_OK;
}

kigpu_status kigpu_set_border(kigpu_tree* tree, const double* U, const double* Psi, int nborder) {
  KI_REQUIRE(tree, "null tree");
  KI_REQUIRE(nborder >= 0, "negative border size");
  TreeData& t = tree->t;
  if (t.ninst > 1 && nborder > 0) {
    set_error("holes are not supported for a batch of instances");
    return KIGPU_ERR_NOTSUPP;
  }
  if (nborder == 0) {
    t.nborder = 0;
    t.h_U.clear();
    t.h_Psi.clear();
    return KIGPU_OK;
  }
  KI_REQUIRE(U && Psi, "null border matrix");
  const size_t len = (size_t)t.ndof * nborder;
  t.h_U.assign(U, U + len);
  t.h_Psi.assign(Psi, Psi + len);
  t.nborder = nborder;
  t.factored = false;
  return KIGPU_OK;
}

kigpu_status kigpu_factor(kigpu_ctx* ctx, kigpu_tree* tree, const kigpu_factor_opts* opts) {
  KI_REQUIRE(ctx && tree && opts, "null argument");
  KI_REQUIRE(opts->id_tol > 0, "id_tol must be greater than zero");
  KI_REQUIRE(opts->pde >= KIGPU_LAPLACE && opts->pde <= KIGPU_STOKES, "unknown pde");
  KI_REQUIRE((opts->pde == KIGPU_STOKES) == (tree->t.sd == 2),
             "solution dimension does not match the pde");
  KI_REQUIRE(opts->precision == KIGPU_FP64 || opts->precision == KIGPU_FP32_REFINE,
             "unknown precision");
  return guard([&] {
    kigpu_factor_opts o = *opts;
    if (o.min_dofs <= 0) o.min_dofs = 16;
    tree->t.factored = false;
    factor_tree(ctx->c, tree->t, o);
  }, &ctx->c);
}

kigpu_status kigpu_factor_ranks(const kigpu_tree* tree, int* rank, int* active) {
  KI_REQUIRE(tree, "null tree");
  const TreeData& t = tree->t;
  if (!t.factored) {
    set_error("tree is not factored");
    return KIGPU_ERR_STATE;
  }
  if (rank) std::memcpy(rank, t.h_rank.data(), sizeof(int) * t.nnodes);
  if (active) std::memcpy(active, t.h_nact.data(), sizeof(int) * t.nnodes);
  return KIGPU_OK;
}

kigpu_status kigpu_factor_export_lists(const kigpu_tree* tree, int64_t* skel_off, int* skel,
                                       int64_t* red_off, int* red) {
  KI_REQUIRE(tree && skel_off && red_off, "null argument");
  const TreeData& t = tree->t;
  if (!t.factored) {
    set_error("tree is not factored");
    return KIGPU_ERR_STATE;
  }
  kigpu_status st = guard([&] { ensure_factor_mirrors(const_cast<kigpu_tree*>(tree)->t); });
  if (st != KIGPU_OK) return st;
  int64_t os = 0, orr = 0;
  for (int x = 0; x < t.nnodes; ++x) {
    skel_off[x] = os;
    red_off[x] = orr;
    const int k = t.h_rank[x];
    if (k < 0) continue;
    const int n = t.h_nact[x];
    const int* sr = &t.h_sr[t.h_act_off[x]];
    if (skel) std::memcpy(skel + os, sr, sizeof(int) * k);
    if (red) std::memcpy(red + orr, sr + k, sizeof(int) * (n - k));
    os += k;
    orr += n - k;
  }
  skel_off[t.nnodes] = os;
  red_off[t.nnodes] = orr;
  return KIGPU_OK;
}

kigpu_status kigpu_factor_export_active(const kigpu_tree* tree, int64_t* act_off, int* active,
                                        int64_t* perm_off, int* perm) {
  KI_REQUIRE(tree && act_off, "null argument");
  const TreeData& t = tree->t;
  if (!t.factored) {
    set_error("tree is not factored");
    return KIGPU_ERR_STATE;
  }
  kigpu_status st = guard([&] { ensure_factor_mirrors(const_cast<kigpu_tree*>(tree)->t); });
  if (st != KIGPU_OK) return st;
  int64_t oa = 0, op = 0;
  for (int x = 0; x < t.nnodes; ++x) {
    act_off[x] = oa;
    if (perm_off) perm_off[x] = op;
    const int n = t.h_nact[x];
    if (active && n) std::memcpy(active + oa, &t.h_act[t.h_act_off[x]], sizeof(int) * n);
    oa += n;
    if (t.h_rank[x] >= 0) {
      if (perm) std::memcpy(perm + op, &t.h_pos[t.h_act_off[x]], sizeof(int) * n);
      op += n;
    }
  }
  act_off[t.nnodes] = oa;
  if (perm_off) perm_off[t.nnodes] = op;
  return KIGPU_OK;
}

kigpu_status kigpu_factor_allskel(const kigpu_tree* tree, int* n, int* allskel) {
  KI_REQUIRE(tree && n, "null argument");
  const TreeData& t = tree->t;
  if (!t.factored) {
    set_error("tree is not factored");
    return KIGPU_ERR_STATE;
  }
  *n = t.nroot;
  kigpu_status st = guard([&] { ensure_factor_mirrors(const_cast<kigpu_tree*>(tree)->t); });
  if (st != KIGPU_OK) return st;
  if (allskel && t.nroot) std::memcpy(allskel, &t.h_act[t.h_act_off[0]], sizeof(int) * t.nroot);
  return KIGPU_OK;
}

kigpu_status kigpu_solve_device(kigpu_ctx* ctx, kigpu_tree* tree, const double* d_b, int nrhs,
                                double* d_x, double* d_alpha, const kigpu_refine_opts* refine,
                                kigpu_solve_info* info) {
  KI_REQUIRE(ctx && tree && d_b && d_x, "null argument");
  KI_REQUIRE(nrhs > 0, "nrhs must be positive");
  if (!tree->t.factored) {
    set_error("solve before factor");
    return KIGPU_ERR_STATE;
  }
  return guard([&] { solve_tree(ctx->c, tree->t, d_b, nrhs, d_x, d_alpha, refine, info); },
               &ctx->c);
}

kigpu_status kigpu_solve_device_noalloc(kigpu_ctx* ctx, kigpu_tree* tree, const double* d_b,
                                        int nrhs, double* d_x) {
  KI_REQUIRE(ctx && tree && d_b && d_x, "null argument");
  KI_REQUIRE(nrhs > 0, "nrhs must be positive");
  if (!tree->t.factored) {
    set_error("solve before factor");
    return KIGPU_ERR_STATE;
  }
  return guard([&] { solve_tree_noalloc(ctx->c, tree->t, d_b, nrhs, d_x); });
}

kigpu_status kigpu_solve(kigpu_ctx* ctx, kigpu_tree* tree, const double* b, int ldb, int nrhs,
                         double* x, int ldx, double* alpha, const kigpu_refine_opts* refine,
                         kigpu_solve_info* info) {
  KI_REQUIRE(ctx && tree && b && x, "null argument");
  KI_REQUIRE(nrhs > 0, "nrhs must be positive");
  const int nd = tree->t.ndof;
  KI_REQUIRE(ldb >= nd && ldx >= nd, "leading dimension smaller than the number of DOFs");
  if (!tree->t.factored) {
    set_error("solve before factor");
    return KIGPU_ERR_STATE;
  }
  return guard([&] {
    Context& c = ctx->c;
    DBuf<double> db, dx, da;
    db.alloc((size_t)nd * nrhs);
    dx.alloc((size_t)nd * nrhs);
    const int h = tree->t.nborder;
    if (h) da.alloc((size_t)h * nrhs);
    KI_HIP(hipMemcpy2DAsync(db.p, sizeof(double) * nd, b, sizeof(double) * ldb,
                            sizeof(double) * nd, nrhs, hipMemcpyHostToDevice, c.stream));
    solve_tree(c, tree->t, db.p, nrhs, dx.p, h ? da.p : nullptr, refine, info);
    KI_HIP(hipMemcpy2DAsync(x, sizeof(double) * ldx, dx.p, sizeof(double) * nd,
                            sizeof(double) * nd, nrhs, hipMemcpyDeviceToHost, c.stream));
    if (h && alpha) da.download(alpha, (size_t)h * nrhs, c.stream);
    stream_wait(c.stream);
  }, &ctx->c);
}

kigpu_status kigpu_geometry_create(kigpu_ctx* ctx, const double* points, const double* normals,
                                   const double* weights, const double* curvatures, int npts,
                                   const double* domain_points, int ndomain, kigpu_geom** out) {
  KI_REQUIRE(ctx && out && npts >= 0 && ndomain >= 0, "invalid argument");
  KI_REQUIRE(npts == 0 || (points && normals && weights && curvatures), "null geometry array");
  KI_REQUIRE(ndomain == 0 || domain_points, "null domain points");
  *out = nullptr;
  auto g = std::make_unique<kigpu_geom>();
  kigpu_status st = guard([&] {
    hipStream_t s = ctx->c.stream;
    g->npts = npts;
    g->ndomain = ndomain;
    std::vector<double> a(npts), b(npts), na(npts), nb(npts), da(ndomain), db(ndomain);
    for (int i = 0; i < npts; ++i) {
      a[i] = points[2 * i];
      b[i] = points[2 * i + 1];
      na[i] = normals[2 * i];
      nb[i] = normals[2 * i + 1];
    }
    for (int i = 0; i < ndomain; ++i) {
      da[i] = domain_points[2 * i];
      db[i] = domain_points[2 * i + 1];
    }
    g->px.upload(a, s);
    g->py.upload(b, s);
    g->nx.upload(na, s);
    g->ny.upload(nb, s);
    g->w.upload(weights, npts, s);
    g->kappa.upload(curvatures, npts, s);
    g->dx.upload(da, s);
    g->dy.upload(db, s);
    stream_wait(s);
  }, &ctx->c);
  if (st == KIGPU_OK) *out = g.release();
  return st;
}

kigpu_status kigpu_geometry_destroy(kigpu_geom* geom) {
  (void)hipDeviceSynchronize();
  delete geom;
  return KIGPU_OK;
}

kigpu_status kigpu_kernel_block(kigpu_ctx* ctx, const kigpu_geom* geom, int pde, int sd,
                                const int* tgt, int ntgt, const int* src, int nsrc, int forward,
                                double* out) {
  KI_REQUIRE(ctx && geom && out && ntgt >= 0 && nsrc >= 0, "invalid argument");
  KI_REQUIRE((ntgt == 0 || tgt) && (nsrc == 0 || src), "null index list");
  KI_REQUIRE(sd == 1 || sd == 2, "solution dimension must be 1 or 2");
  const int lim_t = (forward ? geom->ndomain : geom->npts) * sd;
  for (int i = 0; i < ntgt; ++i) KI_REQUIRE(tgt[i] >= 0 && tgt[i] < lim_t, "target index out of range");
  for (int j = 0; j < nsrc; ++j)
    KI_REQUIRE(src[j] >= 0 && src[j] < geom->npts * sd, "source index out of range");
  return guard([&] {
    const GeomView g{geom->px.p, geom->py.p, geom->nx.p, geom->ny.p, geom->w.p,
                     geom->kappa.p, geom->npts, sd};
    kernel_block(ctx->c, g, geom->dx.p, geom->dy.p, pde, sd, tgt, ntgt, src, nsrc, forward, out);
  });
}

kigpu_status kigpu_proxy_block(kigpu_ctx* ctx, const kigpu_geom* geom, int pde,
                               const double* pxy_p, const double* pxy_n, int np, double pxy_w,
                               const int* box, int nbox, double* out) {
  KI_REQUIRE(ctx && geom && pxy_p && pxy_n && out && np > 0 && nbox >= 0, "invalid argument");
  KI_REQUIRE(nbox == 0 || box, "null index list");
  KI_REQUIRE(pde >= KIGPU_LAPLACE && pde <= KIGPU_STOKES, "unknown pde");
  KI_REQUIRE(nbox <= 65535, "at most 65535 box DOFs");
  const int sd = pde == KIGPU_STOKES ? 2 : 1;
  for (int j = 0; j < nbox; ++j)
    KI_REQUIRE(box[j] >= 0 && box[j] < geom->npts * sd, "box index out of range");
  return guard([&] {
    const GeomView g{geom->px.p, geom->py.p, geom->nx.p, geom->ny.p, geom->w.p,
                     geom->kappa.p, geom->npts, sd};
    proxy_block(ctx->c, g, pde, pxy_p, pxy_n, np, pxy_w, box, nbox, out);
  }, &ctx->c);
}

kigpu_status kigpu_interp_decomp(kigpu_ctx* ctx, const double* A, int m, int n, double tol,
                                 int* k, int* perm, double* T) {
  KI_REQUIRE(ctx && A && k && m > 0 && n > 0 && tol > 0, "invalid argument");
  KI_REQUIRE((int64_t)m * n < (1ll << 31), "matrix too large");
  return guard([&] { interp_decomp(ctx->c, A, m, n, tol, k, perm, T); }, &ctx->c);
}

kigpu_status kigpu_schur_eliminate(kigpu_ctx* ctx, const double* K_BN, int n, int k,
                                   const int* perm, const double* T, double* X_rr,
                                   double* X_rr_lu, int* X_rr_piv, double* L, double* U,
                                   double* schur_update) {
  KI_REQUIRE(ctx && K_BN && perm && T && n > 1 && k >= 1 && k < n, "invalid argument");
  KI_REQUIRE(n - k <= 1024, "at most 1024 redundant DOFs");
  std::vector<int> seen(n, 0);
  for (int i = 0; i < n; ++i) {
    KI_REQUIRE(perm[i] >= 0 && perm[i] < n && !seen[perm[i]], "perm is not a permutation");
    seen[perm[i]] = 1;
  }
  int singular = 0;
  kigpu_status st = guard([&] {
    schur_eliminate(ctx->c, K_BN, n, k, perm, T, X_rr, X_rr_lu, X_rr_piv, L, U, schur_update,
                    &singular);
  }, &ctx->c);
  if (st == KIGPU_OK && singular) {
    set_error("zero or non-finite pivot in the LU of X_rr");
    return KIGPU_ERR_SINGULAR;
  }
  return st;
}

kigpu_status kigpu_kernel_apply_device(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                       const double* d_mu, int nrhs, double* d_y) {
  KI_REQUIRE(ctx && tree && d_mu && d_y && nrhs > 0, "invalid argument");
  KI_REQUIRE(pde_matches(tree, pde), "pde does not match the tree's solution dimension");
  return guard([&] { kernel_apply(ctx->c, tree->t, pde, d_mu, nrhs, d_y); });
}

kigpu_status kigpu_kernel_apply_treecode(kigpu_ctx* ctx, const kigpu_tree* tree,
                                         const double* mu, int nrhs, double* y) {
  KI_REQUIRE(ctx && tree && mu && y && nrhs > 0, "invalid argument");
  KI_REQUIRE(tree->t.sd == 1 && treecode_usable(tree->t),
             "the treecode apply takes the Laplace double layer (one DOF per node)");
  return guard([&] {
    Context& c = ctx->c;
    const size_t len = (size_t)tree->t.ndof * nrhs;
    DBuf<double> dm, dy;
    dm.upload(mu, len, c.stream);
    dy.alloc(len);
    const hipEvent_t tev[3] = {c.ev[12], c.ev[14], c.ev[13]};
    treecode_apply(c, tree->t, dm.p, nrhs, dy.p, tev);
    dy.download(y, len, c.stream);
    stream_wait(c.stream);
    treecode_timing(c, tev);
  }, &ctx->c);
}

kigpu_status kigpu_set_refine_apply(kigpu_ctx* ctx, int mode) {
  KI_REQUIRE(ctx && (mode == 0 || mode == 1), "invalid argument");
  ctx->c.refine_apply = mode;
  return KIGPU_OK;
}

kigpu_status kigpu_set_id_method(kigpu_ctx* ctx, int method) {
  KI_REQUIRE(ctx && method >= 0 && method <= 2, "invalid argument");
  ctx->c.id_method = method;
  return KIGPU_OK;
}

kigpu_status kigpu_kernel_apply(kigpu_ctx* ctx, const kigpu_tree* tree, int pde, const double* mu,
                                int nrhs, double* y) {
  KI_REQUIRE(ctx && tree && mu && y && nrhs > 0, "invalid argument");
  KI_REQUIRE(pde_matches(tree, pde), "pde does not match the tree's solution dimension");
  return guard([&] {
    Context& c = ctx->c;
    const size_t len = (size_t)tree->t.ndof * nrhs;
    DBuf<double> dm, dy;
    dm.upload(mu, len, c.stream);
    dy.alloc(len);
    kernel_apply(c, tree->t, pde, dm.p, nrhs, dy.p);
    dy.download(y, len, c.stream);
    stream_wait(c.stream);
  }, &ctx->c);
}

kigpu_status kigpu_forward_apply(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                 const double* targets, int ntgt, const double* mu, int nrhs,
                                 double* u) {
  KI_REQUIRE(ctx && tree && targets && mu && u && ntgt > 0 && nrhs > 0, "invalid argument");
  KI_REQUIRE(pde_matches(tree, pde), "pde does not match the tree's solution dimension");
  if (tree->t.ninst > 1) {
    set_error("domain targets and sensitivities need a single instance");
    return KIGPU_ERR_NOTSUPP;
  }
  return guard([&] {
    Context& c = ctx->c;
    const TreeData& t = tree->t;
    const int sd = pde == KIGPU_STOKES ? 2 : 1;
    DBuf<double> dt, dm, du;
    dt.upload(targets, 2 * (size_t)ntgt, c.stream);
    dm.upload(mu, (size_t)t.ndof * nrhs, c.stream);
    du.alloc((size_t)sd * ntgt * nrhs);
    forward_apply(c, t, pde, dt.p, ntgt, dm.p, nrhs, du.p);
    du.download(u, (size_t)sd * ntgt * nrhs, c.stream);
    stream_wait(c.stream);
  }, &ctx->c);
}

static const char* check_region(const kigpu_region* r) {
  if (!(r->eps >= 0)) return "region eps must be >= 0";
  if (r->nholes < 0) return "negative hole count";
  switch (r->kind) {
    case KIGPU_REGION_DISC:
      if (r->nholes && !r->holes) return "disc region: holes missing";
      return nullptr;
    case KIGPU_REGION_WINDING:
      if (!r->nodes || r->num_outer < 1) return "winding region: nodes missing";
      if (r->nholes && !r->hole_nodes) return "winding region: hole_nodes missing";
      for (int k = 0; k < r->nholes; ++k)
        if (r->hole_nodes[k] < 1) return "winding region: empty hole section";
      return nullptr;
    case KIGPU_REGION_STAR:
      if (r->nterms < 0 || (r->nterms && !(r->amps && r->freqs && r->phases)))
        return "star region: terms missing";
      return nullptr;
    default:
      return "unknown region kind";
  }
}

kigpu_status kigpu_domain_mask(kigpu_ctx* ctx, const kigpu_region* region, const double* targets,
                               int ntgt, unsigned char* inside) {
  KI_REQUIRE(ctx && region && ntgt >= 0 && (ntgt == 0 || (targets && inside)),
             "invalid argument");
  if (const char* e = check_region(region)) {
    set_error(e);
    return KIGPU_ERR_ARG;
  }
  if (ntgt == 0) return KIGPU_OK;
  return guard([&] {
    Context& c = ctx->c;
    DBuf<double> dt;
    DBuf<unsigned char> din;
    dt.upload(targets, 2 * (size_t)ntgt, c.stream);
    din.alloc(ntgt);
    domain_mask(c, *region, dt.p, ntgt, din.p);
    din.download(inside, ntgt, c.stream);
    stream_wait(c.stream);
  }, &ctx->c);
}

kigpu_status kigpu_domain_solution(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                   const double* targets, int ntgt, const double* mu,
                                   const double* alpha, const double* hole_centers, int nholes,
                                   int nrhs, const kigpu_region* region, double* u) {
  KI_REQUIRE(ctx && tree && targets && mu && u && ntgt > 0 && nrhs > 0 && nholes >= 0,
             "invalid argument");
  KI_REQUIRE(!alpha || nholes == 0 || hole_centers, "hole centres missing");
  KI_REQUIRE(pde_matches(tree, pde), "pde does not match the tree's solution dimension");
  if (region) {
    if (const char* e = check_region(region)) {
      set_error(e);
      return KIGPU_ERR_ARG;
    }
  }
  if (tree->t.ninst > 1) {
    set_error("domain targets need a single instance");
    return KIGPU_ERR_NOTSUPP;
  }
  return guard([&] {
    Context& c = ctx->c;
    const TreeData& t = tree->t;
    const int sd = pde == KIGPU_STOKES ? 2 : 1;
    const size_t na = (size_t)(sd == 2 ? 3 : 1) * nholes;
    DBuf<double> dt, dm, da, du;
    dt.upload(targets, 2 * (size_t)ntgt, c.stream);
    dm.upload(mu, (size_t)t.ndof * nrhs, c.stream);
    if (alpha && nholes) da.upload(alpha, na * nrhs, c.stream);
    du.alloc((size_t)sd * ntgt * nrhs);
    domain_solution(c, t, pde, dt.p, ntgt, dm.p, nrhs, alpha && nholes ? da.p : nullptr,
                    hole_centers, nholes, region, du.p);
    du.download(u, (size_t)sd * ntgt * nrhs, c.stream);
    stream_wait(c.stream);
  }, &ctx->c);
}

kigpu_status kigpu_solve_transpose_device(kigpu_ctx* ctx, kigpu_tree* tree, const double* d_b,
                                          int nrhs, double* d_x, const kigpu_refine_opts* refine,
                                          kigpu_solve_info* info) {
  KI_REQUIRE(ctx && tree && d_b && d_x, "null argument");
  KI_REQUIRE(nrhs > 0, "nrhs must be positive");
  if (!tree->t.factored) {
    set_error("solve before factor");
    return KIGPU_ERR_STATE;
  }
  return guard(
      [&] { solve_tree(ctx->c, tree->t, d_b, nrhs, d_x, nullptr, refine, info, true); },
      &ctx->c);
}

kigpu_status kigpu_solve_transpose(kigpu_ctx* ctx, kigpu_tree* tree, const double* b, int ldb,
                                   int nrhs, double* x, int ldx, const kigpu_refine_opts* refine,
                                   kigpu_solve_info* info) {
  KI_REQUIRE(ctx && tree && b && x, "null argument");
  KI_REQUIRE(nrhs > 0, "nrhs must be positive");
  const int nd = tree->t.ndof;
  KI_REQUIRE(ldb >= nd && ldx >= nd, "leading dimension smaller than the number of DOFs");
  if (!tree->t.factored) {
    set_error("solve before factor");
    return KIGPU_ERR_STATE;
  }
  return guard([&] {
    Context& c = ctx->c;
    DBuf<double> db, dx;
    db.alloc((size_t)nd * nrhs);
    dx.alloc((size_t)nd * nrhs);
    KI_HIP(hipMemcpy2DAsync(db.p, sizeof(double) * nd, b, sizeof(double) * ldb,
                            sizeof(double) * nd, nrhs, hipMemcpyHostToDevice, c.stream));
    solve_tree(c, tree->t, db.p, nrhs, dx.p, nullptr, refine, info, true);
    KI_HIP(hipMemcpy2DAsync(x, sizeof(double) * ldx, dx.p, sizeof(double) * nd,
                            sizeof(double) * nd, nrhs, hipMemcpyDeviceToHost, c.stream));
    stream_wait(c.stream);
  }, &ctx->c);
}

kigpu_status kigpu_kernel_apply_transpose(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                          const double* lam, int nrhs, double* y) {
  KI_REQUIRE(ctx && tree && lam && y && nrhs > 0, "invalid argument");
  KI_REQUIRE(pde_matches(tree, pde), "pde does not match the tree's solution dimension");
  return guard([&] {
    Context& c = ctx->c;
    const size_t len = (size_t)tree->t.ndof * nrhs;
    DBuf<double> dl, dy;
    dl.upload(lam, len, c.stream);
    dy.alloc(len);
    kernel_apply_transpose(c, tree->t, pde, dl.p, nrhs, dy.p);
    dy.download(y, len, c.stream);
    stream_wait(c.stream);
  }, &ctx->c);
}

kigpu_status kigpu_forward_apply_transpose(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                          const double* targets, int ntgt, const double* v,
                                          int nrhs, double* y) {
  KI_REQUIRE(ctx && tree && targets && v && y && ntgt > 0 && nrhs > 0, "invalid argument");
  KI_REQUIRE(pde_matches(tree, pde), "pde does not match the tree's solution dimension");
  if (tree->t.ninst > 1) {
    set_error("domain targets and sensitivities need a single instance");
    return KIGPU_ERR_NOTSUPP;
  }
  return guard([&] {
    Context& c = ctx->c;
    const TreeData& t = tree->t;
    const int sd = pde == KIGPU_STOKES ? 2 : 1;
    DBuf<double> dt, dv, dy;
    dt.upload(targets, 2 * (size_t)ntgt, c.stream);
    dv.upload(v, (size_t)sd * ntgt * nrhs, c.stream);
    dy.alloc((size_t)t.ndof * nrhs);
    forward_apply_transpose(c, t, pde, dt.p, ntgt, dv.p, nrhs, dy.p);
    dy.download(y, (size_t)t.ndof * nrhs, c.stream);
    stream_wait(c.stream);
  }, &ctx->c);
}

kigpu_status kigpu_kernel_sensitivity(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                      const double* lam, const double* mu, double* sens) {
  KI_REQUIRE(ctx && tree && lam && mu && sens, "invalid argument");
  if (pde != KIGPU_LAPLACE || tree->t.sd != 1) {
    set_error("shape sensitivities are implemented for the Laplace double layer only");
    return KIGPU_ERR_NOTSUPP;
  }
  if (tree->t.ninst > 1) {
    set_error("domain targets and sensitivities need a single instance");
    return KIGPU_ERR_NOTSUPP;
  }
  return guard([&] {
    Context& c = ctx->c;
    const TreeData& t = tree->t;
    DBuf<double> dl, dm, ds;
    dl.upload(lam, t.npts, c.stream);
    dm.upload(mu, t.npts, c.stream);
    ds.alloc(6 * (size_t)t.npts);
    kernel_sensitivity(c, t, dl.p, dm.p, ds.p);
    ds.download(sens, 6 * (size_t)t.npts, c.stream);
    stream_wait(c.stream);
  }, &ctx->c);
}

kigpu_status kigpu_forward_sensitivity(kigpu_ctx* ctx, const kigpu_tree* tree, int pde,
                                       const double* targets, int ntgt, const double* v,
                                       const double* mu, double* sens) {
  KI_REQUIRE(ctx && tree && targets && v && mu && sens && ntgt > 0, "invalid argument");
  if (pde != KIGPU_LAPLACE || tree->t.sd != 1) {
    set_error("shape sensitivities are implemented for the Laplace double layer only");
    return KIGPU_ERR_NOTSUPP;
  }
  if (tree->t.ninst > 1) {
    set_error("domain targets and sensitivities need a single instance");
    return KIGPU_ERR_NOTSUPP;
  }
  return guard([&] {
    Context& c = ctx->c;
    const TreeData& t = tree->t;
    DBuf<double> dt, dv, dm, ds;
    dt.upload(targets, 2 * (size_t)ntgt, c.stream);
    dv.upload(v, ntgt, c.stream);
    dm.upload(mu, t.npts, c.stream);
    ds.alloc(6 * (size_t)t.npts);
    forward_sensitivity(c, t, dt.p, ntgt, dv.p, dm.p, ds.p);
    ds.download(sens, 6 * (size_t)t.npts, c.stream);
    stream_wait(c.stream);
  }, &ctx->c);
}

}  // extern "C"
