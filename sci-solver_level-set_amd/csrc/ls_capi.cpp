// C-ABI entry points (include/levelset_hip.h). Host-side state: the mesh as
// the caller gave it, the patch layout, and the device engine.
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/levelset_hip.h"
#include "ls_engine.h"
#include "ls_layout.h"

struct ls_ctx {
  int device = 0;
  int dim = 0;
  int64_t N = 0, E = 0;
  std::vector<double> xyz;
  std::vector<int32_t> cells;
  std::vector<double> vel;   // per element, kept to re-project after re-partitioning
  int patch_sigma = 0;
  int precision = LS_F32;  // tet storage precision
  int exec = LS_EXEC_PERSISTENT;
  bool have_vel = false;
  lsamd::Layout layout;
  bool have_layout = false;
  std::unique_ptr<lsamd::Engine> engine;
  std::string err;
};

namespace {

int fail(ls_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

// Projects the per-element velocity into every copy's local frame and uploads
// it. patch_sigma reproduces the reference's use of the patch's first copy
// velocity for all of the patch's elements (meshFIM_kernels.h:51-53).
int apply_velocity(ls_ctx* c) {
  if (!c->have_layout || !c->engine) return LS_OK;
  const lsamd::Layout& L = c->layout;
  std::vector<double> sig;
  if (c->patch_sigma) {
    const int nf = L.dim == 3 ? 9 : 6, ns = L.dim == 3 ? 3 : 2;
    sig.assign(L.F * ns, 0.0);
    for (int p = 0; p < L.P; ++p) {
      const double* s = &c->vel[3 * (int64_t)L.copy_src[L.ele_off[p]]];
      for (int64_t k = L.ele_off[p]; k < L.ele_off[p + 1]; ++k)
        for (int a = 0; a < ns; ++a) {
          const double* ax = &L.frame[k * nf + 3 * a];
          sig[k * ns + a] = s[0] * ax[0] + s[1] * ax[1] + s[2] * ax[2];
        }
    }
  } else {
    lsamd::project_velocity(L, c->vel.data(), sig);
  }
  std::string err;
  const int rc = c->engine->set_velocity(sig, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int finish_layout(ls_ctx* c, const std::vector<int32_t>& labels, int nparts) {
  std::string err;
  lsamd::Layout L;
  if (lsamd::build_layout(c->dim, c->N, c->xyz.data(), c->E, c->cells.data(), labels.data(), nparts,
                          L, err) != 0)
    return fail(c, LS_ERR_ARG, err);
  if (c->device < 0) {  // host-only context: layout queries, no device state
    c->layout = std::move(L);
    c->engine.reset();
    c->have_layout = true;
    return LS_OK;
  }
  std::unique_ptr<lsamd::Engine> eng(
      lsamd::make_engine(c->dim, c->dim == 2 ? LS_F64 : c->precision, c->device, err));
  if (!eng) return fail(c, LS_ERR_ARG, err);
  int rc = eng->upload(L, err);
  if (rc != LS_OK) return fail(c, rc, err);
  c->layout = std::move(L);
  c->engine = std::move(eng);
  c->engine->set_exec(c->exec);
  c->have_layout = true;
  return c->have_vel ? apply_velocity(c) : LS_OK;
}

}  // namespace

extern "C" {

int ls_create(ls_ctx** out, int device) {
  if (!out) return LS_ERR_ARG;
  ls_ctx* c = new (std::nothrow) ls_ctx();
  if (!c) return LS_ERR_INTERNAL;
  c->device = device;
  *out = c;
  return LS_OK;
}

void ls_destroy(ls_ctx* c) { delete c; }

const char* ls_last_error(const ls_ctx* c) { return c ? c->err.c_str() : "null context"; }

int ls_set_mesh(ls_ctx* c, int dim, int64_t n_verts, const double* xyz, int64_t n_cells,
                const int32_t* cells) {
  if (!c) return LS_ERR_ARG;
  if (dim != 2 && dim != 3) return fail(c, LS_ERR_ARG, "dim must be 2 or 3");
  if (n_verts <= 0 || n_cells <= 0 || !xyz || !cells) return fail(c, LS_ERR_ARG, "empty mesh");
  const int nc = dim + 1;
  for (int64_t t = 0; t < n_cells * nc; ++t)
    if (cells[t] < 0 || cells[t] >= n_verts) return fail(c, LS_ERR_ARG, "cell references a vertex out of range");
  c->dim = dim;
  c->N = n_verts;
  c->E = n_cells;
  c->xyz.assign(xyz, xyz + 3 * n_verts);
  c->cells.assign(cells, cells + n_cells * nc);
  c->have_layout = false;
  c->have_vel = false;
  c->engine.reset();
  return LS_OK;
}

int ls_set_precision(ls_ctx* c, int dtype) {
  if (!c) return LS_ERR_ARG;
  if (dtype != LS_F32 && dtype != LS_F64) return fail(c, LS_ERR_ARG, "dtype must be LS_F32 or LS_F64");
  if (c->have_layout) return fail(c, LS_ERR_STATE, "set the precision before partitioning");
  c->precision = dtype;
  return LS_OK;
}

int ls_set_execution(ls_ctx* c, int mode) {
  if (!c) return LS_ERR_ARG;
  if (mode != LS_EXEC_PERSISTENT && mode != LS_EXEC_KERNELS && mode != LS_EXEC_GRAPH)
    return fail(c, LS_ERR_ARG, "mode must be LS_EXEC_PERSISTENT, LS_EXEC_KERNELS or LS_EXEC_GRAPH");
  c->exec = mode;
  if (c->engine) c->engine->set_exec(mode);
  return LS_OK;
}

int ls_partition_grid(ls_ctx* c, int L, int W, int H, int bl, int bw, int bh) {
  if (!c) return LS_ERR_ARG;
  if (!c->N) return fail(c, LS_ERR_STATE, "set the mesh first");
  std::vector<int32_t> labels;
  int nparts = 0;
  std::string err;
  if (lsamd::partition_grid_blocks(c->N, L, W, H, bl, bw, bh, labels, nparts, err) != 0)
    return fail(c, LS_ERR_ARG, err);
  return finish_layout(c, labels, nparts);
}

int ls_partition_rcb(ls_ctx* c, int n_parts) {
  if (!c) return LS_ERR_ARG;
  if (!c->N) return fail(c, LS_ERR_STATE, "set the mesh first");
  std::vector<int32_t> labels;
  std::string err;
  if (lsamd::partition_rcb(c->N, c->xyz.data(), n_parts, labels, err) != 0) return fail(c, LS_ERR_ARG, err);
  return finish_layout(c, labels, n_parts);
}

int ls_partition_labels(ls_ctx* c, const int32_t* labels, int n_parts) {
  if (!c) return LS_ERR_ARG;
  if (!c->N) return fail(c, LS_ERR_STATE, "set the mesh first");
  if (!labels || n_parts <= 0) return fail(c, LS_ERR_ARG, "labels and n_parts > 0 required");
  std::vector<int32_t> lab(labels, labels + c->N);
  return finish_layout(c, lab, n_parts);
}

int ls_get_layout_info(ls_ctx* c, ls_layout_info* info) {
  if (!c || !info) return LS_ERR_ARG;
  if (!c->have_layout) return fail(c, LS_ERR_STATE, "layout not built");
  const lsamd::Layout& L = c->layout;
  info->n_verts = L.N;
  info->n_cells = L.E;
  info->n_copies = L.F;
  info->n_patches = L.P;
  info->max_patch_copies = L.max_ep;
  info->max_patch_verts = L.max_vp;
  info->max_patch_halo = L.max_hp;
  info->max_vertex_slots = L.max_vslots;
  info->dim = L.dim;
  return LS_OK;
}

int ls_get_layout(ls_ctx* c, int32_t* vert_perm, int32_t* copy_src, int32_t* vert_off, int32_t* ele_off,
                  int32_t* copy_verts) {
  if (!c) return LS_ERR_ARG;
  if (!c->have_layout) return fail(c, LS_ERR_STATE, "layout not built");
  const lsamd::Layout& L = c->layout;
  if (vert_perm) std::copy(L.vert_perm.begin(), L.vert_perm.end(), vert_perm);
  if (copy_src) std::copy(L.copy_src.begin(), L.copy_src.end(), copy_src);
  if (vert_off) std::copy(L.vert_off.begin(), L.vert_off.end(), vert_off);
  if (ele_off) std::copy(L.ele_off.begin(), L.ele_off.end(), ele_off);
  if (copy_verts) std::copy(L.copy_verts.begin(), L.copy_verts.end(), copy_verts);
  return LS_OK;
}

int ls_set_phi(ls_ctx* c, const void* phi, int dtype) {
  if (!c) return LS_ERR_ARG;
  if (!c->have_layout) return fail(c, LS_ERR_STATE, "partition the mesh first");
  if (!c->engine) return fail(c, LS_ERR_STATE, "host-only context (device < 0) has no device state");
  if (!phi || (dtype != LS_F32 && dtype != LS_F64)) return fail(c, LS_ERR_ARG, "bad phi or dtype");
  std::vector<double> v(c->N);
  if (dtype == LS_F32) {
    const float* f = static_cast<const float*>(phi);
    for (int64_t i = 0; i < c->N; ++i) v[i] = f[i];
  } else {
    std::memcpy(v.data(), phi, c->N * sizeof(double));
  }
  std::string err;
  int rc = c->engine->set_phi(v.data(), err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_set_velocity(ls_ctx* c, const double* vel, int patch_sigma) {
  if (!c) return LS_ERR_ARG;
  if (!c->N) return fail(c, LS_ERR_STATE, "set the mesh first");
  if (!vel) return fail(c, LS_ERR_ARG, "null velocity");
  c->vel.assign(vel, vel + 3 * c->E);
  c->patch_sigma = patch_sigma ? 1 : 0;
  c->have_vel = true;
  return apply_velocity(c);
}

int ls_step(ls_ctx* c, const ls_step_params* prm, ls_stats* stats) {
  if (!c || !prm) return LS_ERR_ARG;
  if (!c->have_layout) return fail(c, LS_ERR_STATE, "partition the mesh first");
  if (!c->engine) return fail(c, LS_ERR_STATE, "host-only context (device < 0) has no device state");
  ls_stats local;
  std::string err;
  int rc = c->engine->step(*prm, stats ? *stats : local, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_get_phi(ls_ctx* c, void* out, int dtype) {
  if (!c || !out) return LS_ERR_ARG;
  if (!c->have_layout) return fail(c, LS_ERR_STATE, "partition the mesh first");
  if (!c->engine) return fail(c, LS_ERR_STATE, "host-only context (device < 0) has no device state");
  if (dtype != LS_F32 && dtype != LS_F64) return fail(c, LS_ERR_ARG, "bad dtype");
  std::vector<double> v(c->N);
  std::string err;
  int rc = c->engine->get_phi(v.data(), err);
  if (rc != LS_OK) return fail(c, rc, err);
  if (dtype == LS_F64) std::memcpy(out, v.data(), c->N * sizeof(double));
  else {
    float* f = static_cast<float*>(out);
    for (int64_t i = 0; i < c->N; ++i) f[i] = (float)v[i];
  }
  return LS_OK;
}

int ls_get_narrowband(ls_ctx* c, int32_t* ids, int32_t* n) {
  if (!c || !n) return LS_ERR_ARG;
  if (!c->have_layout) return fail(c, LS_ERR_STATE, "partition the mesh first");
  if (!c->engine) return fail(c, LS_ERR_STATE, "host-only context (device < 0) has no device state");
  std::vector<int32_t> v;
  std::string err;
  int rc = c->engine->get_narrowband(v, err);
  if (rc != LS_OK) return fail(c, rc, err);
  if (ids) std::copy(v.begin(), v.end(), ids);
  *n = (int32_t)v.size();
  return LS_OK;
}

}  // extern "C"
