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
  int patch_sigma = LS_VEL_ELEMENT;  // velocity mode (LS_VEL_*)
  int precision = LS_F32;  // tet storage precision
  int exec = LS_EXEC_PERSISTENT;
  int sweep = LS_SWEEP_AUTO;
  int layout_builder = LS_LAYOUT_DEVICE;
  bool have_vel = false;
  lsamd::Layout layout;
  bool have_layout = false;
  // Batch of independent meshes (n_meshes > 0): mesh m owns vertices
  // [mesh_voff[m], mesh_voff[m+1]) and cells [mesh_coff[m], mesh_coff[m+1]).
  int n_meshes = 0;
  std::vector<int64_t> mesh_voff, mesh_coff;
  std::vector<double> mesh_dt, mesh_bw;  // per mesh; empty: the step's uniform values
  std::vector<int32_t> patch_mesh;       // P, mesh of each patch
  std::unique_ptr<lsamd::Engine> engine;
  std::string err;
};

namespace {

int fail(ls_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

// Uploads the per-element velocity; the engine projects it into every copy's
// local frame. The mode (LS_VEL_*) selects each element's own velocity, the
// patch's first copy velocity for all of the patch's elements
// (meshFIM_kernels.h:51-53), or the 2D advection's velocity from phi
// (meshFIM2d_kernels.h:178-181).
int apply_velocity(ls_ctx* c) {
  if (!c->have_layout || !c->engine) return LS_OK;
  std::string err;
  const int rc = c->engine->set_velocity(c->vel.data(), c->E, c->patch_sigma, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

// Per-patch timestep and band from the per-mesh values, and the mesh of
// every patch (warm starts are accepted per mesh).
int apply_batch_params(ls_ctx* c) {
  if (!c->engine || !c->have_layout) return LS_OK;
  std::vector<double> dt, bw;
  if (c->n_meshes > 0 && !c->mesh_dt.empty()) {
    const int P = c->layout.P;
    dt.resize(P);
    bw.resize(P);
    for (int p = 0; p < P; ++p) {
      dt[p] = c->mesh_dt[c->patch_mesh[p]];
      bw[p] = c->mesh_bw[c->patch_mesh[p]];
    }
  }
  std::string err;
  int rc = c->engine->set_patch_params(dt, bw, err);
  if (rc == LS_OK) rc = c->engine->set_patch_groups(c->patch_mesh, std::max(c->n_meshes, 1), err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

// Mesh of every patch; a patch must not span two meshes of a batch.
int map_patches(ls_ctx* c, const std::vector<int32_t>& labels, int nparts) {
  c->patch_mesh.assign(nparts, -1);
  if (c->n_meshes == 0) {
    std::fill(c->patch_mesh.begin(), c->patch_mesh.end(), 0);
    return LS_OK;
  }
  for (int m = 0; m < c->n_meshes; ++m)
    for (int64_t v = c->mesh_voff[m]; v < c->mesh_voff[m + 1]; ++v) {
      int32_t& pm = c->patch_mesh[labels[v]];
      if (pm < 0) pm = m;
      else if (pm != m)
        return fail(c, LS_ERR_ARG, "patch " + std::to_string(labels[v]) + " spans meshes " + std::to_string(pm) +
                                       " and " + std::to_string(m));
    }
  for (int p = 0; p < nparts; ++p)
    if (c->patch_mesh[p] < 0) return fail(c, LS_ERR_ARG, "patch " + std::to_string(p) + " has no vertex");
  return LS_OK;
}

int finish_layout(ls_ctx* c, const std::vector<int32_t>& labels, int nparts) {
  std::string err;
  for (int32_t l : labels)
    if (l < 0 || l >= nparts) return fail(c, LS_ERR_ARG, "patch label out of range");
  if (map_patches(c, labels, nparts) != LS_OK) return LS_ERR_ARG;
  lsamd::Layout L;
  const bool on_device = c->device >= 0 && c->layout_builder == LS_LAYOUT_DEVICE;
  struct Held {
    lsamd::DeviceLayout d;
    ~Held() { d.release(); }
  } held;
  const int brc = on_device ? lsamd::build_layout_device(c->dim, c->N, c->xyz.data(), c->E, c->cells.data(),
                                                         labels.data(), nparts, c->device, L, held.d, err)
                            : lsamd::build_layout(c->dim, c->N, c->xyz.data(), c->E, c->cells.data(),
                                                  labels.data(), nparts, L, err);
  if (brc == -2) return fail(c, LS_ERR_HIP, err);
  if (brc != 0) return fail(c, LS_ERR_ARG, err);
  if (c->device < 0) {  // host-only context: layout queries, no device state
    c->layout = std::move(L);
    c->engine.reset();
    c->have_layout = true;
    return LS_OK;
  }
  std::unique_ptr<lsamd::Engine> eng(
      lsamd::make_engine(c->dim, c->dim == 2 ? LS_F64 : c->precision, c->device, err));
  if (!eng) return fail(c, LS_ERR_ARG, err);
  int rc = eng->upload(L, on_device ? &held.d : nullptr, err);
  if (rc != LS_OK) return fail(c, rc, err);
  c->layout = std::move(L);
  c->engine = std::move(eng);
  c->engine->set_exec(c->exec);
  rc = c->engine->set_sweep_mode(c->sweep, err);
  if (rc != LS_OK) return fail(c, rc, err);
  c->have_layout = true;
  rc = apply_batch_params(c);
  if (rc != LS_OK) return rc;
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
  c->n_meshes = 0;
  c->mesh_voff.clear();
  c->mesh_coff.clear();
  c->mesh_dt.clear();
  c->mesh_bw.clear();
  return LS_OK;
}

int ls_set_mesh_batch(ls_ctx* c, int dim, int n_meshes, const int64_t* vert_off, const double* xyz,
                      const int64_t* cell_off, const int32_t* cells) {
  if (!c) return LS_ERR_ARG;
  if (dim != 2 && dim != 3) return fail(c, LS_ERR_ARG, "dim must be 2 or 3");
  if (n_meshes <= 0 || !vert_off || !cell_off) return fail(c, LS_ERR_ARG, "n_meshes > 0 and offsets required");
  if (!xyz || !cells) return fail(c, LS_ERR_ARG, "empty mesh");
  if (vert_off[0] != 0 || cell_off[0] != 0) return fail(c, LS_ERR_ARG, "offsets must start at 0");
  for (int m = 0; m < n_meshes; ++m)
    if (vert_off[m + 1] <= vert_off[m] || cell_off[m + 1] <= cell_off[m])
      return fail(c, LS_ERR_ARG, "mesh " + std::to_string(m) + " is empty");
  const int nc = dim + 1;
  for (int m = 0; m < n_meshes; ++m)
    for (int64_t t = cell_off[m] * nc; t < cell_off[m + 1] * nc; ++t)
      if (cells[t] < vert_off[m] || cells[t] >= vert_off[m + 1])
        return fail(c, LS_ERR_ARG, "a cell of mesh " + std::to_string(m) + " names a vertex of another mesh");
  const int rc = ls_set_mesh(c, dim, vert_off[n_meshes], xyz, cell_off[n_meshes], cells);
  if (rc != LS_OK) return rc;
  c->n_meshes = n_meshes;
  c->mesh_voff.assign(vert_off, vert_off + n_meshes + 1);
  c->mesh_coff.assign(cell_off, cell_off + n_meshes + 1);
  return LS_OK;
}

int ls_partition_batch(ls_ctx* c, int patch_verts) {
  if (!c) return LS_ERR_ARG;
  if (c->n_meshes == 0) return fail(c, LS_ERR_STATE, "set a mesh batch first");
  if (patch_verts <= 0) return fail(c, LS_ERR_ARG, "patch_verts must be > 0");
  const int nm = c->n_meshes;
  std::vector<int> parts(nm), base(nm + 1, 0);
  for (int m = 0; m < nm; ++m) {
    const int64_t n = c->mesh_voff[m + 1] - c->mesh_voff[m];
    parts[m] = (int)((n + patch_verts - 1) / patch_verts);
    base[m + 1] = base[m] + parts[m];
  }
  std::vector<int32_t> labels(c->N);
  std::vector<std::string> errs(nm);
  #pragma omp parallel for schedule(dynamic, 1)
  for (int m = 0; m < nm; ++m) {
    const int64_t v0 = c->mesh_voff[m], n = c->mesh_voff[m + 1] - v0;
    std::vector<int32_t> lab;
    if (lsamd::partition_rcb(n, c->xyz.data() + 3 * v0, parts[m], lab, errs[m]) == 0)
      for (int64_t v = 0; v < n; ++v) labels[v0 + v] = base[m] + lab[v];
  }
  for (int m = 0; m < nm; ++m)
    if (!errs[m].empty()) return fail(c, LS_ERR_ARG, "mesh " + std::to_string(m) + ": " + errs[m]);
  return finish_layout(c, labels, base[nm]);
}

int ls_set_batch_params(ls_ctx* c, const double* dt, const double* bandwidth) {
  if (!c) return LS_ERR_ARG;
  if (c->n_meshes == 0) return fail(c, LS_ERR_STATE, "set a mesh batch first");
  if (!dt && !bandwidth) {
    c->mesh_dt.clear();
    c->mesh_bw.clear();
  } else {
    if (!dt || !bandwidth) return fail(c, LS_ERR_ARG, "give both dt and bandwidth, or neither");
    c->mesh_dt.assign(dt, dt + c->n_meshes);
    c->mesh_bw.assign(bandwidth, bandwidth + c->n_meshes);
  }
  return apply_batch_params(c);
}

int ls_get_batch_nb_counts(ls_ctx* c, int32_t* counts) {
  if (!c || !counts) return LS_ERR_ARG;
  if (c->n_meshes == 0) return fail(c, LS_ERR_STATE, "set a mesh batch first");
  if (!c->have_layout || !c->engine) return fail(c, LS_ERR_STATE, "partition the mesh first");
  std::vector<int32_t> v;
  std::string err;
  const int rc = c->engine->get_narrowband(v, err);
  if (rc != LS_OK) return fail(c, rc, err);
  std::fill(counts, counts + c->n_meshes, 0);
  for (int32_t p : v) counts[c->patch_mesh[p]]++;
  return LS_OK;
}

int ls_set_precision(ls_ctx* c, int dtype) {
  if (!c) return LS_ERR_ARG;
  if (dtype != LS_F32 && dtype != LS_F64) return fail(c, LS_ERR_ARG, "dtype must be LS_F32 or LS_F64");
  if (c->have_layout) return fail(c, LS_ERR_STATE, "set the precision before partitioning");
  c->precision = dtype;
  return LS_OK;
}

int ls_set_layout_builder(ls_ctx* c, int builder) {
  if (!c) return LS_ERR_ARG;
  if (builder != LS_LAYOUT_DEVICE && builder != LS_LAYOUT_HOST)
    return fail(c, LS_ERR_ARG, "builder must be LS_LAYOUT_DEVICE or LS_LAYOUT_HOST");
  c->layout_builder = builder;
  return LS_OK;
}

int ls_set_sweep(ls_ctx* c, int mode) {
  if (!c) return LS_ERR_ARG;
  const int org = mode & ~(LS_SWEEP_UNSTAGED | LS_SWEEP_NO_PLANE_BOUND);
  if (mode < 0 || (org != LS_SWEEP_AUTO && org != LS_SWEEP_ELEMENTS && org != LS_SWEEP_TASKS))
    return fail(c, LS_ERR_ARG,
                "mode must be LS_SWEEP_AUTO, LS_SWEEP_ELEMENTS or LS_SWEEP_TASKS, optionally with "
                "LS_SWEEP_UNSTAGED and LS_SWEEP_NO_PLANE_BOUND");
  c->sweep = mode;
  if (!c->engine) return LS_OK;
  std::string err;
  const int rc = c->engine->set_sweep_mode(mode, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
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
  info->n_halo = (int64_t)L.halo_ids.size();
  info->n_slots = (int64_t)L.vslots.size();
  info->n_adjacent = (int64_t)L.padj.size();
  return LS_OK;
}

int ls_get_patch_lists(ls_ctx* c, int32_t* halo_off, int32_t* halo_ids, uint16_t* copy_local, int32_t* vslot_off,
                       uint16_t* vslots, int32_t* padj_off, int32_t* padj) {
  if (!c) return LS_ERR_ARG;
  if (!c->have_layout) return fail(c, LS_ERR_STATE, "layout not built");
  const lsamd::Layout& L = c->layout;
  if (halo_off) std::copy(L.halo_off.begin(), L.halo_off.end(), halo_off);
  if (halo_ids) std::copy(L.halo_ids.begin(), L.halo_ids.end(), halo_ids);
  if (copy_local) std::copy(L.copy_local.begin(), L.copy_local.end(), copy_local);
  if (vslot_off) std::copy(L.vslot_off.begin(), L.vslot_off.end(), vslot_off);
  if (vslots) std::copy(L.vslots.begin(), L.vslots.end(), vslots);
  if (padj_off) std::copy(L.padj_off.begin(), L.padj_off.end(), padj_off);
  if (padj) std::copy(L.padj.begin(), L.padj.end(), padj);
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

int ls_set_velocity(ls_ctx* c, const double* vel, int mode) {
  if (!c) return LS_ERR_ARG;
  if (!c->N) return fail(c, LS_ERR_STATE, "set the mesh first");
  if (!vel) return fail(c, LS_ERR_ARG, "null velocity");
  if (mode != LS_VEL_ELEMENT && mode != LS_VEL_PATCH_FIRST && mode != LS_VEL_PHI_2D)
    return fail(c, LS_ERR_ARG, "unknown velocity mode");
  if (mode == LS_VEL_PHI_2D && c->dim != 2) return fail(c, LS_ERR_ARG, "LS_VEL_PHI_2D is for triangle meshes");
  c->vel.assign(vel, vel + 3 * c->E);
  c->patch_sigma = mode;
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

// Learned warm start: network parameters, training pairs, training,
// gradients and prediction (the engine's tetrahedral path).
namespace {
int engine_ready(ls_ctx* c) {
  if (!c->have_layout) return fail(c, LS_ERR_STATE, "partition the mesh first");
  if (!c->engine) return fail(c, LS_ERR_STATE, "host-only context (device < 0) has no device state");
  return LS_OK;
}
}  // namespace

int ls_mlp_set_params(ls_ctx* c, const float* params, int n) {
  if (!c || !params) return LS_ERR_ARG;
  if (n != LS_MLP_PARAMS) return fail(c, LS_ERR_ARG, "expected LS_MLP_PARAMS parameters");
  if (int rc = engine_ready(c)) return rc;
  std::string err;
  const int rc = c->engine->mlp_set_params(params, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_mlp_get_params(ls_ctx* c, float* params, int n) {
  if (!c || !params) return LS_ERR_ARG;
  if (n != LS_MLP_PARAMS) return fail(c, LS_ERR_ARG, "expected LS_MLP_PARAMS parameters");
  if (int rc = engine_ready(c)) return rc;
  std::string err;
  const int rc = c->engine->mlp_get_params(params, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_mlp_predict(ls_ctx* c) {
  if (!c) return LS_ERR_ARG;
  if (int rc = engine_ready(c)) return rc;
  std::string err;
  const int rc = c->engine->mlp_predict(err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_mlp_collect(ls_ctx* c, int phase, int64_t capacity) {
  if (!c) return LS_ERR_ARG;
  if (phase != 0 && phase != 1) return fail(c, LS_ERR_ARG, "phase must be 0 (before a step) or 1 (after it)");
  if (int rc = engine_ready(c)) return rc;
  std::string err;
  const int rc = c->engine->mlp_collect(phase, capacity, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_mlp_dataset(ls_ctx* c, float* X, float* Y, int64_t* n) {
  if (!c || !n) return LS_ERR_ARG;
  if (int rc = engine_ready(c)) return rc;
  *n = c->engine->mlp_dataset_size();
  if (!X && !Y) return LS_OK;
  std::string err;
  const int rc = c->engine->mlp_get_dataset(X, Y, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_mlp_gradient(ls_ctx* c, const float* X, const float* Y, int64_t n, float* grad, double* loss) {
  if (!c || !X || !Y || !grad || !loss) return LS_ERR_ARG;
  if (int rc = engine_ready(c)) return rc;
  std::string err;
  const int rc = c->engine->mlp_gradient(X, Y, n, grad, loss, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_mlp_gradient_noalloc(ls_ctx* c, const float* X, const float* Y, int64_t n, float* grad, double* loss) {
  if (!c || !X || !Y || !grad || !loss) return LS_ERR_ARG;
  if (int rc = engine_ready(c)) return rc;
  std::string err;
  const int rc = c->engine->mlp_gradient_noalloc(X, Y, n, grad, loss, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_mlp_train(ls_ctx* c, int epochs, int batch, double lr, uint64_t seed, double* loss) {
  if (!c || !loss) return LS_ERR_ARG;
  if (int rc = engine_ready(c)) return rc;
  std::string err;
  const int rc = c->engine->mlp_train(epochs, batch, lr, seed, loss, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_get_distances(ls_ctx* c, double* out) {
  if (!c || !out) return LS_ERR_ARG;
  if (int rc = engine_ready(c)) return rc;
  std::string err;
  const int rc = c->engine->get_distances(out, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_set_warm(ls_ctx* c, const double* warm) {
  if (!c || !warm) return LS_ERR_ARG;
  if (int rc = engine_ready(c)) return rc;
  std::string err;
  const int rc = c->engine->set_warm(warm, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_get_warm(ls_ctx* c, double* out) {
  if (!c || !out) return LS_ERR_ARG;
  if (int rc = engine_ready(c)) return rc;
  std::string err;
  const int rc = c->engine->get_warm(out, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_set_cfl_coefficients(ls_ctx* c, const double* ceik, const double* ccurv) {
  if (!c) return LS_ERR_ARG;
  if (int rc = engine_ready(c)) return rc;
  std::string err;
  const int rc = c->engine->set_cfl_coefficients(ceik, ccurv, c->E, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

int ls_compute_dt(ls_ctx* c, double* dt, double* max_rin) {
  if (!c || !dt || !max_rin) return LS_ERR_ARG;
  if (int rc = engine_ready(c)) return rc;
  std::string err;
  const int rc = c->engine->compute_dt(dt, max_rin, err);
  return rc == LS_OK ? rc : fail(c, rc, err);
}

}  // extern "C"
