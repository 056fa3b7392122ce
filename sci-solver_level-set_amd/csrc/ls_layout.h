// Patch layout for the level-set hot path (host side, built once per mesh).
//
// The layout fixes every ordering the per-timestep kernels depend on:
//   * vertices are renumbered patch-contiguously (stable sort by patch label),
//     following InitPatches2 in src/core3d/cuda/meshFIM3d.cu:495-512;
//   * every element is copied once per distinct patch among its vertices, in
//     first-appearance order, and the copies are stably sorted by patch, as in
//     InitPatches src/core3d/cuda/meshFIM3d.cu:423-493 with the labelling of
//     src/core3d/include/meshFIM_kernels.h:502-561;
//   * each owned vertex lists its element-copy slots ordered by (corner, copy),
//     which is the order the reference's gather CSR produces
//     (src/core3d/cuda/meshFIM3d.cu:537-542, filtered per patch at
//     src/core3d/include/redistance_kernels3d.h:311-329).
// Unlike the reference, each patch also gets a sorted halo list so a kernel
// fetches every outside vertex once per patch and element corners are stored
// as 16-bit patch-local indices. Per-copy local coordinates are computed on
// the device from the renumbered coordinates and corner ids.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lsamd {

constexpr int kMaxLocal = 65535;  // patch-local indices and slots are uint16

struct Layout {
  int dim = 0;        // 2 = triangles, 3 = tetrahedra
  int nc = 0;         // corners per element (dim + 1)
  int64_t N = 0;      // vertices
  int64_t E = 0;      // elements
  int64_t F = 0;      // element copies
  int P = 0;          // patches
  std::vector<int32_t> labels;      // N, patch of each ORIGINAL vertex
  std::vector<int32_t> vert_perm;   // N, new -> original vertex id
  std::vector<int32_t> vert_iperm;  // N, original -> new vertex id
  std::vector<int32_t> vert_off;    // P+1
  std::vector<int32_t> ele_off;     // P+1
  std::vector<int32_t> copy_src;    // F, original element of each copy
  std::vector<int32_t> copy_verts;  // F*nc, renumbered vertex ids (AoS)
  std::vector<uint16_t> copy_local; // F*4, patch-local corner ids (padded to 4)
  std::vector<int32_t> halo_off;    // P+1
  std::vector<int32_t> halo_ids;    // renumbered ids of halo vertices, sorted per patch
  std::vector<int32_t> vslot_off;   // N+1, by renumbered vertex (global offsets)
  std::vector<uint16_t> vslots;     // patch-local slot = local_copy*nc + corner
  std::vector<double> xyz;          // N*3, renumbered order
  std::vector<int32_t> padj_off;    // P+1 patch adjacency (sorted, unique)
  std::vector<int32_t> padj;
  int max_ep = 0, max_vp = 0, max_hp = 0, max_vslots = 0;
};

// GraphPartition_Square (src/core3d/cuda/meshFIM3d.cu:269-352): vertex index
// k*L*W + i*L + j of an L x W x H grid goes to block (k/bh, i/bw, j/bl).
int partition_grid_blocks(int64_t N, int L, int W, int H, int bl, int bw, int bh,
                          std::vector<int32_t>& labels, int& nparts, std::string& err);

// Recursive coordinate bisection into exactly `nparts` parts: the build's
// deterministic stand-in for METIS_PartGraphKway (src/core3d/cuda/meshFIM3d.cu:401).
int partition_rcb(int64_t N, const double* xyz, int nparts, std::vector<int32_t>& labels,
                  std::string& err);

int build_layout(int dim, int64_t N, const double* xyz, int64_t E, const int32_t* cells,
                 const int32_t* labels, int nparts, Layout& L, std::string& err);

// Device copies of the layout arrays (hipMalloc'd), as build_layout_device
// leaves them for the engine to adopt; release() frees what is still held.
struct DeviceLayout {
  int32_t *vert_off = nullptr, *ele_off = nullptr, *halo_off = nullptr, *halo_ids = nullptr;
  int32_t *vslot_off = nullptr, *padj_off = nullptr, *padj = nullptr, *copy_verts = nullptr, *copy_src = nullptr;
  uint16_t *copy_local = nullptr, *vslots = nullptr;
  double* xyz = nullptr;
  void release();
};

// The same layout built on GPU `device` (ls_layout_gpu.hip): identical
// arrays and orders to build_layout, also copied into L. Returns 0, -1 for
// invalid input (err says why) or -2 for a HIP failure.
int build_layout_device(int dim, int64_t N, const double* xyz, int64_t E, const int32_t* cells,
                        const int32_t* labels, int nparts, int device, Layout& L, DeviceLayout& D,
                        std::string& err);


}  // namespace lsamd
