// Device-side data views and kernel launchers shared by the C-ABI layer.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lsamd {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr double kLargeNum = 1e9;    // LARGENUM: unreached distance (exact in FP32)

enum : int8_t { kSeedPoint = 0, kActivePoint = 1, kAlivePoint = 2, kFarPoint = 3 };

// Patch layout resident in HBM (see ls_layout.h for the orderings).
template <typename T>
struct DevMesh {
  int P;
  int64_t N;
  int64_t F;
  const int32_t* vert_off;     // P+1
  const int32_t* ele_off;      // P+1
  const int32_t* halo_off;     // P+1
  const int32_t* halo_ids;     // renumbered global ids
  const int32_t* halo_patch;   // patch owning each halo entry's vertex (vpatch[halo_ids[i]])
  // Per patch, two int4: (vert_off[p], vert_off[p+1], ele_off[p], ele_off[p+1]),
  // (halo_off[p], halo_off[p+1], vslot_off[vert_off[p]], vslot_off[vert_off[p+1]]).
  const int4* phdr;
  const ushort4* copy_local;   // F, patch-local corner ids
  const int32_t* vslot_off;    // N+1
  const uint16_t* vslots;      // patch-local slot ids per owned vertex
  const T* coords;             // F x (6 | 3), local coordinates
  const T* sigma;              // F x (3 | 2), velocity in the local frame
  const __half* hgt;           // F x 4 corner heights over the opposite face, rounded toward zero (tets), or null
  const float* pscale;         // P: largest coordinate scale of a patch's tet copies, rounded up (tets), or null
  const int32_t* padj_off;     // P+1
  const int32_t* padj;
  const int32_t* vpatch;       // N, patch of each renumbered vertex
  // Mesh of each patch in a batch, or null (one mesh): a warm start is
  // accepted or rejected per mesh.
  const int32_t* pgroup;       // P or null
  int ngroups;                 // meshes (1 when pgroup is null)
  // Per-patch timestep and narrowband width for batches of meshes with
  // their own parameters; null: the step's uniform values.
  const double* patch_dt;      // P or null
  const T* patch_bw;           // P or null
  // CFL coefficients (kernel_compute_timestep2d): per original element,
  // looked up through copy_src, or the uniform values when null.
  const int32_t* copy_src;     // F, original element of each copy
  const double* ceik;          // E or null
  const double* ccurv;         // E or null
  double ceik0, ccurv0;
  // Triangles: advect with (phi_A, phi_B) / |grad phi| in each copy's local
  // frame instead of sigma (the reference's 2D advection).
  int sigma_phi;
  // Tets: every element copy's least height is at least 2^-350 and 2^-20 of
  // its coordinate scale, so the solve's length roots use sqrt_len.
  int fast_len;
  // Tets: the need-mask pass also applies the supporting-plane bound
  // (ElemSolve<T, 4>::plane_mask; element sizes allow its FP32 evaluation).
  int plane;
};

template <typename T>
__device__ __forceinline__ T band_of(const DevMesh<T>& m, int p, T bw) {
  return m.patch_bw ? m.patch_bw[p] : bw;
}
template <typename T>
__device__ __forceinline__ double dt_of(const DevMesh<T>& m, int p, double dt) {
  return m.patch_dt ? m.patch_dt[p] : dt;
}
template <typename T>
__device__ __forceinline__ int group_of(const DevMesh<T>& m, int p) {
  return m.pgroup ? m.pgroup[p] : 0;
}

// Per-timestep state (renumbered vertex order).
template <typename T>
struct DevState {
  T* phi;
  T* phi_out;
  T* DT;
  T* DT_out;
  int8_t* label;
  int8_t* tsign;
  int32_t* bcon;       // P, patch converged
  T* bmin;             // P, patch min distance after a sweep
  int32_t* act;        // P, member of the current active list
  int32_t* fresh;      // P, activation flag for neighbour probes
  int32_t* seedflag;   // P
  int32_t* counters;   // small array of device counters
  unsigned long long* sizes;  // kNumSizes patch-size sums for byte accounting
  // Persistent-step state: epoch-tagged list membership and probe marks,
  // the lists, and counters that survive across steps.
  int32_t* inA;        // P, epoch of the active list holding the patch
  int32_t* freshE;     // P, epoch in which the patch was marked for a probe
  int32_t* listA;      // P
  int32_t* listB;      // P
  int32_t* listF;      // P
  int32_t* nb[2];      // P each, narrowband lists (previous / new)
  int32_t* pc;         // kNumPc persistent-step counters
  unsigned long long* ticks;  // kNumTicks phase times (device wall clock)
  unsigned int* bar;          // grid barrier: [0] arrivals, [1] generation
  const T* warm;              // N warm-start distances (learned predictor), or null
  // A probe's state after its sweep (slot values, vertex values, dirty
  // flags, at the patch's ele_off / vert_off positions), kept when the
  // probe leaves the patch unconverged: the patch's visit in the next FIM
  // iteration reads the same distances, so its first sweep would repeat the
  // probe and it resumes from here instead. rs_epoch[p] names the iteration
  // whose visit may resume (0: none).
  T* rs_slot;                 // F x NC
  T* rs_vval;                 // N
  uint8_t* rs_dirty;          // F
  int32_t* rs_epoch;          // P
  int32_t* visitE;            // P, FIM iteration in which the patch was last visited
  // After a visit, ve_ok[copy] = 1 for each element copy that ended the
  // visit clean (no corner changed since its last check: a fixed point of
  // its solves) with every corner owned and holding the value the patch
  // writes out. The next fresh load of the patch in the same FIM run
  // (ve_tag[p] != 0) starts such elements clean instead of dirty.
  uint8_t* ve_ok;             // F
  int32_t* ve_tag;            // P
  // CFL bound of the last narrowband as order-preserving keys: [0] the least
  // timestep, [1] the largest inscribed radius.
  unsigned long long* cfl;
  // Warm start: 1 for each mesh whose acceptance check failed this step.
  int32_t* gfail;             // ngroups
  // Overlapped FIM iterations of the persistent step (one grid barrier per
  // iteration): a second DT_out, visit tags and convergence flags, each
  // pair indexed by iteration parity, and nine P-long lists indexed by
  // iteration mod 3: probe marks [0..2], early members [3..5] (unconverged
  // members of the previous iteration), members to copy back [6..8].
  T* DT_out2;
  int32_t* visitE2;
  int32_t* bcon2;
  int32_t* ov_lists;
};

enum TickIndex : int { kTkSeed = 0, kTkVisit, kTkCopyback, kTkProbe, kTkNext, kTkNb, kTkAdvect, kNumTicks = 8 };

// Persistent-step counters.
enum PcIndex : int {
  kPcNb = 0,       // [0],[1]: narrowband sizes of nb[0], nb[1]
  kPcSeed = 2,     // initial active-list size
  kPcFresh = 3,    // [3],[4]: probe-list sizes by iteration parity
  kPcNext = 5,     // [5],[6]: next active-list sizes by iteration parity
  kPcNbVerts = 7,  // vertices of the new narrowband
  kPcIters = 8,    // FIM iterations
  kPcVisits = 9,   // patch visits
  kPcProbes = 10,  // probes
  kPcOverflow = 11,// FIM iteration cap hit
  kPcResid = 12,   // warm start: max fixed-point residual over band vertices (float bits)
  kPcFallback = 13,// warm start: the residual check failed and the step re-ran plain
  kPcWarm = 14,    // warm start: patches that received warm values
  kPcWorkV = 15,   // [15],[16]: next list entry to visit, by iteration parity
  kPcWorkP = 17,   // [17],[18]: next probe-list entry, by iteration parity
  kPcWorkS = 19,   // next patch to scan for seeds
  kPcWorkA = 20,   // next narrowband patch to advect (per pass)
  // Overlapped FIM iterations, each a triple indexed by iteration mod 3:
  kPcOvF = 24,     // probe marks made in the iteration (patches visited earlier in the step)
  kPcOvEl = 27,    // early members of the iteration (unconverged members of the one before)
  kPcOvMl = 30,    // members of the iteration, copied back in the next one
  kPcOvW = 33,     // next work item of the iteration's phase (early visits, then probes)
  kPcOvVis = 36,   // visits of the iteration
  kPcOvFn = 39,    // probe marks made in the iteration (patches not yet visited in the step)
  kNumPc = 42
};

// Patch-size sums (element copies, owned, halo) per launch kind.
enum SizeCounter : int {
  kSzVisit = 0,   // FIM visits: copies, verts, halo
  kSzProbe = 3,   // neighbour probes
  kSzAdvect = 6,  // advection
  kSzSeed = 9,    // seed scan
  kSzElem = 12,   // element updates executed (FIM visits + probes)
  kSzSolve = 13,  // local solves executed
  kNumSizes = 14
};

// Device counters layout.
enum Counter : int {
  kCntActive = 0,   // entries in the current active list
  kCntFresh = 1,    // entries in the neighbour-probe list
  kCntNext = 2,     // entries in the next active list
  kCntNB = 3,       // narrowband patches
  kCntSeed = 4,     // seed patches
  kCntNBVerts = 5,  // vertices owned by narrowband patches
  kNumCounters = 8
};

}  // namespace lsamd
