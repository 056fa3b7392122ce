// Patch layout built on the GPU: the same arrays, in the same orders, as the
// host builder in ls_layout.cpp (src/core3d/cuda/meshFIM3d.cu:423-570), from
// stable radix sorts on composite keys:
//   * vertices by patch label, ties by vertex id (InitPatches2 renumbering);
//   * element copies by patch, ties by element id, each element contributing
//     its distinct patches in first-appearance order (InitPatches);
//   * halo vertices per patch by (patch, renumbered id), duplicates removed;
//   * owned-vertex slots by (vertex, corner, copy), the order of the gather
//     CSR (meshFIM3d.cu:537-542);
//   * patch adjacency by (patch, neighbour), duplicates removed.
// The device arrays are handed to the engine; the host Layout receives copies
// for the C-ABI's layout queries.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "ls_layout.h"

namespace lsamd {

namespace {

#define LG_HIP(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      err = std::string(#call) + " failed: " + hipGetErrorString(e_);         \
      return -2;                                                              \
    }                                                                         \
  } while (0)

template <typename U>
struct DevBuf {
  U* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t n) {
    if (p) { (void)hipFree(p); p = nullptr; }
    return hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(U));
  }
  U* release() { U* q = p; p = nullptr; return q; }
  void reset() { if (p) (void)hipFree(p); p = nullptr; }
};

int bits_for(uint64_t maxval) {
  int b = 1;
  while (b < 64 && (maxval >> b) != 0) ++b;
  return b;
}

constexpr int kB = 256;
inline unsigned grid(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kB - 1) / kB); }

__global__ void k_check(const int32_t* labels, int64_t N, int P, const int32_t* cells, int64_t EN, int32_t* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N && (labels[i] < 0 || labels[i] >= P)) atomicOr(flag, 1);
  if (i < EN && (cells[i] < 0 || cells[i] >= N)) atomicOr(flag, 2);
}

__global__ void k_iota(int32_t* a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (int32_t)i;
}

// Vertex renumbering products: inverse permutation and renumbered coordinates.
__global__ void k_renumber(const int32_t* perm, const double* xyz, int64_t N, int32_t* iperm, double* xyz_new) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int32_t v = perm[i];
  iperm[v] = (int32_t)i;
#pragma unroll
  for (int a = 0; a < 3; ++a) xyz_new[3 * i + a] = xyz[3 * (int64_t)v + a];
}

// Distinct patches of element e's corners in first-appearance order.
__device__ __forceinline__ int element_patches(const int32_t* cells, const int32_t* labels, int64_t e, int nc,
                                               int* seen) {
  int ns = 0;
  for (int i = 0; i < nc; ++i) {
    const int q = labels[cells[e * nc + i]];
    bool dup = false;
    for (int s = 0; s < ns; ++s) dup |= (seen[s] == q);
    if (!dup) seen[ns++] = q;
  }
  return ns;
}

__global__ void k_copy_count(const int32_t* cells, const int32_t* labels, int64_t E, int nc, int32_t* cnt) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  int seen[4];
  cnt[e] = element_patches(cells, labels, e, nc, seen);
}

__global__ void k_copy_emit(const int32_t* cells, const int32_t* labels, int64_t E, int nc, const int64_t* off,
                            int32_t* key, int32_t* val) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  int seen[4];
  const int ns = element_patches(cells, labels, e, nc, seen);
  for (int s = 0; s < ns; ++s) {
    key[off[e] + s] = seen[s];
    val[off[e] + s] = (int32_t)e;
  }
}

__global__ void k_copy_verts(const int32_t* copy_src, const int32_t* cells, const int32_t* iperm, int64_t F, int nc,
                             int32_t* cv) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= F) return;
  const int64_t e = copy_src[c];
  for (int i = 0; i < nc; ++i) cv[c * nc + i] = iperm[cells[e * nc + i]];
}

// Per copy: number of corners outside the patch (halo) and owned.
__global__ void k_corner_count(const int32_t* cv, const int32_t* cpatch, const int32_t* vert_off, int64_t F, int nc,
                               int32_t* nhalo, int32_t* nown, int32_t* vcnt) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= F) return;
  const int p = cpatch[c];
  const int32_t vs = vert_off[p], ve = vert_off[p + 1];
  int h = 0, o = 0;
  for (int i = 0; i < nc; ++i) {
    const int32_t v = cv[c * nc + i];
    if (v >= vs && v < ve) { ++o; atomicAdd(&vcnt[v], 1); }
    else ++h;
  }
  nhalo[c] = h;
  nown[c] = o;
}

// Halo keys (patch << 32 | vertex) and owned-slot keys
// (vertex << 20 | corner << 17 | patch-local copy).
__global__ void k_corner_emit(const int32_t* cv, const int32_t* cpatch, const int32_t* vert_off, const int32_t* ele_off,
                              int64_t F, int nc, const int64_t* hoff, const int64_t* ooff, uint64_t* hkey,
                              uint64_t* okey) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= F) return;
  const int p = cpatch[c];
  const int32_t vs = vert_off[p], ve = vert_off[p + 1];
  const uint64_t lc = (uint64_t)(c - ele_off[p]);
  int64_t h = hoff[c], o = ooff[c];
  for (int i = 0; i < nc; ++i) {
    const int32_t v = cv[c * nc + i];
    if (v >= vs && v < ve) okey[o++] = ((uint64_t)v << 20) | ((uint64_t)i << 17) | (lc & 0x1FFFFu);
    else hkey[h++] = ((uint64_t)p << 32) | (uint32_t)v;
  }
}

// off[p] = first index k with (key[k] >> 32) >= p, for p = 0..P.
__global__ void k_segment_offsets(const uint64_t* key, int64_t n, int P, int32_t* off) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > P) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)(key[mid] >> 32) < p) lo = mid + 1; else hi = mid;
  }
  off[p] = (int32_t)lo;
}

// off[p] = first index k with key[k] >= p, for p = 0..P (key sorted).
__global__ void k_sorted_offsets(const int32_t* key, int64_t n, int P, int32_t* off) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > P) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (key[mid] < p) lo = mid + 1; else hi = mid;
  }
  off[p] = (int32_t)lo;
}

__global__ void k_low_words(const uint64_t* key, int64_t n, int32_t* out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = (int32_t)(uint32_t)key[k];
}

__global__ void k_copy_local(const int32_t* cv, const int32_t* cpatch, const int32_t* vert_off, const int32_t* halo_off,
                             const int32_t* halo_ids, int64_t F, int nc, uint16_t* cl) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= F) return;
  const int p = cpatch[c];
  const int32_t vs = vert_off[p], ve = vert_off[p + 1];
  const int32_t hs = halo_off[p], he = halo_off[p + 1];
  for (int i = 0; i < 4; ++i) {
    int32_t loc = 0;
    if (i < nc) {
      const int32_t v = cv[c * nc + i];
      if (v >= vs && v < ve) {
        loc = v - vs;
      } else {
        int32_t lo = hs, hi = he;
        while (lo < hi) {
          const int32_t mid = (lo + hi) >> 1;
          if (halo_ids[mid] < v) lo = mid + 1; else hi = mid;
        }
        loc = (ve - vs) + (lo - hs);
      }
    }
    cl[c * 4 + i] = (uint16_t)loc;
  }
}

__global__ void k_slots(const uint64_t* okey, int64_t n, int nc, uint16_t* vslots) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t key = okey[k];
  vslots[k] = (uint16_t)((key & 0x1FFFFu) * (uint64_t)nc + ((key >> 17) & 7u));
}

// Adjacency keys (patch << 32 | patch of the halo vertex).
__global__ void k_adj_emit(const uint64_t* hkey, int64_t n, const int32_t* labels, const int32_t* perm, uint64_t* akey) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t v = (uint32_t)hkey[k];
  akey[k] = (hkey[k] & 0xFFFFFFFF00000000ull) | (uint32_t)labels[perm[v]];
}

// Temporary storage shared by the hipCUB calls.
struct Temp {
  DevBuf<unsigned char> buf;
  size_t n = 0;
  // Never hands hipCUB a null buffer: with one it only reports the size.
  hipError_t ensure(size_t need) {
    if (buf.p && need <= n) return hipSuccess;
    n = std::max<size_t>(need, 256);
    return buf.alloc(n);
  }
};

template <typename K>
int sort_keys(Temp& t, K* in, K* out, int64_t n, int bits, hipStream_t q, std::string& err) {
  size_t need = 0;
  LG_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, need, in, out, (int)n, 0, bits, q));
  LG_HIP(t.ensure(need));
  LG_HIP(hipcub::DeviceRadixSort::SortKeys(t.buf.p, need, in, out, (int)n, 0, bits, q));
  return 0;
}

template <typename K, typename V>
int sort_pairs(Temp& t, K* kin, K* kout, V* vin, V* vout, int64_t n, int bits, hipStream_t q, std::string& err) {
  size_t need = 0;
  LG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, vin, vout, (int)n, 0, bits, q));
  LG_HIP(t.ensure(need));
  LG_HIP(hipcub::DeviceRadixSort::SortPairs(t.buf.p, need, kin, kout, vin, vout, (int)n, 0, bits, q));
  return 0;
}

struct ToI64 {
  __host__ __device__ int64_t operator()(int32_t x) const { return x; }
};

// out[0..n] = exclusive prefix sums of in[0..n-1], accumulated in int64;
// out[n] = total.
int scan64(Temp& t, const int32_t* in, int64_t* out, int64_t n, hipStream_t q, std::string& err) {
  LG_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), q));
  if (n == 0) return 0;
  hipcub::TransformInputIterator<int64_t, ToI64, const int32_t*> it(in, ToI64());
  size_t need = 0;
  LG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, need, it, out + 1, (int)n, q));
  LG_HIP(t.ensure(need));
  LG_HIP(hipcub::DeviceScan::InclusiveSum(t.buf.p, need, it, out + 1, (int)n, q));
  return 0;
}

int scan32(Temp& t, const int32_t* in, int32_t* out, int64_t n, hipStream_t q, std::string& err) {
  LG_HIP(hipMemsetAsync(out, 0, sizeof(int32_t), q));
  if (n == 0) return 0;
  size_t need = 0;
  LG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, need, in, out + 1, (int)n, q));
  LG_HIP(t.ensure(need));
  LG_HIP(hipcub::DeviceScan::InclusiveSum(t.buf.p, need, in, out + 1, (int)n, q));
  return 0;
}

int unique64(Temp& t, const uint64_t* in, uint64_t* out, int64_t n, int64_t& nout, hipStream_t q, std::string& err) {
  DevBuf<int64_t> cnt;
  LG_HIP(cnt.alloc(1));
  size_t need = 0;
  LG_HIP(hipcub::DeviceSelect::Unique(nullptr, need, in, out, cnt.p, (int)n, q));
  LG_HIP(t.ensure(need));
  LG_HIP(hipcub::DeviceSelect::Unique(t.buf.p, need, in, out, cnt.p, (int)n, q));
  LG_HIP(hipMemcpyAsync(&nout, cnt.p, sizeof(int64_t), hipMemcpyDeviceToHost, q));
  LG_HIP(hipStreamSynchronize(q));
  return 0;
}

template <typename U>
int download(std::vector<U>& h, const U* d, int64_t n, hipStream_t q, std::string& err) {
  h.resize((size_t)n);
  if (n) LG_HIP(hipMemcpyAsync(h.data(), d, (size_t)n * sizeof(U), hipMemcpyDeviceToHost, q));
  return 0;
}

}  // namespace

void DeviceLayout::release() {
  void* ptrs[] = {vert_off, ele_off, halo_off, halo_ids, vslot_off, padj_off, padj, copy_verts, copy_src,
                  copy_local, vslots, xyz};
  for (void* p : ptrs) if (p) (void)hipFree(p);
  *this = DeviceLayout();
}

#define LG_RUN(call)             \
  do {                           \
    const int r_ = (call);       \
    if (r_ != 0) return r_;      \
  } while (0)

int build_layout_device(int dim, int64_t N, const double* xyz, int64_t E, const int32_t* cells,
                        const int32_t* labels, int nparts, int device, Layout& L, DeviceLayout& D,
                        std::string& err) {
  if (dim != 2 && dim != 3) { err = "dim must be 2 or 3"; return -1; }
  if (N <= 0 || E <= 0 || nparts <= 0) { err = "empty mesh or no patches"; return -1; }
  if (N >= (int64_t)1 << 31 || E >= (int64_t)1 << 31) { err = "mesh too large for int32 ids"; return -1; }
  const int nc = dim + 1;
  const int P = nparts;
  LG_HIP(hipSetDevice(device));
  hipStream_t q = nullptr;
  LG_HIP(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
  struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } sg{q};
  Temp tmp;

  DevBuf<int32_t> d_cells, d_labels, d_flag;
  DevBuf<double> d_xyz_in;
  LG_HIP(d_cells.alloc((size_t)E * nc));
  LG_HIP(d_labels.alloc((size_t)N));
  LG_HIP(d_xyz_in.alloc((size_t)N * 3));
  LG_HIP(d_flag.alloc(1));
  LG_HIP(hipMemcpyAsync(d_cells.p, cells, (size_t)E * nc * sizeof(int32_t), hipMemcpyHostToDevice, q));
  LG_HIP(hipMemcpyAsync(d_labels.p, labels, (size_t)N * sizeof(int32_t), hipMemcpyHostToDevice, q));
  LG_HIP(hipMemcpyAsync(d_xyz_in.p, xyz, (size_t)N * 3 * sizeof(double), hipMemcpyHostToDevice, q));
  LG_HIP(hipMemsetAsync(d_flag.p, 0, sizeof(int32_t), q));
  k_check<<<grid(std::max<int64_t>(N, E * nc)), kB, 0, q>>>(d_labels.p, N, P, d_cells.p, E * nc, d_flag.p);
  LG_HIP(hipGetLastError());
  int32_t flag = 0;
  LG_HIP(hipMemcpyAsync(&flag, d_flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, q));
  LG_HIP(hipStreamSynchronize(q));
  if (flag & 1) { err = "partition label out of range"; return -1; }
  if (flag & 2) { err = "element references a vertex out of range"; return -1; }

  // Vertex renumbering: labels sorted stably with the vertex ids as values.
  const int pbits = bits_for((uint64_t)P);
  DevBuf<int32_t> d_iota, d_sorted_lab, d_perm, d_iperm, d_vert_off;
  DevBuf<double> d_xyz;
  LG_HIP(d_iota.alloc((size_t)N));
  LG_HIP(d_sorted_lab.alloc((size_t)N));
  LG_HIP(d_perm.alloc((size_t)N));
  LG_HIP(d_iperm.alloc((size_t)N));
  LG_HIP(d_xyz.alloc((size_t)N * 3));
  LG_HIP(d_vert_off.alloc((size_t)P + 1));
  k_iota<<<grid(N), kB, 0, q>>>(d_iota.p, N);
  LG_RUN(sort_pairs(tmp, d_labels.p, d_sorted_lab.p, d_iota.p, d_perm.p, N, pbits, q, err));
  k_renumber<<<grid(N), kB, 0, q>>>(d_perm.p, d_xyz_in.p, N, d_iperm.p, d_xyz.p);
  k_sorted_offsets<<<grid(P + 1), kB, 0, q>>>(d_sorted_lab.p, N, P, d_vert_off.p);
  LG_HIP(hipGetLastError());
  d_xyz_in.reset();

  // Element copies: per element its distinct patches, then a stable sort by patch.
  DevBuf<int32_t> d_ecnt, d_ckey, d_cval, d_cpatch, d_copy_src, d_ele_off;
  DevBuf<int64_t> d_eoff;
  LG_HIP(d_ecnt.alloc((size_t)E));
  LG_HIP(d_eoff.alloc((size_t)E + 1));
  k_copy_count<<<grid(E), kB, 0, q>>>(d_cells.p, d_labels.p, E, nc, d_ecnt.p);
  LG_RUN(scan64(tmp, d_ecnt.p, d_eoff.p, E, q, err));
  int64_t F = 0;
  LG_HIP(hipMemcpyAsync(&F, d_eoff.p + E, sizeof(int64_t), hipMemcpyDeviceToHost, q));
  LG_HIP(hipStreamSynchronize(q));
  if (F >= (int64_t)1 << 31) { err = "too many element copies for int32"; return -1; }
  LG_HIP(d_ckey.alloc((size_t)F));
  LG_HIP(d_cval.alloc((size_t)F));
  LG_HIP(d_cpatch.alloc((size_t)F));
  LG_HIP(d_copy_src.alloc((size_t)F));
  k_copy_emit<<<grid(E), kB, 0, q>>>(d_cells.p, d_labels.p, E, nc, d_eoff.p, d_ckey.p, d_cval.p);
  LG_RUN(sort_pairs(tmp, d_ckey.p, d_cpatch.p, d_cval.p, d_copy_src.p, F, pbits, q, err));
  d_ckey.reset();
  d_cval.reset();
  d_ecnt.reset();
  d_eoff.reset();
  LG_HIP(d_ele_off.alloc((size_t)P + 1));
  k_sorted_offsets<<<grid(P + 1), kB, 0, q>>>(d_cpatch.p, F, P, d_ele_off.p);

  DevBuf<int32_t> d_cv;
  LG_HIP(d_cv.alloc((size_t)F * nc));
  k_copy_verts<<<grid(F), kB, 0, q>>>(d_copy_src.p, d_cells.p, d_iperm.p, F, nc, d_cv.p);
  LG_HIP(hipGetLastError());
  d_cells.reset();

  // Corners: halo keys per patch and owned-slot keys per vertex.
  DevBuf<int32_t> d_nh, d_no, d_slotcnt;
  DevBuf<int64_t> d_hoff, d_ooff;
  LG_HIP(d_nh.alloc((size_t)F));
  LG_HIP(d_no.alloc((size_t)F));
  LG_HIP(d_slotcnt.alloc((size_t)N));
  LG_HIP(d_hoff.alloc((size_t)F + 1));
  LG_HIP(d_ooff.alloc((size_t)F + 1));
  LG_HIP(hipMemsetAsync(d_slotcnt.p, 0, (size_t)N * sizeof(int32_t), q));
  k_corner_count<<<grid(F), kB, 0, q>>>(d_cv.p, d_cpatch.p, d_vert_off.p, F, nc, d_nh.p, d_no.p, d_slotcnt.p);
  LG_RUN(scan64(tmp, d_nh.p, d_hoff.p, F, q, err));
  LG_RUN(scan64(tmp, d_no.p, d_ooff.p, F, q, err));
  int64_t nhk = 0, nok = 0;
  LG_HIP(hipMemcpyAsync(&nhk, d_hoff.p + F, sizeof(int64_t), hipMemcpyDeviceToHost, q));
  LG_HIP(hipMemcpyAsync(&nok, d_ooff.p + F, sizeof(int64_t), hipMemcpyDeviceToHost, q));
  LG_HIP(hipStreamSynchronize(q));
  if (nok >= (int64_t)1 << 31) { err = "too many vertex slots for int32"; return -1; }
  DevBuf<uint64_t> d_hkey, d_hsort, d_okey, d_osort;
  LG_HIP(d_hkey.alloc((size_t)nhk));
  LG_HIP(d_hsort.alloc((size_t)nhk));
  LG_HIP(d_okey.alloc((size_t)nok));
  LG_HIP(d_osort.alloc((size_t)nok));
  k_corner_emit<<<grid(F), kB, 0, q>>>(d_cv.p, d_cpatch.p, d_vert_off.p, d_ele_off.p, F, nc, d_hoff.p, d_ooff.p,
                                       d_hkey.p, d_okey.p);
  LG_HIP(hipGetLastError());
  d_nh.reset();
  d_no.reset();
  d_hoff.reset();
  d_ooff.reset();

  // Halo lists: sorted unique (patch, vertex).
  int64_t H = 0;
  DevBuf<int32_t> d_halo_off, d_halo_ids;
  if (nhk > 0) {
    LG_RUN(sort_keys(tmp, d_hkey.p, d_hsort.p, nhk, 32 + pbits, q, err));
    LG_RUN(unique64(tmp, d_hsort.p, d_hkey.p, nhk, H, q, err));
  }
  LG_HIP(d_halo_off.alloc((size_t)P + 1));
  LG_HIP(d_halo_ids.alloc((size_t)H));
  k_segment_offsets<<<grid(P + 1), kB, 0, q>>>(d_hkey.p, H, P, d_halo_off.p);
  if (H) k_low_words<<<grid(H), kB, 0, q>>>(d_hkey.p, H, d_halo_ids.p);

  // Patch-local corner ids.
  DevBuf<uint16_t> d_cl;
  LG_HIP(d_cl.alloc((size_t)F * 4));
  k_copy_local<<<grid(F), kB, 0, q>>>(d_cv.p, d_cpatch.p, d_vert_off.p, d_halo_off.p, d_halo_ids.p, F, nc, d_cl.p);

  // Owned-vertex slot lists by (vertex, corner, copy).
  DevBuf<int32_t> d_vslot_off;
  DevBuf<uint16_t> d_vslots;
  LG_HIP(d_vslot_off.alloc((size_t)N + 1));
  LG_HIP(d_vslots.alloc((size_t)nok));
  LG_RUN(scan32(tmp, d_slotcnt.p, d_vslot_off.p, N, q, err));
  if (nok > 0) {
    LG_RUN(sort_keys(tmp, d_okey.p, d_osort.p, nok, 20 + bits_for((uint64_t)N), q, err));
    k_slots<<<grid(nok), kB, 0, q>>>(d_osort.p, nok, nc, d_vslots.p);
  }
  d_okey.reset();
  d_osort.reset();

  // Patch adjacency: patches owning the patch's halo vertices.
  int64_t A = 0;
  DevBuf<int32_t> d_padj_off, d_padj;
  LG_HIP(d_padj_off.alloc((size_t)P + 1));
  if (H > 0) {
    DevBuf<uint64_t> d_akey;
    LG_HIP(d_akey.alloc((size_t)H));
    k_adj_emit<<<grid(H), kB, 0, q>>>(d_hkey.p, H, d_labels.p, d_perm.p, d_hsort.p);
    LG_RUN(sort_keys(tmp, d_hsort.p, d_akey.p, H, 32 + pbits, q, err));
    LG_RUN(unique64(tmp, d_akey.p, d_hsort.p, H, A, q, err));
  }
  LG_HIP(d_padj.alloc((size_t)A));
  k_segment_offsets<<<grid(P + 1), kB, 0, q>>>(d_hsort.p, A, P, d_padj_off.p);
  if (A) k_low_words<<<grid(A), kB, 0, q>>>(d_hsort.p, A, d_padj.p);
  LG_HIP(hipGetLastError());

  // Host copies.
  L = Layout();
  L.dim = dim; L.nc = nc; L.N = N; L.E = E; L.F = F; L.P = P;
  L.labels.assign(labels, labels + N);
  LG_RUN(download(L.vert_perm, d_perm.p, N, q, err));
  LG_RUN(download(L.vert_iperm, d_iperm.p, N, q, err));
  LG_RUN(download(L.vert_off, d_vert_off.p, P + 1, q, err));
  LG_RUN(download(L.ele_off, d_ele_off.p, P + 1, q, err));
  LG_RUN(download(L.copy_src, d_copy_src.p, F, q, err));
  LG_RUN(download(L.copy_verts, d_cv.p, F * nc, q, err));
  LG_RUN(download(L.copy_local, d_cl.p, F * 4, q, err));
  LG_RUN(download(L.halo_off, d_halo_off.p, P + 1, q, err));
  LG_RUN(download(L.halo_ids, d_halo_ids.p, H, q, err));
  LG_RUN(download(L.vslot_off, d_vslot_off.p, N + 1, q, err));
  LG_RUN(download(L.vslots, d_vslots.p, nok, q, err));
  LG_RUN(download(L.xyz, d_xyz.p, N * 3, q, err));
  LG_RUN(download(L.padj_off, d_padj_off.p, P + 1, q, err));
  LG_RUN(download(L.padj, d_padj.p, A, q, err));
  LG_HIP(hipStreamSynchronize(q));
  for (int p = 0; p < P; ++p) {
    const int64_t nv = L.vert_off[p + 1] - L.vert_off[p], ne = L.ele_off[p + 1] - L.ele_off[p];
    const int64_t nh = L.halo_off[p + 1] - L.halo_off[p];
    if (nv + nh > kMaxLocal || ne * nc > kMaxLocal) {
      err = "a patch exceeds 65535 local vertices or element-copy slots; use smaller patches";
      return -1;
    }
    L.max_ep = std::max(L.max_ep, (int)ne);
    L.max_vp = std::max(L.max_vp, (int)nv);
    L.max_hp = std::max(L.max_hp, (int)nh);
  }
  for (int64_t v = 0; v < N; ++v) L.max_vslots = std::max(L.max_vslots, L.vslot_off[v + 1] - L.vslot_off[v]);

  D.release();
  D.vert_off = d_vert_off.release();
  D.ele_off = d_ele_off.release();
  D.halo_off = d_halo_off.release();
  D.halo_ids = d_halo_ids.release();
  D.vslot_off = d_vslot_off.release();
  D.padj_off = d_padj_off.release();
  D.padj = d_padj.release();
  D.copy_verts = d_cv.release();
  D.copy_src = d_copy_src.release();
  D.copy_local = d_cl.release();
  D.vslots = d_vslots.release();
  D.xyz = d_xyz.release();
  return 0;
}

}  // namespace lsamd
