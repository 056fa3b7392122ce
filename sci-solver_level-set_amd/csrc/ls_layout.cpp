// Host patch-layout builder; orderings are documented in ls_layout.h.
#include "ls_layout.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace lsamd {

int partition_grid_blocks(int64_t N, int L, int W, int H, int bl, int bw, int bh,
                          std::vector<int32_t>& labels, int& nparts, std::string& err) {
  if ((int64_t)L * W * H != N || bl <= 0 || bw <= 0 || bh <= 0) {
    err = "grid partition: L*W*H must equal the vertex count and block sizes must be positive";
    return -1;
  }
  const int nbl = (L + bl - 1) / bl, nbw = (W + bw - 1) / bw, nbh = (H + bh - 1) / bh;
  nparts = nbl * nbw * nbh;
  labels.assign(N, 0);
  for (int k = 0; k < H; ++k)
    for (int i = 0; i < W; ++i)
      for (int j = 0; j < L; ++j)
        labels[(int64_t)k * L * W + (int64_t)i * L + j] =
            (k / bh) * nbl * nbw + (i / bw) * nbl + (j / bl);
  return 0;
}

namespace {
void rcb_split(const double* xyz, int32_t* ids, int64_t n, int k, int base, int32_t* labels) {
  if (k <= 1 || n <= 1) {
    for (int64_t t = 0; t < n; ++t) labels[ids[t]] = base;
    return;
  }
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t t = 0; t < n; ++t)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], xyz[3 * (int64_t)ids[t] + a]);
      hi[a] = std::max(hi[a], xyz[3 * (int64_t)ids[t] + a]);
    }
  int axis = 0;
  for (int a = 1; a < 3; ++a)
    if (hi[a] - lo[a] > hi[axis] - lo[axis]) axis = a;
  const int kl = k / 2;
  const int64_t m = (n * kl) / k;
  auto less = [&](int32_t a, int32_t b) {
    const double ca = xyz[3 * (int64_t)a + axis], cb = xyz[3 * (int64_t)b + axis];
    return ca < cb || (ca == cb && a < b);
  };
  std::nth_element(ids, ids + m, ids + n, less);
  rcb_split(xyz, ids, m, kl, base, labels);
  rcb_split(xyz, ids + m, n - m, k - kl, base + kl, labels);
}
}  // namespace

int partition_rcb(int64_t N, const double* xyz, int nparts, std::vector<int32_t>& labels,
                  std::string& err) {
  if (nparts <= 0 || nparts > N) {
    err = "rcb partition: need 1 <= nparts <= vertex count";
    return -1;
  }
  std::vector<int32_t> ids(N);
  std::iota(ids.begin(), ids.end(), 0);
  labels.assign(N, 0);
  rcb_split(xyz, ids.data(), N, nparts, 0, labels.data());
  return 0;
}

int build_layout(int dim, int64_t N, const double* xyz, int64_t E, const int32_t* cells,
                 const int32_t* labels, int nparts, Layout& L, std::string& err) {
  if (dim != 2 && dim != 3) { err = "dim must be 2 or 3"; return -1; }
  if (N <= 0 || E <= 0 || nparts <= 0) { err = "empty mesh or no patches"; return -1; }
  if (N >= (int64_t)1 << 31 || E >= (int64_t)1 << 31) { err = "mesh too large for int32 ids"; return -1; }
  const int nc = dim + 1;
  L = Layout();
  L.dim = dim; L.nc = nc; L.N = N; L.E = E; L.P = nparts;
  L.labels.assign(labels, labels + N);
  for (int64_t v = 0; v < N; ++v)
    if (labels[v] < 0 || labels[v] >= nparts) { err = "partition label out of range"; return -1; }
  for (int64_t t = 0; t < E * nc; ++t)
    if (cells[t] < 0 || cells[t] >= N) { err = "element references a vertex out of range"; return -1; }

  // Element copies: distinct patches of each element in first-appearance order,
  // then a stable counting sort by patch (elements stay ascending inside a patch).
  std::vector<int64_t> ecnt(nparts + 1, 0);
  int64_t F = 0;
  for (int64_t e = 0; e < E; ++e) {
    int seen[4]; int ns = 0;
    for (int i = 0; i < nc; ++i) {
      const int q = labels[cells[e * nc + i]];
      bool dup = false;
      for (int s = 0; s < ns; ++s) dup |= (seen[s] == q);
      if (!dup) { seen[ns++] = q; ecnt[q + 1]++; }
    }
    F += ns;
  }
  if (F >= (int64_t)1 << 31) { err = "too many element copies for int32"; return -1; }
  L.F = F;
  L.ele_off.assign(nparts + 1, 0);
  for (int p = 0; p < nparts; ++p) ecnt[p + 1] += ecnt[p];
  for (int p = 0; p <= nparts; ++p) L.ele_off[p] = (int32_t)ecnt[p];
  L.copy_src.assign(F, 0);
  {
    std::vector<int64_t> pos(ecnt.begin(), ecnt.end() - 1);
    for (int64_t e = 0; e < E; ++e) {
      int seen[4]; int ns = 0;
      for (int i = 0; i < nc; ++i) {
        const int q = labels[cells[e * nc + i]];
        bool dup = false;
        for (int s = 0; s < ns; ++s) dup |= (seen[s] == q);
        if (!dup) { seen[ns++] = q; L.copy_src[pos[q]++] = (int32_t)e; }
      }
    }
  }

  // Vertex renumbering: stable counting sort by patch label.
  L.vert_off.assign(nparts + 1, 0);
  for (int64_t v = 0; v < N; ++v) L.vert_off[labels[v] + 1]++;
  for (int p = 0; p < nparts; ++p) L.vert_off[p + 1] += L.vert_off[p];
  L.vert_perm.assign(N, 0);
  L.vert_iperm.assign(N, 0);
  {
    std::vector<int32_t> pos(L.vert_off.begin(), L.vert_off.end() - 1);
    for (int64_t v = 0; v < N; ++v) {
      const int32_t nv = pos[labels[v]]++;
      L.vert_perm[nv] = (int32_t)v;
      L.vert_iperm[v] = nv;
    }
  }
  L.xyz.assign(3 * N, 0.0);
  for (int64_t v = 0; v < N; ++v)
    for (int a = 0; a < 3; ++a) L.xyz[3 * v + a] = xyz[3 * (int64_t)L.vert_perm[v] + a];

  L.copy_verts.assign(F * nc, 0);
  #pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < F; ++c)
    for (int i = 0; i < nc; ++i)
      L.copy_verts[c * nc + i] = L.vert_iperm[cells[(int64_t)L.copy_src[c] * nc + i]];

  // Per-patch halo lists and patch-local corner ids.
  std::vector<std::vector<int32_t>> halos(nparts);
  std::vector<int32_t> new_label(N);
  for (int p = 0; p < nparts; ++p)
    for (int32_t v = L.vert_off[p]; v < L.vert_off[p + 1]; ++v) new_label[v] = p;
  L.copy_local.assign(F * 4, 0);
  int bad = 0;
  #pragma omp parallel for schedule(dynamic, 16) reduction(+ : bad)
  for (int p = 0; p < nparts; ++p) {
    const int32_t vs = L.vert_off[p], ve = L.vert_off[p + 1];
    const int64_t es = L.ele_off[p], ee = L.ele_off[p + 1];
    std::vector<int32_t>& h = halos[p];
    for (int64_t c = es; c < ee; ++c)
      for (int i = 0; i < nc; ++i) {
        const int32_t v = L.copy_verts[c * nc + i];
        if (v < vs || v >= ve) h.push_back(v);
      }
    std::sort(h.begin(), h.end());
    h.erase(std::unique(h.begin(), h.end()), h.end());
    const int64_t nv = ve - vs, nh = (int64_t)h.size(), ne = ee - es;
    if (nv + nh > kMaxLocal || ne * nc > kMaxLocal) { bad++; continue; }
    for (int64_t c = es; c < ee; ++c)
      for (int i = 0; i < nc; ++i) {
        const int32_t v = L.copy_verts[c * nc + i];
        int32_t loc;
        if (v >= vs && v < ve) loc = v - vs;
        else loc = (int32_t)nv + (int32_t)(std::lower_bound(h.begin(), h.end(), v) - h.begin());
        L.copy_local[c * 4 + i] = (uint16_t)loc;
      }
  }
  if (bad) { err = "a patch exceeds 65535 local vertices or element-copy slots; use smaller patches"; return -1; }
  L.halo_off.assign(nparts + 1, 0);
  for (int p = 0; p < nparts; ++p) L.halo_off[p + 1] = L.halo_off[p] + (int32_t)halos[p].size();
  L.halo_ids.resize(L.halo_off[nparts]);
  for (int p = 0; p < nparts; ++p)
    std::copy(halos[p].begin(), halos[p].end(), L.halo_ids.begin() + L.halo_off[p]);

  // Owned-vertex slot lists, ordered by (corner, copy).
  std::vector<int32_t> vcount(N, 0);
  for (int p = 0; p < nparts; ++p) {
    const int32_t vs = L.vert_off[p], ve = L.vert_off[p + 1];
    for (int64_t c = L.ele_off[p]; c < L.ele_off[p + 1]; ++c)
      for (int i = 0; i < nc; ++i) {
        const int32_t v = L.copy_verts[c * nc + i];
        if (v >= vs && v < ve) vcount[v]++;
      }
  }
  L.vslot_off.assign(N + 1, 0);
  for (int64_t v = 0; v < N; ++v) L.vslot_off[v + 1] = L.vslot_off[v] + vcount[v];
  L.vslots.assign(L.vslot_off[N], 0);
  #pragma omp parallel for schedule(dynamic, 16)
  for (int p = 0; p < nparts; ++p) {
    const int32_t vs = L.vert_off[p], ve = L.vert_off[p + 1];
    const int64_t es = L.ele_off[p], ee = L.ele_off[p + 1];
    std::vector<int32_t> fill(ve - vs, 0);
    for (int i = 0; i < nc; ++i)
      for (int64_t c = es; c < ee; ++c) {
        const int32_t v = L.copy_verts[c * nc + i];
        if (v >= vs && v < ve)
          L.vslots[L.vslot_off[v] + fill[v - vs]++] = (uint16_t)((c - es) * nc + i);
      }
  }

  // Patch adjacency: patches owning a halo vertex of the patch.
  L.padj_off.assign(nparts + 1, 0);
  std::vector<std::vector<int32_t>> adj(nparts);
  #pragma omp parallel for schedule(dynamic, 16)
  for (int p = 0; p < nparts; ++p) {
    std::vector<int32_t>& a = adj[p];
    for (int32_t t = L.halo_off[p]; t < L.halo_off[p + 1]; ++t) a.push_back(new_label[L.halo_ids[t]]);
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  for (int p = 0; p < nparts; ++p) L.padj_off[p + 1] = L.padj_off[p] + (int32_t)adj[p].size();
  L.padj.resize(L.padj_off[nparts]);
  for (int p = 0; p < nparts; ++p) std::copy(adj[p].begin(), adj[p].end(), L.padj.begin() + L.padj_off[p]);

  for (int p = 0; p < nparts; ++p) {
    L.max_ep = std::max(L.max_ep, L.ele_off[p + 1] - L.ele_off[p]);
    L.max_vp = std::max(L.max_vp, L.vert_off[p + 1] - L.vert_off[p]);
    L.max_hp = std::max(L.max_hp, L.halo_off[p + 1] - L.halo_off[p]);
  }
  for (int64_t v = 0; v < N; ++v) L.max_vslots = std::max(L.max_vslots, vcount[v]);
  return 0;
}

}  // namespace lsamd
