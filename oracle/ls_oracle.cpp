// CPU oracle for the level-set per-timestep hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: tests, the smoke check
// and the CPU-baseline leg of the benchmark load it; the product never does.
//
// A plain C++ restatement of the reference scheme, templated on the real type:
//   * Solver<double> is the FP64 host build (the parity oracle);
//   * Solver<float> stores distances and level-set values in FP32 exactly as
//     the GPU does and evaluates every element update in FP64 with the GPU's
//     operation order, so GPU results can be compared with it bit for bit.
// Every stage cites the reference function it restates. The patch layout is
// rebuilt here with the reference's own steps (stable sorts over SoA arrays)
// so the product's layout builder is checked independently.
//
// Parity pinning (see tests/): local solvers against brute-force Hopf-Lax
// minimisation and textbook point/segment/triangle distances; FIM against
// exact distances to planes; advection against exact translation of linear
// data and the discrete maximum principle; layout against brute-force
// definitions of the orderings.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

constexpr double kLarge = 1e9;  // LARGENUM (value external to the reference; 1e9 is exact in FP32)
// Relative window for threshold-ambiguous advection decisions: well above
// the FP32 storage error of distances (about 1e-7 relative).
constexpr double kAmbRel = 1e-6;
// A quadratic's discriminant within this fraction of q^2 + |4pr| of zero is
// taken as zero (a double root). Equal values at both ends of an edge, or a
// front parallel to a face, make the discriminant exactly zero; rounding
// alone would otherwise decide whether the candidate exists.
constexpr double kDiscRel = 1e-12;

enum Label { SeedPoint = 0, ActivePoint = 1, AlivePoint = 2, FarPoint = 3 };  // redistance3d.h:22-25

// Discriminant/root type: FP64 for both instances.
template <typename T> struct Prec { using D = double; };

// Triangle Hopf-Lax local solve at C from A, B (src/core2d/include/redistance_kernels.h:287-334),
// unit speed. The reference calls sqrtf on doubles there; this uses the type's own sqrt.
template <typename T>
T tri_solve(T TA, T TB, T LenAB, T LenBC, T LenAC) {
  const T big = (T)kLarge;
  if (TA >= big && TB >= big) return big;
  T TC1 = big, TC2 = big;
  const T TAB = TB - TA;
  const T a = (LenAB * LenAB - TAB * TAB) * LenAB * LenAB;
  const T EdgeTA = TA + LenAC;
  const T EdgeTB = TB + LenBC;
  if (a > (T)0) {
    const T cosA = (LenAC * LenAC + LenAB * LenAB - LenBC * LenBC) / ((T)2 * LenAC * LenAB);
    const T b = (T)2 * LenAB * LenAC * cosA * (TAB * TAB - LenAB * LenAB);
    const T delta = (T)4 * LenAC * LenAC * a * TAB * TAB * ((T)1 - cosA * cosA);
    const T sq = std::sqrt(delta);
    const T l1 = (-b + sq) / ((T)2 * a);
    const T l2 = (-b - sq) / ((T)2 * a);
    if (l1 >= (T)0 && l1 <= (T)1) {
      const T LenAD = l1 * LenAB;
      const T LenCD = std::sqrt(LenAC * LenAC + LenAD * LenAD - (T)2 * LenAC * LenAD * cosA);
      TC1 = l1 * TAB + TA + LenCD;
    }
    if (l2 >= (T)0 && l2 <= (T)1) {
      const T LenAD = l2 * LenAB;
      const T LenCD = std::sqrt(LenAC * LenAC + LenAD * LenAD - (T)2 * LenAC * LenAD * cosA);
      TC2 = l2 * TAB + TA + LenCD;
    }
    return std::fmin(TC2, std::fmin(TC1, std::fmin(EdgeTA, EdgeTB)));
  }
  return std::fmin(EdgeTA, EdgeTB);
}

// Roots of p*x^2 + q*x + r in the discriminant type; false when delta < 0
// beyond the rounding window kDiscRel. Every root in range yields a point of
// the face whose Hopf-Lax value is an admissible candidate, so accepting a
// double root that rounding pushed below zero never lowers the minimum below
// the true one.
template <typename T>
bool roots(T p, T q, T r, T& x1, T& x2) {
  using D = typename Prec<T>::D;
  const D dp = p, dq = q, dr = r;
  D delta = dq * dq - (D)4 * dp * dr;
  if (delta < 0.0 && delta >= -kDiscRel * (dq * dq + std::fabs((D)4 * dp * dr))) delta = 0.0;
  if (!(delta >= 0.0)) return false;
  const D sq = std::sqrt(delta);
  const D inv = (D)1 / ((D)2 * dp);  // one division for both roots
  x1 = (T)((-dq + sq) * inv);
  x2 = (T)((-dq - sq) * inv);
  return true;
}

// Tet Hopf-Lax local solve for D from A, B, C given six edge dot products
// (src/core3d/include/redistance_kernels3d.h:124-285): three face (edge)
// quadratics, three vertex candidates, and one interior quadratic whose second
// barycentric coordinate is (s*l1 + t) / (h + leps), evaluated as a product
// with the reciprocal.
template <typename T>
T tet_solve(T TA, T TB, T TC, T ACAC, T ACBC, T ACCD, T BCBC, T BCCD, T CDCD, T leps) {
  const T big = (T)kLarge;
  if (TA >= big && TB >= big && TC >= big) return big;
  T FaceTAC = big, FaceTAB = big, FaceTBC = big;
  T TD = big;
  const T TAC = TC - TA, TBC = TC - TB, TAB = TB - TA;
  T x1, x2;

  // Edge B-C (lambda1 = 0).
  {
    const T p = BCBC * TBC * TBC - BCBC * BCBC;
    const T q = (BCCD + BCCD) * TBC * TBC - (T)2 * BCBC * BCCD;
    const T r = TBC * TBC * CDCD - BCCD * BCCD;
    if (roots(p, q, r, x1, x2)) {
      const T cand[2] = {x1, x2};
      for (int k = 0; k < 2; ++k) {
        const T l2 = cand[k];
        const T l3 = (T)1 - l2;
        if (l2 >= (T)0 && l2 <= (T)1) {
          const T TE = TB * l2 + TC * l3;
          FaceTBC = std::fmin(FaceTBC, TE + std::sqrt((l2 * BCBC + BCCD) * l2 + (l2 * BCCD + CDCD)));
        }
      }
    }
    FaceTBC = std::fmin(FaceTBC, std::fmin(TB + std::sqrt((BCBC + BCCD) + (BCCD + CDCD)), TC + std::sqrt(CDCD)));
  }
  // Edge A-B (lambda3 = 0).
  {
    const T gx = ACAC - ACBC, gy = ACBC - BCBC, gz = ACCD - BCCD;
    const T t2 = TAB * TAB;
    const T p = (t2 * ACAC - gx * gx) + (BCBC * t2 - gy * gy) - ((ACBC + ACBC) * t2 - (T)2 * gx * gy);
    const T q = -(BCBC * t2 - gy * gy) * (T)2 + ((ACBC + ACBC) * t2 - (T)2 * gx * gy) +
                ((ACCD + ACCD) * t2 - (T)2 * gx * gz) - ((BCCD + BCCD) * t2 - (T)2 * gy * gz);
    const T r = (t2 * BCBC - gy * gy) + ((BCCD + BCCD) * t2 - (T)2 * gy * gz) + (t2 * CDCD - gz * gz);
    if (roots(p, q, r, x1, x2)) {
      const T cand[2] = {x1, x2};
      for (int k = 0; k < 2; ++k) {
        const T l1 = cand[k];
        const T l2 = (T)1 - l1;
        if (l1 >= (T)0 && l1 <= (T)1) {
          const T TE = TA * l1 + TB * l2;
          FaceTAB = std::fmin(FaceTAB, TE + std::sqrt((l1 * ACAC + l2 * ACBC + ACCD) * l1 +
                                                      (l1 * ACBC + l2 * BCBC + BCCD) * l2 +
                                                      (l1 * ACCD + l2 * BCCD + CDCD)));
        }
      }
    }
    FaceTAB = std::fmin(FaceTAB, std::fmin(TB + std::sqrt((BCBC + BCCD) + (BCCD + CDCD)),
                                           TA + std::sqrt((ACAC + ACCD) + (ACCD + CDCD))));
  }
  // Edge A-C (lambda2 = 0).
  {
    const T p = ACAC * TAC * TAC - ACAC * ACAC;
    const T q = (ACCD + ACCD) * TAC * TAC - (T)2 * ACAC * ACCD;
    const T r = TAC * TAC * CDCD - ACCD * ACCD;
    if (roots(p, q, r, x1, x2)) {
      const T cand[2] = {x1, x2};
      for (int k = 0; k < 2; ++k) {
        const T l1 = cand[k];
        const T l3 = (T)1 - l1;
        if (l1 >= (T)0 && l1 <= (T)1) {
          const T TE = TA * l1 + TC * l3;
          FaceTAC = std::fmin(FaceTAC, TE + std::sqrt((l1 * ACAC + ACCD) * l1 + (l1 * ACCD + CDCD)));
        }
      }
    }
    FaceTAC = std::fmin(FaceTAC, std::fmin(TA + std::sqrt((ACAC + ACCD) + (ACCD + CDCD)), TC + std::sqrt(CDCD)));
  }
  // Interior of face ABC.
  {
    const T s = TAC * ACBC - TBC * ACAC;
    const T t = TAC * BCCD - TBC * ACCD;
    const T h = -(TAC * BCBC - TBC * ACBC);
    const T t2 = TAC * TAC;
    const T c1 = t2 * ACAC - ACAC * ACAC;
    const T c2 = BCBC * t2 - ACBC * ACBC;
    const T c3 = (ACBC + ACBC) * t2 - (T)2 * ACAC * ACBC;
    const T c4 = (ACCD + ACCD) * t2 - (T)2 * ACAC * ACCD;
    const T c5 = (BCCD + BCCD) * t2 - (T)2 * ACBC * ACCD;
    const T c6 = t2 * CDCD - ACCD * ACCD;
    const T p = c1 * h * h + c2 * s * s + c3 * s * h;
    const T q = c2 * (T)2 * s * t + c3 * t * h + c4 * h * h + c5 * s * h;
    const T r = c2 * t * t + c5 * t * h + c6 * h * h;
    if (roots(p, q, r, x1, x2)) {
      const T cand[2] = {x1, x2};
      const T ihd = (T)1 / (h + leps);
      for (int k = 0; k < 2; ++k) {
        const T l1 = cand[k];
        const T l2 = (s * l1 + t) * ihd;
        const T l3 = (T)1 - l1 - l2;
        if (l1 >= (T)0 && l1 <= (T)1 && l2 >= (T)0 && l2 <= (T)1 && l3 >= (T)0 && l3 <= (T)1) {
          const T TE = TA * l1 + TB * l2 + TC * l3;
          TD = std::fmin(TD, TE + std::sqrt((l1 * ACAC + l2 * ACBC + ACCD) * l1 +
                                            (l1 * ACBC + l2 * BCBC + BCCD) * l2 +
                                            (l1 * ACCD + l2 * BCCD + CDCD)));
        }
      }
    }
  }
  return std::fmin(TD, std::fmin(FaceTBC, std::fmin(FaceTAB, FaceTAC)));
}

// Edge dot products from the six local coordinates and the four permuted
// solves of one element (src/core3d/include/redistance_kernels3d.h:373-414).
// The element update is evaluated in FP64 whatever the storage type T: the
// reference solver's quadratic coefficients cancel badly in FP32 and its
// candidate acceptance is discontinuous, so FP32 arithmetic there drifts
// 1e-4 from the FP64 scheme while FP32 storage alone stays within 1e-7.
struct TetGeom {
  double ACAC, ACBC, ACCD, BCBC, BCCD, CDCD, ADAD, ADBD, ADCD, BDBD, CDBD;
  template <typename T>
  explicit TetGeom(const T* Lt) {
    double L[6];
    for (int k = 0; k < 6; ++k) L[k] = (double)Lt[k];
    ACBC = L[1] * (L[1] - L[0]) + L[2] * L[2];
    BCCD = (L[1] - L[0]) * (L[3] - L[1]) + L[2] * (L[4] - L[2]);
    ACCD = L[1] * (L[3] - L[1]) + L[2] * (L[4] - L[2]);
    ADBD = L[3] * (L[3] - L[0]) + L[4] * L[4] + L[5] * L[5];
    const double ACAD = L[1] * L[3] + L[2] * L[4];
    const double BCBD = (L[1] - L[0]) * (L[3] - L[0]) + L[2] * L[4];
    const double ADBC = ACBC + BCCD;
    const double ACBD = ACBC + ACCD;
    CDBD = ADBD - ACBD;
    ADCD = ADBD - ADBC;
    CDCD = ADCD - ACCD;
    ADAD = ADCD + ACAD;
    ACAC = ACAD - ACCD;
    BDBD = BCBD + CDBD;
    BCBC = BCBD - BCCD;
  }
  template <typename T>
  void update(T* v, double leps) const {  // v = slot values of corners A,B,C,D
    const double TA = v[0], TB = v[1], TC = v[2], TD = v[3];
    v[3] = (T)std::fmin(TD, tet_solve(TA, TB, TC, ACAC, ACBC, ACCD, BCBC, BCCD, CDCD, leps));
    v[0] = (T)std::fmin(TA, tet_solve(TB, TD, TC, BCBC, -BCCD, -ACBC, CDCD, ACCD, ACAC, leps));
    v[1] = (T)std::fmin(TB, tet_solve(TA, TD, TC, ACAC, -ACCD, -ACBC, CDCD, BCCD, BCBC, leps));
    v[2] = (T)std::fmin(TC, tet_solve(TA, TB, TD, ADAD, ADBD, -ADCD, BDBD, -CDBD, CDCD, leps));
  }
  // Corner k's value from the other three corners alone (the same solve as
  // update(), without the min with corner k's own value).
  template <typename T>
  T corner_value(const T* v, int k, double leps) const {
    const double TA = v[0], TB = v[1], TC = v[2], TD = v[3];
    double r;
    if (k == 3) r = tet_solve(TA, TB, TC, ACAC, ACBC, ACCD, BCBC, BCCD, CDCD, leps);
    else if (k == 0) r = tet_solve(TB, TD, TC, BCBC, -BCCD, -ACBC, CDCD, ACCD, ACAC, leps);
    else if (k == 1) r = tet_solve(TA, TD, TC, ACAC, -ACCD, -ACBC, CDCD, BCCD, BCBC, leps);
    else r = tet_solve(TA, TB, TD, ADAD, ADBD, -ADCD, BDBD, -CDBD, CDCD, leps);
    return (T)std::fmin(kLarge, r);
  }
};

// Triangle edge lengths from the three local coordinates and the three
// permuted solves (src/core2d/include/redistance_kernels.h:441-462), FP64.
struct TriGeom {
  double LenAB, LenBC, LenCA;
  template <typename T>
  explicit TriGeom(const T* ct) {
    const double c[3] = {(double)ct[0], (double)ct[1], (double)ct[2]};
    LenAB = std::fabs(c[0]);
    LenBC = std::sqrt(c[2] * c[2] + (c[1] - c[0]) * (c[1] - c[0]));
    LenCA = std::sqrt(c[1] * c[1] + c[2] * c[2]);
  }
  template <typename T>
  void update(T* v, double) const {
    const double TA = v[0], TB = v[1], TC = v[2];
    v[2] = (T)std::fmin(TC, tri_solve(TA, TB, LenAB, LenBC, LenCA));
    v[0] = (T)std::fmin(TA, tri_solve(TB, TC, LenBC, LenCA, LenAB));
    v[1] = (T)std::fmin(TB, tri_solve(TC, TA, LenCA, LenAB, LenBC));
  }
  template <typename T>
  T corner_value(const T* v, int k, double) const {
    const double TA = v[0], TB = v[1], TC = v[2];
    double r;
    if (k == 2) r = tri_solve(TA, TB, LenAB, LenBC, LenCA);
    else if (k == 0) r = tri_solve(TB, TC, LenBC, LenCA, LenAB);
    else r = tri_solve(TC, TA, LenCA, LenAB, LenBC);
    return (T)std::fmin(kLarge, r);
  }
};

}  // namespace

extern "C" {
struct lso_params {
  double dt;            // timestep
  double bandwidth;     // narrowband width
  double tol;           // FIM convergence tolerance (SMALLNUM)
  double lambda_eps;    // interior-solve regulariser (SMALLNUM at redistance_kernels3d.h:258)
  int niter;            // sweeps per patch visit (NUM_ITER)
  int inside_niter;     // advection passes per step
  int mode;             // 0: advection, 1: mean-curvature flow
  int nthreads;         // OpenMP threads (0 = runtime default)
  int max_fim_iters;    // 0: FIM to convergence; > 0: stop after that many iterations
  int warm;             // 0: plain; 1: warm start from |phi|; 2: from the values given by set_warm
  double warm_margin;   // warm value = prediction * (1 + warm_margin) + warm_abs_margin
  double warm_res_tol;  // acceptance threshold of a warm step (0: tol); above it the step re-runs plain
  double warm_abs_margin;
  int cfl;              // 1: advect with min(dt, the new narrowband's CFL bound)
  int pad;
};
struct lso_stats {
  int fim_iters;
  int nb_count;
  int64_t patch_visits;
  int64_t probes;
  int64_t nb_verts;
  int n_seed_patches;
  int warm_patches;
  double warm_residual;
  int warm_fallback;
  int pad;
  double dt_used;       // the advection's timestep
};
}

namespace {

struct Base {
  virtual ~Base() {}
  virtual int set_phi(const double*) = 0;
  virtual int set_velocity(const double*, int) = 0;
  virtual int step(const lso_params&, lso_stats&) = 0;
  virtual void get_phi(double*) const = 0;
  virtual int get_nb(int32_t*) const = 0;
  virtual void get_layout(int32_t*, int32_t*, int32_t*, int32_t*, int32_t*) const = 0;
  virtual int64_t ncopies() const = 0;
  virtual void set_nb(const int32_t*, int) = 0;
  virtual void get_amb(uint8_t*) const = 0;
  virtual void get_coords(double*) const = 0;
  virtual void advect_all(double dt) = 0;
  virtual int nparts() const = 0;
  virtual void set_warm(const double*) = 0;
  virtual void get_resid(float*) const = 0;
  virtual void set_cfl_coefficients(const double*, const double*) = 0;
  virtual void compute_dt(double*) = 0;
  virtual void get_labels(int32_t*) const = 0;
};

template <typename T>
struct Solver : Base {
  int dim, nc;
  int64_t N, E, F;
  int P;
  std::vector<double> xyz_perm;       // renumbered coordinates
  std::vector<int32_t> vert_permute;  // new -> old
  std::vector<int32_t> vert_ipermute; // old -> new
  std::vector<int32_t> vert_offsets, ele_offsets;
  std::vector<int32_t> ele_permute;   // copy -> original element
  std::vector<int32_t> ele;           // SoA: ele[i*F + c], renumbered vertex ids
  std::vector<T> coords;              // SoA local coordinates: coords[k*F + c]
  std::vector<double> frames;         // per copy, for velocity projection
  std::vector<T> sigma;               // SoA per-copy velocity in the local frame
  bool sigma_phi = false;             // 2D: advect with the velocity from phi (mode 2)
  std::vector<std::vector<int32_t>> lmem;  // per renumbered vertex: patch-local slots
  std::vector<int32_t> bxadj, badj;   // patch adjacency
  // State (renumbered order).
  std::vector<T> phi, DT, DT_out, phi_out;
  std::vector<int32_t> label, con;
  std::vector<char> tsign;
  std::vector<int32_t> narrowband;
  // Vertices whose last advection update depended on a threshold decision
  // (|H| > 1e-8 per element, |sum alpha~ vol| > 1e-8 per vertex) that lies
  // inside the window FP32 storage of the inputs can move it across.
  std::vector<uint8_t> amb;

  Solver(int dim_, int64_t N_, const double* xyz, int64_t E_, const int32_t* cells,
         const int32_t* labels, int P_)
      : dim(dim_), nc(dim_ + 1), N(N_), E(E_), F(0), P(P_) {
    // InitPatches: per-element distinct patch count and first-appearance labels
    // (meshFIM_kernels.h:502-561), then sort copies by patch (meshFIM3d.cu:469).
    std::vector<int32_t> npart(E);
    for (int64_t e = 0; e < E; ++e) {
      int n = 1;
      const int p0 = labels[cells[e * nc]];
      const int p1 = labels[cells[e * nc + 1]];
      const int p2 = labels[cells[e * nc + 2]];
      if (p1 != p0) n++;
      if (p2 != p0 && p2 != p1) n++;
      if (nc == 4) {
        const int p3 = labels[cells[e * nc + 3]];
        if (p3 != p0 && p3 != p1 && p3 != p2) n++;
      }
      npart[e] = n;
    }
    std::vector<int64_t> eoff(E + 1, 0);
    for (int64_t e = 0; e < E; ++e) eoff[e + 1] = eoff[e] + npart[e];
    F = eoff[E];
    std::vector<int32_t> full_label(F), perm_pre(F);
    for (int64_t e = 0; e < E; ++e) {
      int64_t s = eoff[e];
      int ps[4];
      for (int i = 0; i < nc; ++i) ps[i] = labels[cells[e * nc + i]];
      for (int64_t j = 0; j < npart[e]; ++j) perm_pre[s + j] = (int32_t)e;
      full_label[s] = ps[0];
      int i = 1;
      if (ps[1] != ps[0]) full_label[s + i++] = ps[1];
      if (ps[2] != ps[0] && ps[2] != ps[1]) full_label[s + i++] = ps[2];
      if (nc == 4 && ps[3] != ps[0] && ps[3] != ps[1] && ps[3] != ps[2]) full_label[s + i++] = ps[3];
    }
    std::vector<int64_t> order(F);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return full_label[a] < full_label[b]; });
    ele_permute.resize(F);
    ele_offsets.assign(P + 1, 0);
    for (int64_t c = 0; c < F; ++c) {
      ele_permute[c] = perm_pre[order[c]];
      ele_offsets[full_label[order[c]] + 1]++;
    }
    for (int p = 0; p < P; ++p) ele_offsets[p + 1] += ele_offsets[p];
    // InitPatches2: stable sort of vertices by label (meshFIM3d.cu:503-511).
    vert_permute.resize(N);
    std::iota(vert_permute.begin(), vert_permute.end(), 0);
    std::stable_sort(vert_permute.begin(), vert_permute.end(),
                     [&](int32_t a, int32_t b) { return labels[a] < labels[b]; });
    vert_ipermute.resize(N);
    for (int64_t v = 0; v < N; ++v) vert_ipermute[vert_permute[v]] = (int32_t)v;
    vert_offsets.assign(P + 1, 0);
    for (int64_t v = 0; v < N; ++v) vert_offsets[labels[v] + 1]++;
    for (int p = 0; p < P; ++p) vert_offsets[p + 1] += vert_offsets[p];
    xyz_perm.resize(3 * N);
    for (int64_t v = 0; v < N; ++v)
      for (int a = 0; a < 3; ++a) xyz_perm[3 * v + a] = xyz[3 * (int64_t)vert_permute[v] + a];
    // kernel_ele_and_vert3d (meshFIM_kernels.h:420-446): renumbered SoA connectivity.
    ele.resize(nc * F);
    for (int64_t c = 0; c < F; ++c)
      for (int i = 0; i < nc; ++i)
        ele[i * F + c] = vert_ipermute[cells[(int64_t)ele_permute[c] * nc + i]];
    // kernel_compute_local_coords3d / 2d (meshFIM_kernels.h:448-500,
    // meshFIM2d_kernels.h:477-520), evaluated in FP64.
    const int ncoord = dim == 3 ? 6 : 3;
    coords.resize(ncoord * F);
    frames.resize((dim == 3 ? 9 : 6) * F);
    for (int64_t c = 0; c < F; ++c) local_frame(c);
    // Gather CSR (meshFIM3d.cu:537-542): slot indices i*F + c stably sorted by
    // vertex, then filtered per patch as in redistance_kernels3d.h:311-329.
    std::vector<int64_t> mem(nc * F);
    std::iota(mem.begin(), mem.end(), 0);
    std::stable_sort(mem.begin(), mem.end(), [&](int64_t a, int64_t b) { return ele[a] < ele[b]; });
    std::vector<int32_t> vpatch(N);
    for (int p = 0; p < P; ++p)
      for (int32_t v = vert_offsets[p]; v < vert_offsets[p + 1]; ++v) vpatch[v] = p;
    lmem.assign(N, {});
    for (int64_t k = 0; k < nc * F; ++k) {
      const int64_t idx = mem[k];
      const int32_t v = ele[idx];
      const int64_t c = idx % F;
      const int i = (int)(idx / F);
      const int p = vpatch[v];
      if (c >= ele_offsets[p] && c < ele_offsets[p + 1])
        lmem[v].push_back((int32_t)((c - ele_offsets[p]) * nc + i));
    }
    // GenerateBlockNeighbors (meshFIM3d.cu:546-570): patch pairs over mesh edges.
    std::vector<std::pair<int32_t, int32_t>> pairs;
    for (int64_t e = 0; e < E; ++e)
      for (int i = 0; i < nc; ++i)
        for (int j = 0; j < nc; ++j) {
          const int a = labels[cells[e * nc + i]], b = labels[cells[e * nc + j]];
          if (a != b) pairs.push_back({a, b});
        }
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    bxadj.assign(P + 1, 0);
    for (auto& pr : pairs) bxadj[pr.first + 1]++;
    for (int p = 0; p < P; ++p) bxadj[p + 1] += bxadj[p];
    badj.resize(pairs.size());
    for (size_t k = 0; k < pairs.size(); ++k) badj[k] = pairs[k].second;

    phi.assign(N, (T)0); DT.assign(N, (T)kLarge); DT_out.assign(N, (T)0); phi_out.assign(N, (T)0);
    label.assign(N, FarPoint); con.assign(N, 0); tsign.assign(N, 1); amb.assign(N, 0);
    sigma.assign((dim == 3 ? 3 : 2) * F, (T)0);
  }

  void local_frame(int64_t c) {
    auto P3 = [&](int i) { return &xyz_perm[3 * (int64_t)ele[i * F + c]]; };
    const double *A = P3(0), *B = P3(1), *C = P3(2);
    double AB[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    double AC[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    auto cross = [](const double* a, const double* b, double* o) {
      o[0] = a[1] * b[2] - a[2] * b[1];
      o[1] = a[2] * b[0] - a[0] * b[2];
      o[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    const double lenAB = std::sqrt(AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2]);
    double nrm[3];
    cross(AB, AC, nrm);
    if (dim == 3) {
      const double* D = P3(3);
      double AD[3] = {D[0] - A[0], D[1] - A[1], D[2] - A[2]};
      const double lenN = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
      double Z[3] = {nrm[0] / lenN, nrm[1] / lenN, nrm[2] / lenN};
      double X[3] = {AB[0] / lenAB, AB[1] / lenAB, AB[2] / lenAB};
      double Y[3];
      cross(Z, X, Y);
      const double v[6] = {lenAB, dot(AC, X), dot(AC, Y), dot(AD, X), dot(AD, Y), dot(AD, Z)};
      for (int k = 0; k < 6; ++k) coords[k * F + c] = (T)v[k];
      for (int a = 0; a < 3; ++a) {
        frames[9 * c + a] = X[a]; frames[9 * c + 3 + a] = Y[a]; frames[9 * c + 6 + a] = Z[a];
      }
    } else {
      double Y[3];
      cross(nrm, AB, Y);
      const double lenY = std::sqrt(Y[0] * Y[0] + Y[1] * Y[1] + Y[2] * Y[2]);
      coords[0 * F + c] = (T)lenAB;
      coords[1 * F + c] = (T)(dot(AC, AB) / lenAB);
      coords[2 * F + c] = (T)(dot(AC, Y) / lenY);
      for (int a = 0; a < 3; ++a) { frames[6 * c + a] = AB[a] / lenAB; frames[6 * c + 3 + a] = Y[a] / lenY; }
    }
  }

  int set_phi(const double* p) override {
    for (int64_t v = 0; v < N; ++v) phi[v] = (T)p[vert_permute[v]];
    narrowband.clear();
    return 0;
  }
  // mode 1 reproduces the reference's use of the patch's first copy velocity
  // for every element of the patch (meshFIM_kernels.h:51-53); mode 2
  // (triangles) its 2D advection's velocity from phi (meshFIM2d_kernels.h:
  // 170-181), applied in advect_patch; the projected velocity still feeds
  // the CFL bound.
  int set_velocity(const double* vel, int mode) override {
    const int patch_sigma = mode == 1;
    sigma_phi = dim == 2 && mode == 2;
    const int ns = dim == 3 ? 3 : 2, nf = dim == 3 ? 9 : 6;
    std::vector<int32_t> first(F);
    for (int p = 0; p < P; ++p)
      for (int32_t c = ele_offsets[p]; c < ele_offsets[p + 1]; ++c) first[c] = ele_offsets[p];
    for (int64_t c = 0; c < F; ++c) {
      const double* s = &vel[3 * (int64_t)ele_permute[patch_sigma ? first[c] : c]];
      for (int a = 0; a < ns; ++a) {
        const double* ax = &frames[nf * c + 3 * a];
        sigma[a * F + c] = (T)(s[0] * ax[0] + s[1] * ax[1] + s[2] * ax[2]);
      }
    }
    return 0;
  }
  void get_phi(double* out) const override {
    for (int64_t v = 0; v < N; ++v) out[vert_permute[v]] = (double)phi[v];
  }
  int get_nb(int32_t* ids) const override {
    if (ids) std::copy(narrowband.begin(), narrowband.end(), ids);
    return (int)narrowband.size();
  }
  void get_layout(int32_t* vp, int32_t* ep, int32_t* vo, int32_t* eo, int32_t* el) const override {
    if (vp) std::copy(vert_permute.begin(), vert_permute.end(), vp);
    if (ep) std::copy(ele_permute.begin(), ele_permute.end(), ep);
    if (vo) std::copy(vert_offsets.begin(), vert_offsets.end(), vo);
    if (eo) std::copy(ele_offsets.begin(), ele_offsets.end(), eo);
    if (el) std::copy(ele.begin(), ele.end(), el);
  }
  int64_t ncopies() const override { return F; }
  void set_nb(const int32_t* ids, int n) override { narrowband.assign(ids, ids + n); }
  // One advection pass over every patch on the current phi (no seeds, no FIM).
  void advect_all(double dt) override {
    lso_params prm{};
    prm.dt = dt;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < P; ++p) advect_patch(p, prm);
    phi = phi_out;
  }
  void get_coords(double* out) const override {
    const int nk = dim == 3 ? 6 : 3;
    for (int64_t c = 0; c < F; ++c)
      for (int k = 0; k < nk; ++k) out[c * nk + k] = (double)coords[k * F + c];
  }
  void get_labels(int32_t* out) const override {
    for (int64_t v = 0; v < N; ++v) out[vert_permute[v]] = label[v];
  }
  void get_amb(uint8_t* out) const override {
    for (int64_t v = 0; v < N; ++v) out[vert_permute[v]] = amb[v];
  }
  int nparts() const override { return P; }

  // Seeds (kernel_seedlabel / kernel_seedlabel_narrowband,
  // redistance_kernels3d.h:623-770; FindSeedPoint redistance3d.cu:21-65).
  // Returns the sorted list of patches holding a seed element.
  std::vector<int32_t> find_seeds() {
    std::fill(DT.begin(), DT.end(), (T)kLarge);
    std::fill(label.begin(), label.end(), (int32_t)FarPoint);
    std::vector<int32_t> scan;
    if (narrowband.empty()) { scan.resize(P); std::iota(scan.begin(), scan.end(), 0); }
    else scan = narrowband;
    std::vector<char> seeded(P, 0);
    for (int32_t p : scan) {
      for (int32_t c = ele_offsets[p]; c < ele_offsets[p + 1]; ++c) {
        int32_t vid[4]; T val[4];
        bool allpos = true, allneg = true;
        for (int i = 0; i < nc; ++i) {
          vid[i] = ele[i * F + c];
          val[i] = phi[vid[i]];
          allpos = allpos && val[i] >= (T)0;
          allneg = allneg && val[i] <= (T)0;
        }
        for (int i = 0; i < nc; ++i)
          if (val[i] == (T)0) { label[vid[i]] = SeedPoint; seeded[p] = 1; DT[vid[i]] = std::fabs(val[i]); }
        if (!(allpos || allneg)) {
          for (int i = 0; i < nc; ++i) { label[vid[i]] = SeedPoint; DT[vid[i]] = std::fabs(val[i]); }
          seeded[p] = 1;
        }
      }
    }
    std::vector<int32_t> act;
    for (int p = 0; p < P; ++p) if (seeded[p]) act.push_back(p);
    return act;
  }

  // FIM patch sweep (kernel_update_values / kernel_run_check_neghbor,
  // redistance_kernels3d.h:287-621). Reads DT, writes con and (if write_out)
  // DT_out for owned non-seed vertices.
  void update_patch(int p, int niter, bool write_out, T tol, double leps) {
    const int32_t vs = vert_offsets[p], ve = vert_offsets[p + 1];
    const int32_t es = ele_offsets[p], ee = ele_offsets[p + 1];
    const int ne = ee - es, nv = ve - vs;
    std::vector<T> s(nc * ne);
    for (int c = 0; c < ne; ++c)
      for (int i = 0; i < nc; ++i) s[c * nc + i] = DT[ele[i * F + es + c]];
    std::vector<T> oldT(nv), newT(nv);
    for (int v = 0; v < nv; ++v) oldT[v] = newT[v] = DT[vs + v];
    const int ncoord = dim == 3 ? 6 : 3;
    for (int it = 0; it < niter; ++it) {
      for (int v = 0; v < nv; ++v)
        if (!lmem[vs + v].empty()) oldT[v] = s[lmem[vs + v][0]];
      for (int c = 0; c < ne; ++c) {
        T Lc[6];
        for (int k = 0; k < ncoord; ++k) Lc[k] = coords[k * F + es + c];
        if (dim == 3) TetGeom(Lc).update(&s[c * nc], leps);
        else TriGeom(Lc).update(&s[c * nc], leps);
      }
      for (int v = 0; v < nv; ++v) {
        const std::vector<int32_t>& lm = lmem[vs + v];
        if (lm.empty()) continue;
        T m = s[lm[0]];
        for (size_t k = 1; k < lm.size(); ++k) m = std::fmin(s[lm[k]], m);
        for (int32_t l : lm) s[l] = m;
        newT[v] = m;
      }
    }
    for (int v = 0; v < nv; ++v) {
      if (label[vs + v] != SeedPoint) {
        if (write_out) DT_out[vs + v] = newT[v];
        con[vs + v] = (oldT[v] - newT[v] < tol) ? 1 : 0;
      } else {
        con[vs + v] = 1;
      }
    }
  }

  // Host active-list loop (redistance3d::GenerateData, redistance3d.cu:67-224).
  void fim(std::vector<int32_t> active, const lso_params& prm, lso_stats& st) {
    const T tol = (T)prm.tol, bw = (T)prm.bandwidth;
    const double leps = prm.lambda_eps;
    DT_out = DT;
    std::fill(con.begin(), con.end(), 0);
    std::vector<char> bcon(P, 0);
    std::vector<T> bmin(P, (T)kLarge);
    std::vector<int> blabel(P, FarPoint);
    while (!active.empty() && (prm.max_fim_iters <= 0 || st.fim_iters < prm.max_fim_iters)) {
      st.fim_iters++;
      st.patch_visits += (int64_t)active.size();
      #pragma omp parallel for schedule(dynamic, 1)
      for (int64_t k = 0; k < (int64_t)active.size(); ++k)
        update_patch(active[k], prm.niter, true, tol, leps);
      for (int32_t p : active) {  // CopyOutBack + run_reduction_bandwidth
        int all = 1;
        T m = (T)kLarge;
        for (int32_t v = vert_offsets[p]; v < vert_offsets[p + 1]; ++v) {
          DT[v] = DT_out[v];
          all &= con[v] ? 1 : 0;
          m = std::fmin(m, DT_out[v]);
        }
        bcon[p] = (char)all;
        bmin[p] = m;
      }
      for (int32_t p : active) blabel[p] = bcon[p] ? FarPoint : ActivePoint;
      std::vector<int32_t> fresh;
      for (int32_t p : active) {
        if (bcon[p] && bmin[p] < bw) {
          for (int32_t k = bxadj[p]; k < bxadj[p + 1]; ++k) {
            const int32_t q = badj[k];
            if (blabel[q] == FarPoint) { blabel[q] = ActivePoint; fresh.push_back(q); }
          }
        }
      }
      std::vector<int32_t> next;
      for (int32_t p : active) if (!bcon[p]) next.push_back(p);
      if (!fresh.empty()) {
        st.probes += (int64_t)fresh.size();
        #pragma omp parallel for schedule(dynamic, 1)
        for (int64_t k = 0; k < (int64_t)fresh.size(); ++k)
          update_patch(fresh[k], 1, false, tol, leps);
        for (int32_t q : fresh) {  // run_reduction3d
          int all = 1;
          for (int32_t v = vert_offsets[q]; v < vert_offsets[q + 1]; ++v) all &= con[v] ? 1 : 0;
          bcon[q] = (char)all;
        }
        for (int32_t q : fresh) {
          if (!bcon[q]) next.push_back(q);
          else blabel[q] = FarPoint;
        }
      }
      std::sort(next.begin(), next.end());
      active.swap(next);
    }
  }

  // Narrowband (kernel_compute_new_narrowband, redistance_kernels3d.h:22-55)
  // and sign restore (redistance3d.cu:262-272).
  void narrowband_and_restore(const lso_params& prm, lso_stats& st) {
    const T bw = (T)prm.bandwidth;
    narrowband.clear();
    for (int p = 0; p < P; ++p) {
      T m = (T)kLarge;
      for (int32_t v = vert_offsets[p]; v < vert_offsets[p + 1]; ++v) m = std::fmin(m, DT[v]);
      if (m <= bw) narrowband.push_back(p);
    }
    for (int64_t v = 0; v < N; ++v) phi[v] = DT[v] * (T)(2 * (int)tsign[v] - 1);
    st.nb_count = (int)narrowband.size();
    st.nb_verts = 0;
    for (int32_t p : narrowband) st.nb_verts += vert_offsets[p + 1] - vert_offsets[p];
  }

  // Upwind residual-distribution advection (kernel_updateT_single_stage3d,
  // meshFIM_kernels.h:36-408; 2D meshFIM2d_kernels.h:37-362). Shape-function
  // gradients are taken in each copy's local frame, where they are
  // closed-form, instead of from the 4x4 cofactor inverse of global
  // coordinates; K_i = vol * sigma . grad N_i is frame independent.
  void advect_patch(int p, const lso_params& prm) {
    if (prm.mode == 1) return curvature_patch(p, prm);
    const int32_t vs = vert_offsets[p], ve = vert_offsets[p + 1];
    const int32_t es = ele_offsets[p], ee = ele_offsets[p + 1];
    const int ne = ee - es;
    const double dt = prm.dt;
    std::vector<T> s_up(nc * ne), s_dn(nc * ne);  // per-slot products in the storage type
    std::vector<uint8_t> eamb(ne, 0);
    std::vector<double> s_dnabs(nc * ne, 0.0);
    for (int c = 0; c < ne; ++c) {
      const int64_t g = es + c;
      double ph[4], at[4] = {0, 0, 0, 0}, H, vol;
      for (int i = 0; i < nc; ++i) ph[i] = (double)phi[ele[i * F + g]];
      if (dim == 3) {
        const double L0 = coords[0 * F + g], L1 = coords[1 * F + g], L2 = coords[2 * F + g];
        const double L3 = coords[3 * F + g], L4 = coords[4 * F + g], L5 = coords[5 * F + g];
        const double sg[3] = {sigma[0 * F + g], sigma[1 * F + g], sigma[2 * F + g]};
        vol = std::fabs(L0 * L2 * L5) / 6.0;
        double gr[4][3];
        gr[1][0] = 1.0 / L0;
        gr[1][1] = -L1 / (L0 * L2);
        gr[1][2] = (L1 * L4 - L2 * L3) / (L0 * L2 * L5);
        gr[2][0] = 0.0;
        gr[2][1] = 1.0 / L2;
        gr[2][2] = -L4 / (L2 * L5);
        gr[3][0] = 0.0;
        gr[3][1] = 0.0;
        gr[3][2] = 1.0 / L5;
        for (int k = 0; k < 3; ++k) gr[0][k] = -((gr[1][k] + gr[2][k]) + gr[3][k]);
        double K[4];
        for (int i = 0; i < 4; ++i) K[i] = vol * (sg[0] * gr[i][0] + sg[1] * gr[i][1] + sg[2] * gr[i][2]);
        nscheme(K, ph, 4, H, at);
        eamb[c] = near_threshold(K, ph, 4, H);
      } else {
        const double c0 = coords[0 * F + g], c1 = coords[1 * F + g], c2 = coords[2 * F + g];
        double sg[2] = {sigma[0 * F + g], sigma[1 * F + g]};
        vol = std::fabs(c0 * c2) / 2.0;
        double gr[3][2];
        gr[1][0] = 1.0 / c0;
        gr[1][1] = -c1 / (c0 * c2);
        gr[2][0] = 0.0;
        gr[2][1] = 1.0 / c2;
        for (int k = 0; k < 2; ++k) gr[0][k] = -(gr[1][k] + gr[2][k]);
        if (sigma_phi) {  // (phi_A, phi_B) / |grad phi|, 0 where grad phi vanishes
          const double gx = gr[0][0] * ph[0] + gr[1][0] * ph[1] + gr[2][0] * ph[2];
          const double gy = gr[0][1] * ph[0] + gr[1][1] * ph[1] + gr[2][1] * ph[2];
          const double n = std::sqrt(gx * gx + gy * gy);
          sg[0] = n > 0.0 ? ph[0] / n : 0.0;
          sg[1] = n > 0.0 ? ph[1] / n : 0.0;
        }
        double K[3];
        for (int i = 0; i < 3; ++i) K[i] = vol * (sg[0] * gr[i][0] + sg[1] * gr[i][1]);
        nscheme(K, ph, 3, H, at);
        eamb[c] = near_threshold(K, ph, 3, H);
      }
      for (int i = 0; i < nc; ++i) {
        s_up[c * nc + i] = (T)(at[i] * H);
        s_dn[c * nc + i] = (T)(at[i] * vol);
      }
    }
    for (int32_t v = vs; v < ve; ++v) {
      double up = 0.0, dn = 0.0, dscale = 0.0;
      uint8_t a = 0;
      for (int32_t l : lmem[v]) up += (double)s_up[l];
      for (int32_t l : lmem[v]) { dn += (double)s_dn[l]; dscale += std::fabs((double)s_dn[l]); a |= eamb[l / nc]; }
      if (std::fabs(std::fabs(dn) - 1e-8) <= kAmbRel * dscale) a = 1;
      amb[v] = a;
      const double old = (double)phi[v];
      phi_out[v] = (T)((std::fabs(dn) > 1e-8) ? old - dt * (up / dn) : old);
    }
  }

  // Shape-function gradients of copy g in its local frame (the closed forms
  // advect_patch uses) and the copy's volume.
  void local_gradients(int64_t g, double (*gr)[3], double& vol) const {
    if (dim == 3) {
      const double L0 = coords[0 * F + g], L1 = coords[1 * F + g], L2 = coords[2 * F + g];
      const double L3 = coords[3 * F + g], L4 = coords[4 * F + g], L5 = coords[5 * F + g];
      vol = std::fabs(L0 * L2 * L5) / 6.0;
      gr[1][0] = 1.0 / L0;
      gr[1][1] = -L1 / (L0 * L2);
      gr[1][2] = (L1 * L4 - L2 * L3) / (L0 * L2 * L5);
      gr[2][0] = 0.0;
      gr[2][1] = 1.0 / L2;
      gr[2][2] = -L4 / (L2 * L5);
      gr[3][0] = 0.0;
      gr[3][1] = 0.0;
      gr[3][2] = 1.0 / L5;
      for (int k = 0; k < 3; ++k) gr[0][k] = -((gr[1][k] + gr[2][k]) + gr[3][k]);
    } else {
      const double c0 = coords[0 * F + g], c1 = coords[1 * F + g], c2 = coords[2 * F + g];
      vol = std::fabs(c0 * c2) / 2.0;
      gr[1][0] = 1.0 / c0;
      gr[1][1] = -c1 / (c0 * c2);
      gr[2][0] = 0.0;
      gr[2][1] = 1.0 / c2;
      for (int k = 0; k < 2; ++k) gr[0][k] = -(gr[1][k] + gr[2][k]);
      for (int i = 0; i < 3; ++i) gr[i][2] = 0.0;
    }
  }

  // Mean-curvature flow phi_t = kappa |grad phi| (the scheme the reference
  // leaves as dead CPU code, meshFIM3d.cu:243-266), with P1 elements and a
  // lumped mass M_i = sum_e vol_e / (d+1):
  //   kappa_i = -sum_e vol_e gradN_i . n_e / M_i,  n_e = grad phi_e / |grad phi_e|,
  //   |grad phi|_i = sum_e vol_e |grad phi_e| / sum_e vol_e,
  //   phi_i += dt * kappa_i * |grad phi|_i.
  // Elements touching an unreached vertex (|phi| >= LARGENUM / 2) and
  // elements with |grad phi_e| <= 1e-12 contribute nothing. Per-slot
  // products are stored in T; sums are FP64.
  void curvature_patch(int p, const lso_params& prm) {
    const int32_t vs = vert_offsets[p], ve = vert_offsets[p + 1];
    const int32_t es = ele_offsets[p], ee = ele_offsets[p + 1];
    const int ne = ee - es;
    std::vector<T> s_cv(nc * ne), s_vol(nc * ne), s_eik(ne);
    for (int c = 0; c < ne; ++c) {
      const int64_t g = es + c;
      double ph[4], gr[4][3], vol;
      bool fin = true;
      for (int i = 0; i < nc; ++i) {
        ph[i] = (double)phi[ele[i * F + g]];
        fin = fin && std::fabs(ph[i]) < 0.5 * kLarge;
      }
      local_gradients(g, gr, vol);
      double gp[3] = {0.0, 0.0, 0.0};
      for (int i = 0; i < nc; ++i)
        for (int k = 0; k < 3; ++k) gp[k] = gp[k] + ph[i] * gr[i][k];
      const double gn = std::sqrt((gp[0] * gp[0] + gp[1] * gp[1]) + gp[2] * gp[2]);
      const bool use = fin && gn > 1e-12;
      for (int i = 0; i < nc; ++i) {
        const double d = (gr[i][0] * gp[0] + gr[i][1] * gp[1]) + gr[i][2] * gp[2];
        s_cv[c * nc + i] = (T)(use ? vol * d / gn : 0.0);
        s_vol[c * nc + i] = (T)(use ? vol : 0.0);
      }
      s_eik[c] = (T)(use ? vol * gn : 0.0);
    }
    for (int32_t v = vs; v < ve; ++v) {
      double cv = 0.0, sv = 0.0, se = 0.0;
      for (int32_t l : lmem[v]) cv += (double)s_cv[l];
      for (int32_t l : lmem[v]) sv += (double)s_vol[l];
      for (int32_t l : lmem[v]) se += (double)s_eik[l / nc];
      amb[v] = 0;
      const double old = (double)phi[v];
      if (sv > 0.0) {
        const double kappa = -cv / (sv / (double)nc);
        phi_out[v] = (T)(old + prm.dt * kappa * (se / sv));
      } else {
        phi_out[v] = (T)old;
      }
    }
  }

  // True when |H| lies within the perturbation FP32 storage of the corner
  // values can cause: sum |K_i| * |phi_i| * kAmbRel around the 1e-8 cut-off.
  static uint8_t near_threshold(const double* K, const double* ph, int n, double H) {
    double sc = 0.0;
    for (int i = 0; i < n; ++i) sc += std::fabs(K[i]) * (std::fabs(ph[i]) + 1e-3);
    return std::fabs(std::fabs(H) - 1e-8) <= kAmbRel * sc ? 1 : 0;
  }

  // N-scheme distribution coefficients (meshFIM_kernels.h:227-273), FP64.
  static void nscheme(const double* K, const double* ph, int n, double& H, double* at) {
    double Kp[4], Km[4], beta = 0.0;
    H = 0.0;
    for (int i = 0; i < n; ++i) {
      H += K[i] * ph[i];
      Kp[i] = std::fmax(K[i], 0.0);
      Km[i] = std::fmin(K[i], 0.0);
      beta += Km[i];
    }
    beta = 1.0 / beta;
    for (int i = 0; i < n; ++i) at[i] = 0.0;
    if (std::fabs(H) > 1e-8) {
      double alpha[4], theta = 0.0;
      for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc += Km[j] * (ph[i] - ph[j]);
        alpha[i] = (Kp[i] * beta * acc) / H;
      }
      for (int i = 0; i < n; ++i) theta += std::fmax(0.0, alpha[i]);
      for (int i = 0; i < n; ++i) at[i] = std::fmax(alpha[i], 0.0) / theta;
    }
  }

  // Warm start (no reference counterpart): over the previous narrowband,
  // every vertex sharing no element with a seed whose warm value
  // prediction * (1 + margin) + abs margin is below LARGENUM/2 starts the
  // FIM there instead of at LARGENUM,
  // and every patch receiving one joins the first active list. The
  // prediction is |phi| (warm 1) or the values given by set_warm (warm 2).
  std::vector<double> warm;     // renumbered order, for warm 2
  std::vector<T> warm_applied;  // the warm value each vertex started from (LARGENUM: none)
  void set_warm(const double* w) override {
    if (!w) { warm.clear(); return; }
    warm.resize(N);
    for (int64_t v = 0; v < N; ++v) warm[vert_ipermute[v]] = w[v];
  }
  // Vertices sharing an element with a seed keep LARGENUM: the FIM solves
  // them from seed slots it lowers inside a visit and never writes back
  // (redistance_kernels3d.h:440-449), which no check from stored distances
  // can reproduce, so they are finished exactly as in the plain method.
  int apply_warm(const std::vector<int32_t>& prev_nb, const lso_params& prm, std::vector<int32_t>& act) {
    std::vector<char> in(P, 0);
    for (int32_t p : act) in[p] = 1;
    warm_applied.assign(N, (T)kLarge);
    int npatch = 0;
    for (int32_t p : prev_nb) {
      bool any = false;
      std::vector<char> near(vert_offsets[p + 1] - vert_offsets[p], 0);
      for (int32_t c = ele_offsets[p]; c < ele_offsets[p + 1]; ++c) {
        bool seeded = false;
        for (int i = 0; i < nc; ++i) seeded = seeded || label[ele[i * F + c]] == SeedPoint;
        if (!seeded) continue;
        for (int i = 0; i < nc; ++i) {
          const int32_t v = ele[i * F + c];
          if (v >= vert_offsets[p] && v < vert_offsets[p + 1]) near[v - vert_offsets[p]] = 1;
        }
      }
      for (int32_t v = vert_offsets[p]; v < vert_offsets[p + 1]; ++v) {
        if (label[v] == SeedPoint || near[v - vert_offsets[p]]) continue;
        const double base = prm.warm == 2 ? (double)(T)warm[v] : std::fabs((double)phi[v]);
        const double w = base * (1.0 + prm.warm_margin) + prm.warm_abs_margin;
        if (w < 0.5 * kLarge) { DT[v] = (T)w; warm_applied[v] = (T)w; any = true; }
      }
      if (any) { npatch++; in[p] = 1; }
    }
    act.clear();
    for (int p = 0; p < P; ++p) if (in[p]) act.push_back(p);
    return npatch;
  }

  // Acceptance of a warm start. The FIM only lowers values, and a value it
  // lowers is an element solve from corner values no lower than their final
  // ones, so (the solve being monotone) it is no lower than what that element
  // gives from the final values. The only values no solve produced are warm
  // values still held at the end: a vertex holding one inside the band must
  // not lie below F_v, the least value any of its element copies gives it
  // from the other corners' final values (no seed is among them). An undercut
  // region of warm values has such a vertex at its minimum (nothing lower
  // around it can support it), so every undercut deeper than the threshold
  // per element layer is caught; the plain method's own output has no warm
  // values and always passes. Returns max(F_v - DT_v) over those vertices
  // (0 when there are none).
  std::vector<float> resid_v;  // per vertex: F_v - DT_v where checked, else 0
  float residual(const lso_params& prm) {
    const T bw = (T)prm.bandwidth;
    float r = 0.f;
    resid_v.assign(N, 0.f);
    for (int p = 0; p < P; ++p) {
      for (int32_t v = vert_offsets[p]; v < vert_offsets[p + 1]; ++v) {
        if (label[v] == SeedPoint || !(DT[v] <= bw) || DT[v] != warm_applied[v]) continue;
        T f = (T)kLarge;
        for (int32_t l : lmem[v]) {  // v's slots in its own patch: copy * nc + corner
          const int64_t c = ele_offsets[p] + l / nc;
          const int k = l % nc;
          T val[4], L[6];
          for (int i = 0; i < nc; ++i) val[i] = DT[ele[i * F_of() + c]];
          for (int q = 0; q < (dim == 3 ? 6 : 3); ++q) L[q] = coords[q * F_of() + c];
          const T g = dim == 3 ? TetGeom(L).corner_value(val, k, prm.lambda_eps)
                               : TriGeom(L).corner_value(val, k, prm.lambda_eps);
          f = std::fmin(f, g);
        }
        resid_v[v] = (float)((double)f - (double)DT[v]);
        r = std::fmax(r, resid_v[v]);
      }
    }
    return r;
  }
  int64_t F_of() const { return F; }

  // CFL bound (kernel_compute_timestep2d, src/core2d/include/meshFIM2d_kernels.h:405-475,
  // reduced by meshFIM2d::compute_deltaT, src/core2d/cuda/meshFIM2d.cu:609-674),
  // for both element kinds: per element copy of the listed patches, with R
  // the inscribed radius from the copy's local coordinates and d the
  // dimension, dt1 = 4R^2/|Ccurv|/2/d if |Ccurv| > eps, dt2 = 2R/(|Ceik| + |Cadv|)/d
  // if |Ceik| + |Cadv| > eps (eps = 1e-12, meshFIM2d.h:15); the least
  // min(dt1, dt2) and the largest R. Coefficients per original element,
  // 10 and 0 by default (meshFIM2d.cu:724-731).
  std::vector<double> ceik_e, ccurv_e;
  void set_cfl_coefficients(const double* ceik, const double* ccurv) override {
    if (ceik) ceik_e.assign(ceik, ceik + E); else ceik_e.clear();
    if (ccurv) ccurv_e.assign(ccurv, ccurv + E); else ccurv_e.clear();
  }
  static double inradius_tet(const double* L) {
    const double v6 = std::fabs(L[0] * L[2] * L[5]);
    const double abc = std::fabs(L[0] * L[2]);
    const double abd = std::fabs(L[0]) * std::sqrt(L[4] * L[4] + L[5] * L[5]);
    const double ax = L[2] * L[5], ay = -L[1] * L[5], az = L[1] * L[4] - L[2] * L[3];
    const double acd = std::sqrt((ax * ax + ay * ay) + az * az);
    const double bx = L[2] * L[5], by = -(L[1] - L[0]) * L[5], bz = (L[1] - L[0]) * L[4] - L[2] * (L[3] - L[0]);
    const double bcd = std::sqrt((bx * bx + by * by) + bz * bz);
    return v6 / (((abc + abd) + acd) + bcd);
  }
  static double inradius_tri(const double* c) {
    const double a2 = std::fabs(c[0] * c[2]);
    const double per = (std::fabs(c[0]) + std::sqrt((c[1] - c[0]) * (c[1] - c[0]) + c[2] * c[2])) +
                       std::sqrt(c[1] * c[1] + c[2] * c[2]);
    return a2 / per;
  }
  void cfl_bound(const std::vector<int32_t>& patches, double& dtmin, double& rmax) const {
    const int d = dim, ncoord = dim == 3 ? 6 : 3;
    dtmin = kLarge;
    rmax = -1.0;
    for (int32_t p : patches)
      for (int32_t c = ele_offsets[p]; c < ele_offsets[p + 1]; ++c) {
        double L[6];
        for (int k = 0; k < ncoord; ++k) L[k] = (double)coords[k * F + c];
        const double rin = dim == 3 ? inradius_tet(L) : inradius_tri(L);
        double a2 = 0.0;
        for (int k = 0; k < d; ++k) a2 += (double)sigma[k * F + c] * (double)sigma[k * F + c];
        const int e = ele_permute[c];
        const double ceik = std::fabs(ceik_e.empty() ? 10.0 : ceik_e[e]);
        const double ccurv = std::fabs(ccurv_e.empty() ? 0.0 : ccurv_e[e]);
        const double cadv = std::sqrt(a2);
        double dt1 = kLarge, dt2 = kLarge;
        if (ccurv > 1e-12) dt1 = (4.0 * rin * rin) / ccurv / 2.0 / (double)d;
        if (ceik + cadv > 1e-12) dt2 = (rin * 2.0) / (ceik + cadv) / (double)d;
        dtmin = std::fmin(dtmin, std::fmin(dt1, dt2));
        rmax = std::fmax(rmax, rin);
      }
  }
  // Over the current narrowband, or every patch before the first step.
  void compute_dt(double* out) override {
    std::vector<int32_t> all;
    if (narrowband.empty()) { all.resize(P); std::iota(all.begin(), all.end(), 0); }
    cfl_bound(narrowband.empty() ? all : narrowband, out[0], out[1]);
  }
  void get_resid(float* out) const override {
    for (int64_t v = 0; v < N; ++v) out[vert_permute[v]] = resid_v.empty() ? 0.f : resid_v[v];
  }

  // One timestep (meshFIM3d::GenerateData loop body, meshFIM3d.cu:639-678).
  int step(const lso_params& prm, lso_stats& st) override {
    if (prm.nthreads > 0) omp_set_num_threads(prm.nthreads);
    std::memset(&st, 0, sizeof(st));
    std::fill(amb.begin(), amb.end(), 0);
    const std::vector<int32_t> prev_nb = narrowband;
    std::vector<int32_t> act = find_seeds();
    st.n_seed_patches = (int)act.size();
    for (int64_t v = 0; v < N; ++v) tsign[v] = (phi[v] < (T)0) ? 0 : 1;  // kernel_reinit_Tsign
    const bool use_warm = prm.warm != 0 && !prev_nb.empty();
    if (use_warm) st.warm_patches = apply_warm(prev_nb, prm, act);
    fim(act, prm, st);
    if (use_warm) {
      const float r = residual(prm);
      st.warm_residual = r;
      const float rt = (float)(prm.warm_res_tol > 0 ? prm.warm_res_tol : prm.tol);
      if (!(r <= rt)) {  // the warm start is rejected: the plain method from the same phi
        st.warm_fallback = 1;
        act = find_seeds();
        fim(act, prm, st);
      }
    }
    narrowband_and_restore(prm, st);
    st.dt_used = prm.dt;
    if (narrowband.empty()) return 0;
    lso_params aprm = prm;
    if (prm.cfl) {
      double b[2];
      cfl_bound(narrowband, b[0], b[1]);
      aprm.dt = std::fmin(prm.dt, b[0]);
      st.dt_used = aprm.dt;
    }
    for (int it = 0; it < prm.inside_niter; ++it) {
      #pragma omp parallel for schedule(dynamic, 1)
      for (int64_t k = 0; k < (int64_t)narrowband.size(); ++k) advect_patch(narrowband[k], aprm);
      for (int32_t p : narrowband)  // CopyOutBack_levelset3d
        for (int32_t v = vert_offsets[p]; v < vert_offsets[p + 1]; ++v) phi[v] = phi_out[v];
    }
    return 0;
  }
};

}  // namespace

extern "C" {

void* lso_create(int dim, int precision, int64_t N, const double* xyz, int64_t E,
                 const int32_t* cells, const int32_t* labels, int nparts) {
  if (dim != 2 && dim != 3) return nullptr;
  if (precision == 64) return new Solver<double>(dim, N, xyz, E, cells, labels, nparts);
  if (precision == 32) return new Solver<float>(dim, N, xyz, E, cells, labels, nparts);
  return nullptr;
}
void lso_destroy(void* h) { delete static_cast<Base*>(h); }
int lso_set_phi(void* h, const double* phi) { return static_cast<Base*>(h)->set_phi(phi); }
int lso_set_velocity(void* h, const double* vel, int mode) {
  return static_cast<Base*>(h)->set_velocity(vel, mode);
}
int lso_step(void* h, const lso_params* prm, lso_stats* st) { return static_cast<Base*>(h)->step(*prm, *st); }
void lso_get_phi(void* h, double* out) { static_cast<Base*>(h)->get_phi(out); }
int lso_get_narrowband(void* h, int32_t* ids) { return static_cast<Base*>(h)->get_nb(ids); }
int64_t lso_num_copies(void* h) { return static_cast<Base*>(h)->ncopies(); }
// Replaces the previous-step narrowband (seeds scan these patches next step).
void lso_set_narrowband(void* h, const int32_t* ids, int n) { static_cast<Base*>(h)->set_nb(ids, n); }
// Per-vertex (caller order) flag of threshold-ambiguous advection updates.
void lso_get_ambiguous(void* h, uint8_t* out) { static_cast<Base*>(h)->get_amb(out); }
// Advection alone over all patches (tests of the advection scheme).
void lso_advect_all(void* h, double dt) { static_cast<Base*>(h)->advect_all(dt); }
// Warm-start distances for the next step only (caller order; >= LARGENUM/2: none).
void lso_set_warm(void* h, const double* warm) { static_cast<Base*>(h)->set_warm(warm); }
// Per-vertex acceptance residuals of the last warm step (caller order).
void lso_get_warm_residuals(void* h, float* out) { static_cast<Base*>(h)->get_resid(out); }
void lso_set_cfl_coefficients(void* h, const double* ceik, const double* ccurv) {
  static_cast<Base*>(h)->set_cfl_coefficients(ceik, ccurv);
}
// Vertex labels of the last step (SeedPoint 0 ... FarPoint 3), caller order.
void lso_get_labels(void* h, int32_t* out) { static_cast<Base*>(h)->get_labels(out); }
// out[0]: CFL bound, out[1]: largest inscribed radius.
void lso_compute_dt(void* h, double* out) { static_cast<Base*>(h)->compute_dt(out); }
// Inscribed radius from local coordinates (6 for a tet, 3 for a triangle).
double lso_inradius(int dim, const double* L) {
  return dim == 3 ? Solver<double>::inradius_tet(L) : Solver<double>::inradius_tri(L);
}
// Per-copy local coordinates in the storage type, copy-major.
void lso_get_coords(void* h, double* out) { static_cast<Base*>(h)->get_coords(out); }
void lso_get_layout(void* h, int32_t* vert_perm, int32_t* ele_perm, int32_t* vert_off,
                    int32_t* ele_off, int32_t* ele_soa) {
  static_cast<Base*>(h)->get_layout(vert_perm, ele_perm, vert_off, ele_off, ele_soa);
}

// Unit-level local solves for tests. T = {TA, TB, TC}; d = {ACAC, ACBC, ACCD, BCBC, BCCD, CDCD}.
double lso_tet_solve(int precision, const double* T, const double* d, double leps) {
  if (precision == 32)
    return tet_solve<float>((float)T[0], (float)T[1], (float)T[2], (float)d[0], (float)d[1],
                            (float)d[2], (float)d[3], (float)d[4], (float)d[5], (float)leps);
  return tet_solve<double>(T[0], T[1], T[2], d[0], d[1], d[2], d[3], d[4], d[5], leps);
}
double lso_tri_solve(int precision, double TA, double TB, double AB, double BC, double AC) {
  if (precision == 32) return tri_solve<float>((float)TA, (float)TB, (float)AB, (float)BC, (float)AC);
  return tri_solve<double>(TA, TB, AB, BC, AC);
}
// Element update of one tet from six local coordinates (v = 4 slot values, in/out).
void lso_tet_element_update(int precision, const double* L, double* v, double leps) {
  if (precision == 32) {
    float Lf[6], vf[4];
    for (int k = 0; k < 6; ++k) Lf[k] = (float)L[k];
    for (int k = 0; k < 4; ++k) vf[k] = (float)v[k];
    TetGeom(Lf).update(vf, leps);
    for (int k = 0; k < 4; ++k) v[k] = vf[k];
  } else {
    TetGeom(L).update(v, leps);
  }
}
int lso_max_threads() { return omp_get_max_threads(); }

}  // extern "C"
