// gfx950 kernels for the level-set per-timestep hot path.
//
// One workgroup owns one patch at a time (512 threads for FIM sweeps, 256
// for the other phases): the patch's owned and halo vertex values, its
// element-copy slots and its vertex slot lists are staged in LDS, so HBM sees
// each patch's arrays once per visit and every halo vertex once per patch
// (the reference re-reads halo values per element copy,
// src/core3d/include/redistance_kernels3d.h:342-358).
//
// Values are stored in the working type T (FP32 or FP64 for tets, FP64 for
// triangles); every element update (local solves, advection, curvature) is
// evaluated in FP64 from the stored values and rounded to T, with
// contraction off and correctly rounded division and square root.
//
// Two execution paths give identical results: one launch per phase with the
// FIM loop on the host (k_seeds, k_sweep, k_copyback, k_activate,
// k_finalize, k_narrowband, k_advect), and one persistent cooperative launch
// per timestep (k_step) that runs seeds, the FIM loop and the narrowband on
// the device, followed by the advection launches (k_advect_dev,
// k_copy_phi_dev).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/levelset_hip.h"
#include "ls_device.h"
#include "ls_engine.h"
#include "ls_mlp.h"

namespace lsamd {
namespace kern {

// sqrtf and '/' are correctly rounded under -fhip-fp32-correctly-rounded-divide-sqrt;
// __fsqrt_rn is not used because it maps to the native (approximate) sqrt.
__device__ __forceinline__ float d_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double d_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float d_div(float a, float b) { return a / b; }
__device__ __forceinline__ double d_div(double a, double b) { return a / b; }
__device__ __forceinline__ float d_min(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double d_min(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float d_max(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double d_max(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float d_abs(float a) { return fabsf(a); }
__device__ __forceinline__ double d_abs(double a) { return fabs(a); }

// Square root of a squared length of a mesh whose elements' least heights
// are at least 2^-350 and 2^-20 of their size (checked when the heights are
// computed, DevMesh::fast_len): the distances the tet solve takes roots of
// are at least a height, computed with errors far below it, so the argument
// is a normal number of at least 2^-700. This is the correctly
// rounded sequence the compiler emits for sqrt on gfx950 (a reciprocal
// square-root seed and two Newton-Raphson refinements, Markstein's scheme)
// without its rescaling of arguments below 2^-767 and its zero/infinity
// selects, which such arguments never take: the same bits as sqrt, in 10
// instead of 18 instructions.
__device__ __forceinline__ double sqrt_len(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = y * 0.5;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}
template <bool FAST>
__device__ __forceinline__ double len_sqrt(double x) {
  if constexpr (FAST) return sqrt_len(x);
  else return sqrt(x);
}

// A discriminant within this fraction of q^2 + |4pr| of zero is a double
// root: equal values at both ends of an edge, or a front parallel to a face,
// make it exactly zero, and rounding alone must not decide whether the
// candidate exists.
constexpr double kDiscRel = 1e-12;

// Roots of p x^2 + q x + r with the discriminant and roots in FP64. Every
// root in range is a point of the face whose Hopf-Lax value is an admissible
// candidate, so accepting a double root never undercuts the true minimum.
template <typename T>
__device__ __forceinline__ bool quad_roots(T p, T q, T r, T& x1, T& x2) {
  const double dp = p, dq = q, dr = r;
  double delta = dq * dq - 4.0 * dp * dr;
  if (delta < 0.0 && delta >= -kDiscRel * (dq * dq + d_abs(4.0 * dp * dr))) delta = 0.0;
  if (!(delta >= 0.0)) return false;
  const double sq = sqrt(delta);
  const double inv = 1.0 / (2.0 * dp);  // one division for both roots
  x1 = (T)((-dq + sq) * inv);
  x2 = (T)((-dq - sq) * inv);
  return true;
}

// Tet Hopf-Lax solve for the fourth corner (scheme of
// src/core3d/include/redistance_kernels3d.h:124-285).
template <bool FAST, typename T>
__device__ T tet_solve(T TA, T TB, T TC, T ACAC, T ACBC, T ACCD, T BCBC, T BCCD, T CDCD, T leps) {
  const T big = (T)kLargeNum;
  if (TA >= big && TB >= big && TC >= big) return big;
  T FaceTAC = big, FaceTAB = big, FaceTBC = big;
  T TD = big;
  const T TAC = TC - TA, TBC = TC - TB, TAB = TB - TA;
  T x1, x2;
  {
    const T p = BCBC * TBC * TBC - BCBC * BCBC;
    const T q = (BCCD + BCCD) * TBC * TBC - (T)2 * BCBC * BCCD;
    const T r = TBC * TBC * CDCD - BCCD * BCCD;
    if (quad_roots(p, q, r, x1, x2)) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const T l2 = k == 0 ? x1 : x2;
        const T l3 = (T)1 - l2;
        if (l2 >= (T)0 && l2 <= (T)1) {
          const T TE = TB * l2 + TC * l3;
          FaceTBC = d_min(FaceTBC, TE + len_sqrt<FAST>((l2 * BCBC + BCCD) * l2 + (l2 * BCCD + CDCD)));
        }
      }
    }
    FaceTBC = d_min(FaceTBC, d_min(TB + len_sqrt<FAST>((BCBC + BCCD) + (BCCD + CDCD)), TC + len_sqrt<FAST>(CDCD)));
  }
  {
    const T gx = ACAC - ACBC, gy = ACBC - BCBC, gz = ACCD - BCCD;
    const T t2 = TAB * TAB;
    const T p = (t2 * ACAC - gx * gx) + (BCBC * t2 - gy * gy) - ((ACBC + ACBC) * t2 - (T)2 * gx * gy);
    const T q = -(BCBC * t2 - gy * gy) * (T)2 + ((ACBC + ACBC) * t2 - (T)2 * gx * gy) +
                ((ACCD + ACCD) * t2 - (T)2 * gx * gz) - ((BCCD + BCCD) * t2 - (T)2 * gy * gz);
    const T r = (t2 * BCBC - gy * gy) + ((BCCD + BCCD) * t2 - (T)2 * gy * gz) + (t2 * CDCD - gz * gz);
    if (quad_roots(p, q, r, x1, x2)) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const T l1 = k == 0 ? x1 : x2;
        const T l2 = (T)1 - l1;
        if (l1 >= (T)0 && l1 <= (T)1) {
          const T TE = TA * l1 + TB * l2;
          FaceTAB = d_min(FaceTAB, TE + len_sqrt<FAST>((l1 * ACAC + l2 * ACBC + ACCD) * l1 +
                                               (l1 * ACBC + l2 * BCBC + BCCD) * l2 +
                                               (l1 * ACCD + l2 * BCCD + CDCD)));
        }
      }
    }
    FaceTAB = d_min(FaceTAB, d_min(TB + len_sqrt<FAST>((BCBC + BCCD) + (BCCD + CDCD)),
                                   TA + len_sqrt<FAST>((ACAC + ACCD) + (ACCD + CDCD))));
  }
  {
    const T p = ACAC * TAC * TAC - ACAC * ACAC;
    const T q = (ACCD + ACCD) * TAC * TAC - (T)2 * ACAC * ACCD;
    const T r = TAC * TAC * CDCD - ACCD * ACCD;
    if (quad_roots(p, q, r, x1, x2)) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const T l1 = k == 0 ? x1 : x2;
        const T l3 = (T)1 - l1;
        if (l1 >= (T)0 && l1 <= (T)1) {
          const T TE = TA * l1 + TC * l3;
          FaceTAC = d_min(FaceTAC, TE + len_sqrt<FAST>((l1 * ACAC + ACCD) * l1 + (l1 * ACCD + CDCD)));
        }
      }
    }
    FaceTAC = d_min(FaceTAC, d_min(TA + len_sqrt<FAST>((ACAC + ACCD) + (ACCD + CDCD)), TC + len_sqrt<FAST>(CDCD)));
  }
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
    if (quad_roots(p, q, r, x1, x2)) {
      const T ihd = d_div((T)1, h + leps);
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const T l1 = k == 0 ? x1 : x2;
        const T l2 = (s * l1 + t) * ihd;
        const T l3 = (T)1 - l1 - l2;
        if (l1 >= (T)0 && l1 <= (T)1 && l2 >= (T)0 && l2 <= (T)1 && l3 >= (T)0 && l3 <= (T)1) {
          const T TE = TA * l1 + TB * l2 + TC * l3;
          TD = d_min(TD, TE + len_sqrt<FAST>((l1 * ACAC + l2 * ACBC + ACCD) * l1 +
                                     (l1 * ACBC + l2 * BCBC + BCCD) * l2 +
                                     (l1 * ACCD + l2 * BCCD + CDCD)));
        }
      }
    }
  }
  return d_min(TD, d_min(FaceTBC, d_min(FaceTAB, FaceTAC)));
}

// Triangle Hopf-Lax solve (scheme of src/core2d/include/redistance_kernels.h:287-334).
template <typename T>
__device__ T tri_solve(T TA, T TB, T LenAB, T LenBC, T LenAC) {
  const T big = (T)kLargeNum;
  if (TA >= big && TB >= big) return big;
  T TC1 = big, TC2 = big;
  const T TAB = TB - TA;
  const T a = (LenAB * LenAB - TAB * TAB) * LenAB * LenAB;
  const T EdgeTA = TA + LenAC;
  const T EdgeTB = TB + LenBC;
  if (a > (T)0) {
    const T cosA = d_div(LenAC * LenAC + LenAB * LenAB - LenBC * LenBC, (T)2 * LenAC * LenAB);
    const T b = (T)2 * LenAB * LenAC * cosA * (TAB * TAB - LenAB * LenAB);
    const T delta = (T)4 * LenAC * LenAC * a * TAB * TAB * ((T)1 - cosA * cosA);
    const T sq = d_sqrt(delta);
    const T l1 = d_div(-b + sq, (T)2 * a);
    const T l2 = d_div(-b - sq, (T)2 * a);
    if (l1 >= (T)0 && l1 <= (T)1) {
      const T LenAD = l1 * LenAB;
      const T LenCD = d_sqrt(LenAC * LenAC + LenAD * LenAD - (T)2 * LenAC * LenAD * cosA);
      TC1 = l1 * TAB + TA + LenCD;
    }
    if (l2 >= (T)0 && l2 <= (T)1) {
      const T LenAD = l2 * LenAB;
      const T LenCD = d_sqrt(LenAC * LenAC + LenAD * LenAD - (T)2 * LenAC * LenAD * cosA);
      TC2 = l2 * TAB + TA + LenCD;
    }
    return d_min(TC2, d_min(TC1, d_min(EdgeTA, EdgeTB)));
  }
  return d_min(EdgeTA, EdgeTB);
}

// One element update: every corner takes min(own, solve from the other corners),
// all solves reading the values from before the update
// (src/core3d/include/redistance_kernels3d.h:373-414; 2D redistance_kernels.h:441-462).
// Evaluated in FP64 from the stored values: the solver's quadratic coefficients
// cancel badly in FP32 and its candidate acceptance is discontinuous.
template <typename T, int NC>
struct ElemSolve;

// Each corner's solve only looks at points of the opposite face, so it is at
// least min(other corners) + height of the corner over that face; when that
// bound, less a margin far above FP64 rounding of the candidates, exceeds the
// corner's value, min(value, solve) is the value and the solve is skipped.
// The result is bit-identical to always solving.
__device__ __forceinline__ bool must_solve(double Tk, double mo, double Mo, double hk, double scale) {
  const double margin = 1e-12 * (Mo + hk + scale + 1.0);
  return !(mo + hk - margin > Tk);
}

// Edge dot products of a tet from its six local coordinates
// (redistance_kernels3d.h:381-396) and the permuted local solve for each
// corner (:411-414), all in FP64.
struct TetDots {
  double ACAC, ACBC, ACCD, BCBC, BCCD, CDCD, ADAD, ADBD, ADCD, BDBD, CDBD;
  __device__ explicit TetDots(const double* L) {
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
  // The solve for corner k as one body with its arguments chosen by selects,
  // so a wave whose lanes solve different corners runs it once per round:
  //   k=3: (TA, TB, TC; ACAC,  ACBC,  ACCD, BCBC,  BCCD, CDCD)
  //   k=0: (TB, TD, TC; BCBC, -BCCD, -ACBC, CDCD,  ACCD, ACAC)
  //   k=1: (TA, TD, TC; ACAC, -ACCD, -ACBC, CDCD,  BCCD, BCBC)
  //   k=2: (TA, TB, TD; ADAD,  ADBD, -ADCD, BDBD, -CDBD, CDCD)
  template <bool FAST>
  __device__ double solve(int k, const double* Tv, double leps) const {
    const bool k0 = k == 0, k1 = k == 1, k2 = k == 2, k01 = k0 || k1;
    const double a = k0 ? Tv[1] : Tv[0];
    const double b = k01 ? Tv[3] : Tv[1];
    const double c = k2 ? Tv[3] : Tv[2];
    const double d1 = k0 ? BCBC : (k2 ? ADAD : ACAC);
    const double d2 = k0 ? -BCCD : (k1 ? -ACCD : (k2 ? ADBD : ACBC));
    const double d3 = k01 ? -ACBC : (k2 ? -ADCD : ACCD);
    const double d4 = k01 ? CDCD : (k2 ? BDBD : BCBC);
    const double d5 = k0 ? ACCD : (k2 ? -CDBD : BCCD);
    const double d6 = k0 ? ACAC : (k1 ? BCBC : CDCD);
    return tet_solve<FAST>(a, b, c, d1, d2, d3, d4, d5, d6, leps);
  }
};

template <typename T>
struct ElemSolve<T, 4> {
  static constexpr int kCoords = 6;
  // Heights of each corner over its opposite face, in the local frame
  // A=(0,0,0), B=(L0,0,0), C=(L1,L2,0), D=(L3,L4,L5).
  __device__ static void heights(const double* L, double* h) {
    const double v6 = d_abs(L[0] * L[2] * L[5]);
    const double nAx = L[2] * L[5], nAy = -(L[1] - L[0]) * L[5], nAz = (L[1] - L[0]) * L[4] - L[2] * (L[3] - L[0]);
    const double nBx = L[2] * L[5], nBy = -L[1] * L[5], nBz = L[1] * L[4] - L[2] * L[3];
    h[0] = v6 / d_sqrt(nAx * nAx + nAy * nAy + nAz * nAz);
    h[1] = v6 / d_sqrt(nBx * nBx + nBy * nBy + nBz * nBz);
    h[2] = d_abs(L[2] * L[5]) / d_sqrt(L[4] * L[4] + L[5] * L[5]);
    h[3] = d_abs(L[5]);
  }
  // Sum of the absolute local coordinates: the size term of the skip's
  // rounding margin.
  __device__ static double scale(const double* L) {
    return L[0] + d_abs(L[1]) + d_abs(L[2]) + d_abs(L[3]) + d_abs(L[4]) + d_abs(L[5]);
  }
  // Bit k set: corner k must be solved (its value is not below the exact
  // lower bound min(other values) + height of corner k over its face).
  // `hs` holds the copy's heights in FP16 rounded toward zero and `scale` is
  // at least the copy's coordinate scale (the patch's largest): a smaller
  // height or a larger margin only solves more corners, so the skip stays
  // exact, and the pass reads 8 bytes per copy instead of the coordinates.
  __device__ static int need_mask(const __half* hs, const T* v, double scale) {
    const double Tv[4] = {(double)v[0], (double)v[1], (double)v[2], (double)v[3]};
    int mask = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double mo = 1e300, Mo = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j != k) { mo = d_min(mo, Tv[j]); Mo = d_max(Mo, Tv[j]); }
      if (must_solve(Tv[k], mo, Mo, (double)__half2float(hs[k]), scale)) mask |= 1 << k;
    }
    return mask;
  }
  // Supporting-plane bound, applied to the corners `mask` leaves. Corner
  // k's solve is the least of f(E) = T(E) + |D - E| over points E of the
  // opposite face (T linear on it, D the corner): every candidate the solve
  // evaluates is f at a point of the face, the roots' rounding only moving
  // that point. f is convex, so for any face point E0 and every E,
  // f(E) >= f(E0) + grad f(E0).(E - E0), which is linear in E; over the
  // face it is least at a face corner c, where it equals T_c + (c - D).w,
  // w the unit vector from D towards E0. With E0 the foot of the
  // perpendicular this is the height bound; here E0 runs over the three
  // face corners and the bound is the largest of the three (the centroid as
  // a fourth point pruned 5 % more solves in a model but cost more VALU time
  // than they take). The products (c - D).(c' - D) come from squared edge lengths (law
  // of cosines), evaluated in FP32: a margin of 1e-5 (max other value +
  // scale + 1) lies far above their rounding (below 1e-6 of the same, as
  // every edge is at most `scale` long) and the FP64 rounding of the
  // candidates, so a corner is dropped only when its solve cannot go below
  // its value and the result is bit-identical to solving it. DevMesh::plane
  // is set when element sizes keep the squared lengths normal FP32 numbers.
  // On the 64^3 torus it removes about a quarter of the solves the height
  // bound leaves.
  __device__ static float plane_bound(float T0, float T1, float T2, float qk0, float qk1, float qk2, float q01,
                                      float q02, float q12, float n0, float n1, float n2, float r0, float r1,
                                      float r2) {
    const float G01 = 0.5f * (qk0 + qk1 - q01), G02 = 0.5f * (qk0 + qk2 - q02), G12 = 0.5f * (qk1 + qk2 - q12);
    const float b0 = fminf(T0 + n0, fminf(fmaf(G01, r0, T1), fmaf(G02, r0, T2)));
    const float b1 = fminf(T1 + n1, fminf(fmaf(G01, r1, T0), fmaf(G12, r1, T2)));
    const float b2 = fminf(T2 + n2, fminf(fmaf(G02, r2, T0), fmaf(G12, r2, T1)));
    return fmaxf(fmaxf(b0, b1), b2);
  }
  __device__ static int plane_mask(const T* Lt, const T* v, int mask, float scale) {
    float L[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) L[k] = (float)Lt[k];
    // Squared edge lengths in the local frame A=(0,0,0), B=(L0,0,0),
    // C=(L1,L2,0), D=(L3,L4,L5), their roots and reciprocal roots.
    const float d10 = L[1] - L[0], d30 = L[3] - L[0], d31 = L[3] - L[1], d42 = L[4] - L[2];
    const float qAB = L[0] * L[0], qAC = L[1] * L[1] + L[2] * L[2], qAD = L[3] * L[3] + L[4] * L[4] + L[5] * L[5];
    const float qBC = d10 * d10 + L[2] * L[2], qBD = d30 * d30 + L[4] * L[4] + L[5] * L[5];
    const float qCD = d31 * d31 + d42 * d42 + L[5] * L[5];
    const float nAB = __builtin_amdgcn_sqrtf(qAB), nAC = __builtin_amdgcn_sqrtf(qAC);
    const float nAD = __builtin_amdgcn_sqrtf(qAD), nBC = __builtin_amdgcn_sqrtf(qBC);
    const float nBD = __builtin_amdgcn_sqrtf(qBD), nCD = __builtin_amdgcn_sqrtf(qCD);
    const float rAB = __builtin_amdgcn_rcpf(nAB), rAC = __builtin_amdgcn_rcpf(nAC);
    const float rAD = __builtin_amdgcn_rcpf(nAD), rBC = __builtin_amdgcn_rcpf(nBC);
    const float rBD = __builtin_amdgcn_rcpf(nBD), rCD = __builtin_amdgcn_rcpf(nCD);
    const float TA = (float)v[0], TB = (float)v[1], TC = (float)v[2], TD = (float)v[3];
    int out = mask;
    if (mask & 1) {  // A over face B, C, D
      const float b = plane_bound(TB, TC, TD, qAB, qAC, qAD, qBC, qBD, qCD, nAB, nAC, nAD, rAB, rAC, rAD);
      if (b - 1e-5f * (fmaxf(TB, fmaxf(TC, TD)) + scale + 1.0f) > TA) out &= ~1;
    }
    if (mask & 2) {  // B over face A, C, D
      const float b = plane_bound(TA, TC, TD, qAB, qBC, qBD, qAC, qAD, qCD, nAB, nBC, nBD, rAB, rBC, rBD);
      if (b - 1e-5f * (fmaxf(TA, fmaxf(TC, TD)) + scale + 1.0f) > TB) out &= ~2;
    }
    if (mask & 4) {  // C over face A, B, D
      const float b = plane_bound(TA, TB, TD, qAC, qBC, qCD, qAB, qAD, qBD, nAC, nBC, nCD, rAC, rBC, rCD);
      if (b - 1e-5f * (fmaxf(TA, fmaxf(TB, TD)) + scale + 1.0f) > TC) out &= ~4;
    }
    if (mask & 8) {  // D over face A, B, C
      const float b = plane_bound(TA, TB, TC, qAD, qBD, qCD, qAB, qAC, qBC, nAD, nBD, nCD, rAD, rBD, rCD);
      if (b - 1e-5f * (fmaxf(TA, fmaxf(TB, TC)) + scale + 1.0f) > TD) out &= ~8;
    }
    return out;
  }
  // The same from the copy's four FP16 heights packed in 8 bytes.
  __device__ static int need_mask(const uint2 hw, const T* v, double scale) {
    const __half hs[4] = {__ushort_as_half((unsigned short)(hw.x & 0xffffu)),
                          __ushort_as_half((unsigned short)(hw.x >> 16)),
                          __ushort_as_half((unsigned short)(hw.y & 0xffffu)),
                          __ushort_as_half((unsigned short)(hw.y >> 16))};
    return need_mask(hs, v, scale);
  }
  // Corner k's updated value min(v_k, solve_k) from the element's values v.
  // fast: the mesh allows sqrt_len for lengths (DevMesh::fast_len); the
  // uniform branch picks one of two copies of the solve, same bits.
  template <bool FAST>
  __device__ static T corner_update_t(const T* Lt, const T* v, int k, double leps) {
    double L[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) L[j] = (double)Lt[j];
    const double Tv[4] = {(double)v[0], (double)v[1], (double)v[2], (double)v[3]};
    // Selects instead of Tv[k]: a runtime index would put Tv in scratch memory.
    const double Tk = k == 0 ? Tv[0] : (k == 1 ? Tv[1] : (k == 2 ? Tv[2] : Tv[3]));
    const TetDots d(L);
    return (T)d_min(Tk, d.template solve<FAST>(k, Tv, leps));
  }
  __device__ static T corner_update(const T* Lt, const T* v, int k, double leps, bool fast) {
    return fast ? corner_update_t<true>(Lt, v, k, leps) : corner_update_t<false>(Lt, v, k, leps);
  }
  // Solves the corners of `mask` (from need_mask) into v.
  template <bool FAST>
  __device__ static void solve_mask_t(const T* Lt, int mask, T* v, double leps) {
    double L[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) L[k] = (double)Lt[k];
    const double Tv[4] = {(double)v[0], (double)v[1], (double)v[2], (double)v[3]};
    const TetDots d(L);
    // Every solve reads the values from before the update, so the order of
    // the corners does not matter; each lane walks only its own corners.
#pragma nounroll
    for (int m = mask; m; m &= m - 1) {
      const int k = __ffs(m) - 1;
      const double Tk = k == 0 ? Tv[0] : (k == 1 ? Tv[1] : (k == 2 ? Tv[2] : Tv[3]));
      const T y = (T)d_min(Tk, d.template solve<FAST>(k, Tv, leps));
      if (k == 0) v[0] = y;
      else if (k == 1) v[1] = y;
      else if (k == 2) v[2] = y;
      else v[3] = y;
    }
  }
  __device__ static void solve_mask(const T* Lt, int mask, T* v, double leps, bool fast) {
    if (fast) solve_mask_t<true>(Lt, mask, v, leps);
    else solve_mask_t<false>(Lt, mask, v, leps);
  }
  // The value the element gives corner k from the other three (no min with
  // corner k's own value, no skip): what update() stores when it solves k.
  __device__ static T corner_value(const T* Lt, const T* v, int k, double leps, bool fast) {
    double L[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) L[j] = (double)Lt[j];
    const double Tv[4] = {(double)v[0], (double)v[1], (double)v[2], (double)v[3]};
    const TetDots d(L);
    return (T)d_min(kLargeNum, fast ? d.template solve<true>(k, Tv, leps) : d.template solve<false>(k, Tv, leps));
  }
};

template <typename T>
struct ElemSolve<T, 3> {
  static constexpr int kCoords = 3;
  __device__ static int need_mask(const T* ct, const T* v) {
    const double c[3] = {(double)ct[0], (double)ct[1], (double)ct[2]};
    const double LenAB = d_abs(c[0]);
    const double LenBC = d_sqrt(c[2] * c[2] + (c[1] - c[0]) * (c[1] - c[0]));
    const double LenCA = d_sqrt(c[1] * c[1] + c[2] * c[2]);
    const double Tv[3] = {(double)v[0], (double)v[1], (double)v[2]};
    // Heights over the opposite edge: A=(0,0), B=(c0,0), C=(c1,c2).
    const double a2 = d_abs(c[0] * c[2]);
    const double h[3] = {a2 / LenBC, a2 / LenCA, d_abs(c[2])};
    const double scale = LenAB + LenBC + LenCA;
    int mask = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double o1 = Tv[(k + 1) % 3], o2 = Tv[(k + 2) % 3];
      if (must_solve(Tv[k], d_min(o1, o2), d_max(o1, o2), h[k], scale)) mask |= 1 << k;
    }
    return mask;
  }
  __device__ static T corner_update(const T* ct, const T* v, int k, double leps, bool) {
    const double vk = k == 0 ? (double)v[0] : (k == 1 ? (double)v[1] : (double)v[2]);
    return (T)d_min(vk, (double)corner_value(ct, v, k, leps, false));
  }
  __device__ static void solve_mask(const T* ct, int mask, T* v, double leps, bool) {
    const double c[3] = {(double)ct[0], (double)ct[1], (double)ct[2]};
    const double LenAB = d_abs(c[0]);
    const double LenBC = d_sqrt(c[2] * c[2] + (c[1] - c[0]) * (c[1] - c[0]));
    const double LenCA = d_sqrt(c[1] * c[1] + c[2] * c[2]);
    const double TA = (double)v[0], TB = (double)v[1], TC = (double)v[2];
    if (mask & 4) v[2] = (T)d_min(TC, tri_solve(TA, TB, LenAB, LenBC, LenCA));
    if (mask & 1) v[0] = (T)d_min(TA, tri_solve(TB, TC, LenBC, LenCA, LenAB));
    if (mask & 2) v[1] = (T)d_min(TB, tri_solve(TC, TA, LenCA, LenAB, LenBC));
  }
  __device__ static T corner_value(const T* ct, const T* v, int k, double, bool) {
    const double c[3] = {(double)ct[0], (double)ct[1], (double)ct[2]};
    const double LenAB = d_abs(c[0]);
    const double LenBC = d_sqrt(c[2] * c[2] + (c[1] - c[0]) * (c[1] - c[0]));
    const double LenCA = d_sqrt(c[1] * c[1] + c[2] * c[2]);
    const double TA = (double)v[0], TB = (double)v[1], TC = (double)v[2];
    double r;
    if (k == 2) r = tri_solve(TA, TB, LenAB, LenBC, LenCA);
    else if (k == 0) r = tri_solve(TB, TC, LenBC, LenCA, LenAB);
    else r = tri_solve(TC, TA, LenCA, LenAB, LenBC);
    return (T)d_min(kLargeNum, r);
  }
};

// Need mask of element copy c of patch p with values x: tets from the
// stored heights and the patch's scale bound, triangles from their
// coordinates.
template <typename T, int NC>
__device__ __forceinline__ int copy_need_mask(const DevMesh<T>& m, int p, int64_t c, const T* x) {
  if constexpr (NC == 4) {
    return ElemSolve<T, 4>::need_mask(m.hgt + c * 4, x, (double)m.pscale[p]);
  } else {
    T L[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) L[q] = m.coords[c * 3 + q];
    return ElemSolve<T, 3>::need_mask(L, x);
  }
}

// copy_need_mask for element e of a patch whose first copy is es, with the
// heights staged in LDS (s_hgt, 8 bytes per copy) when `stage` is set.
template <typename T, int NC>
__device__ __forceinline__ int staged_need_mask(const DevMesh<T>& m, int p, int64_t es, int e, const T* x,
                                                bool stage, const uint2* s_hgt) {
  if constexpr (NC == 4) {
    if (stage) return ElemSolve<T, 4>::need_mask(s_hgt[e], x, (double)m.pscale[p]);
  }
  return copy_need_mask<T, NC>(m, p, es + e, x);
}

// Local frame and coordinates of one element copy in FP64 (scheme of
// kernel_compute_local_coords3d / 2d, src/core3d/include/meshFIM_kernels.h:448-500):
// tets X along AB, Z along AB x AC, Y = Z x X, coordinates
// (|AB|, AC.X, AC.Y, AD.X, AD.Y, AD.Z); triangles X along AB, Y along
// (AB x AC) x AB, coordinates (|AB|, AC.AB / |AB|, AC.Y / |Y|). `fr` gets the
// axes (3 or 2 of them, 3 components each) for velocity projection.
__device__ inline void g_cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double g_dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

template <int NC>
__device__ void copy_frame(const double* xyz, const int32_t* cv, double* lc, double* fr) {
  const double* A = xyz + 3 * (int64_t)cv[0];
  const double* B = xyz + 3 * (int64_t)cv[1];
  const double* C = xyz + 3 * (int64_t)cv[2];
  const double AB[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
  const double AC[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
  const double lenAB = sqrt(g_dot3(AB, AB));
  double nrm[3];
  g_cross3(AB, AC, nrm);
  if constexpr (NC == 4) {
    const double* D = xyz + 3 * (int64_t)cv[3];
    const double AD[3] = {D[0] - A[0], D[1] - A[1], D[2] - A[2]};
    const double lenN = sqrt(g_dot3(nrm, nrm));
    const double Z[3] = {nrm[0] / lenN, nrm[1] / lenN, nrm[2] / lenN};
    const double X[3] = {AB[0] / lenAB, AB[1] / lenAB, AB[2] / lenAB};
    double Y[3];
    g_cross3(Z, X, Y);
    lc[0] = lenAB; lc[1] = g_dot3(AC, X); lc[2] = g_dot3(AC, Y);
    lc[3] = g_dot3(AD, X); lc[4] = g_dot3(AD, Y); lc[5] = g_dot3(AD, Z);
#pragma unroll
    for (int a = 0; a < 3; ++a) { fr[a] = X[a]; fr[3 + a] = Y[a]; fr[6 + a] = Z[a]; }
  } else {
    double Y[3];
    g_cross3(nrm, AB, Y);
    const double lenY = sqrt(g_dot3(Y, Y));
    lc[0] = lenAB; lc[1] = g_dot3(AC, AB) / lenAB; lc[2] = g_dot3(AC, Y) / lenY;
#pragma unroll
    for (int a = 0; a < 3; ++a) { fr[a] = AB[a] / lenAB; fr[3 + a] = Y[a] / lenY; }
  }
}

// Local coordinates of every copy (cv: F x NC renumbered corner ids, xyz:
// renumbered vertex coordinates), rounded to the storage type.
template <typename T, int NC>
__global__ void k_local_coords(const int32_t* cv, const double* xyz, int64_t F, T* coords) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= F) return;
  constexpr int K = NC == 4 ? 6 : 3;
  double lc[6], fr[9];
  copy_frame<NC>(xyz, cv + c * NC, lc, fr);
#pragma unroll
  for (int k = 0; k < K; ++k) coords[c * K + k] = (T)lc[k];
}

// Per-copy velocity in the copy's local frame from one velocity per original
// element (vel: E x 3). With patch_sigma every copy of a patch uses the
// velocity of the patch's first copy, as the reference's advection reads it
// (src/core3d/include/meshFIM_kernels.h:51-53).
template <typename T, int NC>
__global__ void k_project_velocity(const int32_t* cv, const double* xyz, const int32_t* copy_src,
                                   const int32_t* ele_off, int P, int64_t F, const double* vel, int patch_sigma,
                                   T* sigma) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= F) return;
  int64_t src = copy_src[c];
  if (patch_sigma) {
    int lo = 0, hi = P;  // largest p with ele_off[p] <= c
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (ele_off[mid] <= c) lo = mid; else hi = mid;
    }
    src = copy_src[ele_off[lo]];
  }
  double lc[6], fr[9];
  copy_frame<NC>(xyz, cv + c * NC, lc, fr);
  const double v[3] = {vel[3 * src], vel[3 * src + 1], vel[3 * src + 2]};
#pragma unroll
  for (int a = 0; a < NC - 1; ++a) sigma[c * (NC - 1) + a] = (T)g_dot3(v, fr + 3 * a);
}

// Corner heights of every tet copy, computed once at upload with the same
// code the sweeps used to evaluate per visit; FP32 storage rounds toward zero.
// `slow` is set when a copy's least height is below 2^-350 or below 2^-20 of
// its coordinate scale (the sum of its coordinates' magnitudes, which bounds
// its edge lengths): the lengths the tet solve takes roots of are at least a
// height, with rounding errors of a few ulps of the squared scale, so on
// meshes without such copies their squares stay normal numbers of at least
// 2^-700 and sqrt_len gives sqrt's bits.
template <typename T>
__global__ void k_heights(const T* coords, int64_t F, __half* hgt, int32_t* slow) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= F) return;
  double L[6], h[4];
#pragma unroll
  for (int k = 0; k < 6; ++k) L[k] = (double)coords[c * 6 + k];
  ElemSolve<T, 4>::heights(L, h);
#pragma unroll
  for (int k = 0; k < 4; ++k) hgt[c * 4 + k] = __float2half_rz(__double2float_rz(h[k]));
  const double hmin = d_min(d_min(h[0], h[1]), d_min(h[2], h[3]));
  const double sc = ElemSolve<T, 4>::scale(L);
  if (!(hmin >= 0x1p-350 && hmin >= 0x1p-20 * sc)) slow[0] = 1;
  // The plane bound's FP32 squared lengths stay normal numbers.
  if (!(hmin >= 0x1p-48 && sc <= 0x1p48)) slow[1] = 1;
}

// Per patch, the largest coordinate scale of its tet copies, rounded up to
// FP32: the need-mask pass uses it as every copy's margin term.
template <typename T>
__global__ void k_patch_scale(const T* coords, const int32_t* ele_off, float* pscale) {
  const int p = blockIdx.x;
  double mx = 0.0;
  for (int c = ele_off[p] + threadIdx.x; c < ele_off[p + 1]; c += blockDim.x) {
    double L[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) L[k] = (double)coords[(int64_t)c * 6 + k];
    mx = fmax(mx, ElemSolve<T, 4>::scale(L));
  }
  __shared__ double red[256];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) pscale[p] = __double2float_ru(red[0]);
}

// Wave reductions by xor butterflies on ds_swizzle (bit mode: lane ^ M
// within each half of the wave, the pattern an immediate) and a final
// readlane of both halves: unlike ds_bpermute they keep no per-lane address
// registers live, which the two-workgroups-per-CU builds would spill. The
// results are exact (min, integer sums), so the order does not matter.
template <int M>
__device__ __forceinline__ int swz_xor(int v) {
  return __builtin_amdgcn_ds_swizzle(v, (M << 10) | 0x1F);
}
template <int M>
__device__ __forceinline__ double swz_xor(double v) {
  const int lo = swz_xor<M>(__double2loint(v)), hi = swz_xor<M>(__double2hiint(v));
  return __hiloint2double(hi, lo);
}
template <int M>
__device__ __forceinline__ float swz_xor(float v) {
  return __int_as_float(swz_xor<M>(__float_as_int(v)));
}
__device__ __forceinline__ double lane_value(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
// Every lane of the wave must be active; the result is uniform.
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
  v = d_min(v, swz_xor<1>(v));
  v = d_min(v, swz_xor<2>(v));
  v = d_min(v, swz_xor<4>(v));
  v = d_min(v, swz_xor<8>(v));
  v = d_min(v, swz_xor<16>(v));
  return d_min(lane_value(v, 0), lane_value(v, 32));
}
__device__ __forceinline__ int wave_sum(int v) {
  v += swz_xor<1>(v);
  v += swz_xor<2>(v);
  v += swz_xor<4>(v);
  v += swz_xor<8>(v);
  v += swz_xor<16>(v);
  return __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 32);
}

// Block-wide min; every thread must call it. Returns the result in all threads.
template <typename T>
__device__ __forceinline__ T block_min(T v, T* scratch) {
  v = wave_min(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  T r = scratch[0];
#pragma unroll
  for (int w = 1; w < kBlock / 64; ++w) r = d_min(r, scratch[w]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ int corner(const ushort4& c, int i) {
  return i == 0 ? c.x : (i == 1 ? c.y : (i == 2 ? c.z : c.w));
}

// ---------------------------------------------------------------------------
// Patch bodies. Each runs on one workgroup of BS threads for one patch and
// starts with a barrier so consecutive patches can reuse the LDS.

// Per-step reset of one vertex range: DT := LARGENUM, Label := Far,
// Tsign := (phi >= 0) (redistance3d.cu:31-35 and kernel_reinit_Tsign,
// redistance_kernels3d.h:785-792).
template <typename T>
__device__ __forceinline__ void init_vertex(DevState<T>& s, int64_t v) {
  s.DT[v] = (T)kLargeNum;
  s.label[v] = kFarPoint;
  s.tsign[v] = (s.phi[v] < (T)0) ? 0 : 1;
}

// Seeds of one patch (kernel_seedlabel / _narrowband,
// redistance_kernels3d.h:623-770): a corner with phi == 0 is a seed, and so
// is every corner of a sign-changing element; seeds get DT = |phi|. Returns
// (to every thread) whether the patch holds such an element.
template <typename T, int NC, int BS>
__device__ int seeds_patch(const DevMesh<T>& m, DevState<T>& s, int p, unsigned char* smem) {
  __syncthreads();
  T* s_vtx = (T*)smem;
  const int vs = m.vert_off[p], ve = m.vert_off[p + 1];
  const int es = m.ele_off[p], ee = m.ele_off[p + 1];
  const int hs = m.halo_off[p], he = m.halo_off[p + 1];
  const int nv = ve - vs, ne = ee - es, nh = he - hs;
  for (int i = threadIdx.x; i < nv; i += BS) s_vtx[i] = s.phi[vs + i];
  for (int i = threadIdx.x; i < nh; i += BS) s_vtx[nv + i] = s.phi[m.halo_ids[hs + i]];
  if (threadIdx.x == 0) {
    atomicAdd(s.sizes + kSzSeed + 0, (unsigned long long)ne);
    atomicAdd(s.sizes + kSzSeed + 1, (unsigned long long)nv);
    atomicAdd(s.sizes + kSzSeed + 2, (unsigned long long)nh);
  }
  __syncthreads();
  int flag = 0;
  for (int e = threadIdx.x; e < ne; e += BS) {
    const ushort4 c = m.copy_local[es + e];
    T val[NC];
    bool allpos = true, allneg = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      val[i] = s_vtx[corner(c, i)];
      allpos = allpos && val[i] >= (T)0;
      allneg = allneg && val[i] <= (T)0;
    }
    const bool mixed = !(allpos || allneg);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      if (mixed || val[i] == (T)0) {
        const int l = corner(c, i);
        const int g = l < nv ? vs + l : m.halo_ids[hs + l - nv];
        s.label[g] = kSeedPoint;
        s.DT[g] = d_abs(val[i]);
        flag = 1;
      }
    }
  }
  return __syncthreads_or(flag);
}

// Block-wide min for BS threads; every thread must call it.
template <typename T, int BS>
__device__ __forceinline__ T block_min_n(T v, T* scratch) {
  v = wave_min(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  T r = scratch[0];
#pragma unroll
  for (int w = 1; w < BS / 64; ++w) r = d_min(r, scratch[w]);
  __syncthreads();
  return r;
}

// FIM patch visit (kernel_update_values / kernel_run_check_neghbor,
// redistance_kernels3d.h:287-621): NITER Jacobi sweeps over the patch's
// element copies in LDS; each sweep updates every element from its corner
// values, then every owned vertex takes the min over its slots and writes it
// back to them. An element whose four inputs are unchanged since its last
// update would reproduce its values exactly, so each sweep only visits the
// elements marked dirty, compacted into an LDS work list; within one, a
// corner whose only changed input is its own value is not solved again. The last sweep
// decides convergence per vertex (old - new < tol; seeds count as converged),
// and the workgroup reduces it to the patch flag plus the patch min of the
// written distances (run_reduction_bandwidth, redistance_kernels3d.h:57-97).
//
// Keys whose unsigned order is the order of the (non-NaN) floating-point
// values: the sign bit is flipped for positive values and every bit for
// negative ones, so an integer atomicMin over keys is a floating-point min.
template <typename T>
struct OrderedBits;
template <>
struct OrderedBits<float> {
  using type = unsigned int;
  __device__ static type of(float x) {
    const type b = __float_as_uint(x);
    return b ^ ((b >> 31) ? 0xffffffffu : 0x80000000u);
  }
  __device__ static float value(type k) { return __uint_as_float(k ^ ((k >> 31) ? 0x80000000u : 0xffffffffu)); }
};
template <>
struct OrderedBits<double> {
  using type = unsigned long long;
  __device__ static type of(double x) {
    const type b = (type)__double_as_longlong(x);
    return b ^ ((b >> 63) ? ~0ull : (1ull << 63));
  }
  __device__ static double value(type k) {
    return __longlong_as_double((long long)(k ^ ((k >> 63) ? (1ull << 63) : ~0ull)));
  }
};

// A draw from a phase counter for the whole workgroup. s_draw[0] is the item
// drawn last, s_draw[1] a second item reserved with it (-1: none). While at
// least 16 draws per workgroup remain a draw reserves two consecutive items,
// halving the counter round trips; near the end of the phase it reserves one,
// so a workgroup done early still finds work. The counter numbers the items
// after the G each workgroup took first (first_item), each value one item,
// taken once.
__device__ __forceinline__ int draw_items(int n, int32_t* counter, int* s_draw) {
  __syncthreads();  // every thread has read the previous item
  if (threadIdx.x == 0) {
    if (s_draw[1] >= 0) {
      s_draw[0] = s_draw[1];
      s_draw[1] = -1;
    } else {
      const bool pair = s_draw[0] + 16 * (int)gridDim.x < n;
      const int i = (int)gridDim.x + atomicAdd(counter, pair ? 2 : 1);
      s_draw[0] = i;
      s_draw[1] = pair && i + 1 < n ? i + 1 : -1;
    }
  }
  __syncthreads();
  return s_draw[0];
}

// Work items of a phase with n items: every workgroup first takes the item
// of its own index (all G indices are taken, so every item below G is),
// which keeps a counter round trip off the start of the phase; with more
// items than workgroups the rest are drawn from the counter, so a workgroup
// done early takes the next one. n means no further item.
__device__ __forceinline__ int first_item(int n, int32_t* counter, int* s_draw) {
  (void)counter;
  if (threadIdx.x == 0) {  // the draw state; draw_items synchronises before reading it
    s_draw[0] = (int)blockIdx.x;
    s_draw[1] = -1;
  }
  return (int)blockIdx.x < n ? (int)blockIdx.x : n;
}
__device__ __forceinline__ int later_item(int n, int32_t* counter, int* s_draw) {
  return n <= (int)gridDim.x ? n : draw_items(n, counter, s_draw);
}

// Copies n items into LDS with U independent loads in flight per thread.
template <int BS, int U, typename V>
__device__ __forceinline__ void copy_to_lds(V* dst, const V* src, int n) {
  for (int base = threadIdx.x; base < n; base += U * BS) {
    V r[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + u * BS < n) r[u] = src[base + u * BS];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + u * BS < n) dst[base + u * BS] = r[u];
  }
}

// The patch's convergence flag and min distance, as sweep_patch stores them in
// bcon / bmin, for every thread of the workgroup; visited: the call wrote the
// patch's outputs (a visit, or a probe that continued as one).
template <typename T>
struct SweepResult {
  int con;
  T bmin;
  int visited;
};

// Buffers of a sweep_patch call: patches whose tag in vtag equals the call's
// `members` iteration are read from din (their values after that iteration,
// not yet copied into DT), all others from DT; outputs go to dout and bcon.
// probe_first: sweep 0 is a neighbour probe (no outputs); a converged probe
// ends the call, an unconverged one continues from its state as the patch's
// visit (needs niter >= 2).
template <typename T>
struct SweepIO {
  const T* din;
  const int32_t* vtag;
  T* dout;
  int32_t* bcon;
  int probe_first;
};
template <typename T>
__device__ __forceinline__ SweepIO<T> default_io(const DevState<T>& s) {
  return SweepIO<T>{s.DT_out, s.visitE, s.DT_out, s.bcon, 0};
}

// Per-workgroup statistics in LDS: kNumSizes patch-size sums, then the
// visit and probe counts of the persistent step.
constexpr int kAccVisits = kNumSizes, kAccProbes = kNumSizes + 1, kNumAcc = kNumSizes + 2;
__device__ __forceinline__ void clear_stats(unsigned long long* acc) {
  if ((int)threadIdx.x < kNumAcc) acc[threadIdx.x] = 0ull;
}
// Adds the workgroup's statistics to the device counters (after a barrier
// that ends the workgroup's last update of them).
template <typename T>
__device__ __forceinline__ void flush_stats(const DevState<T>& s, const unsigned long long* acc) {
  const int t = threadIdx.x;
  if (t < kNumSizes && acc[t]) atomicAdd(s.sizes + t, acc[t]);
  if (t == kAccVisits && acc[t]) atomicAdd(&s.pc[kPcVisits], (int32_t)acc[t]);
  if (t == kAccProbes && acc[t]) atomicAdd(&s.pc[kPcProbes], (int32_t)acc[t]);
}

// Marks corner i of element e changed (s_dirty is 4-byte aligned in LDS;
// lanes updating other elements of the same word do so atomically too).
__device__ __forceinline__ void mark_changed(uint8_t* s_dirty, int e, int i) {
  atomicOr((unsigned int*)(s_dirty + (e & ~3)), 1u << (8 * (e & 3) + i));
}

// Corners whose solve can still lower their value, from the corners changed
// since the element's last need-mask check: corner k's solve reads only the
// other corners, so if none of them changed it returns what it returned (or
// would have returned, when the lower-bound skip left it) at that check,
// which is not below corner k's value then or now.
template <int NC>
__device__ __forceinline__ int unchanged_input_filter(int chg) {
  constexpr int full = (1 << NC) - 1;
  chg &= full;
  return __popc(chg) >= 2 ? full : (full & ~chg);
}

// Sweep organisations (bit-identical, same per-corner arithmetic):
// kOrgElements runs one thread per element with a corner to solve, the
// elements listed by their number of solves; kOrgTasks lists the
// (element, corner) solves the exact lower-bound skip leaves and spreads
// them over the workgroup one per thread.
constexpr int kOrgElements = 0, kOrgTasks = 1;

// epoch_in > 0 (a visit): resume after the first sweep from the state a
// probe saved for this iteration, when there is one. epoch_out > 0 (a probe):
// save the state after the sweep for the next iteration's visit when the
// patch stays unconverged. The resumed visit computes the same bits.
// members > 0: vertices of patches visited in FIM iteration `members` are
// read from DT_out (their copy into DT runs concurrently), all others from DT.
// acc: the workgroup's statistics in LDS (kNumSizes patch-size sums, then
// visits and probes), added to the device counters once per workgroup by
// flush_stats, so visits contend on no global counter for them.
// STG fixes the staging choice at compile time (0 or 1); -1 takes `stage_arg`.
template <typename T, int NC, int BS, int ORG, int STG = -1>
__device__ SweepResult<T> sweep_patch(const DevMesh<T>& m, DevState<T>& s, int p, int niter, int write_out,
                                      T tol, double leps, unsigned char* smem, T* red, int* s_ndirty /*[10]*/,
                                      int epoch_in, int epoch_out, int members, const SweepIO<T> io,
                                      unsigned long long* acc, int stage_arg) {
  constexpr bool tasks = ORG == kOrgTasks;
  const int stage = STG < 0 ? stage_arg : STG;
  constexpr int K = ElemSolve<T, NC>::kCoords;
  __syncthreads();
  // The patch's offsets in one load (no dependent vslot_off lookup).
  const int4 h0 = m.phdr[2 * p], h1 = m.phdr[2 * p + 1];
  const int vs = h0.x, ve = h0.y, es = h0.z, ee = h0.w, hs = h1.x, he = h1.y;
  const int nv = ve - vs, ne = ee - es, nh = he - hs;
  const int sbase = h1.z;
  const int nslots = h1.w - sbase;
  T* s_slot = (T*)smem;
  T* s_vtx = s_slot + NC * ne;
  T* s_vval = s_vtx + nv + nh;
  uint16_t* s_voff = (uint16_t*)(s_vval + nv);
  // The slot lists and staged element data start at their sources' offsets
  // modulo 16 bytes (device arrays are 256-byte aligned), so 16-byte loads
  // and stores copy them.
  const int o_vsl = ((int)((unsigned char*)(s_voff + nv + 1) - smem) + 15 & ~15) + ((2 * sbase) & 15);
  uint16_t* s_vsl = (uint16_t*)(smem + o_vsl);
  uint16_t* s_work = s_vsl + nslots;            // compacted dirty elements
  // Aligned regions are carved by byte offsets from smem so that every
  // access stays an LDS access. s_dirty[e]: bits 0..NC-1 mark the corner
  // slots of element e changed since its last need-mask check; bits 4..7
  // hold the need mask between that check and the solves (elements
  // organisation). Whole words are updated with atomicOr.
  const int o_dirty = ((int)((unsigned char*)(s_work + ne) - smem) + 3) & ~3;
  uint8_t* s_dirty = smem + o_dirty;
  // Running minimum of each owned vertex's slots, as ordered keys.
  using Bits = typename OrderedBits<T>::type;
  const int o_vmin = (o_dirty + ne + 7) & ~7;
  Bits* s_vmin = (Bits*)(smem + o_vmin);
  // Staged element data (stage set): each copy's corner ids and FP16
  // heights, read by every sweep of the visit from LDS instead of memory.
  const int o_end = (o_vmin + nv * (int)sizeof(Bits) + 7) & ~7;
  const int o_stage = ((o_end + 15) & ~15) + 8 * (es & 1);
  ushort4* s_cl = (ushort4*)(smem + o_stage);
  const int o_hgt = ((o_stage + 8 * ne + 15) & ~15) + 8 * (es & 1);
  uint2* s_hgt = (uint2*)(smem + o_hgt);
  // Task mode: (element, corner) solves and their results.
  uint32_t* s_task = (uint32_t*)(smem + (stage ? o_hgt + 8 * ne : o_end));
  T* s_new = (T*)(s_task + NC * ne + (NC * ne) % 2);
  const int tid = threadIdx.x;

  // The patch load issues every independent load of a thread before its
  // stores (offsets; owned values, halo ids and owners, 16-byte pieces of
  // the slot lists and staged element data, clean flags; then the selected
  // halo values), so it waits for memory about twice instead of once per
  // array. The unaligned ends of the 16-byte copies go element by element.
  const int tag_p = io.vtag[p], ve_tag_p = s.ve_tag[p];
  const T* own = (members > 0 && tag_p == members) ? io.din : s.DT;
  const uint16_t* vsl_src = m.vslots + sbase;
  const int vsl_head = min(nslots, (int)(((16u - ((2u * (unsigned)sbase) & 15u)) & 15u) / 2u));
  const int vsl_nvec = (nslots - vsl_head) / 8;
  const int vsl_tail0 = vsl_head + 8 * vsl_nvec;  // fewer than 8 elements follow
  const int st_head = stage ? min(ne, es & 1) : 0;
  const int st_nvec = stage ? (ne - st_head) / 2 : 0;
  const int st_tail0 = st_head + 2 * st_nvec;     // at most one element follows
  const uint2* hgt_src = NC == 4 ? (const uint2*)m.hgt + es : nullptr;
  uint16_t e_vh = 0, e_vt = 0;
  ushort4 e_ch, e_ct;
  uint2 e_hh, e_ht;
  const bool do_vh = tid < vsl_head, do_vt = tid < nslots - vsl_tail0;
  const bool do_sh = tid == 0 && st_head > 0, do_st = tid == 0 && stage && st_tail0 < ne;
  if (do_vh) e_vh = vsl_src[tid];
  if (do_vt) e_vt = vsl_src[vsl_tail0 + tid];
  if (do_sh) {
    e_ch = m.copy_local[es];
    if constexpr (NC == 4) e_hh = hgt_src[0];
  }
  if (do_st) {
    e_ct = m.copy_local[es + st_tail0];
    if constexpr (NC == 4) e_ht = hgt_src[st_tail0];
  }
  // The first chunk's clean flags (read by the slot initialisation below).
  // (A visit that may resume from a probe's state reads them later.)
  const bool okb_valid = !(niter > 1 && epoch_in > 0);
  uint8_t okb[4] = {0, 0, 0, 0};
  if (okb_valid && ve_tag_p != 0) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (tid + u * BS < ne) okb[u] = s.ve_ok[es + tid + u * BS];
  }
  const int nwork = max(max(nv, nh), max(vsl_nvec, st_nvec));
  for (int i = tid; i < nwork; i += BS) {
    int h = 0, q = 0, vo = 0;
    T ov = (T)0;
    uint4 w0, w1, w2;
    if (i < nh) {
      h = m.halo_ids[hs + i];
      q = m.halo_patch[hs + i];
    }
    if (i < nv) {
      ov = own[vs + i];
      vo = m.vslot_off[vs + i];
    }
    if (i < vsl_nvec) w0 = reinterpret_cast<const uint4*>(vsl_src + vsl_head)[i];
    if (i < st_nvec) {
      w1 = reinterpret_cast<const uint4*>(m.copy_local + es + st_head)[i];
      if constexpr (NC == 4) w2 = reinterpret_cast<const uint4*>(hgt_src + st_head)[i];
    }
    if (i < nh) {
      // both candidates loaded beside the owner's tag, then selected
      const T a = io.din[h], d = s.DT[h];
      const int t = io.vtag[q];
      s_vtx[nv + i] = (members > 0 && t == members) ? a : d;
    }
    if (i < nv) {
      s_vtx[i] = ov;
      s_vval[i] = ov;
      s_voff[i] = (uint16_t)(vo - sbase);
    }
    if (i < vsl_nvec) reinterpret_cast<uint4*>(s_vsl + vsl_head)[i] = w0;
    if (i < st_nvec) {
      reinterpret_cast<uint4*>(s_cl + st_head)[i] = w1;
      if constexpr (NC == 4) reinterpret_cast<uint4*>(s_hgt + st_head)[i] = w2;
    }
  }
  if (tid == 0) s_voff[nv] = (uint16_t)nslots;
  if (do_vh) s_vsl[tid] = e_vh;
  if (do_vt) s_vsl[vsl_tail0 + tid] = e_vt;
  if (do_sh) {
    s_cl[0] = e_ch;
    if constexpr (NC == 4) s_hgt[0] = e_hh;
  }
  if (do_st) {
    s_cl[st_tail0] = e_ct;
    if constexpr (NC == 4) s_hgt[st_tail0] = e_ht;
  }
  // (A single-sweep visit recomputes the sweep: resuming would leave it none.)
  const bool resume = niter > 1 && epoch_in > 0 && s.rs_epoch[p] == epoch_in;
  if (resume) {
    copy_to_lds<BS, 4>(s_slot, s.rs_slot + (int64_t)es * NC, NC * ne);
    copy_to_lds<BS, 4>(s_dirty, s.rs_dirty + es, ne);
    copy_to_lds<BS, 4>(s_vval, s.rs_vval + vs, nv);
  }
  __syncthreads();
  const bool saved = !resume && ve_tag_p != 0;
  if (!resume)
  for (int base = tid; base < ne; base += 4 * BS) {
    ushort4 c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (base + u * BS < ne) c[u] = stage ? s_cl[base + u * BS] : m.copy_local[es + base + u * BS];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = base + u * BS;
      if (e < ne) {
#pragma unroll
        for (int i = 0; i < NC; ++i) s_slot[e * NC + i] = s_vtx[corner(c[u], i)];
        const bool ok = base == tid && okb_valid ? okb[u] != 0 : (saved && s.ve_ok[es + e]);
        s_dirty[e] = ok ? (uint8_t)0 : (uint8_t)((1 << NC) - 1);
      }
    }
  }

  for (int i = tid; i < nv; i += BS) s_vmin[i] = OrderedBits<T>::of(s_vval[i]);
  // A visit records which elements end it clean (see DevState::ve_ok); the
  // per-vertex flags of its last vertex pass live in the listing area.
  const bool save_clean = write_out && nv <= 2 * ne;
  uint8_t* s_same = (uint8_t*)s_work;
  int vsh = 0;  // log2 of the lanes per vertex in the vertex pass
  while (vsh < 3 && (nv << (vsh + 1)) <= BS) ++vsh;
  int con_all = 1;
  T vmin = (T)kLargeNum;
  int n_upd = 0, n_solve = 0;
  int probing = io.probe_first && niter > 1;  // sweep 0 decides a probe
  int visited = write_out && !probing;
  for (int it = resume ? 1 : 0; it < niter; ++it) {
    if constexpr (tasks) {
      if (tid == 0) *s_ndirty = 0;
      __syncthreads();
      // Wave-aggregated compaction; the loop bound is uniform across the block.
      for (int base = 0; base < ne; base += BS) {
        const int e = base + tid;
        const bool d = e < ne && s_dirty[e];
        const unsigned long long mask = __ballot(d);
        const int lane = tid & 63;
        int off = 0;
        if (lane == 0 && mask) off = atomicAdd(s_ndirty, __popcll(mask));
        off = __builtin_amdgcn_readlane(off, 0);
        if (d) s_work[off + __popcll(mask & ((1ull << lane) - 1ull))] = (uint16_t)e;
      }
      if (tid == 0) s_ndirty[1] = 0;
      __syncthreads();
      const int nd = *s_ndirty;
      // Need masks of the dirty elements, then their solves as tasks.
      for (int base = 0; base < nd; base += BS) {
        const int k = base + tid;
        int e = 0, mask = 0;
        if (k < nd) {
          e = s_work[k];
          T x[NC];
#pragma unroll
          for (int i = 0; i < NC; ++i) x[i] = s_slot[e * NC + i];
          mask = staged_need_mask<T, NC>(m, p, es, e, x, stage, s_hgt) & unchanged_input_filter<NC>(s_dirty[e]);
          if constexpr (NC == 4)
            if (m.plane && mask) mask = ElemSolve<T, 4>::plane_mask(m.coords + (int64_t)(es + e) * K, x, mask, m.pscale[p]);
          s_dirty[e] = 0;
          ++n_upd;
        }
        const int c = __popc(mask);
        n_solve += c;
        const int lane = tid & 63;
        int incl = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const int t = __shfl_up(incl, off, 64);
          if (lane >= off) incl += t;
        }
        const int total = __builtin_amdgcn_readlane(incl, 63);
        int wbase = 0;
        if (lane == 0 && total) wbase = atomicAdd(&s_ndirty[1], total);
        wbase = __builtin_amdgcn_readlane(wbase, 0);
        int o = wbase + incl - c;
#pragma unroll
        for (int j = 0; j < NC; ++j)
          if (mask & (1 << j)) s_task[o++] = (uint32_t)(e * 4 + j);
      }
      __syncthreads();
      const int nt = s_ndirty[1];
      for (int t = tid; t < nt; t += BS) {
        const uint32_t tk = s_task[t];
        const int e = (int)(tk >> 2), j = (int)(tk & 3u);
        T x[NC], L[K];
#pragma unroll
        for (int i = 0; i < NC; ++i) x[i] = s_slot[e * NC + i];
        const T* gc = m.coords + (int64_t)(es + e) * K;
#pragma unroll
        for (int q = 0; q < K; ++q) L[q] = gc[q];
        s_new[t] = ElemSolve<T, NC>::corner_update(L, x, j, leps, m.fast_len != 0);
      }
      __syncthreads();
      for (int t = tid; t < nt; t += BS) {
        const uint32_t tk = s_task[t];
        const int e = (int)(tk >> 2), j = (int)(tk & 3u);
        const T y = s_new[t];
        if (y != s_slot[e * NC + j]) {
          s_slot[e * NC + j] = y;
          mark_changed(s_dirty, e, j);
          const int lv = corner(stage ? s_cl[e] : m.copy_local[es + e], j);
          if (lv < nv) atomicMin(&s_vmin[lv], OrderedBits<T>::of(y));
        }
      }
    } else {
      // Need masks of the dirty elements (kept in s_dirty bits 4..7), then
      // the elements with a corner to solve listed by their number of
      // solves, most first, so the lanes of a wave walk the same number of
      // corners and the chunks that wrap around to the first waves are the
      // cheapest. Counts per solve number c = 1..NC in s_ndirty[1 + c], list
      // cursors in s_ndirty[5 + c].
      if (tid < 8) s_ndirty[2 + tid] = 0;
      __syncthreads();
      for (int base = 0; base < ne; base += BS) {
        const int e = base + tid;
        int c = 0;
        if (e < ne && s_dirty[e]) {
          T x[NC];
#pragma unroll
          for (int i = 0; i < NC; ++i) x[i] = s_slot[e * NC + i];
          int mask = staged_need_mask<T, NC>(m, p, es, e, x, stage, s_hgt) & unchanged_input_filter<NC>(s_dirty[e]);
          if constexpr (NC == 4)
            if (m.plane && mask) mask = ElemSolve<T, 4>::plane_mask(m.coords + (int64_t)(es + e) * K, x, mask, m.pscale[p]);
          s_dirty[e] = (uint8_t)(mask << 4);
          c = __popc(mask);
          n_solve += c;
          ++n_upd;
        }
        // The wave's counts per solve number c = 1..NC, 16 bits each (a
        // patch has far fewer than 65536 element copies), in one atomic.
        unsigned long long pk = 0;
#pragma unroll
        for (int j = 1; j <= NC; ++j) pk |= (unsigned long long)__popcll(__ballot(c == j)) << (16 * (j - 1));
        if ((tid & 63) == 0 && pk) atomicAdd(reinterpret_cast<unsigned long long*>(s_ndirty + 2), pk);
      }
      __syncthreads();
      const unsigned long long cnt = *reinterpret_cast<const unsigned long long*>(s_ndirty + 2);
      int start[NC + 2];
      start[NC] = 0;
#pragma unroll
      for (int j = NC - 1; j >= 1; --j) start[j] = start[j + 1] + (int)((cnt >> (16 * j)) & 0xFFFFull);
      const int nd = start[1] + (int)(cnt & 0xFFFFull);
      for (int base = 0; base < ne; base += BS) {
        const int e = base + tid;
        const int c = e < ne ? __popc(s_dirty[e] >> 4) : 0;
        const int lane = tid & 63;
        // Cursors per solve number, packed like the counts: one atomic per wave.
        unsigned long long bj[NC];
        unsigned long long pc = 0;
#pragma unroll
        for (int j = 1; j <= NC; ++j) {
          bj[j - 1] = __ballot(c == j);
          pc |= (unsigned long long)__popcll(bj[j - 1]) << (16 * (j - 1));
        }
        unsigned long long old = 0;
        if (lane == 0 && pc) old = atomicAdd(reinterpret_cast<unsigned long long*>(s_ndirty + 6), pc);
        const unsigned int olo = __builtin_amdgcn_readlane((unsigned int)old, 0);
        const unsigned int ohi = __builtin_amdgcn_readlane((unsigned int)(old >> 32), 0);
        old = ((unsigned long long)ohi << 32) | olo;
#pragma unroll
        for (int j = 1; j <= NC; ++j)
          if (c == j) {
            const int off = (int)((old >> (16 * (j - 1))) & 0xFFFFull);
            s_work[start[j] + off + __popcll(bj[j - 1] & ((1ull << lane) - 1ull))] = (uint16_t)e;
          }
      }
      __syncthreads();
      for (int k = tid; k < nd; k += BS) {
        const int e = s_work[k];
        T x[NC], y[NC], L[K];
#pragma unroll
        for (int i = 0; i < NC; ++i) x[i] = y[i] = s_slot[e * NC + i];
        const T* gc = m.coords + (int64_t)(es + e) * K;
#pragma unroll
        for (int q = 0; q < K; ++q) L[q] = gc[q];
        ElemSolve<T, NC>::solve_mask(L, s_dirty[e] >> 4, y, leps, m.fast_len != 0);
        int changed = 0;
        const ushort4 cl = stage ? s_cl[e] : m.copy_local[es + e];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
          if (y[i] != x[i]) {
            s_slot[e * NC + i] = y[i];
            changed |= 1 << i;
            const int lv = corner(cl, i);
            if (lv < nv) atomicMin(&s_vmin[lv], OrderedBits<T>::of(y[i]));
          }
        }
        s_dirty[e] = (uint8_t)changed;
      }
    }
    __syncthreads();
    const bool last = (it == niter - 1) || probing;
    // Every slot of an owned vertex equals the vertex value after this pass
    // and slots only decrease, so the minimum over the vertex's slots is the
    // running minimum the solves above left in s_vmin; only a vertex whose
    // minimum dropped lowers its slots. A vertex's slots are split over the
    // 2^vsh adjacent lanes of one wave (as many as the workgroup holds, up to
    // 8); all of them read s_vval[v] before the first lane rewrites it.
    for (int w = tid; w < (nv << vsh); w += BS) {
      const int v = w >> vsh, h = w & ((1 << vsh) - 1);
      const int o0 = s_voff[v], o1 = s_voff[v + 1];
      const T old = s_vval[v];
      const T mn = OrderedBits<T>::value(s_vmin[v]);
      if (mn < old) {
        for (int q = o0 + h; q < o1; q += 1 << vsh) {
          const int sl = s_vsl[q];
          if (mn < s_slot[sl]) {
            s_slot[sl] = mn;
            mark_changed(s_dirty, sl / NC, sl % NC);
          }
        }
        if (h == 0) s_vval[v] = mn;
      }
      if (last && h == 0) {
        T outv;
        int c;
        if (s.label[vs + v] != kSeedPoint) {
          outv = mn;
          c = (old - mn < tol) ? 1 : 0;
        } else {
          outv = s_vtx[v];
          c = 1;
        }
        // The patch's slots of v all hold s_vval[v] now; a fresh load gives
        // them outv.
        if (!probing) {
          if (save_clean) s_same[v] = (uint8_t)(d_min(old, mn) == outv);
          if (write_out) io.dout[vs + v] = outv;
        }
        con_all &= c;
        vmin = d_min(vmin, outv);
      }
    }
    if (probing) {
      // The probe sweep decided: a converged probe ends here; an unconverged
      // one continues as the visit from this state (what a visit resuming
      // from the probe's saved state computes).
      if (__syncthreads_and(con_all)) break;
      probing = 0;
      visited = write_out;
      con_all = 1;
      vmin = (T)kLargeNum;
    }
  }
  con_all = __syncthreads_and(con_all);
  if (save_clean && visited) {
    for (int e = tid; e < ne; e += BS) {
      const ushort4 cl = stage ? s_cl[e] : m.copy_local[es + e];
      int ok = s_dirty[e] == 0;
#pragma unroll
      for (int i = 0; i < NC; ++i) {
        const int lv = corner(cl, i);
        ok = ok && lv < nv && s_same[lv];
      }
      s.ve_ok[es + e] = (uint8_t)ok;
    }
    if (tid == 0) s.ve_tag[p] = 1;
  }
  if (epoch_out > 0 && !con_all) {
    for (int i = tid; i < NC * ne; i += BS) s.rs_slot[(int64_t)es * NC + i] = s_slot[i];
    for (int i = tid; i < ne; i += BS) s.rs_dirty[es + i] = s_dirty[i];
    for (int i = tid; i < nv; i += BS) s.rs_vval[vs + i] = s_vval[i];
    if (tid == 0) s.rs_epoch[p] = epoch_out;
  }
  const T bm = block_min_n<T, BS>(vmin, red);
  n_upd = wave_sum(n_upd);
  n_solve = wave_sum(n_solve);
  if ((tid & 63) == 0 && (n_upd | n_solve)) {
    atomicAdd(acc + kSzElem, (unsigned long long)n_upd);
    atomicAdd(acc + kSzSolve, (unsigned long long)n_solve);
  }
  if (tid == 0) {
    io.bcon[p] = con_all;
    if (visited) s.bmin[p] = bm;
    // byte accounting: a probe and, when it continued, the visit
    if (io.probe_first || !write_out) {
      acc[kSzProbe + 0] += (unsigned long long)ne;
      acc[kSzProbe + 1] += (unsigned long long)nv;
      acc[kSzProbe + 2] += (unsigned long long)nh;
    }
    if (visited) {
      acc[kSzVisit + 0] += (unsigned long long)ne;
      acc[kSzVisit + 1] += (unsigned long long)nv;
      acc[kSzVisit + 2] += (unsigned long long)nh;
    }
  }
  return {con_all, bm, visited};
}

// Narrowband test of one patch (kernel_compute_new_narrowband,
// redistance_kernels3d.h:22-55) fused with the sign restore
// phi = DT * (2*Tsign - 1) (kernel_recover_Tsign_whole, :807-814).
// Returns the patch min to every thread.
template <typename T, int BS>
__device__ T narrowband_patch(const DevMesh<T>& m, DevState<T>& s, int p, T* red) {
  __syncthreads();
  const int vs = m.vert_off[p], ve = m.vert_off[p + 1];
  T mn = (T)kLargeNum;
  for (int v = vs + threadIdx.x; v < ve; v += BS) {
    const T d = s.DT[v];
    mn = d_min(mn, d);
    s.phi[v] = d * (T)(2 * (int)s.tsign[v] - 1);
  }
  return block_min_n<T, BS>(mn, red);
}

// Warm value of vertex v: the predictor's distance times (1 + margin) plus
// an absolute margin; the prediction is the learned predictor's (s.warm) or
// |phi| (identity predictor).
template <typename T>
__device__ __forceinline__ double warm_value(const DevState<T>& s, int64_t v, int mode, double margin,
                                             double abs_margin) {
  const double base = mode == 2 ? (double)s.warm[v] : d_abs((double)s.phi[v]);
  return base * (1.0 + margin) + abs_margin;
}

// Warm start of one patch of the previous narrowband: every vertex sharing
// no element with a seed whose warm value is below LARGENUM/2 starts the FIM
// there instead of at LARGENUM. Vertices next to a seed stay cold: the FIM
// solves them from seed slots it lowers inside a visit and never writes back
// (redistance_kernels3d.h:440-449), which no check from stored distances can
// reproduce, so they are finished exactly as in the plain method. Returns (to
// every thread) whether the patch received any warm value.
// s_near[i] = 1 for each owned vertex i of patch p that shares an element
// with a seed (every element holding it has a copy in p). Starts and ends
// with a workgroup barrier.
template <typename T, int NC, int BS>
__device__ void mark_seed_adjacent(const DevMesh<T>& m, const DevState<T>& s, int p, uint8_t* s_near) {
  const int vs = m.vert_off[p], nv = m.vert_off[p + 1] - vs;
  const int es = m.ele_off[p], ee = m.ele_off[p + 1];
  const int hs = m.halo_off[p];
  __syncthreads();
  for (int i = threadIdx.x; i < nv; i += BS) s_near[i] = 0;
  __syncthreads();
  for (int e = es + threadIdx.x; e < ee; e += BS) {
    const ushort4 c = m.copy_local[e];
    bool seeded = false;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      const int l = corner(c, i);
      seeded = seeded || s.label[l < nv ? vs + l : m.halo_ids[hs + l - nv]] == kSeedPoint;
    }
    if (seeded)
#pragma unroll
      for (int i = 0; i < NC; ++i)
        if (corner(c, i) < nv) s_near[corner(c, i)] = 1;
  }
  __syncthreads();
}

template <typename T, int NC, int BS>
__device__ bool warm_patch(const DevMesh<T>& m, DevState<T>& s, int p, int mode, double margin, double abs_margin,
                           unsigned char* smem) {
  const int vs = m.vert_off[p], ve = m.vert_off[p + 1];
  uint8_t* s_near = smem;
  mark_seed_adjacent<T, NC, BS>(m, s, p, s_near);
  int any = 0;
  for (int v = vs + threadIdx.x; v < ve; v += BS) {
    if (s.label[v] == kSeedPoint || s_near[v - vs]) continue;
    const double w = warm_value(s, v, mode, margin, abs_margin);
    if (w < 0.5 * kLargeNum) {
      s.DT[v] = (T)w;
      any = 1;
    }
  }
  return __syncthreads_or(any) != 0;
}

// Acceptance of a warm start, over one patch that received warm values. The
// FIM only lowers values, and a value it lowers is an element solve from
// corner values no lower than their final ones, so it is no lower than what
// that element gives from the final values. The only values no solve
// produced are warm values still held at the end: an owned non-seed vertex
// holding one inside the band must not lie below F_v, the least value any of
// its element copies gives it from the other corners' final distances (seeds
// at |phi|). An undercut region of warm values has such a vertex at its
// minimum, since nothing lower around it can support it; the plain method's
// output holds no warm values and always passes. The patch maximum of
// F_v - DT_v is merged into pc[kPcResid] (non-negative floats order as their
// bit patterns); above res_tol it fails the patch's mesh (gfail), and the
// first patch to fail a mesh counts it in pc[kPcFallback].
template <typename T, int NC, int BS>
__device__ void residual_patch(const DevMesh<T>& m, DevState<T>& s, int p, T bw, double leps, int mode,
                               double margin, double abs_margin, double res_tol, unsigned char* smem, T* red) {
  constexpr int K = ElemSolve<T, NC>::kCoords;
  const int vs = m.vert_off[p], ve = m.vert_off[p + 1];
  const int es = m.ele_off[p], hs = m.halo_off[p];
  const int nv = ve - vs;
  uint8_t* s_near = smem;  // vertices next to a seed received no warm value
  mark_seed_adjacent<T, NC, BS>(m, s, p, s_near);
  float r = 0.f;
  for (int v = vs + threadIdx.x; v < ve; v += BS) {
    const T dv = s.DT[v];
    if (s.label[v] == kSeedPoint || s_near[v - vs] || !(dv <= bw) ||
        dv != (T)warm_value(s, v, mode, margin, abs_margin))
      continue;
    double f = kLargeNum;
    for (int k = m.vslot_off[v]; k < m.vslot_off[v + 1]; ++k) {
      const int sl = m.vslots[k];
      const int64_t c = es + sl / NC;
      const ushort4 cl = m.copy_local[c];
      T val[NC], L[K];
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const int l = corner(cl, j);
        val[j] = s.DT[l < nv ? vs + l : m.halo_ids[hs + l - nv]];
      }
      const T* gc = m.coords + c * K;
#pragma unroll
      for (int q = 0; q < K; ++q) L[q] = gc[q];
      f = d_min(f, (double)ElemSolve<T, NC>::corner_value(L, val, sl % NC, leps, m.fast_len != 0));
    }
    r = fmaxf(r, (float)(f - (double)dv));
  }
  r = -block_min_n<float, BS>(-r, (float*)red);
  if (threadIdx.x == 0 && r > 0.f) atomicMax((unsigned int*)&s.pc[kPcResid], __float_as_uint(r));
  if (threadIdx.x == 0 && (double)r > res_tol && atomicExch(&s.gfail[group_of(m, p)], 1) == 0)
    atomicAdd(&s.pc[kPcFallback], 1);
}

// Shape-function gradients of a copy in its local frame (FP64).
template <int NC>
struct Grad;

template <>
struct Grad<4> {
  // A=(0,0,0), B=(L0,0,0), C=(L1,L2,0), D=(L3,L4,L5).
  __device__ static void eval(const double* L, double (*g)[3], double& vol) {
    vol = d_abs(L[0] * L[2] * L[5]) / 6.0;
    g[1][0] = 1.0 / L[0];
    g[1][1] = -L[1] / (L[0] * L[2]);
    g[1][2] = (L[1] * L[4] - L[2] * L[3]) / (L[0] * L[2] * L[5]);
    g[2][0] = 0.0;
    g[2][1] = 1.0 / L[2];
    g[2][2] = -L[4] / (L[2] * L[5]);
    g[3][0] = 0.0;
    g[3][1] = 0.0;
    g[3][2] = 1.0 / L[5];
#pragma unroll
    for (int k = 0; k < 3; ++k) g[0][k] = -((g[1][k] + g[2][k]) + g[3][k]);
  }
};

template <>
struct Grad<3> {
  // A=(0,0), B=(c0,0), C=(c1,c2).
  __device__ static void eval(const double* c, double (*g)[3], double& vol) {
    vol = d_abs(c[0] * c[2]) / 2.0;
    g[1][0] = 1.0 / c[0];
    g[1][1] = -c[1] / (c[0] * c[2]);
    g[2][0] = 0.0;
    g[2][1] = 1.0 / c[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) g[0][k] = -(g[1][k] + g[2][k]);
#pragma unroll
    for (int i = 0; i < 3; ++i) g[i][2] = 0.0;
  }
};

// Mean-curvature flow phi_t = kappa |grad phi| on one narrowband patch (the
// scheme the reference keeps as dead CPU code, meshFIM3d.cu:243-266), with
// P1 elements and a lumped mass M_i = sum_e vol_e / (d+1):
//   kappa_i = -sum_e vol_e gradN_i . n_e / M_i,  n_e = grad phi_e / |grad phi_e|,
//   |grad phi|_i = sum_e vol_e |grad phi_e| / sum_e vol_e,
//   phi_i += dt * kappa_i * |grad phi|_i.
// Elements touching an unreached vertex (|phi| >= LARGENUM / 2) or with
// |grad phi_e| <= 1e-12 contribute nothing. Gradients are taken in each
// copy's local frame; only frame-independent quantities are summed.
// Per-slot products are stored as T, sums are FP64.
template <typename T, int NC, int BS>
__device__ void curvature_patch(const DevMesh<T>& m, DevState<T>& s, int p, double dt, unsigned char* smem) {
  __syncthreads();
  constexpr int K = ElemSolve<T, NC>::kCoords;
  const int vs = m.vert_off[p], ve = m.vert_off[p + 1];
  const int es = m.ele_off[p], ee = m.ele_off[p + 1];
  const int hs = m.halo_off[p], he = m.halo_off[p + 1];
  const int nv = ve - vs, ne = ee - es, nh = he - hs;
  // Per slot: vol gradN_i . n_e; per element: vol and vol |grad phi_e| (both
  // the same for every corner of the element).
  T* s_cv = (T*)smem;
  T* s_vol = s_cv + NC * ne;
  T* s_eik = s_vol + ne;
  T* s_vtx = s_eik + ne;
  for (int i = threadIdx.x; i < nv; i += BS) s_vtx[i] = s.phi[vs + i];
  for (int i = threadIdx.x; i < nh; i += BS) s_vtx[nv + i] = s.phi[m.halo_ids[hs + i]];
  if (threadIdx.x == 0) {
    atomicAdd(s.sizes + kSzAdvect + 0, (unsigned long long)ne);
    atomicAdd(s.sizes + kSzAdvect + 1, (unsigned long long)nv);
    atomicAdd(s.sizes + kSzAdvect + 2, (unsigned long long)nh);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ne; e += BS) {
    const ushort4 c = m.copy_local[es + e];
    double ph[NC], L[K], g[NC][3], vol;
    bool fin = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      ph[i] = (double)s_vtx[corner(c, i)];
      fin = fin && d_abs(ph[i]) < 0.5 * kLargeNum;
    }
    const T* gc = m.coords + (int64_t)(es + e) * K;
#pragma unroll
    for (int k = 0; k < K; ++k) L[k] = (double)gc[k];
    Grad<NC>::eval(L, g, vol);
    double gp[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) gp[k] = gp[k] + ph[i] * g[i][k];
    const double gn = d_sqrt((gp[0] * gp[0] + gp[1] * gp[1]) + gp[2] * gp[2]);
    const bool use = fin && gn > 1e-12;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      const double d = (g[i][0] * gp[0] + g[i][1] * gp[1]) + g[i][2] * gp[2];
      s_cv[e * NC + i] = (T)(use ? vol * d / gn : 0.0);
    }
    s_vol[e] = (T)(use ? vol : 0.0);
    s_eik[e] = (T)(use ? vol * gn : 0.0);
  }
  __syncthreads();
  for (int v = threadIdx.x; v < nv; v += BS) {
    const int o0 = m.vslot_off[vs + v], o1 = m.vslot_off[vs + v + 1];
    double cv = 0.0, sv = 0.0, se = 0.0;
    for (int k = o0; k < o1; ++k) cv += (double)s_cv[m.vslots[k]];
    for (int k = o0; k < o1; ++k) sv += (double)s_vol[m.vslots[k] / NC];
    for (int k = o0; k < o1; ++k) se += (double)s_eik[m.vslots[k] / NC];
    const double old = (double)s_vtx[v];
    if (sv > 0.0) {
      const double kappa = -cv / (sv / (double)NC);
      s.phi_out[vs + v] = (T)(old + dt * kappa * (se / sv));
    } else {
      s.phi_out[vs + v] = (T)old;
    }
  }
}

// Narrowband update of one patch: advection (LS_MODE_ADVECT) or mean-curvature
// flow (LS_MODE_MCF).
template <typename T, int NC, int BS>
__device__ void advect_patch(const DevMesh<T>& m, DevState<T>& s, int p, double dt, unsigned char* smem);

// dt is the patch's timestep (dt_of, bounded by the CFL bound when asked).
template <typename T, int NC, int BS>
__device__ __forceinline__ void update_patch(const DevMesh<T>& m, DevState<T>& s, int p, double dt, int mode,
                                             unsigned char* smem) {
  if (mode == LS_MODE_MCF)
    curvature_patch<T, NC, BS>(m, s, p, dt, smem);
  else
    advect_patch<T, NC, BS>(m, s, p, dt, smem);
}

// Upwind residual-distribution (N-scheme) advection of one narrowband patch
// (kernel_updateT_single_stage3d, meshFIM_kernels.h:36-408): per element copy
// K_i = vol sigma.gradN_i, H = sum K_i phi_i, distribution coefficients
// alpha~_i; per owned vertex phi -= dt * sum(alpha~ H) / sum(alpha~ vol).
// Evaluated in FP64 from the stored values; per-slot products and the result
// are stored as T. With FP64 storage the LDS holds H, vol and the NC
// coefficients per element instead of the 2 NC per-slot products (75 % of
// the bytes, so 256-vertex tet patches fit): the vertex pass forms the same
// FP64 products alpha~_i H and alpha~_i vol, so the bits are unchanged.
template <typename T, int NC, int BS>
__device__ void advect_patch(const DevMesh<T>& m, DevState<T>& s, int p, double dt, unsigned char* smem) {
  __syncthreads();
  constexpr int K = ElemSolve<T, NC>::kCoords;
  constexpr int D = NC - 1;
  constexpr bool per_elem = sizeof(T) == 8;
  constexpr int EW = NC + 2;  // per-element record: H, vol, alpha~_0..NC-1
  const int vs = m.vert_off[p], ve = m.vert_off[p + 1];
  const int es = m.ele_off[p], ee = m.ele_off[p + 1];
  const int hs = m.halo_off[p], he = m.halo_off[p + 1];
  const int nv = ve - vs, ne = ee - es, nh = he - hs;
  T* s_up = (T*)smem;
  T* s_dn = s_up + NC * ne;
  double* s_el = (double*)smem;
  T* s_vtx = per_elem ? (T*)(s_el + EW * ne) : s_dn + NC * ne;
  for (int i = threadIdx.x; i < nv; i += BS) s_vtx[i] = s.phi[vs + i];
  for (int i = threadIdx.x; i < nh; i += BS) s_vtx[nv + i] = s.phi[m.halo_ids[hs + i]];
  if (threadIdx.x == 0) {
    atomicAdd(s.sizes + kSzAdvect + 0, (unsigned long long)ne);
    atomicAdd(s.sizes + kSzAdvect + 1, (unsigned long long)nv);
    atomicAdd(s.sizes + kSzAdvect + 2, (unsigned long long)nh);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ne; e += BS) {
    const ushort4 c = m.copy_local[es + e];
    double ph[NC], L[K], sg[D], g[NC][3], vol;
#pragma unroll
    for (int i = 0; i < NC; ++i) ph[i] = (double)s_vtx[corner(c, i)];
    const T* gc = m.coords + (int64_t)(es + e) * K;
#pragma unroll
    for (int k = 0; k < K; ++k) L[k] = (double)gc[k];
    const T* gs = m.sigma + (int64_t)(es + e) * D;
#pragma unroll
    for (int k = 0; k < D; ++k) sg[k] = (double)gs[k];
    Grad<NC>::eval(L, g, vol);
    if constexpr (NC == 3) {
      if (m.sigma_phi) {
        // (phi_A, phi_B) / |grad phi| (src/core2d/include/meshFIM2d_kernels.h:170-181)
        const double gx = g[0][0] * ph[0] + g[1][0] * ph[1] + g[2][0] * ph[2];
        const double gy = g[0][1] * ph[0] + g[1][1] * ph[1] + g[2][1] * ph[2];
        const double n = sqrt(gx * gx + gy * gy);
        sg[0] = n > 0.0 ? ph[0] / n : 0.0;
        sg[1] = n > 0.0 ? ph[1] / n : 0.0;
      }
    }
    double Kc[NC], Kp[NC], Km[NC], at[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      double d = sg[0] * g[i][0] + sg[1] * g[i][1];
      if (D == 3) d = d + sg[D - 1] * g[i][2];
      Kc[i] = vol * d;
    }
    double H = 0.0, beta = 0.0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      H += Kc[i] * ph[i];
      Kp[i] = d_max(Kc[i], 0.0);
      Km[i] = d_min(Kc[i], 0.0);
      beta += Km[i];
      at[i] = 0.0;
    }
    beta = 1.0 / beta;
    if (d_abs(H) > 1e-8) {
      double alpha[NC], theta = 0.0;
#pragma unroll
      for (int i = 0; i < NC; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < NC; ++j) acc += Km[j] * (ph[i] - ph[j]);
        alpha[i] = (Kp[i] * beta * acc) / H;
      }
#pragma unroll
      for (int i = 0; i < NC; ++i) theta += d_max(0.0, alpha[i]);
#pragma unroll
      for (int i = 0; i < NC; ++i) at[i] = d_max(alpha[i], 0.0) / theta;
    }
    if constexpr (per_elem) {
      s_el[e * EW + 0] = H;
      s_el[e * EW + 1] = vol;
#pragma unroll
      for (int i = 0; i < NC; ++i) s_el[e * EW + 2 + i] = at[i];
    } else {
#pragma unroll
      for (int i = 0; i < NC; ++i) {
        s_up[e * NC + i] = (T)(at[i] * H);
        s_dn[e * NC + i] = (T)(at[i] * vol);
      }
    }
  }
  __syncthreads();
  for (int v = threadIdx.x; v < nv; v += BS) {
    const int o0 = m.vslot_off[vs + v], o1 = m.vslot_off[vs + v + 1];
    double up = 0.0, dn = 0.0;
    if constexpr (per_elem) {
      for (int k = o0; k < o1; ++k) {
        const int sl = m.vslots[k], e = sl / NC;
        up += s_el[e * EW + 2 + sl % NC] * s_el[e * EW];
      }
      for (int k = o0; k < o1; ++k) {
        const int sl = m.vslots[k], e = sl / NC;
        dn += s_el[e * EW + 2 + sl % NC] * s_el[e * EW + 1];
      }
    } else {
      for (int k = o0; k < o1; ++k) up += (double)s_up[m.vslots[k]];
      for (int k = o0; k < o1; ++k) dn += (double)s_dn[m.vslots[k]];
    }
    const double old = (double)s_vtx[v];
    s.phi_out[vs + v] = (T)((d_abs(dn) > 1e-8) ? old - dt * (up / dn) : old);
  }
}

// ---------------------------------------------------------------------------
// CFL timestep bound (kernel_compute_timestep2d, meshFIM2d_kernels.h:405-475,
// reduced by meshFIM2d::compute_deltaT, meshFIM2d.cu:609-674): per element copy
// of the narrowband, with R its inscribed radius, d the dimension and the
// element's eikonal, curvature and advection coefficients,
//   dt1 = 4 R^2 / |Ccurv| / 2 / d     when |Ccurv| > eps,
//   dt2 = 2 R / (|Ceik| + |Cadv|) / d  when |Ceik| + |Cadv| > eps,
// and the bound is the least min(dt1, dt2); the largest R is reported beside
// it. The reference computes this for triangles only and discards it; here
// it is computed for both element kinds and can bound the step's advection.

constexpr double kCflEps = 1e-12;  // _EPS, src/core2d/include/meshFIM2d.h:15
// The reference fills every element's coefficients with Ceik = 10 and
// Ccurv = 0 (src/core2d/cuda/meshFIM2d.cu:724-731).
constexpr double kCeikDefault = 10.0, kCcurvDefault = 0.0;

// Inscribed radius of a copy from its local coordinates (FP64): 3 V / (sum of
// the face areas) for a tet, 2 A / perimeter for a triangle.
template <int NC>
struct Inradius;
template <>
struct Inradius<4> {
  // A=(0,0,0), B=(L0,0,0), C=(L1,L2,0), D=(L3,L4,L5).
  __device__ static double of(const double* L) {
    const double v6 = d_abs(L[0] * L[2] * L[5]);
    const double abc = d_abs(L[0] * L[2]);
    const double abd = d_abs(L[0]) * d_sqrt(L[4] * L[4] + L[5] * L[5]);
    const double ax = L[2] * L[5], ay = -L[1] * L[5], az = L[1] * L[4] - L[2] * L[3];
    const double acd = d_sqrt((ax * ax + ay * ay) + az * az);
    const double bx = L[2] * L[5], by = -(L[1] - L[0]) * L[5], bz = (L[1] - L[0]) * L[4] - L[2] * (L[3] - L[0]);
    const double bcd = d_sqrt((bx * bx + by * by) + bz * bz);
    // 3 V / (sum of areas) = (v6 / 2) / ((abc + abd + acd + bcd) / 2)
    return v6 / (((abc + abd) + acd) + bcd);
  }
};
template <>
struct Inradius<3> {
  // A=(0,0), B=(c0,0), C=(c1,c2).
  __device__ static double of(const double* c) {
    const double a2 = d_abs(c[0] * c[2]);  // twice the area
    const double per = (d_abs(c[0]) + d_sqrt((c[1] - c[0]) * (c[1] - c[0]) + c[2] * c[2])) +
                       d_sqrt(c[1] * c[1] + c[2] * c[2]);
    return a2 / per;
  }
};

// CFL bound and inscribed radius of element copy c.
template <typename T, int NC>
__device__ __forceinline__ void cfl_copy(const DevMesh<T>& m, int64_t c, double& dt, double& rin) {
  constexpr int K = ElemSolve<T, NC>::kCoords;
  constexpr int D = NC - 1;
  double L[K];
  const T* gc = m.coords + c * K;
#pragma unroll
  for (int k = 0; k < K; ++k) L[k] = (double)gc[k];
  rin = Inradius<NC>::of(L);
  const T* gs = m.sigma + c * D;
  double a2 = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) a2 += (double)gs[k] * (double)gs[k];
  const int e = m.copy_src[c];
  const double ceik = d_abs(m.ceik ? m.ceik[e] : m.ceik0);
  const double ccurv = d_abs(m.ccurv ? m.ccurv[e] : m.ccurv0);
  const double cadv = d_sqrt(a2);
  double dt1 = kLargeNum, dt2 = kLargeNum;
  if (ccurv > kCflEps) dt1 = (4.0 * rin * rin) / ccurv / 2.0 / (double)D;
  if (ceik + cadv > kCflEps) dt2 = (rin * 2.0) / (ceik + cadv) / (double)D;
  dt = d_min(dt1, dt2);
}

// One patch's bound, merged into s.cfl with order-preserving atomics. Every
// thread of the workgroup calls it; red holds BS/64 doubles.
template <typename T, int NC, int BS>
__device__ void cfl_patch(const DevMesh<T>& m, DevState<T>& s, int p, double* red) {
  const int es = m.ele_off[p], ee = m.ele_off[p + 1];
  double dt = kLargeNum, rin = -1.0;
  for (int64_t c = es + threadIdx.x; c < ee; c += BS) {
    double d, r;
    cfl_copy<T, NC>(m, c, d, r);
    dt = d_min(dt, d);
    rin = d_max(rin, r);
  }
  __syncthreads();
  dt = block_min_n<double, BS>(dt, red);
  rin = -block_min_n<double, BS>(-rin, red);
  if (threadIdx.x == 0) {
    atomicMin(&s.cfl[0], OrderedBits<double>::of(dt));
    atomicMax(&s.cfl[1], OrderedBits<double>::of(rin));
  }
}

__global__ void k_cfl_init(unsigned long long* cfl) {
  if (threadIdx.x == 0) {
    cfl[0] = OrderedBits<double>::of(kLargeNum);
    cfl[1] = OrderedBits<double>::of(-1.0);
  }
}

// One workgroup per listed patch (the narrowband), or per patch when list is null.
template <typename T, int NC>
__global__ __launch_bounds__(kBlock) void k_cfl(DevMesh<T> m, DevState<T> s, const int32_t* list) {
  __shared__ double red[kBlock / 64];
  cfl_patch<T, NC, kBlock>(m, s, list ? list[blockIdx.x] : (int)blockIdx.x, red);
}

// The advection timestep of patch p: the step's (or the patch's mesh's) dt,
// bounded by the narrowband's CFL bound when the step asks for it.
template <typename T>
__device__ __forceinline__ double advect_dt(const DevMesh<T>& m, const DevState<T>& s, int p, double dt, int cfl) {
  const double d = dt_of(m, p, dt);
  return cfl ? d_min(d, OrderedBits<double>::value(s.cfl[0])) : d;
}

// ---------------------------------------------------------------------------
// Host-driven path: one launch per phase.

// 8 waves per patch: two patches share a CU (LDS permitting), so one
// patch's barrier and FP64 latency stalls overlap the other's work. Measured
// against 1024 threads on MI355X: 64^3 torus 3.72 vs 3.84 ms per step,
// 192^3 MCF 15.1 vs 21.8 ms (per-phase path); with the persistent kernel,
// torus 3.36 vs 3.38 ms and the 512-mesh batch 166 vs 257 ms per step.
constexpr int kFimBlock = 512;
// The persistent step kernel at one workgroup per CU runs 768 threads: three
// waves per SIMD instead of two hide more of the sweep's dependent FP64
// latency, and its ~160 VGPRs still fit without spilling (64^3 torus, FP64:
// 2.28 against 2.36 ms per step with 512 threads).
constexpr int kStepBlock1 = 768;
constexpr int kFimIterCap = 100000;  // FIM iteration cap when none is requested

template <typename T>
__global__ __launch_bounds__(kBlock) void k_init(int64_t N, DevState<T> s) {
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < N; v += (int64_t)gridDim.x * kBlock)
    init_vertex(s, v);
}

__global__ void k_set_counters(int32_t* c, int n, int32_t val) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = val;
}

__global__ void k_fill(int32_t* a, int64_t n, int32_t val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = val;
}

// vpatch[v] = p for the renumbered vertices v of patch p.
__global__ void k_vertex_patch(const int32_t* vert_off, int P, int32_t* vpatch) {
  const int p = blockIdx.x;
  if (p >= P) return;
  for (int v = vert_off[p] + threadIdx.x; v < vert_off[p + 1]; v += blockDim.x) vpatch[v] = p;
}

// Owner patch of every halo entry, so a visit's halo gather needs no
// dependent vpatch lookup.
__global__ void k_halo_patch(const int32_t* halo_off, const int32_t* halo_ids, const int32_t* vpatch, int P,
                             int32_t* halo_patch) {
  const int p = blockIdx.x;
  if (p >= P) return;
  for (int i = halo_off[p] + threadIdx.x; i < halo_off[p + 1]; i += blockDim.x) halo_patch[i] = vpatch[halo_ids[i]];
}

// Per-patch offset header (DevMesh::phdr).
__global__ void k_patch_hdr(const int32_t* vert_off, const int32_t* ele_off, const int32_t* halo_off,
                            const int32_t* vslot_off, int P, int4* hdr) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
    hdr[2 * p] = make_int4(vert_off[p], vert_off[p + 1], ele_off[p], ele_off[p + 1]);
    hdr[2 * p + 1] = make_int4(halo_off[p], halo_off[p + 1], vslot_off[vert_off[p]], vslot_off[vert_off[p + 1]]);
  }
}

__global__ void k_clear_sizes(unsigned long long* c) {
  if ((int)threadIdx.x < kNumSizes) c[threadIdx.x] = 0ull;
}

template <typename T, int NC>
__global__ __launch_bounds__(kBlock) void k_seeds(DevMesh<T> m, DevState<T> s, const int32_t* scan,
                                                  int32_t* active, int32_t* counters) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int p = scan[blockIdx.x];
  const int flag = seeds_patch<T, NC, kBlock>(m, s, p, smem);
  if (threadIdx.x == 0 && flag) active[atomicAdd(&counters[kCntSeed], 1)] = p;
}

template <typename T, int NC, int BS, int ORG>
__global__ __launch_bounds__(BS, 4) void k_sweep(DevMesh<T> m, DevState<T> s, const int32_t* list,
                                              int niter, int write_out, T tol, double leps, int epoch_in,
                                              int epoch_out, int stage) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ T red[BS / 64];
  __shared__ __align__(8) int s_ndirty[10];
  __shared__ unsigned long long s_acc[kNumAcc];
  clear_stats(s_acc);  // sweep_patch starts with a barrier
  sweep_patch<T, NC, BS, ORG>(m, s, list[blockIdx.x], niter, write_out, tol, leps, smem, red, s_ndirty, epoch_in,
                              epoch_out, 0, default_io(s), s_acc, stage);
  __syncthreads();
  flush_stats(s, s_acc);
}

// CopyOutBack (redistance_kernels3d.h:772-783) plus the active-list label of
// the patch: it stays active iff it did not converge.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_copyback(DevMesh<T> m, DevState<T> s, const int32_t* list) {
  const int p = list[blockIdx.x];
  const int vs = m.vert_off[p], ve = m.vert_off[p + 1];
  for (int v = vs + threadIdx.x; v < ve; v += kBlock) s.DT[v] = s.DT_out[v];
  if (threadIdx.x == 0) s.act[p] = s.bcon[p] ? 0 : 1;
}

// Neighbour activation (redistance3d.cu:141-181): a converged patch whose
// min is below the band activates every neighbour not currently active.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_activate(DevMesh<T> m, DevState<T> s, const int32_t* list, int n,
                                                     T bw, int32_t* fresh_list) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int p = list[i];
  if (!(s.bcon[p] && s.bmin[p] < band_of(m, p, bw))) return;
  for (int k = m.padj_off[p]; k < m.padj_off[p + 1]; ++k) {
    const int q = m.padj[k];
    if (s.act[q] == 0 && atomicExch(&s.fresh[q], 1) == 0)
      fresh_list[atomicAdd(&s.counters[kCntFresh], 1)] = q;
  }
}

// Next active list (redistance3d.cu:183-222): unconverged patches of the
// current list, then probed neighbours that did not converge.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_finalize(DevState<T> s, const int32_t* list, int n,
                                                     const int32_t* fresh_list, int nfresh, int32_t* next) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    const int p = list[i];
    if (s.act[p]) next[atomicAdd(&s.counters[kCntNext], 1)] = p;
  } else if (i < n + nfresh) {
    const int q = fresh_list[i - n];
    s.fresh[q] = 0;
    if (!s.bcon[q]) next[atomicAdd(&s.counters[kCntNext], 1)] = q;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_narrowband(DevMesh<T> m, DevState<T> s, T bw, int32_t* nb) {
  __shared__ T red[kBlock / 64];
  const int p = blockIdx.x;
  const T mn = narrowband_patch<T, kBlock>(m, s, p, red);
  if (threadIdx.x == 0 && mn <= band_of(m, p, bw)) {
    nb[atomicAdd(&s.counters[kCntNB], 1)] = p;
    atomicAdd(&s.counters[kCntNBVerts], m.vert_off[p + 1] - m.vert_off[p]);
  }
}

template <typename T, int NC>
__global__ __launch_bounds__(kBlock) void k_advect(DevMesh<T> m, DevState<T> s, const int32_t* nb, double dt,
                                                   int mode, int cfl) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int p = nb[blockIdx.x];
  update_patch<T, NC, kBlock>(m, s, p, advect_dt(m, s, p, dt, cfl), mode, smem);
}

// CopyOutBack_levelset3d (meshFIM_kernels.h:22-34).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy_phi(DevMesh<T> m, DevState<T> s, const int32_t* nb) {
  const int p = nb[blockIdx.x];
  for (int v = m.vert_off[p] + threadIdx.x; v < m.vert_off[p + 1]; v += kBlock) s.phi[v] = s.phi_out[v];
}

// The same over a narrowband list whose length is read on the device (after
// the persistent step): blocks take list entries grid-stride.
template <typename T, int NC>
__global__ __launch_bounds__(kBlock) void k_advect_dev(DevMesh<T> m, DevState<T> s, const int32_t* nb,
                                                       const int32_t* count, double dt, int mode, int cfl) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = *count;
  for (int i = blockIdx.x; i < n; i += gridDim.x)
    update_patch<T, NC, kBlock>(m, s, nb[i], advect_dt(m, s, nb[i], dt, cfl), mode, smem);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy_phi_dev(DevMesh<T> m, DevState<T> s, const int32_t* nb,
                                                         const int32_t* count) {
  const int n = *count;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int p = nb[i];
    for (int v = m.vert_off[p] + threadIdx.x; v < m.vert_off[p + 1]; v += kBlock) s.phi[v] = s.phi_out[v];
  }
}

inline int grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  return (int)(g < 1 ? 1 : g);
}

// ---------------------------------------------------------------------------
// Learned warm start: inference over the previous narrowband, training-pair
// collection, gradients and Adam (ls_mlp.h describes the network).

constexpr int kMlpBlock = 256;  // 4 waves, each 16 samples per tile
static_assert(mlp::kNumParams == LS_MLP_PARAMS, "parameter count of the C-ABI");

// Patch of element copy c (copies are sorted by patch).
template <typename T>
__device__ inline int copy_patch(const DevMesh<T>& m, int64_t c) {
  int lo = 0, hi = m.P;  // ele_off[lo] <= c < ele_off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (m.ele_off[mid] <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Warm values of one narrowband patch: per element copy the predicted
// corner distances, per owned vertex their mean over its copies.
template <typename T, int NC>
__global__ __launch_bounds__(kMlpBlock) void k_mlp_predict(DevMesh<T> m, DevState<T> s, const int32_t* nb,
                                                           mlp::Weights w, T* warm) {
  constexpr int K = ElemSolve<T, NC>::kCoords;
  extern __shared__ __align__(16) unsigned char smem[];
  const int p = nb[blockIdx.x];
  const int vs = m.vert_off[p], ve = m.vert_off[p + 1];
  const int es = m.ele_off[p], ee = m.ele_off[p + 1];
  const int hs = m.halo_off[p], he = m.halo_off[p + 1];
  const int nv = ve - vs, ne = ee - es, nh = he - hs;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15;
  float* tile = (float*)smem + wave * mlp::kWaveTile;
  float* X = tile;
  float* H1 = X + 16 * mlp::kInPad;
  float* H2 = H1 + 16 * mlp::kHid;
  float* O = H2 + 16 * mlp::kHid;
  float* s_pred = (float*)smem + (kMlpBlock / 64) * mlp::kWaveTile;  // NC per copy (one per slot)
  float* s_vtx = s_pred + NC * ne;
  for (int i = threadIdx.x; i < nv; i += kMlpBlock) s_vtx[i] = (float)s.phi[vs + i];
  for (int i = threadIdx.x; i < nh; i += kMlpBlock) s_vtx[nv + i] = (float)s.phi[m.halo_ids[hs + i]];
  __syncthreads();
  const int ntiles = (ne + 15) / 16;
  for (int t = wave; t < ntiles; t += kMlpBlock / 64) {
    const int e = 16 * t + r;
    float ph[4] = {0.f, 0.f, 0.f, 0.f}, sc = 1.f;
    if (lane < 16) {
      float x[mlp::kIn];
      if (e < ne) {
        const ushort4 c = m.copy_local[es + e];
        float L[6];
#pragma unroll
        for (int i = 0; i < NC; ++i) ph[i] = s_vtx[corner(c, i)];
#pragma unroll
        for (int k = 0; k < K; ++k) L[k] = (float)m.coords[(int64_t)(es + e) * K + k];
        sc = mlp::copy_features<NC>(ph, L, x);
      } else {
#pragma unroll
        for (int k = 0; k < mlp::kIn; ++k) x[k] = 0.f;
      }
#pragma unroll
      for (int k = 0; k < mlp::kInPad; ++k) X[r * mlp::kInPad + k] = k < mlp::kIn ? x[k] : 0.f;
    }
    mlp::wave_sync();
    mlp::forward16(w, X, H1, H2, O, lane);
    if (lane < 16 && e < ne) {
#pragma unroll
      for (int i = 0; i < NC; ++i) {
        const float d = fabsf(ph[i]) + sc * O[r * mlp::kOutPad + i];
        s_pred[e * NC + i] = d > 0.f ? d : 0.f;
      }
    }
    mlp::wave_sync();
  }
  __syncthreads();
  for (int v = threadIdx.x; v < nv; v += kMlpBlock) {
    const int o0 = m.vslot_off[vs + v], o1 = m.vslot_off[vs + v + 1];
    float acc = 0.f;
    for (int k = o0; k < o1; ++k) acc += s_pred[m.vslots[k]];
    warm[vs + v] = o1 > o0 ? (T)(acc / (float)(o1 - o0)) : (T)kLargeNum;
  }
}

// Training inputs of the copies listed in `copies`: padded features from
// the current phi (before the step) into X[i][kInPad].
template <typename T, int NC>
__global__ __launch_bounds__(kBlock) void k_mlp_collect_x(DevMesh<T> m, DevState<T> s, const int32_t* copies,
                                                          int64_t n, float* X) {
  constexpr int K = ElemSolve<T, NC>::kCoords;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int64_t c = copies[i];
  // The copy's patch-local corners index its patch's vertices, then halo.
  const int p = copy_patch(m, c);
  const int vs = m.vert_off[p], nv = m.vert_off[p + 1] - vs, hs = m.halo_off[p];
  const ushort4 cl = m.copy_local[c];
  float ph[4], L[6], x[mlp::kIn];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int lv = corner(cl, k);
    ph[k] = (float)s.phi[lv < nv ? vs + lv : m.halo_ids[hs + lv - nv]];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) L[k] = (float)m.coords[c * K + k];
  mlp::copy_features<NC>(ph, L, x);
#pragma unroll
  for (int k = 0; k < mlp::kInPad; ++k) X[i * mlp::kInPad + k] = k < mlp::kIn ? x[k] : 0.f;
}

// Training targets after the step: y_k = D_k / s - |x_k| with D the
// redistanced distances (DT) at the copy's corners.
template <typename T, int NC>
__global__ __launch_bounds__(kBlock) void k_mlp_collect_y(DevMesh<T> m, DevState<T> s, const int32_t* copies,
                                                          int64_t n, const float* X, float* Y) {
  constexpr int K = ElemSolve<T, NC>::kCoords;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int64_t c = copies[i];
  const int p = copy_patch(m, c);
  const int vs = m.vert_off[p], nv = m.vert_off[p + 1] - vs, hs = m.halo_off[p];
  const ushort4 cl = m.copy_local[c];
  const float inv = 1.0f / (float)m.coords[c * K];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k >= NC) {  // triangles: no fourth corner
      Y[i * mlp::kOut + k] = 0.f;
      continue;
    }
    const int lv = corner(cl, k);
    const float D = (float)s.DT[lv < nv ? vs + lv : m.halo_ids[hs + lv - nv]];
    Y[i * mlp::kOut + k] = D < 0.5f * (float)kLargeNum ? D * inv - fabsf(X[i * mlp::kInPad + k]) : 0.f;
  }
}

// Loss sum (O - Y)^2 over valid samples and its gradient for every parameter
// (flat layout of ls_mlp.h, scaled by 1/n_total), accumulated with atomics.
// Samples are rows idx[i] of X/Y when idx is given, else rows i.
__global__ __launch_bounds__(kMlpBlock) void k_mlp_grad(const float* X, const float* Y, const int32_t* idx,
                                                        int64_t n, float inv_total, mlp::Weights w, float* grad,
                                                        float* loss) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, r = lane & 15;
  float* tile = (float*)smem + wave * mlp::kWaveTile;
  float* Xs = tile;
  float* H1 = Xs + 16 * mlp::kInPad;
  float* H2 = H1 + 16 * mlp::kHid;
  float* O = H2 + 16 * mlp::kHid;
  float* red = (float*)smem + (kMlpBlock / 64) * mlp::kWaveTile;  // kNumParams
  for (int i = threadIdx.x; i < mlp::kNumParams; i += kMlpBlock) red[i] = 0.f;
  __syncthreads();
  mlp::f32x4 a1[2][4], a2[4][4], a3[4];
  float gb1[4] = {0.f, 0.f, 0.f, 0.f}, gb2[4] = {0.f, 0.f, 0.f, 0.f}, gb3 = 0.f, lsum = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    a3[i] = mlp::f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 4; ++j) a2[i][j] = mlp::f32x4{0.f, 0.f, 0.f, 0.f};
    a1[0][i] = a1[1][i] = mlp::f32x4{0.f, 0.f, 0.f, 0.f};
  }
  const int64_t ntiles = (n + 15) / 16;
  for (int64_t t = (int64_t)blockIdx.x * (kMlpBlock / 64) + wave; t < ntiles; t += (int64_t)gridDim.x * (kMlpBlock / 64)) {
    const int64_t i = 16 * t + r;
    const bool ok = i < n;
    const int64_t row = ok ? (idx ? (int64_t)idx[i] : i) : 0;
    // Inputs: lane group g copies columns 8g..8g+7 of the sample's row.
#pragma unroll
    for (int k = 0; k < 8; ++k) Xs[r * mlp::kInPad + 8 * g + k] = ok ? X[row * mlp::kInPad + 8 * g + k] : 0.f;
    mlp::wave_sync();
    mlp::forward16(w, Xs, H1, H2, O, lane);
    // dO = 2 (O - Y) / n_total over the 4 outputs of valid samples; the
    // output tile is overwritten with dO. Lane (g, r) handles O[r][g].
    {
      const float y = ok ? Y[row * mlp::kOut + g] : 0.f;
      const float o = O[r * mlp::kOutPad + g];
      const float d = ok ? o - y : 0.f;
      lsum += d * d;
      mlp::wave_sync();
      O[r * mlp::kOutPad + g] = 2.f * d * inv_total;
#pragma unroll
      for (int k = 4; k < mlp::kOutPad; k += 4) O[r * mlp::kOutPad + g + k] = 0.f;
    }
    mlp::wave_sync();
    // db3, dW3 = H2^T dO, dH2 = (dO W3^T) * [H2 > 0].
    if (lane < 16) {
      float acc = 0.f;
      for (int mm = 0; mm < 16; ++mm) acc += O[mm * mlp::kOutPad + r];
      gb3 += acc;
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const mlp::bf16x4 A = mlp::pack4(H2[(4 * g) * mlp::kHid + 16 * mt + r], H2[(4 * g + 1) * mlp::kHid + 16 * mt + r],
                                       H2[(4 * g + 2) * mlp::kHid + 16 * mt + r], H2[(4 * g + 3) * mlp::kHid + 16 * mt + r]);
      const mlp::bf16x4 B = mlp::pack4(O[(4 * g) * mlp::kOutPad + r], O[(4 * g + 1) * mlp::kOutPad + r],
                                       O[(4 * g + 2) * mlp::kOutPad + r], O[(4 * g + 3) * mlp::kOutPad + r]);
      a3[mt] = mlp::mfma(A, B, a3[mt]);
    }
    {
      const mlp::bf16x4 A = mlp::load_a(O, mlp::kOutPad, 0, lane);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const uint16_t* p = w.w3 + (16 * nt + r) * mlp::kOutPad + 4 * g;
        mlp::bf16x4 B;
        B[0] = (short)p[0]; B[1] = (short)p[1]; B[2] = (short)p[2]; B[3] = (short)p[3];
        const mlp::f32x4 d = mlp::mfma(A, B, mlp::f32x4{0.f, 0.f, 0.f, 0.f});
        mlp::wave_sync();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float* h = &H2[(4 * g + j) * mlp::kHid + 16 * nt + r];
          *h = *h > 0.f ? d[j] : 0.f;
        }
      }
    }
    mlp::wave_sync();
    // db2, dW2 = H1^T dH2, dH1 = (dH2 W2^T) * [H1 > 0].
#pragma unroll
    for (int q = 0; q < 4; ++q)
      for (int mm = 0; mm < 4; ++mm) gb2[q] += H2[(4 * g + mm) * mlp::kHid + 16 * q + r];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const mlp::bf16x4 A = mlp::pack4(H1[(4 * g) * mlp::kHid + 16 * mt + r], H1[(4 * g + 1) * mlp::kHid + 16 * mt + r],
                                       H1[(4 * g + 2) * mlp::kHid + 16 * mt + r], H1[(4 * g + 3) * mlp::kHid + 16 * mt + r]);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const mlp::bf16x4 B = mlp::pack4(H2[(4 * g) * mlp::kHid + 16 * nt + r], H2[(4 * g + 1) * mlp::kHid + 16 * nt + r],
                                         H2[(4 * g + 2) * mlp::kHid + 16 * nt + r], H2[(4 * g + 3) * mlp::kHid + 16 * nt + r]);
        a2[mt][nt] = mlp::mfma(A, B, a2[mt][nt]);
      }
    }
    {
      mlp::f32x4 d[4];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        d[nt] = mlp::f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
          const uint16_t* p = w.w2 + (16 * nt + r) * mlp::kHid + 16 * ks + 4 * g;
          mlp::bf16x4 B;
          B[0] = (short)p[0]; B[1] = (short)p[1]; B[2] = (short)p[2]; B[3] = (short)p[3];
          d[nt] = mlp::mfma(mlp::load_a(H2, mlp::kHid, ks, lane), B, d[nt]);
        }
      }
      mlp::wave_sync();
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float* h = &H1[(4 * g + j) * mlp::kHid + 16 * nt + r];
          *h = *h > 0.f ? d[nt][j] : 0.f;
        }
    }
    mlp::wave_sync();
    // db1, dW1 = X^T dH1.
#pragma unroll
    for (int q = 0; q < 4; ++q)
      for (int mm = 0; mm < 4; ++mm) gb1[q] += H1[(4 * g + mm) * mlp::kHid + 16 * q + r];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const mlp::bf16x4 A = mlp::pack4(Xs[(4 * g) * mlp::kInPad + 16 * mt + r], Xs[(4 * g + 1) * mlp::kInPad + 16 * mt + r],
                                       Xs[(4 * g + 2) * mlp::kInPad + 16 * mt + r], Xs[(4 * g + 3) * mlp::kInPad + 16 * mt + r]);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const mlp::bf16x4 B = mlp::pack4(H1[(4 * g) * mlp::kHid + 16 * nt + r], H1[(4 * g + 1) * mlp::kHid + 16 * nt + r],
                                         H1[(4 * g + 2) * mlp::kHid + 16 * nt + r], H1[(4 * g + 3) * mlp::kHid + 16 * nt + r]);
        a1[mt][nt] = mlp::mfma(A, B, a1[mt][nt]);
      }
    }
    mlp::wave_sync();
  }
  // Merge the wave accumulators in LDS, then into global memory.
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int in = 16 * mt + 4 * g + j, hid = 16 * nt + r;
        if (in < mlp::kIn) atomicAdd(&red[mlp::kOffW1 + in * mlp::kHid + hid], a1[mt][nt][j]);
      }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        atomicAdd(&red[mlp::kOffW2 + (16 * mt + 4 * g + j) * mlp::kHid + 16 * nt + r], a2[mt][nt][j]);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (r < mlp::kOut) atomicAdd(&red[mlp::kOffW3 + (16 * mt + 4 * g + j) * mlp::kOut + r], a3[mt][j]);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    atomicAdd(&red[mlp::kOffB1 + 16 * q + r], gb1[q]);
    atomicAdd(&red[mlp::kOffB2 + 16 * q + r], gb2[q]);
  }
  if (lane < mlp::kOut) atomicAdd(&red[mlp::kOffB3 + lane], gb3);
  __syncthreads();
  for (int i = threadIdx.x; i < mlp::kNumParams; i += kMlpBlock)
    if (red[i] != 0.f) atomicAdd(&grad[i], red[i]);
  // Loss: lanes hold partial sums of squared errors.
  float l = lsum;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
  if (lane == 0 && l != 0.f) atomicAdd(loss, l);
}

// Adam step on the FP32 parameters, then the BF16/padded copies the MFMA
// kernels read.
__global__ void k_mlp_adam(float* prm, float* mom, float* vel, const float* grad, int n, float lr, float beta1,
                           float beta2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gr = grad[i];
  const float mo = beta1 * mom[i] + (1.f - beta1) * gr;
  const float ve = beta2 * vel[i] + (1.f - beta2) * gr * gr;
  mom[i] = mo;
  vel[i] = ve;
  prm[i] -= lr * (mo / c1) / (sqrtf(ve / c2) + eps);
}

__global__ void k_mlp_pack(const float* prm, uint16_t* w1, uint16_t* w2, uint16_t* w3, float* b1, float* b2,
                           float* b3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < mlp::kInPad * mlp::kHid) {
    const int in = i / mlp::kHid, h = i % mlp::kHid;
    w1[i] = in < mlp::kIn ? mlp::to_bf16(prm[mlp::kOffW1 + in * mlp::kHid + h]) : (uint16_t)0;
  }
  if (i < mlp::kHid * mlp::kHid) w2[i] = mlp::to_bf16(prm[mlp::kOffW2 + i]);
  if (i < mlp::kHid * mlp::kOutPad) {
    const int h = i / mlp::kOutPad, o = i % mlp::kOutPad;
    w3[i] = o < mlp::kOut ? mlp::to_bf16(prm[mlp::kOffW3 + h * mlp::kOut + o]) : (uint16_t)0;
  }
  if (i < mlp::kHid) {
    b1[i] = prm[mlp::kOffB1 + i];
    b2[i] = prm[mlp::kOffB2 + i];
  }
  if (i < mlp::kOutPad) b3[i] = i < mlp::kOut ? prm[mlp::kOffB3 + i] : 0.f;
}

// Scratch of the allocation-free gradient entry point
// (ls_mlp_gradient_noalloc): the staged samples and the gradient
// accumulator, in one file-scope device array. Each GPU the library runs on
// holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for that device
// (scratch_address). Calls that overlap on one GPU share it and race.
constexpr int64_t kScratchSamples = 16384;
constexpr size_t kScratchX = (size_t)kScratchSamples * mlp::kInPad;  // padded inputs
constexpr size_t kScratchY = (size_t)kScratchSamples * mlp::kOut;    // targets
constexpr size_t kScratchG = (size_t)mlp::kNumParams + 1;             // gradient sums + loss
static __device__ __align__(256) float scratch[kScratchX + kScratchY + kScratchG];

__global__ void k_zero_f(float* a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.f;
}

// ---------------------------------------------------------------------------
// Persistent path: one cooperative launch per timestep. Phases are separated
// by grid barriers; the FIM active-list loop runs on the device, so a step
// needs no host round trip.

template <typename T>
struct StepArgs {
  int first;         // 1: seeds scan every patch; 0: the previous narrowband
  int nb_in;         // nb buffer holding the previous narrowband
  int niter;
  int inside_niter;
  int max_iters;     // FIM iteration cap (0 = none)
  T tol;
  T bw;
  double leps;
  double dt;
  int mode;          // LS_MODE_ADVECT or LS_MODE_MCF
  int warm;          // 0: plain; 1: warm start from |phi|; 2: from s.warm
  double warm_margin;// warm value = prediction * (1 + warm_margin) + warm_abs
  double warm_abs;
  float res_tol;     // acceptance threshold above which a warm start falls back
  int cfl;           // 1: the advection's dt is bounded by the new narrowband's CFL bound
  int stage;         // 1: sweeps stage each copy's corner ids and heights in LDS
};

// Counters are read right after a grid barrier, whose acquire fence already
// dropped stale lines; a relaxed agent-scope load suffices.
__device__ __forceinline__ int ld_acquire(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier for a co-resident grid (cooperative launch): one arrival
// atomic per workgroup and a generation word the waiting workgroups poll.
// The workgroup barrier makes every wave's stores visible to wave 0, whose
// agent-scope fences then publish them (release) and drop stale cache lines
// for the whole CU (acquire); one wave per workgroup pays for the fences.
__device__ __forceinline__ void grid_barrier(unsigned int* bar, unsigned int nblocks) {
  __syncthreads();
  if (threadIdx.x == 0) {
    // Release this workgroup's writes (L2 write-back; a full fence would
    // also invalidate L2 here, which the acquire below repeats).
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned int gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int arrived =
        __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    if (arrived == nblocks) {
      __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen)
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the other workgroups' writes
  }
  __syncthreads();
}

// Block 0 accumulates barrier-to-barrier phase times.
// Phase timer kept in LDS: tk[0..kNumTicks) accumulate, tk[kNumTicks] holds
// the last stamp. Only the timing thread touches it, so no thread carries
// timer registers through the launch.
struct PhaseClock {
  unsigned long long* tk;
  __device__ void start(bool on) {
    if (on) {
      for (int i = 0; i < kNumTicks; ++i) tk[i] = 0ull;
      tk[kNumTicks] = wall_clock64();
    }
  }
  __device__ void lap(bool on, int i) {
    if (on) {
      const unsigned long long n = wall_clock64();
      tk[i] += n - tk[kNumTicks];
      tk[kNumTicks] = n;
    }
  }
};

// ORG selects the sweep organisation (kOrg*) and WARM compiles the
// warm-start phases in: each variant carries only the code it runs. WPE is
// the waves per SIMD the variant is compiled for: 4 keeps the register
// demand within the 128 VGPRs that let two 512-thread workgroups share a CU;
// 3 (one 768-thread workgroup per CU) leaves 168, so nothing is spilled to
// scratch, whose dirty lines every grid barrier would otherwise write back.
// STG >= 0 compiles one staging choice only (a.stage must equal it), which
// frees the registers the other path's addresses held.
template <typename T, int NC, int BS, int ORG, bool WARM, int WPE, int STG = -1>
__global__ __launch_bounds__(BS, WPE) void k_step(DevMesh<T> m, DevState<T> s, StepArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ T red[BS / 64];
  __shared__ __align__(8) int s_ndirty[10];
  __shared__ int s_work[2];  // work draws: item, reserved second item
  __shared__ double cred[BS / 64];  // CFL reductions
  __shared__ unsigned long long s_acc[kNumAcc];  // statistics, flushed at the end
  const int G = gridDim.x, b = blockIdx.x, tid = threadIdx.x;
  const int64_t gtid = (int64_t)b * BS + tid, gsize = (int64_t)G * BS;
  const int nb_out = 1 - a.nb_in;
  const bool clk = (gtid == 0);
  __shared__ unsigned long long tk[kNumTicks + 1];
  PhaseClock pcl{tk};
  pcl.start(clk);
  if (tid == 0) s_work[0] = s_work[1] = -1;  // visible after the first grid barrier
  clear_stats(s_acc);

  // Phases 0-2 run once, or twice when a warm start fails its residual
  // check: the second attempt is the plain method from the same phi over
  // the meshes that failed (every mesh of a batch is independent; the others
  // keep their warm results).
  for (int attempt = 0; attempt < 2; ++attempt) {
    const bool warm = WARM && a.warm != 0 && !a.first && attempt == 0;
    const bool redo = attempt == 1;

  // Phase 0: vertex reset and sign save; patch marks; counters.
  for (int64_t v = gtid; v < m.N; v += gsize)
    if (!redo || s.gfail[group_of(m, m.vpatch[v])]) init_vertex(s, v);
  if (!redo)
    for (int64_t g = gtid; g < m.ngroups; g += gsize) s.gfail[g] = 0;
  for (int64_t q = gtid; q < m.P; q += gsize) {
    s.inA[q] = 0;
    s.freshE[q] = 0;
    s.rs_epoch[q] = 0;
    s.visitE[q] = 0;
    s.visitE2[q] = 0;
    s.ve_tag[q] = 0;
  }
  if (gtid < kNumSizes && attempt == 0) s.sizes[gtid] = 0ull;
  if (gtid == 0) {
    s.pc[kPcSeed] = 0;
    s.pc[kPcFresh] = s.pc[kPcFresh + 1] = 0;
    s.pc[kPcNext] = s.pc[kPcNext + 1] = 0;
    s.pc[kPcWorkV] = s.pc[kPcWorkV + 1] = s.pc[kPcWorkP] = s.pc[kPcWorkP + 1] = 0;
    s.pc[kPcWorkS] = 0;
    s.pc[kPcNbVerts] = 0;
    if (attempt == 0) {
      s.pc[kPcIters] = s.pc[kPcVisits] = s.pc[kPcProbes] = 0;
      s.pc[kPcFallback] = 0;
    }
    s.pc[kPcOverflow] = 0;
    s.pc[kPcNb + nb_out] = 0;
    for (int k = kPcOvF; k < kNumPc; ++k) s.pc[k] = 0;
    if (a.cfl) {
      s.cfl[0] = OrderedBits<double>::of(kLargeNum);
      s.cfl[1] = OrderedBits<double>::of(-1.0);
    }
    if (attempt == 0) {  // the first attempt's warm-start figures are reported
      s.pc[kPcResid] = 0;
      s.pc[kPcWarm] = 0;
    }
  }
  grid_barrier(s.bar, G);

  // Phase 1: seeds over every patch (first step) or the previous narrowband.
  const int nscan = a.first ? m.P : ld_acquire(&s.pc[kPcNb + a.nb_in]);
  for (int i = first_item(nscan, &s.pc[kPcWorkS], s_work); i < nscan;
       i = later_item(nscan, &s.pc[kPcWorkS], s_work)) {
    const int p = a.first ? i : s.nb[a.nb_in][i];
    if (redo && !s.gfail[group_of(m, p)]) continue;
    if (seeds_patch<T, NC, BS>(m, s, p, smem) && tid == 0) {
      s.listA[atomicAdd(&s.pc[kPcSeed], 1)] = p;
      s.inA[p] = 1;
    }
  }
  grid_barrier(s.bar, G);
  if constexpr (WARM) if (warm) {
    // Phase 1b: warm values over the previous narrowband; every patch that
    // receives one joins the first active list.
    for (int i = b; i < nscan; i += G) {
      const int p = s.nb[a.nb_in][i];
      if (warm_patch<T, NC, BS>(m, s, p, a.warm, a.warm_margin, a.warm_abs, smem) && tid == 0) {
        atomicAdd(&s.pc[kPcWarm], 1);
        if (s.inA[p] == 0) {
          s.inA[p] = 1;
          s.listA[atomicAdd(&s.pc[kPcSeed], 1)] = p;
        }
      }
    }
    grid_barrier(s.bar, G);
  }
  pcl.lap(clk, kTkSeed);
  // Phase 2: FIM (redistance3d.cu:85-224) with overlapped iterations and one
  // grid barrier per iteration. The phase of iteration E (after the barrier
  // that ends iteration E-1) holds, as independent work items: the copy-back
  // of E-1's members into DT; the visits of E's early members (E-1's
  // unconverged members, or the first list); and the probes of the patches
  // E-1's visits marked, each continuing at once as the patch's visit of E
  // when it finds the patch unconverged. Every item reads the state after
  // E-1: patches visited in E-1 from that iteration's DT_out (by parity),
  // all others from DT, so the copy-back can run beside them; E's visits
  // write the other DT_out. The sets of visits, probes and their inputs are
  // those of the reference's loop; only the overlap differs.
  {
    const int P = m.P;
    int32_t* ovF = s.ov_lists;          // [3][P] probe marks, patches visited earlier in the step
    int32_t* ovE = s.ov_lists + 3 * P;  // [3][P] early members
    int32_t* ovM = s.ov_lists + 6 * P;  // [3][P] members to copy back
    int32_t* ovN = s.ov_lists + 9 * P;  // [3][P] probe marks, patches not yet visited in the step
    int iters = 0;
    for (int E = 1;; ++E) {
      const int sE = E % 3, sP = (E + 2) % 3, sN = (E + 1) % 3;
      // parity buffers by selects (a runtime-indexed array would live in scratch)
      const bool odd = E & 1;
      T* const dcur = odd ? s.DT_out2 : s.DT_out;
      const T* const dprev = odd ? s.DT_out : s.DT_out2;
      int32_t* const vcur = odd ? s.visitE2 : s.visitE;
      const int32_t* const vprev = odd ? s.visitE : s.visitE2;
      int32_t* const bcur = odd ? s.bcon2 : s.bcon;
      const int32_t* const bprev = odd ? s.bcon : s.bcon2;
      // E-1's phase counted as an iteration when it visited a patch (read
      // here with this phase's counters, not as a load of its own).
      const int visPrev = E == 1 ? 0 : ld_acquire(&s.pc[kPcOvVis + sP]);
      const int nEarly = E == 1 ? ld_acquire(&s.pc[kPcSeed]) : ld_acquire(&s.pc[kPcOvEl + sE]);
      const int nNew = ld_acquire(&s.pc[kPcOvFn + sP]);
      const int nProbe = ld_acquire(&s.pc[kPcOvF + sP]) + nNew;
      const int nCopy = ld_acquire(&s.pc[kPcOvMl + sP]);
      if (gtid == 0) {  // slots this phase neither reads nor writes; the next one uses them
        s.pc[kPcOvF + sN] = 0;
        s.pc[kPcOvFn + sN] = 0;
        s.pc[kPcOvEl + sP] = 0;
        s.pc[kPcOvMl + sN] = 0;
        s.pc[kPcOvW + sN] = 0;
        s.pc[kPcOvVis + sN] = 0;
      }
      if (visPrev > 0) {
        ++iters;
        if (gtid == 0) s.pc[kPcIters] += 1;
      }
      const bool capped = a.max_iters > 0 && iters >= a.max_iters;
      const bool done = (nEarly == 0 && nProbe == 0) || capped;
      const int n = done ? 0 : nEarly + nProbe;
      // CopyOutBack of E-1's members (this phase's visits read them from
      // dprev; the next phase reads DT). With fewer items than workgroups the
      // workgroups without an item do it, so it delays no item; otherwise
      // every workgroup does its share first, while dprev is still in cache.
      for (int i = n < G ? b - n : b; i >= 0 && i < nCopy; i += n < G ? G - n : G) {
        const int p = ovM[sP * P + i];
        const int vs = m.vert_off[p], ve = m.vert_off[p + 1];
        for (int v = vs + tid; v < ve; v += BS) s.DT[v] = dprev[v];
      }
      if (done) {
        if (capped && nEarly + nProbe > 0 && gtid == 0) s.pc[kPcOverflow] = 1;
        grid_barrier(s.bar, G);
        pcl.lap(clk, kTkVisit);
        break;
      }
      const int members = E - 1;
      const int32_t* early = E == 1 ? s.listA : ovE + sE * P;
      // With under two items per workgroup the slowest workgroup's items set
      // the phase's length, so the first-visit probes go first (see below);
      // with more, the phase is work-bound and the early visits go first.
      const int nFirst = n < 2 * G ? nNew : 0;
      for (int i = first_item(n, &s.pc[kPcOvW + sE], s_work); i < n; i = later_item(n, &s.pc[kPcOvW + sE], s_work)) {
        // kind 0: a visit; 1: a probe that continues as the visit when the
        // patch is unconverged; 2: a one-sweep probe (a one-sweep visit then
        // recomputes its sweep from a fresh load). In a short phase items go
        // longest first as far as that is known: probes of patches the front
        // reaches for the first time in the step (they continue as full
        // visits), the early visits, then the other probes (most end after
        // one sweep); otherwise the early visits, the first-visit probes, the
        // others. The order decides only which workgroup runs an item and when.
        int p, kind;
        const int j = i < nFirst ? nEarly + i : (i < nFirst + nEarly ? i - nFirst : i);
        if (j < nEarly) {
          p = early[j];
          kind = 0;
        } else {
          p = j < nEarly + nNew ? ovN[sP * P + (j - nEarly)] : ovF[sP * P + (j - nEarly - nNew)];
          // an unconverged member of E-1 is one of E's early members
          if (members > 0 && vprev[p] == members && !bprev[p]) continue;
          if (tid == 0) s_acc[kAccProbes] += 1ull;
          kind = a.niter > 1 ? 1 : 2;
        }
        SweepResult<T> r;
        for (;;) {
          r = sweep_patch<T, NC, BS, ORG, STG>(m, s, p, kind == 2 ? 1 : a.niter, kind == 2 ? 0 : 1, a.tol, a.leps,
                                               smem, red, s_ndirty, 0, 0, members,
                                               SweepIO<T>{dprev, vprev, dcur, bcur, kind == 1 ? 1 : 0},
                                               s_acc, a.stage);
          if (kind == 2 && !r.con) {
            kind = 0;
            continue;
          }
          break;
        }
        if (!r.visited) continue;
        // p is a member of E: mark its neighbours for probes when it
        // converged below the band, keep it for E+1 when it did not, and
        // copy it back in E+1's phase.
        if (r.con && r.bmin < band_of(m, p, a.bw)) {
          for (int k = m.padj_off[p] + tid; k < m.padj_off[p + 1]; k += BS) {
            const int q = m.padj[k];
            if (atomicMax(&s.freshE[q], E) < E) {
              // visit tags are zero until the patch's first visit in the step
              // (a tag being written concurrently only changes the list)
              if (s.visitE[q] == 0 && s.visitE2[q] == 0) ovN[sE * P + atomicAdd(&s.pc[kPcOvFn + sE], 1)] = q;
              else ovF[sE * P + atomicAdd(&s.pc[kPcOvF + sE], 1)] = q;
            }
          }
        }
        if (tid == 0) {
          vcur[p] = E;
          ovM[sE * P + atomicAdd(&s.pc[kPcOvMl + sE], 1)] = p;
          if (!r.con) ovE[sN * P + atomicAdd(&s.pc[kPcOvEl + sN], 1)] = p;
          atomicAdd(&s.pc[kPcOvVis + sE], 1);
          s_acc[kAccVisits] += 1ull;
        }
      }
      grid_barrier(s.bar, G);
      pcl.lap(clk, kTkVisit);
    }
  }

  if (!warm) break;
  // Phase 2b: acceptance over the patches that received warm values (the
  // previous narrowband); a mesh with a patch above the threshold rejects its
  // warm start and re-runs plain.
  if constexpr (WARM)
    for (int i = b; i < nscan; i += G) {
      const int p = s.nb[a.nb_in][i];
      residual_patch<T, NC, BS>(m, s, p, band_of(m, p, a.bw), a.leps, a.warm, a.warm_margin, a.warm_abs,
                                a.res_tol, smem, red);
    }
  grid_barrier(s.bar, G);
  const float resid = __uint_as_float((unsigned int)ld_acquire(&s.pc[kPcResid]));
  if (resid <= a.res_tol) break;
  }

  // Phase 3: narrowband + sign restore over every patch.
  for (int p = b; p < m.P; p += G) {
    const T mn = narrowband_patch<T, BS>(m, s, p, red);
    if (tid == 0 && mn <= band_of(m, p, a.bw)) {
      s.nb[nb_out][atomicAdd(&s.pc[kPcNb + nb_out], 1)] = p;
      atomicAdd(&s.pc[kPcNbVerts], m.vert_off[p + 1] - m.vert_off[p]);
    }
  }
  grid_barrier(s.bar, G);
  if (a.cfl) {
    // Phase 4: CFL bound over the new narrowband, read by the advection launches.
    const int nnb = ld_acquire(&s.pc[kPcNb + nb_out]);
    for (int i = b; i < nnb; i += G) cfl_patch<T, NC, BS>(m, s, s.nb[nb_out][i], cred);
    grid_barrier(s.bar, G);
  }
  pcl.lap(clk, kTkNb);

  // Advection (or MCF) over the new narrowband follows as its own launches
  // (k_advect_dev, k_copy_phi_dev), so this kernel's LDS is the sweep's.
  pcl.lap(clk, kTkAdvect);
  if (clk)
    for (int i = 0; i < kNumTicks; ++i) s.ticks[i] = tk[i];
  __syncthreads();
  flush_stats(s, s_acc);
}

}  // namespace kern

using namespace kern;


// ---------------------------------------------------------------------------
// Host driver.

#define LS_HIP(expr)                                                         \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      err = std::string(#expr) + ": " + hipGetErrorString(e_);               \
      return LS_ERR_HIP;                                                     \
    }                                                                        \
  } while (0)

template <typename U>
static hipError_t dev_upload(U** dptr, const U* host, size_t n) {
  if (*dptr) { (void)hipFree(*dptr); *dptr = nullptr; }
  hipError_t e = hipMalloc((void**)dptr, std::max<size_t>(n, 1) * sizeof(U));
  if (e != hipSuccess) return e;
  if (n) e = hipMemcpy(*dptr, host, n * sizeof(U), hipMemcpyHostToDevice);
  return e;
}
// Takes over a device allocation (freeing what *dptr held) and clears the source.
template <typename U>
static void adopt(U** dptr, U*& src) {
  if (*dptr) (void)hipFree(*dptr);
  *dptr = src;
  src = nullptr;
}
template <typename U>
static hipError_t dev_alloc(U** dptr, size_t n) {
  if (*dptr) { (void)hipFree(*dptr); *dptr = nullptr; }
  return hipMalloc((void**)dptr, std::max<size_t>(n, 1) * sizeof(U));
}

// Device address of `scratch` on each GPU, resolved once per device with
// that device current (the array's name in host code is only a stand-in).
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[256] = {};
static hipError_t scratch_address(int device, float** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= 256) return hipErrorInvalidDevice;
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *out = static_cast<float*>(g_scratch_ptr[device]);
  return hipSuccess;
}

template <typename T, int NC>
class EngineImpl final : public Engine {
 public:
  explicit EngineImpl(int device) : device_(device) {}
  ~EngineImpl() override { release(); }

  int working_bytes() const override { return (int)sizeof(T); }
  void set_exec(int mode) override { exec_ = mode; }

  int upload(const Layout& L, DeviceLayout* D, std::string& err) override {
    LS_HIP(hipSetDevice(device_));
    LS_HIP(scratch_address(device_, &scratch_));
    if (!stream_) LS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    if (!ev_[0]) for (auto& e : ev_) LS_HIP(hipEventCreate(&e));
    P_ = L.P; N_ = L.N; F_ = L.F;
    vert_off_h_ = L.vert_off;
    vert_perm_h_ = L.vert_perm;
    ele_off_h_ = L.ele_off;
    max_ep_ = L.max_ep;
    max_vp_ = L.max_vp;
    max_hp_ = L.max_hp;
    int lds_max = 0;
    LS_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device_));
    // Exact per-patch LDS needs.
    fim_lds_ = adv_lds_ = seed_lds_ = 0;
    size_t fim_task_lds = 0, fim_stage_lds = 0, fim_task_stage_lds = 0;
    for (int p = 0; p < P_; ++p) {
      const size_t ne = L.ele_off[p + 1] - L.ele_off[p], nv = L.vert_off[p + 1] - L.vert_off[p];
      const size_t nh = L.halo_off[p + 1] - L.halo_off[p];
      const size_t ns = L.vslot_off[L.vert_off[p + 1]] - L.vslot_off[L.vert_off[p]];
      // (+32: the slot lists' 16-byte alignment shift)
      const size_t fim = ((NC * ne * sizeof(T) + (2 * nv + nh) * sizeof(T) + 2 * (nv + 1 + ns) + 32 + 2 * ne + 3 +
                           ne + 7) & ~size_t(7)) + nv * sizeof(T);  // + running vertex minima
      const size_t fim_tasks = ((fim + 7) & ~size_t(7)) + 4 * (NC * ne + (NC * ne) % 2) + NC * ne * sizeof(T);
      fim_task_lds = std::max(fim_task_lds, fim_tasks);
      // Staged corner ids and heights: 16 bytes per copy after the vertex
      // minima, each array shifted to its source's 16-byte alignment.
      fim_stage_lds = std::max(fim_stage_lds, ((fim + 7) & ~size_t(7)) + 16 * ne + 64);
      fim_task_stage_lds = std::max(fim_task_stage_lds, fim_tasks + 16 * ne + 64);
      // Advection: per-slot products (FP32) or per-element records (FP64);
      // curvature: per-slot terms plus two per-element values.
      const size_t adv = ((sizeof(T) == 8 ? (NC + 2) * ne : 2 * NC * ne) + nv + nh) * sizeof(T);
      fim_lds_ = std::max(fim_lds_, fim);
      adv_lds_ = std::max(adv_lds_, adv);
      seed_lds_ = std::max(seed_lds_, (nv + nh) * sizeof(T));
    }
    // The task-compacted sweep needs more LDS; size_kernels reports it when
    // it does not fit.
    fim_elem_lds_ = (fim_lds_ + 15) & ~size_t(15);
    fim_task_lds_ = (fim_task_lds + 15) & ~size_t(15);
    fim_elem_stage_lds_ = (fim_stage_lds + 15) & ~size_t(15);
    fim_task_stage_lds_ = (fim_task_stage_lds + 15) & ~size_t(15);
    adv_lds_ = (adv_lds_ + 15) & ~size_t(15);
    seed_lds_ = (seed_lds_ + 15) & ~size_t(15);
    lds_max_ = lds_max;
    {
      const int rc = size_kernels(err);
      if (rc != LS_OK) return rc;
    }
    // The renumbered coordinates and corner ids stay resident for the local
    // coordinates and velocity projection.
    if (D) {
      adopt(&d_vert_off_, D->vert_off);
      adopt(&d_ele_off_, D->ele_off);
      adopt(&d_halo_off_, D->halo_off);
      adopt(&d_halo_ids_, D->halo_ids);
      adopt(&d_copy_local_, reinterpret_cast<ushort4*&>(D->copy_local));
      adopt(&d_vslot_off_, D->vslot_off);
      adopt(&d_vslots_, D->vslots);
      adopt(&d_padj_off_, D->padj_off);
      adopt(&d_padj_, D->padj);
      adopt(&d_xyz_, D->xyz);
      adopt(&d_copy_verts_, D->copy_verts);
      adopt(&d_copy_src_, D->copy_src);
    } else {
      LS_HIP(dev_upload(&d_vert_off_, L.vert_off.data(), L.vert_off.size()));
      LS_HIP(dev_upload(&d_ele_off_, L.ele_off.data(), L.ele_off.size()));
      LS_HIP(dev_upload(&d_halo_off_, L.halo_off.data(), L.halo_off.size()));
      LS_HIP(dev_upload(&d_halo_ids_, L.halo_ids.data(), L.halo_ids.size()));
      LS_HIP(dev_upload(&d_copy_local_, reinterpret_cast<const ushort4*>(L.copy_local.data()), (size_t)F_));
      LS_HIP(dev_upload(&d_vslot_off_, L.vslot_off.data(), L.vslot_off.size()));
      LS_HIP(dev_upload(&d_vslots_, L.vslots.data(), L.vslots.size()));
      LS_HIP(dev_upload(&d_padj_off_, L.padj_off.data(), L.padj_off.size()));
      LS_HIP(dev_upload(&d_padj_, L.padj.data(), L.padj.size()));
      LS_HIP(dev_upload(&d_xyz_, L.xyz.data(), L.xyz.size()));
      LS_HIP(dev_upload(&d_copy_verts_, L.copy_verts.data(), L.copy_verts.size()));
      LS_HIP(dev_upload(&d_copy_src_, L.copy_src.data(), L.copy_src.size()));
    }
    {
      constexpr int K = ElemSolve<T, NC>::kCoords;
      LS_HIP(dev_alloc(&d_coords_, (size_t)F_ * K));
      if (F_ > 0)
        k_local_coords<T, NC><<<(int)((F_ + 255) / 256), 256, 0, stream_>>>(d_copy_verts_, d_xyz_, F_, d_coords_);
      LS_HIP(hipGetLastError());
      if (NC == 4 && F_ > 0) {
        LS_HIP(dev_alloc(&d_hgt_, (size_t)F_ * 4));
        int32_t* d_slow = nullptr;
        LS_HIP(dev_alloc(&d_slow, 2));
        hipError_t e = hipMemsetAsync(d_slow, 0, 2 * sizeof(int32_t), stream_);
        if (e == hipSuccess) {
          k_heights<T><<<(int)((F_ + 255) / 256), 256, 0, stream_>>>(d_coords_, F_, d_hgt_, d_slow);
          e = hipGetLastError();
        }
        int32_t slow[2] = {1, 1};
        if (e == hipSuccess) e = hipMemcpyAsync(slow, d_slow, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, stream_);
        if (e == hipSuccess) e = hipStreamSynchronize(stream_);
        (void)hipFree(d_slow);
        LS_HIP(e);
        fast_len_ = slow[0] == 0;
        plane_ok_ = slow[1] == 0;
        LS_HIP(dev_alloc(&d_pscale_, (size_t)P_));
        k_patch_scale<T><<<P_, 256, 0, stream_>>>(d_coords_, d_ele_off_, d_pscale_);
        LS_HIP(hipGetLastError());
        LS_HIP(hipStreamSynchronize(stream_));
      }
    }
    LS_HIP(dev_alloc(&d_rs_slot_, (size_t)F_ * NC));
    LS_HIP(dev_alloc(&d_rs_vval_, (size_t)N_));
    LS_HIP(dev_alloc(&d_rs_dirty_, (size_t)F_));
    LS_HIP(dev_alloc(&d_rs_epoch_, (size_t)P_));
    LS_HIP(hipMemset(d_rs_epoch_, 0, (size_t)P_ * sizeof(int32_t)));
    LS_HIP(dev_alloc(&d_visitE_, (size_t)P_));
    LS_HIP(hipMemset(d_visitE_, 0, (size_t)P_ * sizeof(int32_t)));
    LS_HIP(dev_alloc(&d_ve_ok_, (size_t)F_));
    LS_HIP(dev_alloc(&d_ve_tag_, (size_t)P_));
    LS_HIP(hipMemset(d_ve_tag_, 0, (size_t)P_ * sizeof(int32_t)));
    LS_HIP(dev_alloc(&d_vpatch_, (size_t)N_));
    if (P_ > 0) k_vertex_patch<<<P_, kBlock, 0, stream_>>>(d_vert_off_, P_, d_vpatch_);
    LS_HIP(hipGetLastError());
    {
      int32_t H = 0;
      if (P_ > 0) LS_HIP(hipMemcpy(&H, d_halo_off_ + P_, sizeof(int32_t), hipMemcpyDeviceToHost));
      LS_HIP(dev_alloc(&d_halo_patch_, (size_t)H));
      if (P_ > 0 && H > 0)
        k_halo_patch<<<P_, kBlock, 0, stream_>>>(d_halo_off_, d_halo_ids_, d_vpatch_, P_, d_halo_patch_);
      LS_HIP(hipGetLastError());
      LS_HIP(dev_alloc(&d_phdr_, (size_t)2 * P_));
      if (P_ > 0)
        k_patch_hdr<<<(P_ + kBlock - 1) / kBlock, kBlock, 0, stream_>>>(d_vert_off_, d_ele_off_, d_halo_off_,
                                                                          d_vslot_off_, P_, d_phdr_);
      LS_HIP(hipGetLastError());
    }
    LS_HIP(hipStreamSynchronize(stream_));
    LS_HIP(dev_alloc(&d_sigma_, (size_t)F_ * (NC - 1)));
    LS_HIP(hipMemset(d_sigma_, 0, (size_t)F_ * (NC - 1) * sizeof(T)));
    LS_HIP(dev_alloc(&d_phi_, N_));
    LS_HIP(dev_alloc(&d_phi_out_, N_));
    LS_HIP(dev_alloc(&d_DT_, N_));
    LS_HIP(dev_alloc(&d_DT_out_, N_));
    LS_HIP(dev_alloc(&d_DT_out2_, N_));
    LS_HIP(dev_alloc(&d_visitE2_, (size_t)P_));
    LS_HIP(hipMemset(d_visitE2_, 0, (size_t)P_ * sizeof(int32_t)));
    LS_HIP(dev_alloc(&d_bcon2_, (size_t)P_));
    LS_HIP(dev_alloc(&d_ov_lists_, (size_t)12 * P_));
    LS_HIP(dev_alloc(&d_label_, N_));
    LS_HIP(dev_alloc(&d_tsign_, N_));
    LS_HIP(dev_alloc(&d_bcon_, P_));
    LS_HIP(dev_alloc(&d_bmin_, P_));
    LS_HIP(dev_alloc(&d_act_, P_));
    LS_HIP(dev_alloc(&d_fresh_, P_));
    LS_HIP(dev_alloc(&d_seedflag_, P_));
    LS_HIP(dev_alloc(&d_counters_, kNumCounters));

    LS_HIP(dev_alloc(&d_inA_, P_));
    LS_HIP(dev_alloc(&d_freshE_, P_));
    LS_HIP(dev_alloc(&d_listF_, P_));
    // Phase ticks, size sums, CFL keys and step counters in one block, read
    // back after a persistent step with one copy into pinned host memory.
    LS_HIP(dev_alloc(&d_stat_, kStatWords));
    d_ticks_ = d_stat_;
    d_sizes_ = d_ticks_ + kNumTicks;
    d_cfl_ = d_sizes_ + kNumSizes;
    d_pc_ = reinterpret_cast<int32_t*>(d_cfl_ + 2);
    if (!h_stat_) LS_HIP(hipHostMalloc((void**)&h_stat_, kStatWords * sizeof(unsigned long long), hipHostMallocDefault));
    LS_HIP(dev_alloc(&d_bar_, 2));
    LS_HIP(dev_alloc(&d_gfail_, 1));
    ngroups_ = 1;
    LS_HIP(hipMemset(d_bar_, 0, 2 * sizeof(unsigned int)));
    {
      int khz = 0;
      LS_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device_));
      wall_khz_ = khz > 0 ? khz : 100000;
    }
    LS_HIP(hipMemset(d_pc_, 0, kNumPc * sizeof(int32_t)));
    for (auto& l : d_lists_) LS_HIP(dev_alloc(&l, P_));
    LS_HIP(dev_alloc(&d_nb_, P_));
    LS_HIP(dev_alloc(&d_scan_all_, P_));
    {
      std::vector<int32_t> iota(P_);
      for (int p = 0; p < P_; ++p) iota[p] = p;
      LS_HIP(hipMemcpy(d_scan_all_, iota.data(), P_ * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    LS_HIP(hipMemset(d_act_, 0, P_ * sizeof(int32_t)));
    LS_HIP(hipMemset(d_fresh_, 0, P_ * sizeof(int32_t)));
    LS_HIP(hipMemset(d_phi_, 0, N_ * sizeof(T)));
    nnb_ = 0;
    uploaded_ = true;
    return LS_OK;
  }

  // -- learned warm start ------------------------------------------------
  int mlp_alloc(std::string& err) {
    if (d_mlp_prm_) return LS_OK;
    LS_HIP(dev_alloc(&d_mlp_prm_, (size_t)mlp::kNumParams));
    LS_HIP(dev_alloc(&d_mlp_m_, (size_t)mlp::kNumParams));
    LS_HIP(dev_alloc(&d_mlp_v_, (size_t)mlp::kNumParams));
    LS_HIP(dev_alloc(&d_mlp_grad_, (size_t)mlp::kNumParams + 1));  // + loss
    LS_HIP(dev_alloc(&d_mlp_w1_, (size_t)mlp::kInPad * mlp::kHid));
    LS_HIP(dev_alloc(&d_mlp_w2_, (size_t)mlp::kHid * mlp::kHid));
    LS_HIP(dev_alloc(&d_mlp_w3_, (size_t)mlp::kHid * mlp::kOutPad));
    LS_HIP(dev_alloc(&d_mlp_b_, (size_t)(2 * mlp::kHid + mlp::kOutPad)));
    LS_HIP(hipMemset(d_mlp_m_, 0, mlp::kNumParams * sizeof(float)));
    LS_HIP(hipMemset(d_mlp_v_, 0, mlp::kNumParams * sizeof(float)));
    return LS_OK;
  }
  mlp::Weights mlp_weights() const {
    return mlp::Weights{d_mlp_w1_, d_mlp_w2_, d_mlp_w3_, d_mlp_b_, d_mlp_b_ + mlp::kHid, d_mlp_b_ + 2 * mlp::kHid};
  }
  int mlp_pack(hipStream_t q, std::string& err) {
    k_mlp_pack<<<(mlp::kHid * mlp::kHid + 255) / 256, 256, 0, q>>>(d_mlp_prm_, d_mlp_w1_, d_mlp_w2_, d_mlp_w3_, d_mlp_b_,
                                                                 d_mlp_b_ + mlp::kHid, d_mlp_b_ + 2 * mlp::kHid);
    LS_HIP(hipGetLastError());
    return LS_OK;
  }
  int mlp_set_params(const float* prm, std::string& err) override {
    if (!uploaded_) { err = "layout not built"; return LS_ERR_STATE; }
    if (int rc = mlp_alloc(err)) return rc;
    LS_HIP(hipMemcpy(d_mlp_prm_, prm, mlp::kNumParams * sizeof(float), hipMemcpyHostToDevice));
    LS_HIP(hipMemset(d_mlp_m_, 0, mlp::kNumParams * sizeof(float)));
    LS_HIP(hipMemset(d_mlp_v_, 0, mlp::kNumParams * sizeof(float)));
    adam_t_ = 0;
    if (int rc = mlp_pack(stream_, err)) return rc;
    LS_HIP(hipStreamSynchronize(stream_));
    mlp_ready_ = true;
    return LS_OK;
  }
  int mlp_get_params(float* prm, std::string& err) override {
    if (!mlp_ready_) { err = "no network parameters set"; return LS_ERR_STATE; }
    LS_HIP(hipMemcpy(prm, d_mlp_prm_, mlp::kNumParams * sizeof(float), hipMemcpyDeviceToHost));
    return LS_OK;
  }
  size_t mlp_predict_lds() const {
    return (size_t)(kMlpBlock / 64) * mlp::kWaveTile * sizeof(float) +
           (size_t)(4 * max_ep_ + max_vp_ + max_hp_) * sizeof(float);
  }
  // Warm values for the next step from the current phi over the current
  // narrowband (the next step's previous narrowband).
  int mlp_predict(std::string& err) override {
    if (!mlp_ready_) { err = "no network parameters set"; return LS_ERR_STATE; }
    if (nnb_ == 0) { err = "no narrowband yet: the first step has nothing to warm-start"; return LS_ERR_STATE; }
    if (!d_warm_) {
      LS_HIP(dev_alloc(&d_warm_, (size_t)N_));
      std::vector<T> big(N_, (T)kLargeNum);
      LS_HIP(hipMemcpy(d_warm_, big.data(), N_ * sizeof(T), hipMemcpyHostToDevice));
    }
    k_mlp_predict<T, NC><<<nnb_, kMlpBlock, mlp_predict_lds(), stream_>>>(mesh(), state(), d_nb_cur(), mlp_weights(),
                                                                      d_warm_);
    LS_HIP(hipGetLastError());
    LS_HIP(hipStreamSynchronize(stream_));
    return LS_OK;
  }
  // Training pairs: phase 0 before a step records the inputs of every copy
  // of the current narrowband; phase 1 after the step records their targets.
  int mlp_collect(int phase, int64_t capacity, std::string& err) override {
    if (phase == 0) {
      if (nnb_ == 0) { err = "no narrowband yet"; return LS_ERR_STATE; }
      if (!d_ds_x_) {
        if (capacity <= 0) { err = "dataset capacity must be > 0"; return LS_ERR_ARG; }
        LS_HIP(dev_alloc(&d_ds_x_, (size_t)capacity * mlp::kInPad));
        LS_HIP(dev_alloc(&d_ds_y_, (size_t)capacity * mlp::kOut));
        ds_cap_ = capacity;
        ds_n_ = 0;
      }
      std::vector<int32_t> nb(nnb_);
      LS_HIP(hipMemcpy(nb.data(), d_nb_cur(), nnb_ * sizeof(int32_t), hipMemcpyDeviceToHost));
      std::sort(nb.begin(), nb.end());
      std::vector<int32_t> copies;
      for (int32_t p : nb)
        for (int32_t c = ele_off_h_[p]; c < ele_off_h_[p + 1]; ++c) copies.push_back(c);
      const int64_t n = std::min<int64_t>((int64_t)copies.size(), ds_cap_ - ds_n_);
      pend_n_ = n;
      if (n <= 0) return LS_OK;
      if (pend_cap_ < n) {
        if (d_pend_) LS_HIP(hipFree(d_pend_));
        d_pend_ = nullptr;
        LS_HIP(dev_alloc(&d_pend_, (size_t)n));
        pend_cap_ = n;
      }
      LS_HIP(hipMemcpy(d_pend_, copies.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
      k_mlp_collect_x<T, NC><<<(int)((n + kBlock - 1) / kBlock), kBlock, 0, stream_>>>(
          mesh(), state(), d_pend_, n, d_ds_x_ + ds_n_ * mlp::kInPad);
      LS_HIP(hipGetLastError());
      LS_HIP(hipStreamSynchronize(stream_));
      return LS_OK;
    }
    if (pend_n_ <= 0) return LS_OK;
    k_mlp_collect_y<T, NC><<<(int)((pend_n_ + kBlock - 1) / kBlock), kBlock, 0, stream_>>>(
        mesh(), state(), d_pend_, pend_n_, d_ds_x_ + ds_n_ * mlp::kInPad, d_ds_y_ + ds_n_ * mlp::kOut);
    LS_HIP(hipGetLastError());
    LS_HIP(hipStreamSynchronize(stream_));
    ds_n_ += pend_n_;
    pend_n_ = 0;
    return LS_OK;
  }
  int64_t mlp_dataset_size() const override { return ds_n_; }
  int mlp_get_dataset(float* X, float* Y, std::string& err) override {
    if (ds_n_ == 0) return LS_OK;
    std::vector<float> xp((size_t)ds_n_ * mlp::kInPad);
    LS_HIP(hipMemcpy(xp.data(), d_ds_x_, xp.size() * sizeof(float), hipMemcpyDeviceToHost));
    if (X)
      for (int64_t i = 0; i < ds_n_; ++i) std::memcpy(X + i * mlp::kIn, &xp[i * mlp::kInPad], mlp::kIn * sizeof(float));
    if (Y) LS_HIP(hipMemcpy(Y, d_ds_y_, (size_t)ds_n_ * mlp::kOut * sizeof(float), hipMemcpyDeviceToHost));
    return LS_OK;
  }
  size_t mlp_grad_lds() const {
    return (size_t)(kMlpBlock / 64) * mlp::kWaveTile * sizeof(float) + (size_t)mlp::kNumParams * sizeof(float);
  }
  // Gradient sums into g[0, kNumParams) and the loss sum into g[kNumParams]
  // (the context's accumulator when g is null).
  int mlp_grad_launch(const float* X, const float* Y, const int32_t* idx, int64_t n, float inv_total, hipStream_t q,
                      std::string& err, float* g = nullptr) {
    if (!g) g = d_mlp_grad_;
    const int64_t tiles = (n + 15) / 16;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, 1024));
    k_mlp_grad<<<grid, kMlpBlock, mlp_grad_lds(), q>>>(X, Y, idx, n, inv_total, mlp_weights(), g,
                                                       g + mlp::kNumParams);
    LS_HIP(hipGetLastError());
    return LS_OK;
  }
  // Loss sum and the gradient of the mean loss over n caller samples.
  int mlp_gradient(const float* X, const float* Y, int64_t n, float* grad, double* loss, std::string& err) override {
    if (!mlp_ready_) { err = "no network parameters set"; return LS_ERR_STATE; }
    if (n <= 0) { err = "n must be > 0"; return LS_ERR_ARG; }
    std::vector<float> xp((size_t)n * mlp::kInPad, 0.f);
    for (int64_t i = 0; i < n; ++i) std::memcpy(&xp[i * mlp::kInPad], X + i * mlp::kIn, mlp::kIn * sizeof(float));
    float *dx = nullptr, *dy = nullptr;
    LS_HIP(dev_alloc(&dx, xp.size()));
    hipError_t e = dev_alloc(&dy, (size_t)n * mlp::kOut);
    if (e == hipSuccess) e = hipMemcpy(dx, xp.data(), xp.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dy, Y, (size_t)n * mlp::kOut * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d_mlp_grad_, 0, (mlp::kNumParams + 1) * sizeof(float));
    int rc = LS_OK;
    if (e == hipSuccess) rc = mlp_grad_launch(dx, dy, nullptr, n, 1.0f / (float)n, stream_, err);
    if (e == hipSuccess && rc == LS_OK) e = hipStreamSynchronize(stream_);
    std::vector<float> g(mlp::kNumParams + 1);
    if (e == hipSuccess && rc == LS_OK)
      e = hipMemcpy(g.data(), d_mlp_grad_, g.size() * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (rc != LS_OK) return rc;
    LS_HIP(e);
    std::memcpy(grad, g.data(), mlp::kNumParams * sizeof(float));
    *loss = g[mlp::kNumParams];
    return LS_OK;
  }
  // The same loss and gradient with the samples staged in `scratch` and the
  // sums accumulated there: no device allocation. The accumulator is cleared
  // by a kernel, so the sequence stays graph-capturable.
  int mlp_gradient_noalloc(const float* X, const float* Y, int64_t n, float* grad, double* loss,
                           std::string& err) override {
    if (!mlp_ready_) { err = "no network parameters set"; return LS_ERR_STATE; }
    if (n <= 0) { err = "n must be > 0"; return LS_ERR_ARG; }
    if (n > kScratchSamples) {
      err = "n = " + std::to_string(n) + " exceeds the " + std::to_string(kScratchSamples) +
            " samples the allocation-free gradient's scratch holds";
      return LS_ERR_LIMIT;
    }
    if (!scratch_) { err = "scratch not resolved"; return LS_ERR_STATE; }
    float* dx = scratch_;
    float* dy = dx + kScratchX;
    float* dg = dy + kScratchY;
    hipStream_t q = stream_;
    k_zero_f<<<(int)((n * mlp::kInPad + 255) / 256), 256, 0, q>>>(dx, n * mlp::kInPad);  // padding columns
    k_zero_f<<<(int)((kScratchG + 255) / 256), 256, 0, q>>>(dg, (int64_t)kScratchG);
    LS_HIP(hipGetLastError());
    LS_HIP(hipMemcpy2DAsync(dx, mlp::kInPad * sizeof(float), X, mlp::kIn * sizeof(float), mlp::kIn * sizeof(float),
                            (size_t)n, hipMemcpyHostToDevice, q));
    LS_HIP(hipMemcpyAsync(dy, Y, (size_t)n * mlp::kOut * sizeof(float), hipMemcpyHostToDevice, q));
    if (int rc = mlp_grad_launch(dx, dy, nullptr, n, 1.0f / (float)n, q, err, dg)) return rc;
    float lsum = 0.f;
    LS_HIP(hipMemcpyAsync(grad, dg, mlp::kNumParams * sizeof(float), hipMemcpyDeviceToHost, q));
    LS_HIP(hipMemcpyAsync(&lsum, dg + mlp::kNumParams, sizeof(float), hipMemcpyDeviceToHost, q));
    LS_HIP(hipStreamSynchronize(q));
    *loss = lsum;
    return LS_OK;
  }

  // Adam over the collected pairs, shuffled every epoch; returns the mean
  // squared error per output of the last epoch.
  int mlp_train(int epochs, int batch, double lr, uint64_t seed, double* loss, std::string& err) override {
    if (!mlp_ready_) { err = "no network parameters set"; return LS_ERR_STATE; }
    if (ds_n_ == 0) { err = "no training pairs collected"; return LS_ERR_STATE; }
    if (epochs <= 0 || batch <= 0) { err = "epochs and batch must be > 0"; return LS_ERR_ARG; }
    int32_t* didx = nullptr;
    LS_HIP(dev_alloc(&didx, (size_t)ds_n_));
    std::vector<int32_t> perm(ds_n_);
    for (int64_t i = 0; i < ds_n_; ++i) perm[i] = (int32_t)i;
    uint64_t st = seed ? seed : 0x9e3779b97f4a7c15ull;
    auto rnd = [&st]() {  // splitmix64
      uint64_t z = (st += 0x9e3779b97f4a7c15ull);
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      return z ^ (z >> 31);
    };
    const float b1 = 0.9f, b2 = 0.999f;
    double last = 0.0;
    hipError_t e = hipSuccess;
    int rc = LS_OK;
    for (int ep = 0; ep < epochs && e == hipSuccess && rc == LS_OK; ++ep) {
      for (int64_t i = ds_n_ - 1; i > 0; --i) std::swap(perm[i], perm[(int64_t)(rnd() % (uint64_t)(i + 1))]);
      e = hipMemcpy(didx, perm.data(), ds_n_ * sizeof(int32_t), hipMemcpyHostToDevice);
      double sum = 0.0;
      for (int64_t b0 = 0; b0 < ds_n_ && e == hipSuccess && rc == LS_OK; b0 += batch) {
        const int64_t nb = std::min<int64_t>(batch, ds_n_ - b0);
        e = hipMemsetAsync(d_mlp_grad_, 0, (mlp::kNumParams + 1) * sizeof(float), stream_);
        if (e != hipSuccess) break;
        rc = mlp_grad_launch(d_ds_x_, d_ds_y_, didx + b0, nb, 1.0f / (float)nb, stream_, err);
        if (rc != LS_OK) break;
        ++adam_t_;
        const float c1 = 1.f - std::pow(b1, (float)adam_t_), c2 = 1.f - std::pow(b2, (float)adam_t_);
        k_mlp_adam<<<(mlp::kNumParams + 255) / 256, 256, 0, stream_>>>(d_mlp_prm_, d_mlp_m_, d_mlp_v_, d_mlp_grad_,
                                                                      mlp::kNumParams, (float)lr, b1, b2, 1e-8f, c1, c2);
        rc = mlp_pack(stream_, err);
        float l = 0.f;
        if (rc == LS_OK) e = hipMemcpyAsync(&l, d_mlp_grad_ + mlp::kNumParams, sizeof(float), hipMemcpyDeviceToHost, stream_);
        if (e == hipSuccess) e = hipStreamSynchronize(stream_);
        sum += l;
      }
      last = sum / (double)(ds_n_ * mlp::kOut);
    }
    (void)hipFree(didx);
    if (rc != LS_OK) return rc;
    LS_HIP(e);
    *loss = last;
    return LS_OK;
  }
  // Unsigned distances after the last step's redistancing (caller order).
  int get_distances(double* out, std::string& err) override {
    std::vector<T> h(N_);
    LS_HIP(hipMemcpy(h.data(), d_DT_, N_ * sizeof(T), hipMemcpyDeviceToHost));
    for (int64_t v = 0; v < N_; ++v) out[vert_perm_h_[v]] = (double)h[v];
    return LS_OK;
  }

  // Caller predictor distances for LS_WARM_LEARNED steps (caller's vertex
  // order), rounded to the working type as ls_mlp_predict's values are.
  int set_warm(const double* warm_user, std::string& err) override {
    if (!uploaded_) { err = "layout not built"; return LS_ERR_STATE; }
    if (!d_warm_) LS_HIP(dev_alloc(&d_warm_, (size_t)N_));
    std::vector<T> h(N_);
    for (int64_t v = 0; v < N_; ++v) h[v] = (T)warm_user[vert_perm_h_[v]];
    LS_HIP(hipMemcpy(d_warm_, h.data(), N_ * sizeof(T), hipMemcpyHostToDevice));
    return LS_OK;
  }

  int get_warm(double* out_user, std::string& err) override {
    if (!d_warm_) { err = "no warm values (ls_mlp_predict or ls_set_warm)"; return LS_ERR_STATE; }
    std::vector<T> h(N_);
    LS_HIP(hipMemcpy(h.data(), d_warm_, N_ * sizeof(T), hipMemcpyDeviceToHost));
    for (int64_t v = 0; v < N_; ++v) out_user[vert_perm_h_[v]] = (double)h[v];
    return LS_OK;
  }

  int set_phi(const double* phi_user, std::string& err) override {
    if (!uploaded_) { err = "layout not built"; return LS_ERR_STATE; }
    std::vector<T> h(N_);
    for (int64_t v = 0; v < N_; ++v) h[v] = (T)phi_user[vert_perm_h_[v]];
    LS_HIP(hipMemcpy(d_phi_, h.data(), N_ * sizeof(T), hipMemcpyHostToDevice));
    nnb_ = 0;  // a new initial condition seeds from every patch
    return LS_OK;
  }

  int set_velocity(const double* vel, int64_t E, int mode, std::string& err) override {
    if (!uploaded_) { err = "layout not built"; return LS_ERR_STATE; }
    sigma_phi_ = NC == 3 && mode == LS_VEL_PHI_2D;
    const int patch_sigma = mode == LS_VEL_PATCH_FIRST;
    if (F_ == 0) return LS_OK;
    double* d_vel = nullptr;
    LS_HIP(dev_upload(&d_vel, vel, (size_t)E * 3));
    k_project_velocity<T, NC><<<(int)((F_ + 255) / 256), 256, 0, stream_>>>(
        d_copy_verts_, d_xyz_, d_copy_src_, d_ele_off_, P_, F_, d_vel, patch_sigma, d_sigma_);
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipStreamSynchronize(stream_);
    (void)hipFree(d_vel);
    LS_HIP(le);
    LS_HIP(se);
    return LS_OK;
  }

  // LDS of the FIM sweep for the sweep mode, and the persistent grid size.
  int size_kernels(std::string& err) {
    const int org_mode = sweep_mode_ & ~(LS_SWEEP_UNSTAGED | LS_SWEEP_NO_PLANE_BOUND);
    const bool fits = fim_task_lds_ <= (size_t)lds_max_;
    if (org_mode == LS_SWEEP_TASKS && !fits) {
      err = "the task-compacted sweep needs " + std::to_string(fim_task_lds_) + " B of LDS, device allows " +
            std::to_string(lds_max_) + " B";
      return LS_ERR_LIMIT;
    }
    // AUTO runs per-element sweeps: measured faster on MI355X than the
    // task-compacted sweep for 256-vertex patches, whose extra listing pass
    // and per-task reloads of the element geometry outweigh the divergence
    // they remove.
    org_ = org_mode == LS_SWEEP_TASKS ? kOrgTasks : kOrgElements;
    fim_lds_ = org_ == kOrgTasks ? fim_task_lds_ : fim_elem_lds_;
    stage_ = 0;
    step_lds_ = std::max(fim_lds_, seed_lds_);
    step_lds_ = (step_lds_ + 15) & ~size_t(15);
    const size_t need = std::max(step_lds_, adv_lds_);
    if (need > (size_t)lds_max_) {
      err = "patch needs " + std::to_string(need) + " B of LDS, device allows " + std::to_string(lds_max_) +
            " B; use smaller patches";
      return LS_ERR_LIMIT;
    }
    int cus = 0, per_cu = 0;
    LS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_));
    per_cu = 1 << 30;
    step_one_per_cu_ = false;
    for (int warm = 0; warm < 2; ++warm) {
      int n = 0;
      LS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, step_kernel(org_, warm, 0), kFimBlock, step_lds_));
      per_cu = std::min(per_cu, n);
    }
    // Two workgroups per CU only pay when patches outnumber CUs several times
    // over; with fewer, workgroups sharing a CU would take patches while other
    // CUs idle (64^3 torus, 1024 patches: 5.0 ms per step at two per CU). One
    // per CU runs the variant compiled for it.
    step_one_per_cu_ = !(P_ >= 8 * cus && per_cu > 1);
    if (step_one_per_cu_) {
      for (int warm = 0; warm < 2; ++warm) {
        int n = 0;
        LS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, step_kernel(org_, warm, 0), step_block(), step_lds_));
        if (n < 1) per_cu = 0;
      }
    }
    // Sweeps stage each copy's corner ids and heights in LDS (16 bytes per
    // copy) unless that lowers the persistent grid's workgroups per CU: the
    // need-mask and solve passes of every sweep then read them from LDS, not
    // again from L2 or HBM.
    if (!(sweep_mode_ & LS_SWEEP_UNSTAGED) && per_cu > 0) {
      const size_t fim_s = org_ == kOrgTasks ? fim_task_stage_lds_ : fim_elem_stage_lds_;
      const size_t step_s = (std::max(fim_s, seed_lds_) + 15) & ~size_t(15);
      bool ok = step_s <= (size_t)lds_max_;
      const int want = step_one_per_cu_ ? 1 : per_cu;
      for (int warm = 0; warm < 2 && ok; ++warm) {
        int n = 0;
        LS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, step_kernel(org_, warm, 1), step_block(), step_s));
        ok = n >= want;
      }
      if (ok) {
        stage_ = 1;
        fim_lds_ = fim_s;
        step_lds_ = step_s;
      }
    }
    coop_grid_ = cus * (step_one_per_cu_ ? std::min(per_cu, 1) : per_cu);
    cus_ = cus;
    adv_grid_ = 4 * cus;
    if (coop_grid_ <= 0) { err = "persistent step kernel does not fit on a CU"; return LS_ERR_LIMIT; }
    if (graph_exec_) { LS_HIP(hipGraphExecDestroy(graph_exec_)); graph_exec_ = nullptr; }
    if (graph_) { LS_HIP(hipGraphDestroy(graph_)); graph_ = nullptr; }
    return LS_OK;
  }
  // Threads per workgroup of the persistent kernel step_kernel() returns.
  int step_block() const { return (step_one_per_cu_ && org_ != kOrgTasks) ? kStepBlock1 : kFimBlock; }
  // The persistent kernel for the sweep organisation, warm-start support,
  // staging and workgroups per CU chosen by size_kernels (the tasks
  // organisation has only the two-per-CU build). The two-per-CU elements
  // build is compiled once per staging choice: at 128 VGPRs, the registers
  // the unused path holds would be spilled.
  const void* step_kernel(int org, int warm, int stage) const {
    if (org == kOrgTasks)
      return warm ? (const void*)k_step<T, NC, kFimBlock, kOrgTasks, true, 4>
                  : (const void*)k_step<T, NC, kFimBlock, kOrgTasks, false, 4>;
    if (step_one_per_cu_)
      return warm ? (const void*)k_step<T, NC, kStepBlock1, kOrgElements, true, 3>
                  : (const void*)k_step<T, NC, kStepBlock1, kOrgElements, false, 3>;
    if (stage)
      return warm ? (const void*)k_step<T, NC, kFimBlock, kOrgElements, true, 4, 1>
                  : (const void*)k_step<T, NC, kFimBlock, kOrgElements, false, 4, 1>;
    return warm ? (const void*)k_step<T, NC, kFimBlock, kOrgElements, true, 4, 0>
                : (const void*)k_step<T, NC, kFimBlock, kOrgElements, false, 4, 0>;
  }
  int set_sweep_mode(int mode, std::string& err) override {
    const int old = sweep_mode_;
    sweep_mode_ = mode;
    if (!uploaded_) return LS_OK;
    const int rc = size_kernels(err);
    if (rc != LS_OK) {
      sweep_mode_ = old;
      std::string e2;
      (void)size_kernels(e2);
    }
    return rc;
  }
  int sweep_tasks() const override { return org_ == kOrgTasks ? 1 : 0; }
  // The sweeps apply the supporting-plane bound (tets, sizes allow, not switched off).
  bool plane_on() const { return NC == 4 && plane_ok_ && !(sweep_mode_ & LS_SWEEP_NO_PLANE_BOUND); }
  void launch_sweep(int n, hipStream_t q, const DevMesh<T>& m, const DevState<T>& s, const int32_t* list, int niter,
                    int write_out, T tol, double leps, int epoch_in, int epoch_out) {
    if (org_ == kOrgTasks)
      k_sweep<T, NC, kFimBlock, kOrgTasks><<<n, kFimBlock, fim_lds_, q>>>(m, s, list, niter, write_out, tol, leps,
                                                                         epoch_in, epoch_out, stage_);
    else
      k_sweep<T, NC, kFimBlock, kOrgElements><<<n, kFimBlock, fim_lds_, q>>>(m, s, list, niter, write_out, tol, leps,
                                                                            epoch_in, epoch_out, stage_);
  }

  // CFL coefficients per original element (null: the uniform defaults).
  int set_cfl_coefficients(const double* ceik, const double* ccurv, int64_t E, std::string& err) override {
    if (!uploaded_) { err = "layout not built"; return LS_ERR_STATE; }
    if (ceik) LS_HIP(dev_upload(&d_ceik_, ceik, (size_t)E));
    else if (d_ceik_) { LS_HIP(hipFree(d_ceik_)); d_ceik_ = nullptr; }
    if (ccurv) LS_HIP(dev_upload(&d_ccurv_, ccurv, (size_t)E));
    else if (d_ccurv_) { LS_HIP(hipFree(d_ccurv_)); d_ccurv_ = nullptr; }
    return LS_OK;
  }
  // CFL bound over the current narrowband (every patch before the first step).
  int compute_dt(double* dt, double* max_rin, std::string& err) override {
    if (!uploaded_) { err = "layout not built"; return LS_ERR_STATE; }
    const int n = nnb_ > 0 ? nnb_ : P_;
    k_cfl_init<<<1, 64, 0, stream_>>>(d_cfl_);
    if (n > 0) k_cfl<T, NC><<<n, kBlock, 0, stream_>>>(mesh(), state(), nnb_ > 0 ? d_nb_cur() : nullptr);
    LS_HIP(hipGetLastError());
    unsigned long long k[2];
    LS_HIP(hipMemcpyAsync(k, d_cfl_, sizeof(k), hipMemcpyDeviceToHost, stream_));
    LS_HIP(hipStreamSynchronize(stream_));
    *dt = key_value(k[0]);
    *max_rin = key_value(k[1]);
    return LS_OK;
  }
  static double key_value(unsigned long long k) {
    const unsigned long long b = k ^ ((k >> 63) ? (1ull << 63) : ~0ull);
    double d;
    std::memcpy(&d, &b, sizeof(d));
    return d;
  }

  int set_patch_groups(const std::vector<int32_t>& group, int ngroups, std::string& err) override {
    if (!uploaded_) { err = "layout not built"; return LS_ERR_STATE; }
    if (ngroups <= 1) {
      ngroups_ = 1;
      return LS_OK;
    }
    if (group.size() != (size_t)P_) { err = "patch groups need P values"; return LS_ERR_ARG; }
    for (int32_t g : group)
      if (g < 0 || g >= ngroups) { err = "patch group out of range"; return LS_ERR_ARG; }
    if (d_pgroup_) LS_HIP(hipFree(d_pgroup_));
    if (d_gfail_) LS_HIP(hipFree(d_gfail_));
    d_pgroup_ = nullptr;
    d_gfail_ = nullptr;
    LS_HIP(dev_upload(&d_pgroup_, group.data(), group.size()));
    LS_HIP(dev_alloc(&d_gfail_, (size_t)ngroups));
    LS_HIP(hipMemset(d_gfail_, 0, (size_t)ngroups * sizeof(int32_t)));
    ngroups_ = ngroups;
    return LS_OK;
  }
  int set_patch_params(const std::vector<double>& dt, const std::vector<double>& bw, std::string& err) override {
    if (!uploaded_) { err = "layout not built"; return LS_ERR_STATE; }
    if (dt.empty() && bw.empty()) {
      use_patch_prm_ = false;
      return LS_OK;
    }
    if (dt.size() != (size_t)P_ || bw.size() != (size_t)P_) { err = "per-patch parameters need P values"; return LS_ERR_ARG; }
    if (!d_patch_dt_) {
      LS_HIP(dev_alloc(&d_patch_dt_, (size_t)P_));
      LS_HIP(dev_alloc(&d_patch_bw_, (size_t)P_));
    }
    std::vector<T> hb(bw.size());
    for (size_t i = 0; i < bw.size(); ++i) hb[i] = (T)bw[i];
    LS_HIP(hipMemcpy(d_patch_dt_, dt.data(), dt.size() * sizeof(double), hipMemcpyHostToDevice));
    LS_HIP(hipMemcpy(d_patch_bw_, hb.data(), hb.size() * sizeof(T), hipMemcpyHostToDevice));
    use_patch_prm_ = true;
    return LS_OK;
  }

  int step(const ls_step_params& prm, ls_stats& st, std::string& err) override {
    if (!uploaded_) { err = "layout not built"; return LS_ERR_STATE; }
    if (exec_ != LS_EXEC_KERNELS && coop_ok_) return step_persistent(prm, st, err);
    return step_kernels(prm, st, err);
  }

  // A cooperative launch is refused while another process holds the
  // device's cooperative queue. The step then runs, and later steps keep
  // running, on the per-phase path, which computes the same results;
  // ls_stats.exec_mode reports the path taken.
  bool coop_refused(hipError_t e) {
    if (e != hipErrorIllegalState && e != hipErrorCooperativeLaunchTooLarge) return false;
    (void)hipGetLastError();
    coop_ok_ = false;
    return true;
  }

  // One cooperative launch per step (optionally replayed from a hipGraph).
  // Advection passes over the narrowband the persistent step just wrote
  // (list nb[1 - nb_sel_], length in the step counters).
  hipError_t launch_nb_advect(const ls_step_params& prm, hipStream_t q) {
    const int nb_out = 1 - nb_sel_;
    const int32_t* nb = nb_out == 0 ? d_nb_ : d_lists_[3];
    const int32_t* cnt = d_pc_ + kPcNb + nb_out;
    const DevMesh<T> m = mesh();
    const DevState<T> s = state();
    for (int pass = 0; pass < prm.inside_niter; ++pass) {
      k_advect_dev<T, NC><<<adv_grid_, kBlock, adv_lds_, q>>>(m, s, nb, cnt, prm.dt, prm.mode, prm.cfl);
      k_copy_phi_dev<T><<<adv_grid_, kBlock, 0, q>>>(m, s, nb, cnt);
    }
    return hipGetLastError();
  }

  int step_persistent(const ls_step_params& prm, ls_stats& st, std::string& err) {
    if (prm.niter <= 0 || prm.inside_niter < 0) { err = "niter must be > 0, inside_niter >= 0"; return LS_ERR_ARG; }
    if (prm.mode != LS_MODE_ADVECT && prm.mode != LS_MODE_MCF) {
      err = "mode must be LS_MODE_ADVECT or LS_MODE_MCF";
      return LS_ERR_ARG;
    }
    std::memset(&st, 0, sizeof(st));
    st.exec_mode = exec_;
    st.sweep_tasks = org_ == kOrgTasks ? 1 : 0;
    st.sweep_staged = stage_;
    st.fast_len_roots = fast_len_ ? 1 : 0;
    st.plane_bound = plane_on() ? 1 : 0;
    st.step_wg_per_cu = step_one_per_cu_ ? 1 : coop_grid_ / std::max(1, cus_);
    DevMesh<T> m = mesh();
    DevState<T> s = state();
    StepArgs<T> a;
    a.first = nnb_ > 0 ? 0 : 1;
    a.nb_in = nb_sel_;
    a.niter = prm.niter;
    a.inside_niter = prm.inside_niter;
    // The device FIM loop always carries a cap so a step cannot run unbounded.
    a.max_iters = prm.max_fim_iters > 0 ? prm.max_fim_iters : kFimIterCap;
    a.tol = (T)prm.tol;
    a.bw = (T)prm.bandwidth;
    a.leps = prm.lambda_eps;
    a.dt = prm.dt;
    a.mode = prm.mode;
    a.warm = prm.warm;
    a.warm_margin = prm.warm_margin;
    a.warm_abs = prm.warm_abs_margin;
    a.cfl = prm.cfl;
    a.stage = stage_;
    a.res_tol = (float)(prm.warm_res_tol > 0 ? prm.warm_res_tol : prm.tol);
    if (a.warm == LS_WARM_LEARNED && !a.first && !d_warm_) {
      err = "no learned warm-start values (run ls_mlp_predict)";
      return LS_ERR_STATE;
    }
    if (a.warm < LS_WARM_NONE || a.warm > LS_WARM_LEARNED) { err = "bad warm mode"; return LS_ERR_ARG; }
    if (int rc = check_cfl(prm, err)) return rc;
    hipStream_t q = stream_;
    void* args[] = {&m, &s, &a};
    if (!pc_synced_) {  // the previous narrowband size, when the host-driven path ran last
      pc_nb_host_ = nnb_;
      LS_HIP(hipMemcpyAsync(d_pc_ + kPcNb + nb_sel_, &pc_nb_host_, sizeof(int32_t), hipMemcpyHostToDevice, q));
      pc_synced_ = true;
    }
    LS_HIP(hipEventRecord(ev_[0], q));
    if (exec_ == LS_EXEC_GRAPH) {
      // The captured arguments hold device pointers: re-capture when one that
      // may be null (per-patch parameters, warm values) appears.
      const int key[3] = {a.first, a.nb_in, (use_patch_prm_ ? 2 : 1) + (d_warm_ ? 4 : 0)};
      const bool same = graph_exec_ && key[0] == graph_key_[0] && key[1] == graph_key_[1] &&
                        key[2] == graph_key_[2] && std::memcmp(&graph_prm_, &prm, sizeof(prm)) == 0;
      if (!same) {
        if (graph_exec_) { LS_HIP(hipGraphExecDestroy(graph_exec_)); graph_exec_ = nullptr; }
        if (graph_) { LS_HIP(hipGraphDestroy(graph_)); graph_ = nullptr; }
        LS_HIP(hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal));
        hipError_t le = hipLaunchCooperativeKernel(step_kernel(org_, a.warm != 0, stage_), dim3(coop_grid_),
                                                   dim3(step_block()), args, step_lds_, q);
        if (le == hipSuccess) le = launch_nb_advect(prm, q);
        hipError_t ce = hipStreamEndCapture(q, &graph_);
        if (le != hipSuccess || ce != hipSuccess) {
          err = std::string("hipGraph capture of the step kernel failed: ") +
                hipGetErrorString(le != hipSuccess ? le : ce);
          return LS_ERR_HIP;
        }
        LS_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
        graph_key_[0] = key[0]; graph_key_[1] = key[1]; graph_key_[2] = key[2];
        graph_prm_ = prm;
      }
      const hipError_t ge = hipGraphLaunch(graph_exec_, q);
      if (coop_refused(ge)) return step_kernels(prm, st, err);
      LS_HIP(ge);
    } else {
      const hipError_t le = hipLaunchCooperativeKernel(step_kernel(org_, a.warm != 0, stage_), dim3(coop_grid_),
                                                       dim3(step_block()), args, step_lds_, q);
      if (coop_refused(le)) return step_kernels(prm, st, err);
      LS_HIP(le);
      LS_HIP(launch_nb_advect(prm, q));
    }
    LS_HIP(hipEventRecord(ev_[4], q));
    LS_HIP(hipMemcpyAsync(h_stat_, d_stat_, kStatWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, q));
    LS_HIP(hipStreamSynchronize(q));
    const unsigned long long* tk = h_stat_;
    const unsigned long long* sz = tk + kNumTicks;
    unsigned long long ck[2] = {0, 0};
    if (prm.cfl) { ck[0] = sz[kNumSizes]; ck[1] = sz[kNumSizes + 1]; }
    int32_t pc[kNumPc];
    std::memcpy(pc, sz + kNumSizes + 2, sizeof(pc));
    if (pc[kPcOverflow] && prm.max_fim_iters <= 0) {
      err = "FIM did not converge within " + std::to_string(kFimIterCap) + " iterations";
      return LS_ERR_INTERNAL;
    }
    const int nb_out = 1 - nb_sel_;
    nb_sel_ = nb_out;
    nnb_ = pc[kPcNb + nb_out];
    st.fim_iters = pc[kPcIters];
    st.patch_visits = pc[kPcVisits];
    st.probes = pc[kPcProbes];
    st.n_seed_patches = pc[kPcSeed];
    st.nb_count = nnb_;
    st.nb_verts = pc[kPcNbVerts];
    fill_cfl(prm, ck, st);
    float ms = 0.f;
    LS_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[4]));
    st.ms_total = ms;
    st.ms_step_kernel = ms;
    fill_sizes(st, sz);
    const double to_ms = 1.0 / (double)wall_khz_;
    st.ms_ph_seed = tk[kTkSeed] * to_ms; st.ms_ph_visit = tk[kTkVisit] * to_ms;
    st.ms_ph_copyback = tk[kTkCopyback] * to_ms; st.ms_ph_probe = tk[kTkProbe] * to_ms;
    st.ms_ph_next = tk[kTkNext] * to_ms; st.ms_ph_nb = tk[kTkNb] * to_ms; st.ms_ph_advect = tk[kTkAdvect] * to_ms;
    unsigned int rb = (unsigned int)pc[kPcResid];
    float rf;
    std::memcpy(&rf, &rb, sizeof(rf));
    st.warm_residual = rf;
    st.warm_patches = pc[kPcWarm];
    st.warm_fallback = pc[kPcFallback];
    return LS_OK;
  }

  int check_cfl(const ls_step_params& prm, std::string& err) const {
    if (prm.cfl != 0 && prm.cfl != 1) { err = "cfl must be 0 or 1"; return LS_ERR_ARG; }
    if (prm.cfl && use_patch_prm_) {
      err = "the CFL bound is per mesh: with per-mesh timesteps call ls_compute_dt per mesh instead";
      return LS_ERR_ARG;
    }
    return LS_OK;
  }
  static void fill_cfl(const ls_step_params& prm, const unsigned long long* ck, ls_stats& st) {
    st.dt_used = prm.dt;
    st.cfl_dt = st.cfl_max_rin = 0.0;
    if (prm.cfl && st.nb_count > 0) {
      st.cfl_dt = key_value(ck[0]);
      st.cfl_max_rin = key_value(ck[1]);
      st.dt_used = std::min(prm.dt, st.cfl_dt);
    }
  }

  static void fill_sizes(ls_stats& st, const unsigned long long* sz) {
    st.visit_copies = sz[kSzVisit]; st.visit_verts = sz[kSzVisit + 1]; st.visit_halo = sz[kSzVisit + 2];
    st.probe_copies = sz[kSzProbe]; st.probe_verts = sz[kSzProbe + 1]; st.probe_halo = sz[kSzProbe + 2];
    st.adv_copies = sz[kSzAdvect]; st.adv_verts = sz[kSzAdvect + 1]; st.adv_halo = sz[kSzAdvect + 2];
    st.seed_copies = sz[kSzSeed]; st.seed_verts = sz[kSzSeed + 1]; st.seed_halo = sz[kSzSeed + 2];
    st.elem_updates = sz[kSzElem]; st.local_solves = sz[kSzSolve];
  }

  // One launch per phase; the FIM loop is driven from the host.
  int step_kernels(const ls_step_params& prm, ls_stats& st, std::string& err) {
    if (prm.niter <= 0 || prm.inside_niter < 0) { err = "niter must be > 0, inside_niter >= 0"; return LS_ERR_ARG; }
    if (int rc = check_cfl(prm, err)) return rc;
    if (prm.mode != LS_MODE_ADVECT && prm.mode != LS_MODE_MCF) {
      err = "mode must be LS_MODE_ADVECT or LS_MODE_MCF";
      return LS_ERR_ARG;
    }
    if (prm.warm != LS_WARM_NONE) {
      err = "warm starts run in the persistent step kernel (LS_EXEC_PERSISTENT or LS_EXEC_GRAPH)";
      return LS_ERR_ARG;
    }
    std::memset(&st, 0, sizeof(st));
    st.exec_mode = LS_EXEC_KERNELS;
    st.sweep_tasks = org_ == kOrgTasks ? 1 : 0;
    st.sweep_staged = stage_;
    st.fast_len_roots = fast_len_ ? 1 : 0;
    st.plane_bound = plane_on() ? 1 : 0;
    st.step_wg_per_cu = 0;
    pc_synced_ = false;
    const DevMesh<T> m = mesh();
    const DevState<T> s = state();
    const T tol = (T)prm.tol, bw = (T)prm.bandwidth;
    const double leps = prm.lambda_eps, dt = prm.dt;
    hipStream_t q = stream_;
    LS_HIP(hipEventRecord(ev_[0], q));
    // Seeds over every patch on the first step, else over the previous narrowband.
    kev_used_ = 0;
    k_init<T><<<grid_for(N_), kBlock, 0, q>>>(N_, s);
    k_fill<<<grid_for(P_), kBlock, 0, q>>>(d_rs_epoch_, P_, 0);
    k_fill<<<grid_for(P_), kBlock, 0, q>>>(d_ve_tag_, P_, 0);
    k_set_counters<<<1, 64, 0, q>>>(d_counters_, kNumCounters, 0);
    k_clear_sizes<<<1, 64, 0, q>>>(d_sizes_);
    const int32_t* scan = nnb_ > 0 ? d_nb_cur() : d_scan_all_;
    const int nscan = nnb_ > 0 ? nnb_ : P_;
    int32_t* cur = d_lists_[0];
    k_seeds<T, NC><<<nscan, kBlock, seed_lds_, q>>>(m, s, scan, cur, d_counters_);
    LS_HIP(hipGetLastError());
    LS_HIP(hipEventRecord(ev_[1], q));
    int32_t cnt[kNumCounters];
    LS_HIP(hipMemcpyAsync(cnt, d_counters_, sizeof(cnt), hipMemcpyDeviceToHost, q));
    LS_HIP(hipStreamSynchronize(q));
    int nact = cnt[kCntSeed];
    st.n_seed_patches = nact;
    int32_t* fresh_list = d_lists_[1];
    int32_t* next = d_lists_[2];
    while (nact > 0 && (prm.max_fim_iters <= 0 || st.fim_iters < prm.max_fim_iters)) {
      st.fim_iters++;
      st.patch_visits += nact;
      int ke = kev_begin(0, q);
      launch_sweep(nact, q, m, s, cur, prm.niter, 1, tol, leps, st.fim_iters, 0);
      kev_end(ke, q);
      k_copyback<T><<<nact, kBlock, 0, q>>>(m, s, cur);
      k_set_counters<<<1, 64, 0, q>>>(d_counters_ + kCntFresh, 2, 0);
      k_activate<T><<<grid_for(nact), kBlock, 0, q>>>(m, s, cur, nact, bw, fresh_list);
      LS_HIP(hipGetLastError());
      int32_t nf = 0;
      LS_HIP(hipMemcpyAsync(&nf, d_counters_ + kCntFresh, sizeof(int32_t), hipMemcpyDeviceToHost, q));
      LS_HIP(hipStreamSynchronize(q));
      if (nf > 0) {
        st.probes += nf;
        int kp = kev_begin(0, q);
        launch_sweep(nf, q, m, s, fresh_list, 1, 0, tol, leps, 0, st.fim_iters + 1);
        kev_end(kp, q);
      }
      k_finalize<T><<<grid_for(nact + nf), kBlock, 0, q>>>(s, cur, nact, fresh_list, nf, next);
      LS_HIP(hipGetLastError());
      LS_HIP(hipMemcpyAsync(&nact, d_counters_ + kCntNext, sizeof(int32_t), hipMemcpyDeviceToHost, q));
      LS_HIP(hipStreamSynchronize(q));
      std::swap(cur, next);
      if (st.fim_iters > kFimIterCap) { err = "FIM did not converge"; return LS_ERR_INTERNAL; }
    }
    LS_HIP(hipEventRecord(ev_[2], q));
    // Narrowband into the buffer not holding the previous one.
    nb_sel_ ^= 1;
    k_narrowband<T><<<P_, kBlock, 0, q>>>(m, s, bw, d_nb_cur());
    LS_HIP(hipGetLastError());
    LS_HIP(hipMemcpyAsync(cnt, d_counters_, sizeof(cnt), hipMemcpyDeviceToHost, q));
    LS_HIP(hipStreamSynchronize(q));
    nnb_ = cnt[kCntNB];
    st.nb_count = nnb_;
    st.nb_verts = cnt[kCntNBVerts];
    LS_HIP(hipEventRecord(ev_[3], q));
    unsigned long long ck[2] = {0, 0};
    if (nnb_ > 0 && prm.cfl) {
      k_cfl_init<<<1, 64, 0, q>>>(d_cfl_);
      k_cfl<T, NC><<<nnb_, kBlock, 0, q>>>(m, s, d_nb_cur());
      LS_HIP(hipGetLastError());
      LS_HIP(hipMemcpyAsync(ck, d_cfl_, sizeof(ck), hipMemcpyDeviceToHost, q));
    }
    if (nnb_ > 0) {
      for (int it = 0; it < prm.inside_niter; ++it) {
        int ka = kev_begin(1, q);
        k_advect<T, NC><<<nnb_, kBlock, adv_lds_, q>>>(m, s, d_nb_cur(), dt, prm.mode, prm.cfl);
        kev_end(ka, q);
        k_copy_phi<T><<<nnb_, kBlock, 0, q>>>(m, s, d_nb_cur());
      }
      LS_HIP(hipGetLastError());
    }
    LS_HIP(hipEventRecord(ev_[4], q));
    LS_HIP(hipEventSynchronize(ev_[4]));
    float ms;
    LS_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[1])); st.ms_seed = ms;
    LS_HIP(hipEventElapsedTime(&ms, ev_[1], ev_[2])); st.ms_fim = ms;
    LS_HIP(hipEventElapsedTime(&ms, ev_[2], ev_[3])); st.ms_nb = ms;
    LS_HIP(hipEventElapsedTime(&ms, ev_[3], ev_[4])); st.ms_advect = ms;
    LS_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[4])); st.ms_total = ms;
    for (size_t i = 0; i < kev_used_; ++i) {
      LS_HIP(hipEventElapsedTime(&ms, kev_[2 * i], kev_[2 * i + 1]));
      if (kev_kind_[i] == 0) { st.ms_sweep_kernels += ms; st.sweep_launches++; }
      else st.ms_advect_kernels += ms;
    }
    unsigned long long sz[kNumSizes];
    LS_HIP(hipMemcpy(sz, d_sizes_, sizeof(sz), hipMemcpyDeviceToHost));
    fill_sizes(st, sz);
    fill_cfl(prm, ck, st);
    return LS_OK;
  }

  int get_phi(double* out, std::string& err) override {
    if (!uploaded_) { err = "layout not built"; return LS_ERR_STATE; }
    std::vector<T> h(N_);
    LS_HIP(hipMemcpy(h.data(), d_phi_, N_ * sizeof(T), hipMemcpyDeviceToHost));
    for (int64_t v = 0; v < N_; ++v) out[vert_perm_h_[v]] = (double)h[v];
    return LS_OK;
  }

  int get_narrowband(std::vector<int32_t>& ids, std::string& err) override {
    ids.assign(nnb_, 0);
    if (nnb_ > 0) LS_HIP(hipMemcpy(ids.data(), d_nb_cur(), nnb_ * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::sort(ids.begin(), ids.end());
    return LS_OK;
  }

 private:
  DevMesh<T> mesh() const {
    DevMesh<T> m;
    m.P = P_; m.N = N_; m.F = F_;
    m.vert_off = d_vert_off_; m.ele_off = d_ele_off_; m.halo_off = d_halo_off_; m.halo_ids = d_halo_ids_;
    m.halo_patch = d_halo_patch_;
    m.phdr = d_phdr_;
    m.copy_local = d_copy_local_; m.vslot_off = d_vslot_off_; m.vslots = d_vslots_;
    m.coords = d_coords_; m.sigma = d_sigma_; m.padj_off = d_padj_off_; m.padj = d_padj_;
    m.vpatch = d_vpatch_;
    m.hgt = d_hgt_;
    m.pscale = d_pscale_;
    m.patch_dt = use_patch_prm_ ? d_patch_dt_ : nullptr;
    m.patch_bw = use_patch_prm_ ? d_patch_bw_ : nullptr;
    m.copy_src = d_copy_src_;
    m.ceik = d_ceik_;
    m.ccurv = d_ccurv_;
    m.pgroup = d_pgroup_;
    m.ngroups = ngroups_;
    m.sigma_phi = sigma_phi_ ? 1 : 0;
    m.fast_len = fast_len_ ? 1 : 0;
    m.plane = plane_on() ? 1 : 0;
    m.ceik0 = kCeikDefault;
    m.ccurv0 = kCcurvDefault;
    return m;
  }
  DevState<T> state() const {
    DevState<T> s;
    s.phi = d_phi_; s.phi_out = d_phi_out_; s.DT = d_DT_; s.DT_out = d_DT_out_;
    s.label = d_label_; s.tsign = d_tsign_; s.bcon = d_bcon_; s.bmin = d_bmin_;
    s.act = d_act_; s.fresh = d_fresh_; s.seedflag = d_seedflag_; s.counters = d_counters_;
    s.sizes = d_sizes_;
    s.inA = d_inA_; s.freshE = d_freshE_;
    s.listA = d_lists_[0]; s.listB = d_lists_[1]; s.listF = d_listF_;
    s.nb[0] = d_nb_; s.nb[1] = d_lists_[3];
    s.pc = d_pc_;
    s.ticks = d_ticks_;
    s.bar = d_bar_;
    s.warm = d_warm_;
    s.rs_slot = d_rs_slot_;
    s.rs_vval = d_rs_vval_;
    s.rs_dirty = d_rs_dirty_;
    s.rs_epoch = d_rs_epoch_;
    s.visitE = d_visitE_;
    s.ve_ok = d_ve_ok_;
    s.ve_tag = d_ve_tag_;
    s.cfl = d_cfl_;
    s.gfail = d_gfail_;
    s.DT_out2 = d_DT_out2_;
    s.visitE2 = d_visitE2_;
    s.bcon2 = d_bcon2_;
    s.ov_lists = d_ov_lists_;
    return s;
  }
  int32_t* d_nb_cur() const { return nb_sel_ ? d_lists_[3] : d_nb_; }

  void release() {
    (void)hipSetDevice(device_);
    void* ptrs[] = {d_vert_off_, d_ele_off_, d_halo_off_, d_halo_ids_, d_copy_local_, d_vslot_off_,
                    d_vslots_, d_padj_off_, d_padj_, d_coords_, d_sigma_, d_phi_, d_phi_out_, d_DT_,
                    d_DT_out_, d_label_, d_tsign_, d_bcon_, d_bmin_, d_act_, d_fresh_, d_seedflag_,
                    d_counters_, d_nb_, d_scan_all_, d_lists_[0], d_lists_[1], d_lists_[2],
                    d_lists_[3], d_inA_, d_freshE_, d_listF_, d_stat_, d_bar_,
                    d_patch_dt_, d_patch_bw_, d_warm_, d_hgt_, d_mlp_prm_, d_mlp_m_, d_mlp_v_, d_mlp_grad_,
                    d_mlp_b_, d_mlp_w1_, d_mlp_w2_, d_mlp_w3_, d_ds_x_, d_ds_y_, d_pend_, d_xyz_,
                    d_copy_verts_, d_copy_src_, d_rs_slot_, d_rs_vval_, d_rs_dirty_, d_rs_epoch_, d_visitE_,
                    d_vpatch_, d_ve_ok_, d_ve_tag_, d_gfail_, d_pgroup_, d_ceik_, d_ccurv_, d_DT_out2_,
                    d_visitE2_,
                    d_bcon2_, d_ov_lists_, d_pscale_, d_halo_patch_, d_phdr_};
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (void* p : ptrs) if (p) (void)hipFree(p);
    if (h_stat_) (void)hipHostFree(h_stat_);
    for (auto& e : ev_) if (e) (void)hipEventDestroy(e);
    for (auto& e : kev_) if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  int device_;
  bool uploaded_ = false;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev_[5] = {};
  int P_ = 0;
  int64_t N_ = 0, F_ = 0;
  size_t fim_lds_ = 0, adv_lds_ = 0, seed_lds_ = 0;
  bool coop_ok_ = true;  // cooperative launches accepted so far
  int org_ = kOrgElements;  // FIM sweep organisation in use (kOrg*)
  int adv_grid_ = 1;          // blocks of the narrowband advection after a persistent step
  int sweep_mode_ = LS_SWEEP_AUTO;
  size_t fim_elem_lds_ = 0, fim_task_lds_ = 0, fim_elem_stage_lds_ = 0, fim_task_stage_lds_ = 0;
  int stage_ = 0;  // sweeps stage corner ids and heights in LDS (size_kernels)
  int lds_max_ = 0;
  std::vector<int32_t> vert_off_h_, vert_perm_h_;
  int nnb_ = 0, nb_sel_ = 0;
  int32_t *d_vert_off_ = nullptr, *d_ele_off_ = nullptr, *d_halo_off_ = nullptr, *d_halo_ids_ = nullptr;
  ushort4* d_copy_local_ = nullptr;
  int32_t* d_vslot_off_ = nullptr;
  uint16_t* d_vslots_ = nullptr;
  int32_t *d_padj_off_ = nullptr, *d_padj_ = nullptr;
  T *d_coords_ = nullptr, *d_sigma_ = nullptr;
  double* d_xyz_ = nullptr;          // N x 3, renumbered vertex coordinates
  T *d_rs_slot_ = nullptr, *d_rs_vval_ = nullptr;  // probe states for resumed visits
  uint8_t* d_rs_dirty_ = nullptr;
  int32_t* d_rs_epoch_ = nullptr;
  int32_t* d_visitE_ = nullptr;  // P, FIM iteration of the last visit (persistent step)
  uint8_t* d_ve_ok_ = nullptr;   // F, element copies a visit left clean
  int32_t* d_ve_tag_ = nullptr;  // P, the patch has ve_ok from this FIM run
  int32_t* d_vpatch_ = nullptr;  // N, patch of each renumbered vertex
  int32_t* d_halo_patch_ = nullptr;  // halo entries, patch owning each one's vertex
  int4* d_phdr_ = nullptr;           // 2P, per-patch offset header
  int32_t* d_copy_verts_ = nullptr;  // F x NC, renumbered corner ids
  int32_t* d_copy_src_ = nullptr;    // F, original element of each copy
  __half* d_hgt_ = nullptr;  // tet copies: corner heights (F x 4), FP16 rounded toward zero
  bool fast_len_ = false;    // the tet solve's length roots use sqrt_len (k_heights)
  bool plane_ok_ = false;    // element sizes allow the FP32 plane bound (k_heights)
  float* d_pscale_ = nullptr;  // tets: per-patch coordinate scale bound (P)
  T* d_warm_ = nullptr;  // learned warm-start distances (N), written by the predictor
  unsigned long long* d_cfl_ = nullptr;  // CFL bound keys (2)
  int32_t* d_gfail_ = nullptr;   // warm-start failure flag per mesh (ngroups_)
  bool sigma_phi_ = false;       // triangles advect with the velocity from phi (LS_VEL_PHI_2D)
  int32_t* d_pgroup_ = nullptr;  // mesh of each patch of a batch (P), or null
  int ngroups_ = 1;
  float* scratch_ = nullptr;  // this device's `scratch` (scratch_address), never freed
  // overlapped FIM iterations of the persistent step (see DevState)
  T* d_DT_out2_ = nullptr;
  int32_t *d_visitE2_ = nullptr, *d_bcon2_ = nullptr, *d_ov_lists_ = nullptr;
  double* d_ceik_ = nullptr;   // per original element, or null (uniform default)
  double* d_ccurv_ = nullptr;
  float *d_mlp_prm_ = nullptr, *d_mlp_m_ = nullptr, *d_mlp_v_ = nullptr, *d_mlp_grad_ = nullptr, *d_mlp_b_ = nullptr;
  uint16_t *d_mlp_w1_ = nullptr, *d_mlp_w2_ = nullptr, *d_mlp_w3_ = nullptr;
  bool mlp_ready_ = false;
  int adam_t_ = 0;
  float *d_ds_x_ = nullptr, *d_ds_y_ = nullptr;  // collected training pairs
  int64_t ds_cap_ = 0, ds_n_ = 0, pend_n_ = 0, pend_cap_ = 0;
  int32_t* d_pend_ = nullptr;  // copies awaiting their targets
  std::vector<int32_t> ele_off_h_;
  int max_ep_ = 0, max_vp_ = 0, max_hp_ = 0;
  double* d_patch_dt_ = nullptr;
  T* d_patch_bw_ = nullptr;
  bool use_patch_prm_ = false;
  T *d_phi_ = nullptr, *d_phi_out_ = nullptr, *d_DT_ = nullptr, *d_DT_out_ = nullptr, *d_bmin_ = nullptr;
  int8_t *d_label_ = nullptr, *d_tsign_ = nullptr;
  int32_t *d_bcon_ = nullptr, *d_act_ = nullptr, *d_fresh_ = nullptr, *d_seedflag_ = nullptr, *d_counters_ = nullptr;
  unsigned long long* d_sizes_ = nullptr;  // in d_stat_
  // d_ticks_, d_sizes_, d_cfl_ and d_pc_ live in one block, d_stat_, so a
  // persistent step reads them back with one copy into h_stat_ (pinned).
  static constexpr int kStatWords = kNumTicks + kNumSizes + 2 + (kNumPc + 1) / 2;
  unsigned long long* d_stat_ = nullptr;
  unsigned long long* h_stat_ = nullptr;
  std::vector<hipEvent_t> kev_;  // kernel-timing events, pairs (start, end)
  std::vector<int> kev_kind_;    // 0 sweep, 1 advect, per pair
  size_t kev_used_ = 0;

  // Records a start event for a timed kernel launch; returns the pair index.
  int kev_begin(int kind, hipStream_t q) {
    if (2 * (kev_used_ + 1) > kev_.size()) {
      for (int k = 0; k < 64; ++k) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        kev_.push_back(e);
      }
      kev_kind_.resize(kev_.size() / 2);
    }
    const int i = (int)kev_used_++;
    kev_kind_[i] = kind;
    (void)hipEventRecord(kev_[2 * i], q);
    return i;
  }
  void kev_end(int i, hipStream_t q) {
    if (i >= 0) (void)hipEventRecord(kev_[2 * i + 1], q);
  }
  int32_t *d_nb_ = nullptr, *d_scan_all_ = nullptr;
  int32_t* d_lists_[4] = {nullptr, nullptr, nullptr, nullptr};
  int32_t *d_inA_ = nullptr, *d_freshE_ = nullptr, *d_listF_ = nullptr, *d_pc_ = nullptr;
  unsigned long long* d_ticks_ = nullptr;
  unsigned int* d_bar_ = nullptr;
  int wall_khz_ = 100000;
  int exec_ = LS_EXEC_PERSISTENT;
  bool pc_synced_ = true;
  int32_t pc_nb_host_ = 0;
  int coop_grid_ = 0;
  bool step_one_per_cu_ = false;  // persistent grid of one workgroup per CU
  int cus_ = 1;                   // compute units of the device
  size_t step_lds_ = 0;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  int graph_key_[3] = {-1, -1, -1};  // (first, nb_in, params hash) the graph was captured for
  ls_step_params graph_prm_{};
};

Engine* make_engine(int dim, int dtype, int device, std::string& err) {
  if (dim == 3 && dtype == LS_F64) return new EngineImpl<double, 4>(device);
  if (dim == 3) return new EngineImpl<float, 4>(device);
  if (dim == 2) return new EngineImpl<double, 3>(device);
  err = "dim must be 2 or 3";
  return nullptr;
}

}  // namespace lsamd
