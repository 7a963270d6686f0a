// elements.h — per-element device math, written in the reference's operand
// order so every FP64 result is the double the reference computes.
#pragma once

#include "fem.h"

namespace cloth {

__device__ __forceinline__ void ld3(const double* x, uint32_t v, double* o) {
  o[0] = x[(size_t)v * 3];
  o[1] = x[(size_t)v * 3 + 1];
  o[2] = x[(size_t)v * 3 + 2];
}

// Deformation quantities of triangleEnergyGradient() (membrane.ts:17-80).
struct TriState {
  double a, b, c, d;
  double F00, F10, F20, F01, F11, F21;
  double E00, E11, E01;
};

__device__ __forceinline__ TriState tri_state(const double* x0, const double* x1, const double* x2,
                                              const float* inv) {
  TriState s;
  s.a = inv[0]; s.b = inv[1]; s.c = inv[2]; s.d = inv[3];
  const double e1x = x1[0] - x0[0], e1y = x1[1] - x0[1], e1z = x1[2] - x0[2];
  const double e2x = x2[0] - x0[0], e2y = x2[1] - x0[1], e2z = x2[2] - x0[2];
  s.F00 = s.a * e1x + s.c * e2x; s.F10 = s.a * e1y + s.c * e2y; s.F20 = s.a * e1z + s.c * e2z;
  s.F01 = s.b * e1x + s.d * e2x; s.F11 = s.b * e1y + s.d * e2y; s.F21 = s.b * e1z + s.d * e2z;
  const double C00 = s.F00 * s.F00 + s.F10 * s.F10 + s.F20 * s.F20;
  const double C11 = s.F01 * s.F01 + s.F11 * s.F11 + s.F21 * s.F21;
  const double C01 = s.F00 * s.F01 + s.F10 * s.F11 + s.F20 * s.F21;
  s.E00 = 0.5 * (C00 - 1);
  s.E11 = 0.5 * (C11 - 1);
  s.E01 = 0.5 * C01;
  return s;
}

// Triangle energy (membrane.ts:38-39).
__device__ __forceinline__ double tri_energy(const TriState& s, double area, const DMat& mat) {
  const double psi = 0.5 * mat.c00 * s.E00 * s.E00 + 0.5 * mat.c11 * s.E11 * s.E11 + mat.c01 * s.E00 * s.E11 +
                     2 * mat.g * s.E01 * s.E01;
  return mat.thick * area * psi;
}

// Gradient contribution of `corner` (membrane.ts:41-74).
__device__ __forceinline__ void tri_grad_corner(const TriState& s, double area, const DMat& mat, int corner,
                                                double* out) {
  const double S00 = mat.c00 * s.E00 + mat.c01 * s.E11;
  const double S11 = mat.c11 * s.E11 + mat.c01 * s.E00;
  const double S01 = 2 * mat.g * s.E01;
  const double P0[3] = {s.F00 * S00 + s.F01 * S01, s.F10 * S00 + s.F11 * S01, s.F20 * S00 + s.F21 * S01};
  const double P1[3] = {s.F00 * S01 + s.F01 * S11, s.F10 * S01 + s.F11 * S11, s.F20 * S01 + s.F21 * S11};
  const double sc = mat.thick * area;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double g1 = sc * (P0[i] * s.a + P1[i] * s.b);
    const double g2 = sc * (P0[i] * s.c + P1[i] * s.d);
    out[i] = corner == 0 ? -(g1 + g2) : corner == 1 ? g1 : g2;
  }
}

// Analytic HVP contribution of `corner` (membrane-hvp.ts:18-75).
// The 2x2 symmetric stress [[s00, s01], [s01, s11]] with its negative
// eigenvalues set to zero (tension only): the geometric term dF S of the
// membrane Hessian becomes positive semidefinite.  Not in the reference; used
// for the Newton direction under CS_FLAG_PSD_NEWTON.
template <class T>
__device__ __forceinline__ void stress_psd(T& s00, T& s11, T& s01) {
  const T mean = T(0.5) * (s00 + s11), half = T(0.5) * (s00 - s11);
  const T rad = sqrt(half * half + s01 * s01);
  const T l1 = mean + rad, l2 = mean - rad;
  if (l2 >= T(0)) return;
  if (l1 <= T(0)) {
    s00 = s11 = s01 = T(0);
    return;
  }
  // S+ = l1 v1 v1^T = l1 / (l1 - l2) (S - l2 I)
  const T k = l1 / (l1 - l2);
  s00 = k * (s00 - l2);
  s11 = k * (s11 - l2);
  s01 = k * s01;
}

__device__ __forceinline__ void tri_hvp_corner(const TriState& s, const double* v0, const double* v1,
                                               const double* v2, double area, const DMat& mat, int corner,
                                               double* out, bool psd = false) {
  const double w1x = v1[0] - v0[0], w1y = v1[1] - v0[1], w1z = v1[2] - v0[2];
  const double w2x = v2[0] - v0[0], w2y = v2[1] - v0[1], w2z = v2[2] - v0[2];
  const double a = s.a, b = s.b, c = s.c, d = s.d;
  const double dF00 = a * w1x + c * w2x, dF10 = a * w1y + c * w2y, dF20 = a * w1z + c * w2z;
  const double dF01 = b * w1x + d * w2x, dF11 = b * w1y + d * w2y, dF21 = b * w1z + d * w2z;
  const double F00 = s.F00, F10 = s.F10, F20 = s.F20, F01 = s.F01, F11 = s.F11, F21 = s.F21;
  // E recomputed with the HVP's own expression (membrane-hvp.ts:38-40)
  const double E00 = 0.5 * (F00 * F00 + F10 * F10 + F20 * F20 - 1);
  const double E11 = 0.5 * (F01 * F01 + F11 * F11 + F21 * F21 - 1);
  const double E01 = 0.5 * (F00 * F01 + F10 * F11 + F20 * F21);
  const double dC00 = 2 * (F00 * dF00 + F10 * dF10 + F20 * dF20);
  const double dC11 = 2 * (F01 * dF01 + F11 * dF11 + F21 * dF21);
  const double dC01 = dF00 * F01 + F00 * dF01 + dF10 * F11 + F10 * dF11 + dF20 * F21 + F20 * dF21;
  const double dE00 = 0.5 * dC00, dE11 = 0.5 * dC11, dE01 = 0.5 * dC01;
  double S00 = mat.c00 * E00 + mat.c01 * E11;
  double S11 = mat.c11 * E11 + mat.c01 * E00;
  double S01 = 2 * mat.g * E01;
  if (psd) stress_psd(S00, S11, S01);
  const double dS00 = mat.c00 * dE00 + mat.c01 * dE11;
  const double dS11 = mat.c11 * dE11 + mat.c01 * dE00;
  const double dS01 = 2 * mat.g * dE01;
  const double dP0[3] = {dF00 * S00 + F00 * dS00 + dF01 * S01 + F01 * dS01,
                         dF10 * S00 + F10 * dS00 + dF11 * S01 + F11 * dS01,
                         dF20 * S00 + F20 * dS00 + dF21 * S01 + F21 * dS01};
  const double dP1[3] = {dF00 * S01 + F00 * dS01 + dF01 * S11 + F01 * dS11,
                         dF10 * S01 + F10 * dS01 + dF11 * S11 + F11 * dS11,
                         dF20 * S01 + F20 * dS01 + dF21 * S11 + F21 * dS11};
  const double sc = mat.thick * area;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double g1 = sc * (dP0[i] * a + dP1[i] * b);
    const double g2 = sc * (dP0[i] * c + dP1[i] * d);
    out[i] = corner == 0 ? -(g1 + g2) : corner == 1 ? g1 : g2;
  }
}

// dihedral() mesh.ts:129-140 on four positions.
__device__ __forceinline__ double dihedral4(const double* p0, const double* p1, const double* p2,
                                            const double* p3) {
  const double e0x = p1[0] - p0[0], e0y = p1[1] - p0[1], e0z = p1[2] - p0[2];
  const double e1x = p2[0] - p0[0], e1y = p2[1] - p0[1], e1z = p2[2] - p0[2];
  const double e2x = p3[0] - p0[0], e2y = p3[1] - p0[1], e2z = p3[2] - p0[2];
  const double n1x = e0y * e1z - e0z * e1y, n1y = e0z * e1x - e0x * e1z, n1z = e0x * e1y - e0y * e1x;
  const double n2x = e0y * e2z - e0z * e2y, n2y = e0z * e2x - e0x * e2z, n2z = e0x * e2y - e0y * e2x;
  const double l1 = jsm::hypot3(n1x, n1y, n1z) + 1e-30;
  const double l2 = jsm::hypot3(n2x, n2y, n2z) + 1e-30;
  const double cosT = jsm::jmin(1, jsm::jmax(-1, (n1x * n2x + n1y * n2y + n1z * n2z) / (l1 * l2)));
  return jsm::acos(cosT);
}

// hingeEnergy() bending.ts:7-16.
__device__ __forceinline__ double hinge_energy4(const double P[4][3], double rest, double len, double asum,
                                                double k) {
  const double th = dihedral4(P[0], P[1], P[2], P[3]);
  const double dd = th - rest;
  const double w = len * len / jsm::jmax(asum, 1e-12);
  return 0.5 * k * dd * dd * w;
}

// FD gradient of one hinge w.r.t. vertex v (addHingeGradient, bending.ts:19-42):
// every slot holding v is perturbed, exactly like the reference's shared copy.
__device__ __forceinline__ void hinge_grad_vertex(const DMesh& me, const double* x, uint32_t hh, uint32_t v,
                                                  double k, double* out) {
  const double eps = 1e-7;
  uint32_t vs[4];
  double P[4][3];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    vs[s] = me.hv[(size_t)hh * 4 + s];
    ld3(x, vs[s], P[s]);
  }
  const double rest = me.hrest[hh], len = me.hlen[hh], asum = me.hasum[hh];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double old = x[(size_t)v * 3 + i];
#pragma unroll
    for (int s = 0; s < 4; s++)
      if (vs[s] == v) P[s][i] = old + eps;
    const double ep = hinge_energy4(P, rest, len, asum, k);
#pragma unroll
    for (int s = 0; s < 4; s++)
      if (vs[s] == v) P[s][i] = old - eps;
    const double em = hinge_energy4(P, rest, len, asum, k);
#pragma unroll
    for (int s = 0; s < 4; s++)
      if (vs[s] == v) P[s][i] = old;
    out[i] = (ep - em) / (2 * eps);
  }
}

// Membrane + bending gradient at vertex v (evalInternal, fem.ts:15-54).
__device__ __forceinline__ void internal_grad_vertex(const DMesh& me, const DMat& mat, const double* x, uint32_t v,
                                                     double* acc) {
  acc[0] = acc[1] = acc[2] = 0;
  const uint32_t b = me.vt_off[v], e = me.vt_off[v + 1];
  for (uint32_t k = b; k < e; k++) {
    const uint32_t rec = me.vt_rec[k];
    const uint32_t t = rec >> 2;
    const int corner = rec & 3;
    double x0[3], x1[3], x2[3], gc[3];
    ld3(x, me.idx[(size_t)t * 3], x0);
    ld3(x, me.idx[(size_t)t * 3 + 1], x1);
    ld3(x, me.idx[(size_t)t * 3 + 2], x2);
    const TriState s = tri_state(x0, x1, x2, me.inv_dm + (size_t)t * 4);
    tri_grad_corner(s, me.area[t], mat, corner, gc);
    acc[0] += gc[0]; acc[1] += gc[1]; acc[2] += gc[2];
  }
  if (mat.kb > 0) {
    const uint32_t hb = me.vh_off[v], he = me.vh_off[v + 1];
    for (uint32_t k = hb; k < he; k++) {
      const uint32_t hh = me.vh_rec[k] >> 2;
      double gc[3];
      hinge_grad_vertex(me, x, hh, v, mat.kb, gc);
      acc[0] += gc[0]; acc[1] += gc[1]; acc[2] += gc[2];
    }
  }
}

}  // namespace cloth
