// collide.h — FP64 collision primitives on the device, in the reference's
// operand order: Ericson closest points (closest-point.ts:14-112), the cubic
// coplanarity solver (cubic.ts:4-66), VT/EE CCD (ccd-vt.ts:20-115,
// ccd-ee.ts:24-101), barrier (barrier.ts:4-24) and Coulomb friction
// (friction.ts:14-23).
#pragma once

#include "common.h"

namespace cloth {

struct Closest {
  double s, t, dist, rx, ry, rz;
};

// kDist = false leaves dist at 0 for callers that decide from |r|^2 first
template <bool kDist = true>
__device__ __forceinline__ Closest pack_cp(double s, double t, double rx, double ry, double rz) {
  return Closest{s, t, kDist ? jsm::hypot3(rx, ry, rz) : 0.0, rx, ry, rz};
}

// closestPointVertexTriangle() (Ericson 5.1.5)
template <bool kDist = true>
__device__ inline Closest closest_vt(const double* P, const double* A, const double* B, const double* Cc) {
  const double px = P[0], py = P[1], pz = P[2], ax = A[0], ay = A[1], az = A[2];
  const double bx = B[0], by = B[1], bz = B[2], cx = Cc[0], cy = Cc[1], cz = Cc[2];
  const double abx = bx - ax, aby = by - ay, abz = bz - az;
  const double acx = cx - ax, acy = cy - ay, acz = cz - az;
  const double apx = px - ax, apy = py - ay, apz = pz - az;
  const double d1 = abx * apx + aby * apy + abz * apz;
  const double d2 = acx * apx + acy * apy + acz * apz;
  if (d1 <= 0 && d2 <= 0) return pack_cp<kDist>(0, 0, px - ax, py - ay, pz - az);
  const double bpx = px - bx, bpy = py - by, bpz = pz - bz;
  const double d3 = abx * bpx + aby * bpy + abz * bpz;
  const double d4 = acx * bpx + acy * bpy + acz * bpz;
  if (d3 >= 0 && d4 <= d3) return pack_cp<kDist>(1, 0, px - bx, py - by, pz - bz);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    const double v = d1 / (d1 - d3);
    return pack_cp<kDist>(v, 0, px - (ax + v * abx), py - (ay + v * aby), pz - (az + v * abz));
  }
  const double cpx = px - cx, cpy = py - cy, cpz = pz - cz;
  const double d5 = abx * cpx + aby * cpy + abz * cpz;
  const double d6 = acx * cpx + acy * cpy + acz * cpz;
  if (d6 >= 0 && d5 <= d6) return pack_cp<kDist>(0, 1, px - cx, py - cy, pz - cz);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) {
    const double w = d2 / (d2 - d6);
    return pack_cp<kDist>(0, w, px - (ax + w * acx), py - (ay + w * acy), pz - (az + w * acz));
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0) {
    const double w = (d4 - d3) / (d4 - d3 + (d5 - d6));
    const double qx = bx + w * (cx - bx), qy = by + w * (cy - by), qz = bz + w * (cz - bz);
    return pack_cp<kDist>(1 - w, w, px - qx, py - qy, pz - qz);
  }
  const double denom = 1 / (va + vb + vc);
  const double v = vb * denom;
  const double w = vc * denom;
  const double qx = ax + abx * v + acx * w;
  const double qy = ay + aby * v + acy * w;
  const double qz = az + abz * v + acz * w;
  return pack_cp<kDist>(v, w, px - qx, py - qy, pz - qz);
}

__device__ __forceinline__ double clamp_(double v, double lo, double hi) { return v < lo ? lo : v > hi ? hi : v; }

// closestPointEdgeEdge() (Ericson 5.1.9)
template <bool kDist = true>
__device__ inline Closest closest_ee(const double* A, const double* B, const double* Cc, const double* D) {
  const double ax = A[0], ay = A[1], az = A[2], bx = B[0], by = B[1], bz = B[2];
  const double cx = Cc[0], cy = Cc[1], cz = Cc[2], dx = D[0], dy = D[1], dz = D[2];
  const double d1x = bx - ax, d1y = by - ay, d1z = bz - az;
  const double d2x = dx - cx, d2y = dy - cy, d2z = dz - cz;
  const double rx = ax - cx, ry = ay - cy, rz = az - cz;
  const double a = d1x * d1x + d1y * d1y + d1z * d1z;
  const double e = d2x * d2x + d2y * d2y + d2z * d2z;
  const double f = d2x * rx + d2y * ry + d2z * rz;
  double s, t;
  if (a <= 1e-30 && e <= 1e-30) {
    s = 0; t = 0;
  } else if (a <= 1e-30) {
    s = 0; t = clamp_(f / e, 0, 1);
  } else {
    const double c = d1x * rx + d1y * ry + d1z * rz;
    if (e <= 1e-30) {
      t = 0; s = clamp_(-c / a, 0, 1);
    } else {
      const double b = d1x * d2x + d1y * d2y + d1z * d2z;
      const double denom = a * e - b * b;
      s = denom > 1e-30 ? clamp_((b * f - c * e) / denom, 0, 1) : 0;
      t = (b * s + f) / e;
      if (t < 0) {
        t = 0; s = clamp_(-c / a, 0, 1);
      } else if (t > 1) {
        t = 1; s = clamp_((b - c) / a, 0, 1);
      }
    }
  }
  const double axp = ax + d1x * s, ayp = ay + d1y * s, azp = az + d1z * s;
  const double cxp = cx + d2x * t, cyp = cy + d2y * t, czp = cz + d2z * t;
  const double rx2 = axp - cxp, ry2 = ayp - cyp, rz2 = azp - czp;
  return Closest{s, t, kDist ? jsm::hypot3(rx2, ry2, rz2) : 0.0, rx2, ry2, rz2};
}

__device__ inline int solve_quadratic01(double a2, double a1, double a0, double* out) {
  const double eps = 1e-12;
  if (jsm::fabs_(a2) < eps) {
    if (jsm::fabs_(a1) < eps) return 0;
    const double r = -a0 / a1;
    if (r >= 0 && r <= 1) { out[0] = r; return 1; }
    return 0;
  }
  const double disc = a1 * a1 - 4 * a2 * a0;
  if (disc < 0) return 0;
  const double sd = jsm::sqrt_(disc);
  int k = 0;
  const double r0 = (-a1 - sd) / (2 * a2), r1 = (-a1 + sd) / (2 * a2);
  if (r0 >= 0 && r0 <= 1) out[k++] = r0;
  if (r1 >= 0 && r1 <= 1) out[k++] = r1;
  if (k == 2 && out[1] < out[0]) { const double t = out[0]; out[0] = out[1]; out[1] = t; }
  return k;
}

// solveCubic01() cubic.ts:4-49
__device__ inline int solve_cubic01(double a3, double a2, double a1, double a0, double* res) {
  const double eps = 1e-12;
  if (jsm::fabs_(a3) < eps) return solve_quadratic01(a2, a1, a0, res);
  const double A = a2 / a3, B = a1 / a3, C = a0 / a3;
  const double sqA = A * A;
  const double p = (3 * B - sqA) / 3;
  const double q = (2 * sqA * A - 9 * A * B + 27 * C) / 27;
  double roots[3];
  int nr = 0;
  const double disc = (q * q) / 4 + (p * p * p) / 27;
  const double shift = A / 3;
  const double PI = 3.141592653589793;
  if (disc > eps) {
    const double sd = jsm::sqrt_(disc);
    const double u = jsm::cbrt(-q / 2 + sd);
    const double v = jsm::cbrt(-q / 2 - sd);
    roots[nr++] = u + v - shift;
  } else if (jsm::fabs_(disc) <= eps) {
    const double u = jsm::cbrt(-q / 2);
    roots[nr++] = 2 * u - shift;
    roots[nr++] = -u - shift;
  } else {
    const double r = jsm::sqrt_(-(p * p * p) / 27);
    const double phi = jsm::acos(jsm::jmin(1, jsm::jmax(-1, -q / (2 * r))));
    const double s = 2 * jsm::cbrt(r);
    roots[nr++] = s * jsm::cos(phi / 3) - shift;
    roots[nr++] = s * jsm::cos((phi + 2 * PI) / 3) - shift;
    roots[nr++] = s * jsm::cos((phi + 4 * PI) / 3) - shift;
  }
  double out[3];
  int no = 0;
  for (int i = 0; i < nr; i++) {
    double r = roots[i];
    if (!jsm::isfinite_(r)) continue;
    for (int k = 0; k < 4; k++) {
      const double f = ((a3 * r + a2) * r + a1) * r + a0;
      const double df = (3 * a3 * r + 2 * a2) * r + a1;
      if (jsm::fabs_(df) < 1e-18) break;
      r -= f / df;
    }
    if (r >= -1e-9 && r <= 1 + 1e-9) out[no++] = jsm::jmin(1, jsm::jmax(0, r));
  }
  // ascending insertion sort (Array.prototype.sort with x - y)
  for (int i = 1; i < no; i++) {
    const double v = out[i];
    int j = i - 1;
    while (j >= 0 && out[j] > v) { out[j + 1] = out[j]; j--; }
    out[j + 1] = v;
  }
  int k = 0;
  for (int i = 0; i < no; i++)
    if (i == 0 || out[i] - out[i - 1] > 1e-9) res[k++] = out[i];
  return k;
}

__device__ __forceinline__ void cross_v(double* o, const double* u, const double* v) {
  o[0] = u[1] * v[2] - u[2] * v[1];
  o[1] = u[2] * v[0] - u[0] * v[2];
  o[2] = u[0] * v[1] - u[1] * v[0];
}
__device__ __forceinline__ double dot_v(const double* u, const double* v) {
  return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}

// Step-start and trial positions of the four stencil vertices.
struct Seg4 {
  double x0[4][3], x1[4][3];
};

__device__ __forceinline__ double stencil_dev(const Seg4& S) {
  double mx = 0, my = 0, mz = 0;
  for (int v = 0; v < 4; v++) {
    mx += S.x1[v][0] - S.x0[v][0];
    my += S.x1[v][1] - S.x0[v][1];
    mz += S.x1[v][2] - S.x0[v][2];
  }
  mx /= 4; my /= 4; mz /= 4;
  double dev = 0;
  for (int v = 0; v < 4; v++) {
    const double dx = S.x1[v][0] - S.x0[v][0] - mx;
    const double dy = S.x1[v][1] - S.x0[v][1] - my;
    const double dz = S.x1[v][2] - S.x0[v][2] - mz;
    const double m = jsm::jmax(jsm::jmax(jsm::fabs_(dx), jsm::fabs_(dy)), jsm::fabs_(dz));
    if (m > dev) dev = m;
  }
  return dev;
}

__device__ __forceinline__ Closest ee_at(const double* a0, const double* da, const double* b0, const double* db,
                                         const double* c0, const double* dc, const double* d0, const double* dd,
                                         double t);

// The head of vtCCD / eeCCD: resting contact at t = 0 (TOI 0) and the
// relative-motion early-out (Infinity), ccd-vt.ts:26-52 and ccd-ee.ts:41-59,
// with the same arithmetic as the full routines below.  Returns false when
// the cubic roots and the sampled fallback must decide.
__device__ inline bool ccd_head(const Seg4& S, int kind, double thickness, double& toi) {
  double d0;
  if (kind == 0) {
    d0 = closest_vt(S.x0[0], S.x0[1], S.x0[2], S.x0[3]).dist;
  } else {
    double a0[3], da[3], b0[3], db[3], c0[3], dc[3], e0[3], de[3];
    for (int i = 0; i < 3; i++) {
      a0[i] = S.x0[0][i]; da[i] = S.x1[0][i] - S.x0[0][i];
      b0[i] = S.x0[1][i]; db[i] = S.x1[1][i] - S.x0[1][i];
      c0[i] = S.x0[2][i]; dc[i] = S.x1[2][i] - S.x0[2][i];
      e0[i] = S.x0[3][i]; de[i] = S.x1[3][i] - S.x0[3][i];
    }
    d0 = ee_at(a0, da, b0, db, c0, dc, e0, de, 0).dist;
  }
  if (d0 <= thickness) {
    toi = 0;
    return true;
  }
  const double dev = stencil_dev(S);
  if (d0 - 4 * dev > thickness) {
    toi = __builtin_inf();
    return true;
  }
  return false;
}

__device__ __forceinline__ double norm2(const Closest& c) { return c.rx * c.rx + c.ry * c.ry + c.rz * c.rz; }

// ccd_head with the two comparisons of the closest distance d0 decided from
// |r|^2 wherever a relative margin of 1e-9 separates them from their
// thresholds (the hypot of the reference is within a few ulp of |r|); the
// exact head runs only in the band around a threshold or for non-finite or
// extreme inputs.  Same results as ccd_head.
__device__ inline bool ccd_head_fast(const Seg4& S, int kind, double thickness, double& toi) {
  Closest c0;
  if (kind == 0) {
    c0 = closest_vt<false>(S.x0[0], S.x0[1], S.x0[2], S.x0[3]);
  } else {
    // eeCCD evaluates its t = 0 configuration as x0 + dx * 0
    double A[4][3];
    for (int v = 0; v < 4; v++)
      for (int i = 0; i < 3; i++) A[v][i] = S.x0[v][i] + (S.x1[v][i] - S.x0[v][i]) * 0;
    c0 = closest_ee<false>(A[0], A[1], A[2], A[3]);
  }
  const double r2 = norm2(c0);
  const double lo = 1 - 1e-9, hi = 1 + 1e-9;
  if (!(r2 >= 1e-200 && r2 <= 1e200) || !(thickness >= 1e-100 && thickness <= 1e100))
    return ccd_head(S, kind, thickness, toi);
  const double t2 = thickness * thickness;
  if (r2 < t2 * lo) {  // d0 < thickness
    toi = 0;
    return true;
  }
  // the second comparison subtracts 4 dev from d0: keep d0 <= 1e6 thickness
  // so that its rounding stays far below the margin
  if (!(r2 > t2 * hi) || r2 > t2 * 1e12) return ccd_head(S, kind, thickness, toi);
  const double dev = stencil_dev(S);
  const double b = thickness + 4 * dev;
  if (!(b <= 1e100)) return ccd_head(S, kind, thickness, toi);
  const double b2 = b * b;
  if (r2 > b2 * hi) {  // d0 - 4 dev > thickness
    toi = __builtin_inf();
    return true;
  }
  if (r2 < b2 * lo) return false;
  return ccd_head(S, kind, thickness, toi);
}

// vtCCD() ccd-vt.ts:20-115; slots 0..3 = p, a, b, c.
__device__ inline double vt_ccd(const Seg4& S, double thickness) {
  const double d0 = closest_vt(S.x0[0], S.x0[1], S.x0[2], S.x0[3]).dist;
  if (d0 <= thickness) return 0;
  const double dev = stencil_dev(S);
  if (d0 - 4 * dev > thickness) return __builtin_inf();
  double p0[3], dp[3], a0[3], da[3], b0[3], db[3], c0[3], dc[3];
  for (int i = 0; i < 3; i++) {
    p0[i] = S.x0[0][i]; dp[i] = S.x1[0][i] - S.x0[0][i];
    a0[i] = S.x0[1][i]; da[i] = S.x1[1][i] - S.x0[1][i];
    b0[i] = S.x0[2][i]; db[i] = S.x1[2][i] - S.x0[2][i];
    c0[i] = S.x0[3][i]; dc[i] = S.x1[3][i] - S.x0[3][i];
  }
  double u0[3], du[3], v0[3], dv[3], w0[3], dw[3];
  for (int i = 0; i < 3; i++) {
    u0[i] = p0[i] - a0[i]; du[i] = dp[i] - da[i];
    v0[i] = b0[i] - a0[i]; dv[i] = db[i] - da[i];
    w0[i] = c0[i] - a0[i]; dw[i] = dc[i] - da[i];
  }
  double n0[3], n1[3], n2[3], t1[3], t2[3];
  cross_v(n0, v0, w0);
  cross_v(t1, v0, dw); cross_v(t2, dv, w0);
  for (int i = 0; i < 3; i++) n1[i] = t1[i] + t2[i];
  cross_v(n2, dv, dw);
  const double c0c = dot_v(u0, n0);
  const double c1c = dot_v(u0, n1) + dot_v(du, n0);
  const double c2c = dot_v(u0, n2) + dot_v(du, n1);
  const double c3c = dot_v(du, n2);
  double roots[3];
  const int nr = solve_cubic01(c3c, c2c, c1c, c0c, roots);
  for (int k = 0; k < nr; k++) {
    const double t = roots[k];
    double pp[3], aa[3], bb[3], cc[3];
    for (int i = 0; i < 3; i++) {
      pp[i] = p0[i] + dp[i] * t; aa[i] = a0[i] + da[i] * t;
      bb[i] = b0[i] + db[i] * t; cc[i] = c0[i] + dc[i] * t;
    }
    if (closest_vt(pp, aa, bb, cc).dist <= thickness) return t;
  }
  double minD = __builtin_inf(), minT = 0;
  const int N = 16;
  for (int k = 0; k <= N; k++) {
    const double t = (double)k / N;
    double pp[3], aa[3], bb[3], cc[3];
    for (int i = 0; i < 3; i++) {
      pp[i] = p0[i] + dp[i] * t; aa[i] = a0[i] + da[i] * t;
      bb[i] = b0[i] + db[i] * t; cc[i] = c0[i] + dc[i] * t;
    }
    const double d = closest_vt(pp, aa, bb, cc).dist;
    if (d < minD) { minD = d; minT = t; }
  }
  if (minD <= thickness) return minT;
  return __builtin_inf();
}

__device__ __forceinline__ Closest ee_at(const double* a0, const double* da, const double* b0, const double* db,
                                         const double* c0, const double* dc, const double* d0, const double* dd,
                                         double t) {
  double A[3], B[3], Cc[3], D[3];
  for (int i = 0; i < 3; i++) {
    A[i] = a0[i] + da[i] * t; B[i] = b0[i] + db[i] * t;
    Cc[i] = c0[i] + dc[i] * t; D[i] = d0[i] + dd[i] * t;
  }
  return closest_ee(A, B, Cc, D);
}

// eeCCD() ccd-ee.ts:24-101; slots 0..3 = a, b, c, d.
__device__ inline double ee_ccd(const Seg4& S, double thickness) {
  double a0[3], da[3], b0[3], db[3], c0[3], dc[3], d0[3], dd[3];
  for (int i = 0; i < 3; i++) {
    a0[i] = S.x0[0][i]; da[i] = S.x1[0][i] - S.x0[0][i];
    b0[i] = S.x0[1][i]; db[i] = S.x1[1][i] - S.x0[1][i];
    c0[i] = S.x0[2][i]; dc[i] = S.x1[2][i] - S.x0[2][i];
    d0[i] = S.x0[3][i]; dd[i] = S.x1[3][i] - S.x0[3][i];
  }
  const double q0 = ee_at(a0, da, b0, db, c0, dc, d0, dd, 0).dist;
  if (q0 <= thickness) return 0;
  const double dev = stencil_dev(S);
  if (q0 - 4 * dev > thickness) return __builtin_inf();
  double e10[3], de1[3], e20[3], de2[3], r0[3], dr[3];
  for (int i = 0; i < 3; i++) {
    e10[i] = b0[i] - a0[i]; de1[i] = db[i] - da[i];
    e20[i] = d0[i] - c0[i]; de2[i] = dd[i] - dc[i];
    r0[i] = c0[i] - a0[i]; dr[i] = dc[i] - da[i];
  }
  double n0[3], n1[3], n2[3], t1[3], t2[3];
  cross_v(n0, e10, e20);
  cross_v(t1, e10, de2); cross_v(t2, de1, e20);
  for (int i = 0; i < 3; i++) n1[i] = t1[i] + t2[i];
  cross_v(n2, de1, de2);
  const double c0c = dot_v(r0, n0);
  const double c1c = dot_v(r0, n1) + dot_v(dr, n0);
  const double c2c = dot_v(r0, n2) + dot_v(dr, n1);
  const double c3c = dot_v(dr, n2);
  double roots[3];
  const int nr = solve_cubic01(c3c, c2c, c1c, c0c, roots);
  for (int k = 0; k < nr; k++) {
    const double t = roots[k];
    const Closest cp = ee_at(a0, da, b0, db, c0, dc, d0, dd, t);
    if (cp.dist <= thickness && cp.s >= -1e-6 && cp.s <= 1 + 1e-6 && cp.t >= -1e-6 && cp.t <= 1 + 1e-6) return t;
  }
  double minD = __builtin_inf(), minT = 0;
  const int N = 16;
  for (int k = 0; k <= N; k++) {
    const double t = (double)k / N;
    const double d2 = ee_at(a0, da, b0, db, c0, dc, d0, dd, t).dist;
    if (d2 < minD) { minD = d2; minT = t; }
  }
  if (minD <= thickness) return minT;
  return __builtin_inf();
}

// barrier.ts:4-24
__device__ __forceinline__ double barrier_value(double d, double dHat) {
  if (d >= dHat) return 0;
  const double dc = jsm::jmax(d, 1e-12);
  const double t = dc - dHat;
  return -t * t * jsm::log(dc / dHat);
}
__device__ __forceinline__ double barrier_grad_scalar(double d, double dHat, double kappa, double w) {
  if (d >= dHat) return 0;
  const double dc = jsm::jmax(d, 1e-12);
  const double t = dc - dHat;
  const double db = -2 * t * jsm::log(dc / dHat) - (t * t) / dc;
  return kappa * w * db;
}
__device__ __forceinline__ double barrier_normal_force(double d, double dHat, double kappa, double w) {
  return jsm::jmax(0, -barrier_grad_scalar(d, dHat, kappa, w));
}

// coulombForce() friction.ts:14-23
__device__ __forceinline__ void coulomb_force(const double* u, double nx, double ny, double nz, double lambdaN,
                                              double mu, double eps, double* f) {
  const double vn = u[0] * nx + u[1] * ny + u[2] * nz;
  const double tx = u[0] - vn * nx, ty = u[1] - vn * ny, tz = u[2] - vn * nz;
  const double mag = jsm::sqrt_(tx * tx + ty * ty + tz * tz + eps * eps);
  const double s = (mu * jsm::jmax(0, lambdaN)) / mag;
  f[0] = -s * tx; f[1] = -s * ty; f[2] = -s * tz;
}

}  // namespace cloth
