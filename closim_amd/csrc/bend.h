// bend.h — analytic gradient and Hessian-vector product of the hinge bending
// energy hingeEnergy (bending.ts:7-16): W = 0.5 k (theta - theta0)^2 l^2 / A
// with theta the dihedral() of mesh.ts:129-140, the unsigned angle
// acos(n1.n2 / |n1||n2|) between n1 = e x f and n2 = e x g, e = x1 - x0,
// f = x2 - x0, g = x3 - x0.  The reference differentiates W by central
// differences (addHingeGradient, bending.ts:19-42) and leaves it out of the
// Newton operator (fem.ts:106-107).
//
// acos has an infinite slope at a flat hinge (theta = pi, the rest angle of a
// flat sheet), so the derivatives go through the signed fold angle
//   phi = atan2(S, C),  S = -|e| det[e, f, g],  C = -((e.e)(f.g) - (e.f)(e.g)),
// (n1 x n2 = det[e,f,g] e and n1.n2 = (e.e)(f.g) - (e.f)(e.g)), for which
// theta = pi - |phi| and which is smooth through the flat state:
//   dphi = (C dS - S dC) / (C^2 + S^2),
//   dS/de = -(e/|e| det + |e| f x g),  dS/df = -|e| g x e,  dS/dg = -|e| e x f,
//   dC/de = -(2 e (f.g) - f (e.g) - g (e.f)),
//   dC/df = -((e.e) g - e (e.g)),  dC/dg = -((e.e) f - e (e.f)),
// and dW = k w (theta - theta0) dtheta with dtheta = -sign(phi) dphi
// (sign(0) = 0: at a flat hinge the gradient is the symmetric limit).  The
// vertex gradients are those of e (x1), f (x2), g (x3), and x0 takes minus
// their sum.  The Hessian-vector product is the same computation in dual
// numbers (value, derivative along p): exact to rounding, no step size.
#pragma once

#include "common.h"

namespace cloth {

// Dual number: value and directional derivative.
struct Dual {
  double v, d;
};
__host__ __device__ __forceinline__ Dual mkd(double v, double d = 0.0) { return Dual{v, d}; }
__host__ __device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
__host__ __device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
__host__ __device__ __forceinline__ Dual operator-(Dual a) { return Dual{-a.v, -a.d}; }
__host__ __device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
__host__ __device__ __forceinline__ Dual operator*(double s, Dual a) { return Dual{s * a.v, s * a.d}; }
__host__ __device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  return Dual{a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)};
}
__host__ __device__ __forceinline__ Dual dsqrt(Dual a) {
  const double r = sqrt(a.v);
  return Dual{r, a.d / (2 * r)};
}
__host__ __device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }
__host__ __device__ __forceinline__ double val(double a) { return a; }
__host__ __device__ __forceinline__ double val(Dual a) { return a.v; }
__host__ __device__ __forceinline__ double der(double) { return 0.0; }
__host__ __device__ __forceinline__ double der(Dual a) { return a.d; }

template <class T>
__host__ __device__ __forceinline__ T dot3(const T* a, const T* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
template <class T>
__host__ __device__ __forceinline__ void cross3(const T* a, const T* b, T* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// Gradient of one hinge's energy (k w (theta - theta0) dtheta) at X[4][3];
// with T = Dual (X carrying p as derivative parts) the derivative parts of G
// are the Hessian-vector product.  A degenerate hinge (zero edge or a
// triangle of zero area) contributes 0.
template <class T>
__host__ __device__ __forceinline__ void hinge_grad_t(const T (&X)[4][3], double rest, double w, double k,
                                                      T (&G)[4][3]) {
  T e[3], f[3], g[3];
  for (int i = 0; i < 3; i++) {
    e[i] = X[1][i] - X[0][i];
    f[i] = X[2][i] - X[0][i];
    g[i] = X[3][i] - X[0][i];
  }
  const T ee = dot3(e, e);
  T fxg[3], gxe[3], exf[3];
  cross3(f, g, fxg);
  cross3(g, e, gxe);
  cross3(e, f, exf);
  const T det = dot3(e, fxg);
  const T fg = dot3(f, g), ef = dot3(e, f), eg = dot3(e, g);
  const T C = -(ee * fg - ef * eg);
  const double lev = sqrt(val(ee));
  if (!(lev > 0)) {
    for (int c = 0; c < 4; c++)
      for (int i = 0; i < 3; i++) G[c][i] = T{};
    return;
  }
  const T le = dsqrt(ee);
  const T S = -(le * det);
  const T den = C * C + S * S;
  if (!(val(den) > 0)) {
    for (int c = 0; c < 4; c++)
      for (int i = 0; i < 3; i++) G[c][i] = T{};
    return;
  }
  const double phi = atan2(val(S), val(C));
  const double dphi = (val(C) * der(S) - val(S) * der(C)) / val(den);
  const double sgn = phi > 0 ? 1.0 : (phi < 0 ? -1.0 : 0.0);
  const double theta = 3.141592653589793 - fabs(phi);  // Math.PI
  // coef = k w (theta - theta0) (-sign phi) = k w (phi - (pi - theta0) sign phi),
  // whose derivative along p is k w dphi wherever it exists (through the flat
  // state when theta0 = pi, where the energy is 0.5 k w phi^2)
  T coef;
  if constexpr (sizeof(T) == sizeof(double)) {
    coef = T(k * w * (theta - rest) * (-sgn));
  } else {
    coef = T{k * w * (theta - rest) * (-sgn), k * w * dphi};
  }
  for (int i = 0; i < 3; i++) {
    const T dSe = -(det / le * e[i] + le * fxg[i]);
    const T dSf = -(le * gxe[i]);
    const T dSg = -(le * exf[i]);
    const T dCe = -(2.0 * e[i] * fg - f[i] * eg - g[i] * ef);
    const T dCf = -(ee * g[i] - e[i] * eg);
    const T dCg = -(ee * f[i] - e[i] * ef);
    const T ge = (C * dSe - S * dCe) / den;
    const T gf = (C * dSf - S * dCf) / den;
    const T gg = (C * dSg - S * dCg) / den;
    G[1][i] = coef * ge;
    G[2][i] = coef * gf;
    G[3][i] = coef * gg;
    G[0][i] = -(G[1][i] + G[2][i] + G[3][i]);
  }
}

// The hinge's weight l^2 / max(A, 1e-12) (bending.ts:14).
__host__ __device__ __forceinline__ double hinge_weight(double len, double asum) {
  return len * len / (asum > 1e-12 ? asum : 1e-12);
}

}  // namespace cloth
