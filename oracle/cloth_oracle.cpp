// cloth_oracle.cpp — CPU FP64 restatement of closim's CpuSolver (the oracle).
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark load
// this library; the product never does.
//
// Each function follows the cited reference file line by line.  Arithmetic is
// written in the same operand order as the TypeScript (JS evaluates + and *
// left to right exactly like C++), compiled without FP contraction or
// reassociation, so every elementwise result is the IEEE-754 double the
// reference computes.  Math.hypot/log/acos/cbrt/cos are restated from the
// algorithms V8 implements them with (jsmath.h).  Every reduction over a whole
// array (dot products, norms, energy totals) is returned correctly rounded
// (double-double accumulation, one final rounding): the reference's naive
// left-to-right sum differs from it by at most its own rounding error, and
// the correctly rounded value does not depend on summation order.
//
// Parity status of each part is recorded in tests/test_oracle_*.py, which pin
// it against the reference's own tests (known answers, FD checks, invariants)
// and against brute force / numpy / closed forms where those tests leave it
// open (block-Jacobi inverses and fallback, pcg() against numpy.linalg, the FD
// hinge gradient against a complex-step derivative, static pairs against
// brute force, killApproachVelocity on one-contact closed forms, commit's
// last write per key, the Armijo decrease of every accepted trial).  What no
// check can fix is the last bits of the central differences (FD bending
// gradient, FD membrane HVP): they are pinned to their truncation and
// round-off bounds only.  orc_set_sum_mode(1) switches every array-wide
// reduction to the reference's plain left-to-right sum.

#include "cloth_oracle.h"
#include "jsmath.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

const double INF = std::numeric_limits<double>::infinity();
const double NaN = std::numeric_limits<double>::quiet_NaN();
thread_local std::string g_err;
int g_fault = ORC_FAULT_NONE;

// JS Math.max / Math.min semantics (NaN propagates, +0 > -0).
inline double jmax(double a, double b) {
  if (std::isnan(a) || std::isnan(b)) return NaN;
  if (a == 0 && b == 0) return std::signbit(a) ? b : a;
  return a > b ? a : b;
}
inline double jmin(double a, double b) {
  if (std::isnan(a) || std::isnan(b)) return NaN;
  if (a == 0 && b == 0) return std::signbit(a) ? a : b;
  return a < b ? a : b;
}

inline double hypot3(double a, double b, double c) { return orcm::hypot3(a, b, c); }

struct Mat {
  double rho, thick, c00, c11, c01, g, bw, bf, damp;
};
Mat from_cs(const cs_material& m) {
  return {m.areal_density, m.thickness, m.stretch_warp, m.stretch_weft, m.stretch_coupling,
          m.shear, m.bend_warp, m.bend_weft, m.damping};
}

struct Hinge {
  uint32_t v0, v1, v2, v3;
  double rest, len, asum;
};

struct Mesh {
  uint32_t n = 0, m = 0;
  std::vector<uint32_t> idx;
  std::vector<float> invDm, areas, masses, restPos;
  std::vector<Hinge> hinges;
};

// ---------------------------------------------------------------- mesh.ts

// cross() mesh.ts:142-144
inline void cross3(double ax, double ay, double az, double bx, double by, double bz, double* o) {
  o[0] = ay * bz - az * by;
  o[1] = az * bx - ax * bz;
  o[2] = ax * by - ay * bx;
}

// dihedral() mesh.ts:129-140 (unsigned acos; rest angle pi for a flat sheet).
template <typename T>
double dihedral(const T* x, uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3) {
  const double e0x = (double)x[v1 * 3] - (double)x[v0 * 3], e0y = (double)x[v1 * 3 + 1] - (double)x[v0 * 3 + 1],
               e0z = (double)x[v1 * 3 + 2] - (double)x[v0 * 3 + 2];
  const double e1x = (double)x[v2 * 3] - (double)x[v0 * 3], e1y = (double)x[v2 * 3 + 1] - (double)x[v0 * 3 + 1],
               e1z = (double)x[v2 * 3 + 2] - (double)x[v0 * 3 + 2];
  const double e2x = (double)x[v3 * 3] - (double)x[v0 * 3], e2y = (double)x[v3 * 3 + 1] - (double)x[v0 * 3 + 1],
               e2z = (double)x[v3 * 3 + 2] - (double)x[v0 * 3 + 2];
  double n1[3], n2[3];
  cross3(e0x, e0y, e0z, e1x, e1y, e1z, n1);
  cross3(e0x, e0y, e0z, e2x, e2y, e2z, n2);
  const double l1 = hypot3(n1[0], n1[1], n1[2]) + 1e-30;
  const double l2 = hypot3(n2[0], n2[1], n2[2]) + 1e-30;
  const double cosT = jmin(1, jmax(-1, (n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2]) / (l1 * l2)));
  return orcm::acos(cosT);
}

// buildHinges() mesh.ts:99-126.  Map insertion order == first occurrence of
// the undirected edge key while scanning triangles/edges in order.
std::vector<Hinge> build_hinges(const std::vector<uint32_t>& indices, const float* pos,
                                const std::vector<float>& areas) {
  const size_t m = indices.size() / 3;
  struct Occ { uint32_t tri, a, b, opp; };
  std::unordered_map<uint64_t, uint32_t> slot;  // key -> order index
  std::vector<std::vector<Occ>> lists;
  slot.reserve(m * 2);
  for (size_t t = 0; t < m; t++) {
    const uint32_t ids[3] = {indices[t * 3], indices[t * 3 + 1], indices[t * 3 + 2]};
    for (int e = 0; e < 3; e++) {
      const uint32_t a = ids[e], b = ids[(e + 1) % 3], opp = ids[(e + 2) % 3];
      const uint64_t k = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
      auto it = slot.find(k);
      if (it == slot.end()) {
        slot.emplace(k, (uint32_t)lists.size());
        lists.push_back({});
        lists.back().push_back({(uint32_t)t, a, b, opp});
      } else {
        lists[it->second].push_back({(uint32_t)t, a, b, opp});
      }
    }
  }
  std::vector<Hinge> hinges;
  for (auto& list : lists) {
    if (list.size() != 2) continue;
    const Occ& e0 = list[0];
    const Occ& e1 = list[1];
    Hinge h;
    h.v0 = e0.a; h.v1 = e0.b; h.v2 = e0.opp; h.v3 = e1.opp;
    h.rest = dihedral(pos, h.v0, h.v1, h.v2, h.v3);
    const double ax = (double)pos[h.v0 * 3] - (double)pos[h.v1 * 3];
    const double ay = (double)pos[h.v0 * 3 + 1] - (double)pos[h.v1 * 3 + 1];
    const double az = (double)pos[h.v0 * 3 + 2] - (double)pos[h.v1 * 3 + 2];
    h.len = hypot3(ax, ay, az);
    h.asum = (double)areas[e0.tri] + (double)areas[e1.tri];
    hinges.push_back(h);
  }
  return hinges;
}

// ---------------------------------------------------------- membrane.ts

struct TriGrad {
  double energy;
  double grad[9];
  double E00, E11, E01;
};

// triangleEnergyGradient() membrane.ts:17-80
TriGrad tri_energy_grad(const double* x0, const double* x1, const double* x2, const double* inv,
                        double area, const Mat& mat) {
  const double a = inv[0], b = inv[1], c = inv[2], d = inv[3];
  const double e1x = x1[0] - x0[0], e1y = x1[1] - x0[1], e1z = x1[2] - x0[2];
  const double e2x = x2[0] - x0[0], e2y = x2[1] - x0[1], e2z = x2[2] - x0[2];
  const double F00 = a * e1x + c * e2x, F10 = a * e1y + c * e2y, F20 = a * e1z + c * e2z;
  const double F01 = b * e1x + d * e2x, F11 = b * e1y + d * e2y, F21 = b * e1z + d * e2z;
  const double C00 = F00 * F00 + F10 * F10 + F20 * F20;
  const double C11 = F01 * F01 + F11 * F11 + F21 * F21;
  const double C01 = F00 * F01 + F10 * F11 + F20 * F21;
  const double E00 = 0.5 * (C00 - 1);
  const double E11 = 0.5 * (C11 - 1);
  const double E01 = 0.5 * C01;
  const double C00k = mat.c00, C11k = mat.c11, C01k = mat.c01, G = mat.g;
  const double psi = 0.5 * C00k * E00 * E00 + 0.5 * C11k * E11 * E11 + C01k * E00 * E11 + 2 * G * E01 * E01;
  TriGrad r;
  r.energy = mat.thick * area * psi;
  const double dPsi_dE00 = C00k * E00 + C01k * E11;
  const double dPsi_dE11 = C11k * E11 + C01k * E00;
  const double dPsi_dE01 = (g_fault == ORC_FAULT_SHEAR_4G ? 4 : 2) * G * E01;
  const double S00 = dPsi_dE00, S11 = dPsi_dE11, S01 = dPsi_dE01;
  const double P00 = F00 * S00 + F01 * S01;
  const double P10 = F10 * S00 + F11 * S01;
  const double P20 = F20 * S00 + F21 * S01;
  const double P01 = F00 * S01 + F01 * S11;
  const double P11 = F10 * S01 + F11 * S11;
  const double P21 = F20 * S01 + F21 * S11;
  const double s = mat.thick * area;
  for (int i = 0; i < 3; i++) {
    const double Pi0 = i == 0 ? P00 : i == 1 ? P10 : P20;
    const double Pi1 = i == 0 ? P01 : i == 1 ? P11 : P21;
    const double g1 = s * (Pi0 * a + Pi1 * b);
    const double g2 = s * (Pi0 * c + Pi1 * d);
    const double g0 = -(g1 + g2);
    r.grad[0 * 3 + i] = g0;
    r.grad[1 * 3 + i] = g1;
    r.grad[2 * 3 + i] = g2;
  }
  r.E00 = E00; r.E11 = E11; r.E01 = E01;
  return r;
}

// triangleHvp() membrane-hvp.ts:18-75
void tri_hvp(const double* x0, const double* x1, const double* x2, const double* v0, const double* v1,
             const double* v2, const double* inv, double area, const Mat& mat, double* out) {
  const double a = inv[0], b = inv[1], c = inv[2], d = inv[3];
  const double e1x = x1[0] - x0[0], e1y = x1[1] - x0[1], e1z = x1[2] - x0[2];
  const double e2x = x2[0] - x0[0], e2y = x2[1] - x0[1], e2z = x2[2] - x0[2];
  const double w1x = v1[0] - v0[0], w1y = v1[1] - v0[1], w1z = v1[2] - v0[2];
  const double w2x = v2[0] - v0[0], w2y = v2[1] - v0[1], w2z = v2[2] - v0[2];
  const double F00 = a * e1x + c * e2x, F10 = a * e1y + c * e2y, F20 = a * e1z + c * e2z;
  const double F01 = b * e1x + d * e2x, F11 = b * e1y + d * e2y, F21 = b * e1z + d * e2z;
  const double dF00 = a * w1x + c * w2x, dF10 = a * w1y + c * w2y, dF20 = a * w1z + c * w2z;
  const double dF01 = b * w1x + d * w2x, dF11 = b * w1y + d * w2y, dF21 = b * w1z + d * w2z;
  const double C00k = mat.c00, C11k = mat.c11, C01k = mat.c01, G = mat.g;
  const double E00 = 0.5 * (F00 * F00 + F10 * F10 + F20 * F20 - 1);
  const double E11 = 0.5 * (F01 * F01 + F11 * F11 + F21 * F21 - 1);
  const double E01 = 0.5 * (F00 * F01 + F10 * F11 + F20 * F21);
  const double dC00 = 2 * (F00 * dF00 + F10 * dF10 + F20 * dF20);
  const double dC11 = 2 * (F01 * dF01 + F11 * dF11 + F21 * dF21);
  const double dC01 = dF00 * F01 + F00 * dF01 + dF10 * F11 + F10 * dF11 + dF20 * F21 + F20 * dF21;
  const double dE00 = 0.5 * dC00, dE11 = 0.5 * dC11, dE01 = 0.5 * dC01;
  const double S00 = C00k * E00 + C01k * E11;
  const double S11 = C11k * E11 + C01k * E00;
  const double S01 = 2 * G * E01;
  const double dS00 = C00k * dE00 + C01k * dE11;
  const double dS11 = C11k * dE11 + C01k * dE00;
  const double dS01 = 2 * G * dE01;
  const double dP00 = dF00 * S00 + F00 * dS00 + dF01 * S01 + F01 * dS01;
  const double dP10 = dF10 * S00 + F10 * dS00 + dF11 * S01 + F11 * dS01;
  const double dP20 = dF20 * S00 + F20 * dS00 + dF21 * S01 + F21 * dS01;
  const double dP01 = dF00 * S01 + F00 * dS01 + dF01 * S11 + F01 * dS11;
  const double dP11 = dF10 * S01 + F10 * dS01 + dF11 * S11 + F11 * dS11;
  const double dP21 = dF20 * S01 + F20 * dS01 + dF21 * S11 + F21 * dS11;
  const double s = mat.thick * area;
  const double P[6] = {dP00, dP10, dP20, dP01, dP11, dP21};
  for (int i = 0; i < 3; i++) {
    const double Pi0 = P[i], Pi1 = P[3 + i];
    const double g1 = s * (Pi0 * a + Pi1 * b);
    const double g2 = s * (Pi0 * c + Pi1 * d);
    const double g0 = -(g1 + g2);
    out[0 * 3 + i] = g0;
    out[1 * 3 + i] = g1;
    out[2 * 3 + i] = g2;
  }
}

// triangleHessian9() membrane-blocks.ts:21-98
void tri_hessian9(const double* x0, const double* x1, const double* x2, const double* inv, double area,
                  const Mat& mat, double* h) {
  const double a = inv[0], b = inv[1], c = inv[2], d = inv[3];
  const double* X[3] = {x0, x1, x2};
  const double K[3][2] = {{-(a + c), -(b + d)}, {a, b}, {c, d}};
  double F[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 2; j++)
      for (int v = 0; v < 3; v++) F[i][j] += K[v][j] * X[v][i];
  double C[2][2] = {{0, 0}, {0, 0}};
  for (int l = 0; l < 2; l++)
    for (int m = 0; m < 2; m++)
      for (int i = 0; i < 3; i++) C[l][m] += F[i][l] * F[i][m];
  const double E00 = 0.5 * (C[0][0] - 1), E11 = 0.5 * (C[1][1] - 1), E01 = 0.5 * C[0][1];
  const double C00k = mat.c00, C11k = mat.c11, C01k = mat.c01, G = mat.g;
  const double S00 = C00k * E00 + C01k * E11;
  const double S11 = C11k * E11 + C01k * E00;
  const double S01 = 2 * G * E01;
  const double S[2][2] = {{S00, S01}, {S01, S11}};
  double D[2][2][2][2];
  std::memset(D, 0, sizeof(D));
  D[0][0][0][0] = C00k;
  D[1][1][1][1] = C11k;
  D[0][0][1][1] = C01k; D[1][1][0][0] = C01k;
  D[0][1][0][1] = G; D[0][1][1][0] = G; D[1][0][0][1] = G; D[1][0][1][0] = G;
  const double s = mat.thick * area;
  double M[9][3][2];
  std::memset(M, 0, sizeof(M));
  for (int v = 0; v < 3; v++)
    for (int ra = 0; ra < 3; ra++)
      for (int j = 0; j < 2; j++) M[v * 3 + ra][ra][j] = K[v][j];
  for (int A = 0; A < 9; A++) {
    for (int B = 0; B < 9; B++) {
      double sum = 0;
      for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 2; j++) {
          const double mA = M[A][i][j];
          if (mA == 0) continue;
          double dP = 0;
          for (int k = 0; k < 2; k++) {
            dP += M[B][i][k] * S[k][j];
            double dE = 0;
            for (int l = 0; l < 2; l++) {
              for (int m2 = 0; m2 < 2; m2++) {
                const double dd = D[k][j][l][m2];
                if (dd == 0) continue;
                double dElm = 0;
                for (int p = 0; p < 3; p++) dElm += M[B][p][l] * F[p][m2] + F[p][l] * M[B][p][m2];
                dE += dd * 0.5 * dElm;
              }
            }
            dP += F[i][k] * dE;
          }
          sum += mA * dP;
        }
      }
      h[A * 9 + B] = s * sum;
    }
  }
}

// ----------------------------------------------------------- bending.ts

// hingeEnergy() bending.ts:7-16
double hinge_energy(const double* x, uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3, double rest,
                    double len, double asum, double k) {
  const double th = dihedral(x, v0, v1, v2, v3);
  const double d = th - rest;
  const double w = len * len / jmax(asum, 1e-12);
  return 0.5 * k * d * d * w;
}

// addHingeGradient() bending.ts:19-42 (FD eps 1e-7 on a private copy of x).
double add_hinge_gradient(std::vector<double>& tmp, double* grad, const Hinge& h, double k) {
  const double eps = 1e-7;
  const uint32_t verts[4] = {h.v0, h.v1, h.v2, h.v3};
  const double e = hinge_energy(tmp.data(), h.v0, h.v1, h.v2, h.v3, h.rest, h.len, h.asum, k);
  for (int vi = 0; vi < 4; vi++) {
    const uint32_t v = verts[vi];
    for (int i = 0; i < 3; i++) {
      const size_t id = (size_t)v * 3 + i;
      const double old = tmp[id];
      tmp[id] = old + eps;
      const double ep = hinge_energy(tmp.data(), h.v0, h.v1, h.v2, h.v3, h.rest, h.len, h.asum, k);
      tmp[id] = old - eps;
      const double em = hinge_energy(tmp.data(), h.v0, h.v1, h.v2, h.v3, h.rest, h.len, h.asum, k);
      tmp[id] = old;
      grad[id] += (ep - em) / (g_fault == ORC_FAULT_HINGE_HALF ? eps : 2 * eps);
    }
  }
  return e;
}

// --------------------------------------------------------------- fem.ts

inline void load3(const double* x, uint32_t i, double* o) {
  o[0] = x[(size_t)i * 3]; o[1] = x[(size_t)i * 3 + 1]; o[2] = x[(size_t)i * 3 + 2];
}
inline void inv4(const Mesh& mesh, size_t t, double* inv) {
  for (int k = 0; k < 4; k++) inv[k] = mesh.invDm[t * 4 + k];
}

// evalInternal() fem.ts:15-54.  `tmp` is the scratch copy addHingeGradient makes.
void eval_internal(const Mesh& mesh, const Mat& mat, const double* x, double* grad, double* energy_out,
                   double* max_strain_out) {
  const size_t n3 = (size_t)mesh.n * 3;
  std::fill(grad, grad + n3, 0.0);
  orcm::dd energy = orcm::dd_zero();
  double maxStrain = 0;
  for (size_t t = 0; t < mesh.m; t++) {
    const uint32_t i0 = mesh.idx[t * 3], i1 = mesh.idx[t * 3 + 1], i2 = mesh.idx[t * 3 + 2];
    double x0[3], x1[3], x2[3], inv[4];
    load3(x, i0, x0); load3(x, i1, x1); load3(x, i2, x2); inv4(mesh, t, inv);
    const TriGrad r = tri_energy_grad(x0, x1, x2, inv, mesh.areas[t], mat);
    orcm::dd_add(energy, r.energy);
    maxStrain = jmax(jmax(jmax(maxStrain, std::fabs(r.E00)), std::fabs(r.E11)), std::fabs(r.E01));
    const uint32_t vs[3] = {i0, i1, i2};
    for (int v = 0; v < 3; v++)
      for (int i = 0; i < 3; i++) grad[(size_t)vs[v] * 3 + i] += r.grad[v * 3 + i];
  }
  const double kb = 0.5 * (mat.bw + mat.bf);
  if (kb > 0) {
    std::vector<double> tmp(x, x + n3);
    for (const Hinge& h : mesh.hinges) orcm::dd_add(energy, add_hinge_gradient(tmp, grad, h, kb));
  }
  if (energy_out) *energy_out = orcm::dd_value(energy);
  if (max_strain_out) *max_strain_out = maxStrain;
}

// internalEnergyOnly() fem.ts:56-75
double internal_energy(const Mesh& mesh, const Mat& mat, const double* x) {
  orcm::dd e = orcm::dd_zero();
  for (size_t t = 0; t < mesh.m; t++) {
    const uint32_t i0 = mesh.idx[t * 3], i1 = mesh.idx[t * 3 + 1], i2 = mesh.idx[t * 3 + 2];
    double x0[3], x1[3], x2[3], inv[4];
    load3(x, i0, x0); load3(x, i1, x1); load3(x, i2, x2); inv4(mesh, t, inv);
    orcm::dd_add(e, tri_energy_grad(x0, x1, x2, inv, mesh.areas[t], mat).energy);
  }
  const double kb = 0.5 * (mat.bw + mat.bf);
  if (kb > 0) {
    for (const Hinge& h : mesh.hinges)
      orcm::dd_add(e, hinge_energy(x, h.v0, h.v1, h.v2, h.v3, h.rest, h.len, h.asum, kb));
  }
  return orcm::dd_value(e);
}

// evalMembrane() fem.ts:78-104 (gradient only; energy unused by callers here).
void eval_membrane(const Mesh& mesh, const Mat& mat, const double* x, double* grad) {
  std::fill(grad, grad + (size_t)mesh.n * 3, 0.0);
  for (size_t t = 0; t < mesh.m; t++) {
    const uint32_t i0 = mesh.idx[t * 3], i1 = mesh.idx[t * 3 + 1], i2 = mesh.idx[t * 3 + 2];
    double x0[3], x1[3], x2[3], inv[4];
    load3(x, i0, x0); load3(x, i1, x1); load3(x, i2, x2); inv4(mesh, t, inv);
    const TriGrad r = tri_energy_grad(x0, x1, x2, inv, mesh.areas[t], mat);
    const uint32_t vs[3] = {i0, i1, i2};
    for (int v = 0; v < 3; v++)
      for (int i = 0; i < 3; i++) grad[(size_t)vs[v] * 3 + i] += r.grad[v * 3 + i];
  }
}

// membraneHvp() fem.ts:108-130 (central FD, h = 1e-8 (1 + |p|)).
void membrane_hvp_fd(const Mesh& mesh, const Mat& mat, const double* x, const double* p, double* out) {
  const size_t n3 = (size_t)mesh.n * 3;
  const double eps = 1e-8;
  orcm::dd pn = orcm::dd_zero();
  for (size_t i = 0; i < n3; i++) orcm::dd_add_prod(pn, p[i], p[i]);
  const double pNorm = std::sqrt(orcm::dd_value(pn));
  const double h = eps * (1 + pNorm);
  std::vector<double> xp(n3), xm(n3), gp(n3), gm(n3);
  for (size_t i = 0; i < n3; i++) {
    xp[i] = x[i] + h * p[i];
    xm[i] = x[i] - h * p[i];
  }
  eval_membrane(mesh, mat, xp.data(), gp.data());
  eval_membrane(mesh, mat, xm.data(), gm.data());
  for (size_t i = 0; i < n3; i++) out[i] = (gp[i] - gm[i]) / (2 * h);
}

// evalMembraneHvp() membrane-hvp.ts:80-108
void membrane_hvp_analytic(const Mesh& mesh, const Mat& mat, const double* x, const double* p, double* out) {
  std::fill(out, out + (size_t)mesh.n * 3, 0.0);
  for (size_t t = 0; t < mesh.m; t++) {
    const uint32_t i0 = mesh.idx[t * 3], i1 = mesh.idx[t * 3 + 1], i2 = mesh.idx[t * 3 + 2];
    double x0[3], x1[3], x2[3], v0[3], v1[3], v2[3], inv[4], r[9];
    load3(x, i0, x0); load3(x, i1, x1); load3(x, i2, x2);
    load3(p, i0, v0); load3(p, i1, v1); load3(p, i2, v2); inv4(mesh, t, inv);
    tri_hvp(x0, x1, x2, v0, v1, v2, inv, mesh.areas[t], mat, r);
    const uint32_t vs[3] = {i0, i1, i2};
    for (int v = 0; v < 3; v++)
      for (int i = 0; i < 3; i++) out[(size_t)vs[v] * 3 + i] += r[v * 3 + i];
  }
}

// evalMembraneBlocks() membrane-blocks.ts:102-130
void membrane_blocks(const Mesh& mesh, const Mat& mat, const double* x, double* blocks) {
  std::fill(blocks, blocks + (size_t)mesh.n * 9, 0.0);
  double h[81];
  for (size_t t = 0; t < mesh.m; t++) {
    const uint32_t i0 = mesh.idx[t * 3], i1 = mesh.idx[t * 3 + 1], i2 = mesh.idx[t * 3 + 2];
    double x0[3], x1[3], x2[3], inv[4];
    load3(x, i0, x0); load3(x, i1, x1); load3(x, i2, x2); inv4(mesh, t, inv);
    tri_hessian9(x0, x1, x2, inv, mesh.areas[t], mat, h);
    const uint32_t vs[3] = {i0, i1, i2};
    for (int vc = 0; vc < 3; vc++)
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) blocks[(size_t)vs[vc] * 9 + a * 3 + b] += h[(vc * 3 + a) * 9 + (vc * 3 + b)];
  }
}

// ---------------------------------------------------- preconditioners

// buildBlockFactors() block-jacobi.ts:30-87
void build_block_factors(const Mesh& mesh, const Mat& mat, const double* x, const double* diag,
                         std::vector<double>& inv, std::vector<uint8_t>& flag) {
  const size_t n = mesh.n;
  std::vector<double> blocks(n * 9);
  membrane_blocks(mesh, mat, x, blocks.data());
  inv.assign(n * 9, 0.0);
  flag.assign(n, 0);
  for (size_t v = 0; v < n; v++) {
    const double* bl = &blocks[v * 9];
    const double h00 = bl[0], h01 = bl[1], h02 = bl[2];
    const double h10 = bl[3], h11 = bl[4], h12 = bl[5];
    const double h20 = bl[6], h21 = bl[7], h22 = bl[8];
    const double d0 = diag[v * 3], d1 = diag[v * 3 + 1], d2 = diag[v * 3 + 2];
    const double B00 = h00 + d0, B11 = h11 + d1, B22 = h22 + d2;
    const double B01 = 0.5 * (h01 + h10);
    const double B02 = 0.5 * (h02 + h20);
    const double B12 = 0.5 * (h12 + h21);
    const double tr = std::fabs(B00) + std::fabs(B11) + std::fabs(B22);
    const double tiny = 1e-20 * (1 + tr);
    bool ok = true;
    if (!(B00 > tiny)) ok = false;
    if (ok && !(B00 * B11 - B01 * B01 > tiny * (1 + std::fabs(B00) + std::fabs(B11)))) ok = false;
    const double det = B00 * (B11 * B22 - B12 * B12) - B01 * (B01 * B22 - B12 * B02) + B02 * (B01 * B12 - B11 * B02);
    if (ok && !(det > tiny * (1 + tr) * (1 + tr))) ok = false;
    double* o = &inv[v * 9];
    if (ok) {
      const double s = 1 / det;
      o[0] = (B11 * B22 - B12 * B12) * s;
      o[1] = (g_fault == ORC_FAULT_BJ_COFACTOR ? B01 * B22 - B02 * B12 : B02 * B12 - B01 * B22) * s;
      o[2] = (B01 * B12 - B11 * B02) * s;
      o[3] = (B02 * B12 - B01 * B22) * s;
      o[4] = (B00 * B22 - B02 * B02) * s;
      o[5] = (B01 * B02 - B00 * B12) * s;
      o[6] = (B01 * B12 - B11 * B02) * s;
      o[7] = (B01 * B02 - B00 * B12) * s;
      o[8] = (B00 * B11 - B01 * B01) * s;
      flag[v] = 1;
    } else {
      const double i0 = d0 > 1e-12 ? 1 / d0 : 0;
      const double i1 = d1 > 1e-12 ? 1 / d1 : 0;
      const double i2 = d2 > 1e-12 ? 1 / d2 : 0;
      o[0] = i0; o[1] = 0; o[2] = 0;
      o[3] = 0; o[4] = i1; o[5] = 0;
      o[6] = 0; o[7] = 0; o[8] = i2;
      flag[v] = 0;
    }
  }
}

// applyBlockFactors() block-jacobi.ts:90-104
void apply_block_factors(const std::vector<double>& inv, const double* r, double* out, size_t n) {
  for (size_t v = 0; v < n; v++) {
    const double r0 = r[v * 3], r1 = r[v * 3 + 1], r2 = r[v * 3 + 2];
    const double* f = &inv[v * 9];
    out[v * 3] = f[0] * r0 + f[1] * r1 + f[2] * r2;
    out[v * 3 + 1] = f[3] * r0 + f[4] * r1 + f[5] * r2;
    out[v * 3 + 2] = f[6] * r0 + f[7] * r1 + f[8] * r2;
  }
}

struct SchwarzDomains {
  std::vector<uint32_t> domainOf, localOf;
  std::vector<std::vector<uint32_t>> members;
};

// morton3() schwarz.ts:27-36
uint32_t morton3(double x, double y, double z) {
  auto q = [](double v) -> int {
    return (int)jmax(0, jmin(1023, std::floor(v * 1024)));
  };
  const int xi = q(x), yi = q(y), zi = q(z);
  uint32_t m = 0;
  for (int b = 9; b >= 0; b--) {
    m = (m << 3) | (uint32_t)((((xi >> b) & 1) << 2) | (((yi >> b) & 1) << 1) | ((zi >> b) & 1));
  }
  return m;
}

// buildSchwarzDomains() schwarz.ts:44-140
SchwarzDomains build_schwarz_domains(const Mesh& mesh, const float* rest, int targetSize, int minSize,
                                     int maxSize) {
  (void)maxSize;  // growth caps at targetSize <= maxSize by construction (schwarz.ts:138)
  const uint32_t n = mesh.n;
  std::vector<std::vector<uint32_t>> adj(n);
  for (size_t t = 0; t < mesh.m; t++) {
    const uint32_t a = mesh.idx[t * 3], b = mesh.idx[t * 3 + 1], c = mesh.idx[t * 3 + 2];
    adj[a].push_back(b); adj[a].push_back(c);
    adj[b].push_back(a); adj[b].push_back(c);
    adj[c].push_back(a); adj[c].push_back(b);
  }
  for (auto& l : adj) {
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
  }
  double mnx = INF, mny = INF, mnz = INF, mxx = -INF, mxy = -INF, mxz = -INF;
  for (uint32_t v = 0; v < n; v++) {
    mnx = jmin(mnx, rest[v * 3]); mny = jmin(mny, rest[v * 3 + 1]); mnz = jmin(mnz, rest[v * 3 + 2]);
    mxx = jmax(mxx, rest[v * 3]); mxy = jmax(mxy, rest[v * 3 + 1]); mxz = jmax(mxz, rest[v * 3 + 2]);
  }
  const double sx = jmax(mxx - mnx, 1e-12), sy = jmax(mxy - mny, 1e-12), sz = jmax(mxz - mnz, 1e-12);
  std::vector<uint32_t> code(n), order(n);
  for (uint32_t v = 0; v < n; v++) {
    order[v] = v;
    code[v] = morton3((rest[v * 3] - mnx) / sx, (rest[v * 3 + 1] - mny) / sy, (rest[v * 3 + 2] - mnz) / sz);
  }
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return code[a] != code[b] ? code[a] < code[b] : a < b;
  });
  const uint32_t NONE = 0xffffffffu;
  std::vector<uint32_t> domainOf(n, NONE);
  std::vector<std::vector<uint32_t>> members;
  for (uint32_t seed : order) {
    if (domainOf[seed] != NONE) continue;
    const uint32_t id = (uint32_t)members.size();
    std::vector<uint32_t> mem{seed};
    domainOf[seed] = id;
    std::vector<uint32_t> queue{seed};
    size_t qh = 0;
    while ((int)mem.size() < targetSize && qh < queue.size()) {
      const uint32_t cur = queue[qh++];
      for (uint32_t w : adj[cur]) {
        if ((int)mem.size() >= targetSize) break;
        if (domainOf[w] != NONE) continue;
        domainOf[w] = id;
        mem.push_back(w);
        queue.push_back(w);
      }
    }
    members.push_back(mem);
  }
  for (;;) {
    long small = -1;
    for (size_t d = 0; d < members.size(); d++) {
      if (!members[d].empty() && (int)members[d].size() < minSize && (int)n >= minSize) {
        if (small < 0 || members[d].size() < members[small].size()) small = (long)d;
      }
    }
    if (small < 0) break;
    std::vector<uint32_t> nb;
    for (size_t t = 0; t < mesh.m; t++) {
      const uint32_t ds[3] = {domainOf[mesh.idx[t * 3]], domainOf[mesh.idx[t * 3 + 1]], domainOf[mesh.idx[t * 3 + 2]]};
      bool has = ds[0] == (uint32_t)small || ds[1] == (uint32_t)small || ds[2] == (uint32_t)small;
      if (!has) continue;
      for (uint32_t b : ds)
        if (b != (uint32_t)small) nb.push_back(b);
    }
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    long best = -1;
    for (uint32_t c : nb) {
      if (members[c].empty()) continue;
      if (best < 0 || members[c].size() < members[best].size()) best = (long)c;
    }
    if (best < 0) break;
    for (uint32_t v : members[small]) {
      domainOf[v] = (uint32_t)best;
      members[best].push_back(v);
    }
    members[small].clear();
  }
  SchwarzDomains out;
  std::vector<uint32_t> remap(members.size(), NONE);
  for (size_t d = 0; d < members.size(); d++) {
    if (members[d].empty()) continue;
    remap[d] = (uint32_t)out.members.size();
    std::vector<uint32_t> mem = members[d];
    std::sort(mem.begin(), mem.end());
    out.members.push_back(mem);
  }
  out.domainOf.resize(n);
  out.localOf.resize(n);
  for (uint32_t v = 0; v < n; v++) out.domainOf[v] = remap[domainOf[v]];
  for (auto& mem : out.members)
    for (size_t l = 0; l < mem.size(); l++) out.localOf[mem[l]] = (uint32_t)l;
  return out;
}

struct SchwarzFactors {
  std::vector<std::vector<double>> data;
  std::vector<uint8_t> flag;
};

// buildSchwarzFactors() schwarz.ts:154-252
SchwarzFactors build_schwarz_factors(const Mesh& mesh, const Mat& mat, const double* x, const double* diag,
                                     const SchwarzDomains& dom, const std::vector<std::vector<uint32_t>>& v2t) {
  SchwarzFactors f;
  const size_t D = dom.members.size();
  f.flag.assign(D, 0);
  f.data.resize(D);
  double h[81];
  for (size_t d = 0; d < D; d++) {
    const auto& mem = dom.members[d];
    const size_t dv = mem.size(), dw = dv * 3;
    std::vector<double> A(dw * dw, 0.0);
    std::unordered_map<uint32_t, uint32_t> loc;
    for (size_t l = 0; l < dv; l++) loc[mem[l]] = (uint32_t)l;
    for (uint32_t v : mem) {
      for (uint32_t t : v2t[v]) {
        const uint32_t cvs[3] = {mesh.idx[t * 3], mesh.idx[t * 3 + 1], mesh.idx[t * 3 + 2]};
        int ci = cvs[0] == v ? 0 : cvs[1] == v ? 1 : cvs[2] == v ? 2 : -1;
        double x0[3], x1[3], x2[3], inv[4];
        load3(x, cvs[0], x0); load3(x, cvs[1], x1); load3(x, cvs[2], x2); inv4(mesh, t, inv);
        tri_hessian9(x0, x1, x2, inv, mesh.areas[t], mat, h);
        for (int cj = 0; cj < 3; cj++) {
          auto it = loc.find(cvs[cj]);
          if (it == loc.end()) continue;
          const uint32_t lw = it->second, lv = loc[v];
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) A[(lv * 3 + a) * dw + (lw * 3 + b)] += h[(ci * 3 + a) * 9 + (cj * 3 + b)];
        }
      }
    }
    for (size_t l = 0; l < dv; l++)
      for (int a = 0; a < 3; a++) A[(l * 3 + a) * dw + (l * 3 + a)] += diag[(size_t)mem[l] * 3 + a];
    for (size_t i = 0; i < dw; i++)
      for (size_t j = i + 1; j < dw; j++) {
        const double s = 0.5 * (A[i * dw + j] + A[j * dw + i]);
        A[i * dw + j] = s; A[j * dw + i] = s;
      }
    double scale = 0;
    for (size_t i = 0; i < dw; i++) scale = jmax(scale, std::fabs(A[i * dw + i]));
    const double tiny = 1e-20 * (1 + scale);
    bool ok = true;
    std::vector<double> L = A;
    for (size_t k = 0; k < dw && ok; k++) {
      double sum = L[k * dw + k];
      for (size_t s = 0; s < k; s++) sum -= L[k * dw + s] * L[k * dw + s];
      if (!(sum > tiny)) { ok = false; break; }
      L[k * dw + k] = std::sqrt(sum);
      for (size_t i = k + 1; i < dw; i++) {
        double s2 = L[i * dw + k];
        for (size_t s = 0; s < k; s++) s2 -= L[i * dw + s] * L[k * dw + s];
        L[i * dw + k] = s2 / L[k * dw + k];
      }
    }
    if (ok) {
      for (size_t i = 0; i < dw; i++)
        for (size_t j = i + 1; j < dw; j++) L[i * dw + j] = 0;
      f.data[d] = std::move(L);
      f.flag[d] = 1;
    } else {
      std::vector<double> J(dw * dw, 0.0);
      for (size_t l = 0; l < dv; l++)
        for (int a = 0; a < 3; a++) {
          const double dd = diag[(size_t)mem[l] * 3 + a];
          J[(l * 3 + a) * dw + (l * 3 + a)] = dd > 1e-12 ? 1 / dd : 0;
        }
      f.data[d] = std::move(J);
      f.flag[d] = 0;
    }
  }
  return f;
}

// solveLocal() + applySchwarz() schwarz.ts:255-309
void apply_schwarz(const SchwarzFactors& f, const SchwarzDomains& dom, const double* r, double* out, size_t n) {
  std::fill(out, out + n * 3, 0.0);
  std::vector<double> rhs, sol, y;
  for (size_t d = 0; d < dom.members.size(); d++) {
    const auto& mem = dom.members[d];
    const size_t dv = mem.size(), dw = dv * 3;
    rhs.assign(dw, 0.0); sol.assign(dw, 0.0); y.assign(dw, 0.0);
    for (size_t l = 0; l < dv; l++)
      for (int a = 0; a < 3; a++) rhs[l * 3 + a] = r[(size_t)mem[l] * 3 + a];
    const std::vector<double>& M = f.data[d];
    if (f.flag[d] == 1) {
      for (size_t i = 0; i < dw; i++) {
        double s = rhs[i];
        for (size_t j = 0; j < i; j++) s -= M[i * dw + j] * y[j];
        y[i] = s / M[i * dw + i];
      }
      for (size_t ii = dw; ii-- > 0;) {
        double s = y[ii];
        for (size_t j = ii + 1; j < dw; j++)
          s -= (g_fault == ORC_FAULT_SCHWARZ_BACK_ROW ? M[ii * dw + j] : M[j * dw + ii]) * sol[j];
        sol[ii] = s / M[ii * dw + ii];
      }
    } else {
      for (size_t i = 0; i < dw; i++) {
        double s = 0;
        for (size_t j = 0; j < dw; j++) s += M[i * dw + j] * rhs[j];
        sol[i] = s;
      }
    }
    for (size_t l = 0; l < dv; l++)
      for (int a = 0; a < 3; a++) out[(size_t)mem[l] * 3 + a] += sol[l * 3 + a];
  }
}

// ------------------------------------------- two-level Schwarz (coarse level)
// Coarse space: one constant vector per Schwarz domain and axis, normalised
// restriction R[(d*3+a)][v*3+a] = 1/sqrt(|d|) for v in d, P = R^T
// (mas.ts:1-13, coarse-csr.ts:1-13).

// restrictCoarse() mas.ts:23-40
void restrict_coarse(const double* r, const SchwarzDomains& dom, double* rc) {
  const size_t D = dom.members.size();
  std::fill(rc, rc + D * 3, 0.0);
  for (size_t d = 0; d < D; d++) {
    const double s = 1 / std::sqrt((double)dom.members[d].size());
    for (uint32_t v : dom.members[d]) {
      rc[d * 3] += r[(size_t)v * 3] * s;
      rc[d * 3 + 1] += r[(size_t)v * 3 + 1] * s;
      rc[d * 3 + 2] += r[(size_t)v * 3 + 2] * s;
    }
  }
}

// prolongateAdd() mas.ts:43-58
void prolongate_add(const double* c, const SchwarzDomains& dom, double* z) {
  for (size_t d = 0; d < dom.members.size(); d++) {
    const double s = 1 / std::sqrt((double)dom.members[d].size());
    for (uint32_t v : dom.members[d]) {
      z[(size_t)v * 3] += c[d * 3] * s;
      z[(size_t)v * 3 + 1] += c[d * 3 + 1] * s;
      z[(size_t)v * 3 + 2] += c[d * 3 + 2] * s;
    }
  }
}

// buildCoarseDiag() mas.ts:68-101: diag(R A P), intra-domain corner pairs of
// the membrane Hessian plus the Jacobi diagonal spread.
std::vector<double> build_coarse_diag(const Mesh& mesh, const Mat& mat, const double* x, const double* diag,
                                      const SchwarzDomains& dom) {
  const size_t D = dom.members.size();
  std::vector<double> dc(D * 3, 0.0);
  double h[81];
  for (size_t t = 0; t < mesh.m; t++) {
    const uint32_t cvs[3] = {mesh.idx[t * 3], mesh.idx[t * 3 + 1], mesh.idx[t * 3 + 2]};
    double x0[3], x1[3], x2[3], inv[4];
    load3(x, cvs[0], x0); load3(x, cvs[1], x1); load3(x, cvs[2], x2); inv4(mesh, t, inv);
    tri_hessian9(x0, x1, x2, inv, mesh.areas[t], mat, h);
    for (int ci = 0; ci < 3; ci++)
      for (int cj = 0; cj < 3; cj++) {
        const uint32_t d1 = dom.domainOf[cvs[ci]], d2 = dom.domainOf[cvs[cj]];
        if (d1 != d2) continue;
        const double sc = 1.0 / (double)dom.members[d1].size();
        for (int a = 0; a < 3; a++) dc[d1 * 3 + a] += sc * h[(ci * 3 + a) * 9 + (cj * 3 + a)];
      }
  }
  for (size_t d = 0; d < D; d++) {
    const double sc = 1.0 / (double)dom.members[d].size();
    for (uint32_t v : dom.members[d])
      for (int a = 0; a < 3; a++) dc[d * 3 + a] += sc * diag[(size_t)v * 3 + a];
  }
  return dc;
}

struct CoarsePattern {
  std::vector<uint32_t> off, col;  // block rows (D + 1), block columns sorted per row
};

// buildCoarsePattern() coarse-csr.ts:30-50: one block per pair of domains
// sharing a triangle.
CoarsePattern build_coarse_pattern(const Mesh& mesh, const SchwarzDomains& dom) {
  const size_t D = dom.members.size();
  std::vector<std::vector<uint32_t>> rows(D);
  for (size_t t = 0; t < mesh.m; t++) {
    const uint32_t ds[3] = {dom.domainOf[mesh.idx[t * 3]], dom.domainOf[mesh.idx[t * 3 + 1]],
                            dom.domainOf[mesh.idx[t * 3 + 2]]};
    for (uint32_t a : ds)
      for (uint32_t b : ds) rows[a].push_back(b);
  }
  CoarsePattern p;
  p.off.assign(D + 1, 0);
  for (size_t d = 0; d < D; d++) {
    auto& r = rows[d];
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    p.off[d + 1] = p.off[d] + (uint32_t)r.size();
    p.col.insert(p.col.end(), r.begin(), r.end());
  }
  return p;
}

// blockPos() coarse-csr.ts:53-62
int64_t block_pos(const CoarsePattern& p, uint32_t d1, uint32_t d2) {
  for (uint32_t q = p.off[d1]; q < p.off[d1 + 1]; q++) {
    if (p.col[q] == d2) return q;
    if (p.col[q] > d2) break;
  }
  return -1;
}

// assembleCoarseRaw() coarse-csr.ts:70-110: membrane corner-pair blocks scaled
// by 1/sqrt(|d1||d2|), then the Jacobi-diagonal spread on the diagonal blocks.
std::vector<double> assemble_coarse_raw(const Mesh& mesh, const Mat& mat, const double* x, const double* diag,
                                        const SchwarzDomains& dom, const CoarsePattern& pat) {
  std::vector<double> blocks(pat.col.size() * 9, 0.0);
  double h[81];
  for (size_t t = 0; t < mesh.m; t++) {
    const uint32_t cvs[3] = {mesh.idx[t * 3], mesh.idx[t * 3 + 1], mesh.idx[t * 3 + 2]};
    double x0[3], x1[3], x2[3], inv[4];
    load3(x, cvs[0], x0); load3(x, cvs[1], x1); load3(x, cvs[2], x2); inv4(mesh, t, inv);
    tri_hessian9(x0, x1, x2, inv, mesh.areas[t], mat, h);
    for (int ci = 0; ci < 3; ci++)
      for (int cj = 0; cj < 3; cj++) {
        const uint32_t d1 = dom.domainOf[cvs[ci]], d2 = dom.domainOf[cvs[cj]];
        const int64_t q = block_pos(pat, d1, d2);
        if (q < 0) throw std::runtime_error("coarse pattern misses a block");
        const double sc = 1 / std::sqrt((double)dom.members[d1].size() * (double)dom.members[d2].size());
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) blocks[q * 9 + a * 3 + b] += sc * h[(ci * 3 + a) * 9 + (cj * 3 + b)];
      }
  }
  for (size_t d = 0; d < dom.members.size(); d++) {
    const int64_t q = block_pos(pat, (uint32_t)d, (uint32_t)d);
    const double sc = 1.0 / (double)dom.members[d].size();
    for (uint32_t v : dom.members[d])
      for (int a = 0; a < 3; a++) blocks[q * 9 + a * 3 + a] += sc * diag[(size_t)v * 3 + a];
  }
  return blocks;
}

// assembleCoarseValues() coarse-csr.ts:117-138: B(d1,d2) = 0.5 (B + B(d2,d1)^T).
std::vector<double> symmetrize_coarse(const CoarsePattern& pat, const std::vector<double>& raw) {
  std::vector<double> out(raw);
  const size_t D = pat.off.size() - 1;
  for (uint32_t d1 = 0; d1 < D; d1++)
    for (uint32_t q = pat.off[d1]; q < pat.off[d1 + 1]; q++) {
      const int64_t w = block_pos(pat, pat.col[q], d1);
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) out[q * 9 + a * 3 + b] = 0.5 * (raw[q * 9 + a * 3 + b] + raw[w * 9 + b * 3 + a]);
    }
  return out;
}

// coarseSpmv() coarse-csr.ts:141-157
void coarse_spmv(const CoarsePattern& pat, const std::vector<double>& B, const double* x, double* out) {
  const size_t D = pat.off.size() - 1;
  std::fill(out, out + D * 3, 0.0);
  for (uint32_t d1 = 0; d1 < D; d1++)
    for (uint32_t q = pat.off[d1]; q < pat.off[d1 + 1]; q++) {
      const uint32_t d2 = pat.col[q];
      for (int a = 0; a < 3; a++) {
        double s = 0;
        for (int b = 0; b < 3; b++) s += B[q * 9 + a * 3 + b] * x[d2 * 3 + b];
        out[d1 * 3 + a] += s;
      }
    }
}

// invertSym33() coarse-csr.ts:170-205 (the block-Jacobi thresholds; diagonal
// fallback 1/fb).  Returns whether the block was positive definite.
bool invert_sym33(double B00, double B01, double B02, double B11, double B12, double B22, double fb0, double fb1,
                  double fb2, double* inv) {
  const double tr = std::fabs(B00) + std::fabs(B11) + std::fabs(B22);
  const double tiny = 1e-20 * (1 + tr);
  bool ok = true;
  if (!(B00 > tiny)) ok = false;
  if (ok && !(B00 * B11 - B01 * B01 > tiny * (1 + std::fabs(B00) + std::fabs(B11)))) ok = false;
  const double det = B00 * (B11 * B22 - B12 * B12) - B01 * (B01 * B22 - B12 * B02) + B02 * (B01 * B12 - B11 * B02);
  if (ok && !(det > tiny * (1 + tr) * (1 + tr))) ok = false;
  std::fill(inv, inv + 9, 0.0);
  if (ok) {
    const double s = 1 / det;
    inv[0] = (B11 * B22 - B12 * B12) * s;
    inv[1] = (B02 * B12 - B01 * B22) * s;
    inv[2] = (B01 * B12 - B11 * B02) * s;
    inv[3] = inv[1];
    inv[4] = (B00 * B22 - B02 * B02) * s;
    inv[5] = (B01 * B02 - B00 * B12) * s;
    inv[6] = inv[2];
    inv[7] = inv[5];
    inv[8] = (B00 * B11 - B01 * B01) * s;
    return true;
  }
  inv[0] = fb0 > 1e-12 ? 1 / fb0 : 0;
  inv[4] = fb1 > 1e-12 ? 1 / fb1 : 0;
  inv[8] = fb2 > 1e-12 ? 1 / fb2 : 0;
  return false;
}

// The diagonal-block inverse of domain d as applyCoarseBlockJacobi and
// coarsePcg read it (coarse-csr.ts:218-224, 257-263).
void coarse_diag_inverse(const CoarsePattern& pat, const std::vector<double>& B, const std::vector<double>& dc,
                         uint32_t d, double* inv) {
  const int64_t q = block_pos(pat, d, d);
  const double* b = &B[q * 9];
  invert_sym33(b[0], 0.5 * (b[1] + b[3]), 0.5 * (b[2] + b[6]), b[4], 0.5 * (b[5] + b[7]), b[8], dc[d * 3],
               dc[d * 3 + 1], dc[d * 3 + 2], inv);
}

void coarse_block_jacobi(const CoarsePattern& pat, const std::vector<double>& B, const std::vector<double>& dc,
                         const double* r, double* out) {
  const size_t D = pat.off.size() - 1;
  double inv[9];
  for (uint32_t d = 0; d < D; d++) {
    coarse_diag_inverse(pat, B, dc, d, inv);
    for (int a = 0; a < 3; a++)
      out[d * 3 + a] = inv[a * 3] * r[d * 3] + inv[a * 3 + 1] * r[d * 3 + 1] + inv[a * 3 + 2] * r[d * 3 + 2];
  }
}

// coarsePcg() coarse-csr.ts:243-303: fixed-K block-Jacobi PCG on Ac from x = 0
// (relative tolerance, pAp breakdown latch); sums left to right as b.reduce.
int coarse_pcg(const CoarsePattern& pat, const std::vector<double>& B, const std::vector<double>& dc,
               const double* b, int maxIters, double tol, double* x, double* residual_out) {
  const size_t n = (pat.off.size() - 1) * 3;
  std::vector<double> r(b, b + n), z(n), p(n), Ap(n);
  std::fill(x, x + n, 0.0);
  coarse_block_jacobi(pat, B, dc, r.data(), z.data());
  p = z;
  double rz = 0;
  for (size_t i = 0; i < n; i++) rz += r[i] * z[i];
  double bb = 0;
  for (size_t i = 0; i < n; i++) bb += b[i] * b[i];
  const double absTol = jmax(tol * jmax(std::sqrt(bb), 1e-12), 1e-12);
  auto norm = [&]() {
    double s = 0;
    for (size_t i = 0; i < n; i++) s += r[i] * r[i];
    return std::sqrt(s);
  };
  double residual = norm();
  if (residual_out) *residual_out = residual;
  if (residual < absTol) return 0;
  for (int k = 0; k < maxIters; k++) {
    coarse_spmv(pat, B, p.data(), Ap.data());
    double pAp = 0;
    for (size_t i = 0; i < n; i++) pAp += p[i] * Ap[i];
    if (!(pAp > 1e-30)) return k;
    const double alpha = rz / pAp;
    for (size_t i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    residual = norm();
    if (residual_out) *residual_out = residual;
    if (residual < absTol) return k + 1;
    coarse_block_jacobi(pat, B, dc, r.data(), z.data());
    double rzNew = 0;
    for (size_t i = 0; i < n; i++) rzNew += r[i] * z[i];
    const double beta = rzNew / rz;
    for (size_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    rz = rzNew;
  }
  return maxIters;
}

// buildCoarseExact() mas.ts:155-226: dense Ac = R A P, symmetrised, and its
// Cholesky factor (flag 1) or the Jacobi fallback (flag 0).
int build_coarse_exact(const Mesh& mesh, const Mat& mat, const double* x, const double* diag,
                       const SchwarzDomains& dom, std::vector<double>& ac, std::vector<double>& chol) {
  const size_t D = dom.members.size(), dc = D * 3;
  ac.assign(dc * dc, 0.0);
  double h[81];
  for (size_t t = 0; t < mesh.m; t++) {
    const uint32_t cvs[3] = {mesh.idx[t * 3], mesh.idx[t * 3 + 1], mesh.idx[t * 3 + 2]};
    double x0[3], x1[3], x2[3], inv[4];
    load3(x, cvs[0], x0); load3(x, cvs[1], x1); load3(x, cvs[2], x2); inv4(mesh, t, inv);
    tri_hessian9(x0, x1, x2, inv, mesh.areas[t], mat, h);
    for (int ci = 0; ci < 3; ci++)
      for (int cj = 0; cj < 3; cj++) {
        const uint32_t d1 = dom.domainOf[cvs[ci]], d2 = dom.domainOf[cvs[cj]];
        const double sc = 1 / std::sqrt((double)dom.members[d1].size() * (double)dom.members[d2].size());
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) ac[(d1 * 3 + a) * dc + (d2 * 3 + b)] += sc * h[(ci * 3 + a) * 9 + (cj * 3 + b)];
      }
  }
  for (size_t d = 0; d < D; d++) {
    const double sc = 1.0 / (double)dom.members[d].size();
    for (uint32_t v : dom.members[d])
      for (int a = 0; a < 3; a++) ac[(d * 3 + a) * dc + (d * 3 + a)] += sc * diag[(size_t)v * 3 + a];
  }
  for (size_t i = 0; i < dc; i++)
    for (size_t j = i + 1; j < dc; j++) {
      const double sv = 0.5 * (ac[i * dc + j] + ac[j * dc + i]);
      ac[i * dc + j] = sv; ac[j * dc + i] = sv;
    }
  chol = ac;
  double scale = 0;
  for (size_t i = 0; i < dc; i++) scale = jmax(scale, std::fabs(chol[i * dc + i]));
  const double tiny = 1e-20 * (1 + scale);
  bool ok = true;
  for (size_t k = 0; k < dc && ok; k++) {
    double sum = chol[k * dc + k];
    for (size_t s2 = 0; s2 < k; s2++) sum -= chol[k * dc + s2] * chol[k * dc + s2];
    if (!(sum > tiny)) { ok = false; break; }
    chol[k * dc + k] = std::sqrt(sum);
    for (size_t i = k + 1; i < dc; i++) {
      double v = chol[i * dc + k];
      for (size_t s2 = 0; s2 < k; s2++) v -= chol[i * dc + s2] * chol[k * dc + s2];
      chol[i * dc + k] = v / chol[k * dc + k];
    }
  }
  if (ok) {
    for (size_t i = 0; i < dc; i++)
      for (size_t j = i + 1; j < dc; j++) chol[i * dc + j] = 0;
    return 1;
  }
  chol.assign(dc * dc, 0.0);
  for (size_t i = 0; i < dc; i++) chol[i * dc + i] = ac[i * dc + i] > 1e-12 ? 1 / ac[i * dc + i] : 0;
  return 0;
}

// solveCoarseExact() mas.ts:229-250
void solve_coarse_exact(const std::vector<double>& chol, int flag, size_t dc, const double* rhs, double* out) {
  if (flag == 1) {
    std::vector<double> y(dc);
    for (size_t i = 0; i < dc; i++) {
      double s = rhs[i];
      for (size_t j = 0; j < i; j++) s -= chol[i * dc + j] * y[j];
      y[i] = s / chol[i * dc + i];
    }
    for (size_t i = dc; i-- > 0;) {
      double s = y[i];
      for (size_t j = i + 1; j < dc; j++) s -= chol[j * dc + i] * out[j];
      out[i] = s / chol[i * dc + i];
    }
  } else {
    for (size_t i = 0; i < dc; i++) {
      double s = 0;
      for (size_t j = 0; j < dc; j++) s += chol[i * dc + j] * rhs[j];
      out[i] = s;
    }
  }
}

// Two-level apply z = Schwarz1(r) + P C(R r) (applyMasTwoLevel mas.ts:126-143
// and the GPU path's C0 / C1 corrections, gpu-newton.ts:1493-1524):
// mode 1 damped Jacobi C = omega Dc^-1, mode 2 block-Jacobi C = Bdd^-1,
// mode 3 coarse PCG of `iters` iterations (tolerance 1e-3).
void apply_two_level(const SchwarzFactors& f, const SchwarzDomains& dom, const CoarsePattern& pat,
                     const std::vector<double>& B, const std::vector<double>& dc, int mode, int iters, double omega,
                     const double* r, double* z, size_t n) {
  apply_schwarz(f, dom, r, z, n);
  const size_t D = dom.members.size();
  std::vector<double> rc(D * 3), c(D * 3, 0.0);
  restrict_coarse(r, dom, rc.data());
  if (mode == 1) {
    for (size_t i = 0; i < D * 3; i++) c[i] = dc[i] > 1e-12 ? (omega * rc[i]) / dc[i] : 0;
  } else if (mode == 2) {
    coarse_block_jacobi(pat, B, dc, rc.data(), c.data());
  } else if (mode == 3) {
    coarse_pcg(pat, B, dc, rc.data(), iters, 1e-3, c.data(), nullptr);
  }
  prolongate_add(c.data(), dom, z);
}

// ------------------------------------------------------------ collision

struct Aabb {
  double minX, minY, minZ, maxX, maxY, maxZ;
};

// sweptTriAabb() aabb.ts:9-31
Aabb swept_tri_aabb(const double* x0, const double* x1, uint32_t i0, uint32_t i1, uint32_t i2, double pad) {
  double minX = INF, minY = INF, minZ = INF, maxX = -INF, maxY = -INF, maxZ = -INF;
  const uint32_t ids[3] = {i0, i1, i2};
  for (uint32_t v : ids) {
    const double* xs[2] = {x0, x1};
    for (const double* x : xs) {
      const double px = x[(size_t)v * 3], py = x[(size_t)v * 3 + 1], pz = x[(size_t)v * 3 + 2];
      if (px < minX) minX = px;
      if (py < minY) minY = py;
      if (pz < minZ) minZ = pz;
      if (px > maxX) maxX = px;
      if (py > maxY) maxY = py;
      if (pz > maxZ) maxZ = pz;
    }
  }
  return {minX - pad, minY - pad, minZ - pad, maxX + pad, maxY + pad, maxZ + pad};
}

// overlaps() aabb.ts:33-39
inline bool overlaps(const Aabb& a, const Aabb& b) {
  return a.minX <= b.maxX && a.maxX >= b.minX && a.minY <= b.maxY && a.maxY >= b.minY && a.minZ <= b.maxZ &&
         a.maxZ >= b.minZ;
}

// TriBvh bvh.ts:14-125 (median split on the longest centroid axis).
struct TriBvh {
  struct Node {
    Aabb aabb;
    int left, right, tri;
  };
  std::vector<Node> nodes;
  int root = -1;
  std::vector<Aabb> boxes;
  std::vector<double> cx, cy, cz;
  std::vector<uint32_t> order;

  void build(const double* x0, const double* x1, const uint32_t* idx, uint32_t triCount, double pad) {
    boxes.resize(triCount); cx.resize(triCount); cy.resize(triCount); cz.resize(triCount);
    for (uint32_t t = 0; t < triCount; t++) {
      const uint32_t i0 = idx[t * 3], i1 = idx[t * 3 + 1], i2 = idx[t * 3 + 2];
      boxes[t] = swept_tri_aabb(x0, x1, i0, i1, i2, pad);
      cx[t] = (x0[i0 * 3] + x0[i1 * 3] + x0[i2 * 3] + x1[i0 * 3] + x1[i1 * 3] + x1[i2 * 3]) / 6;
      cy[t] = (x0[i0 * 3 + 1] + x0[i1 * 3 + 1] + x0[i2 * 3 + 1] + x1[i0 * 3 + 1] + x1[i1 * 3 + 1] + x1[i2 * 3 + 1]) / 6;
      cz[t] = (x0[i0 * 3 + 2] + x0[i1 * 3 + 2] + x0[i2 * 3 + 2] + x1[i0 * 3 + 2] + x1[i1 * 3 + 2] + x1[i2 * 3 + 2]) / 6;
    }
    nodes.clear();
    if (triCount == 0) { root = -1; return; }
    order.resize(triCount);
    for (uint32_t i = 0; i < triCount; i++) order[i] = i;
    nodes.reserve(2 * triCount);
    root = build_rec(0, triCount);
  }

  int build_rec(size_t lo, size_t hi) {
    if (hi - lo == 1) {
      nodes.push_back({boxes[order[lo]], -1, -1, (int)order[lo]});
      return (int)nodes.size() - 1;
    }
    double minX = INF, maxX = -INF, minY = INF, maxY = -INF, minZ = INF, maxZ = -INF;
    for (size_t k = lo; k < hi; k++) {
      const uint32_t t = order[k];
      if (cx[t] < minX) minX = cx[t];
      if (cx[t] > maxX) maxX = cx[t];
      if (cy[t] < minY) minY = cy[t];
      if (cy[t] > maxY) maxY = cy[t];
      if (cz[t] < minZ) minZ = cz[t];
      if (cz[t] > maxZ) maxZ = cz[t];
    }
    const double sx = maxX - minX, sy = maxY - minY, sz = maxZ - minZ;
    const int axis = sx >= sy && sx >= sz ? 0 : sy >= sz ? 1 : 2;
    const std::vector<double>& key = axis == 0 ? cx : axis == 1 ? cy : cz;
    std::sort(order.begin() + lo, order.begin() + hi, [&](uint32_t p, uint32_t q) {
      return key[p] != key[q] ? key[p] < key[q] : p < q;
    });
    const size_t len = hi - lo;
    const size_t mid = std::max<size_t>(1, len / 2);
    const int left = build_rec(lo, lo + mid);
    const int right = build_rec(lo + mid, hi);
    const Aabb a = nodes[left].aabb, b = nodes[right].aabb;
    Aabb u{jmin(a.minX, b.minX), jmin(a.minY, b.minY), jmin(a.minZ, b.minZ),
           jmax(a.maxX, b.maxX), jmax(a.maxY, b.maxY), jmax(a.maxZ, b.maxZ)};
    nodes.push_back({u, left, right, -1});
    return (int)nodes.size() - 1;
  }

  // selfPairs() bvh.ts:79-124.  Exclusion = triangle pairs sharing a vertex
  // (buildAdjacencyExclusions bvh.ts:133-151; identical to the key set for
  // fewer than 1000003 triangles).
  std::vector<std::pair<uint32_t, uint32_t>> self_pairs(const uint32_t* idx) const {
    std::vector<std::pair<uint32_t, uint32_t>> out;
    if (root < 0) return out;
    std::vector<std::pair<int, int>> stack{{root, root}};
    while (!stack.empty()) {
      const auto [na, nb] = stack.back();
      stack.pop_back();
      const Node& A = nodes[na];
      const Node& B = nodes[nb];
      if (!overlaps(A.aabb, B.aabb)) continue;
      const bool leafA = A.tri >= 0, leafB = B.tri >= 0;
      if (leafA && leafB) {
        const int a = A.tri, b = B.tri;
        if (a == b) continue;
        const uint32_t tA = (uint32_t)std::min(a, b), tB = (uint32_t)std::max(a, b);
        bool shared = false;
        for (int i = 0; i < 3 && !shared; i++)
          for (int j = 0; j < 3; j++)
            if (idx[tA * 3 + i] == idx[tB * 3 + j]) { shared = true; break; }
        if (shared) continue;
        out.push_back({tA, tB});
        continue;
      }
      if (na == nb) {
        stack.push_back({A.left, A.left});
        stack.push_back({A.left, A.right});
        stack.push_back({A.right, A.right});
      } else if (leafA) {
        stack.push_back({na, B.left});
        stack.push_back({na, B.right});
      } else if (leafB) {
        stack.push_back({A.left, nb});
        stack.push_back({A.right, nb});
      } else {
        stack.push_back({A.left, B.left});
        stack.push_back({A.left, B.right});
        stack.push_back({A.right, B.left});
        stack.push_back({A.right, B.right});
      }
    }
    std::sort(out.begin(), out.end());
    return out;
  }
};

struct Closest {
  double s, t, dist, rx, ry, rz;
};

// closestPointVertexTriangle() closest-point.ts:14-62 (Ericson 5.1.5).
Closest closest_vt(double px, double py, double pz, double ax, double ay, double az, double bx, double by,
                   double bz, double cx, double cy, double cz) {
  auto pack = [](double s, double t, double rx, double ry, double rz) {
    return Closest{s, t, hypot3(rx, ry, rz), rx, ry, rz};
  };
  const double abx = bx - ax, aby = by - ay, abz = bz - az;
  const double acx = cx - ax, acy = cy - ay, acz = cz - az;
  const double apx = px - ax, apy = py - ay, apz = pz - az;
  const double d1 = abx * apx + aby * apy + abz * apz;
  const double d2 = acx * apx + acy * apy + acz * apz;
  if (d1 <= 0 && d2 <= 0) return pack(0, 0, px - ax, py - ay, pz - az);
  const double bpx = px - bx, bpy = py - by, bpz = pz - bz;
  const double d3 = abx * bpx + aby * bpy + abz * bpz;
  const double d4 = acx * bpx + acy * bpy + acz * bpz;
  if (d3 >= 0 && d4 <= d3) return pack(1, 0, px - bx, py - by, pz - bz);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    const double v = d1 / (d1 - d3);
    return pack(v, 0, px - (ax + v * abx), py - (ay + v * aby), pz - (az + v * abz));
  }
  const double cpx = px - cx, cpy = py - cy, cpz = pz - cz;
  const double d5 = abx * cpx + aby * cpy + abz * cpz;
  const double d6 = acx * cpx + acy * cpy + acz * cpz;
  if (d6 >= 0 && d5 <= d6) return pack(0, 1, px - cx, py - cy, pz - cz);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) {
    const double w = d2 / (d2 - d6);
    return pack(0, w, px - (ax + w * acx), py - (ay + w * acy), pz - (az + w * acz));
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0) {
    const double w = (d4 - d3) / (d4 - d3 + (d5 - d6));
    const double qx = bx + w * (cx - bx), qy = by + w * (cy - by), qz = bz + w * (cz - bz);
    return pack(1 - w, w, px - qx, py - qy, pz - qz);
  }
  const double denom = 1 / (va + vb + vc);
  const double v = vb * denom;
  const double w = vc * denom;
  const double qx = ax + abx * v + acx * w;
  const double qy = ay + aby * v + acy * w;
  const double qz = az + abz * v + acz * w;
  return pack(v, w, px - qx, py - qy, pz - qz);
}

inline double clamp01(double v, double lo, double hi) { return v < lo ? lo : v > hi ? hi : v; }

// closestPointEdgeEdge() closest-point.ts:73-108 (Ericson 5.1.9).
Closest closest_ee(double ax, double ay, double az, double bx, double by, double bz, double cx, double cy,
                   double cz, double dx, double dy, double dz) {
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
    s = 0; t = clamp01(f / e, 0, 1);
  } else {
    const double c = d1x * rx + d1y * ry + d1z * rz;
    if (e <= 1e-30) {
      t = 0; s = clamp01(-c / a, 0, 1);
    } else {
      const double b = g_fault == ORC_FAULT_EE_TRANSPOSE ? d1x * d1x + d1y * d2y + d1z * d2z
                                                         : d1x * d2x + d1y * d2y + d1z * d2z;
      const double denom = a * e - b * b;
      s = denom > 1e-30 ? clamp01((b * f - c * e) / denom, 0, 1) : 0;
      t = (b * s + f) / e;
      if (t < 0) { t = 0; s = clamp01(-c / a, 0, 1); }
      else if (t > 1) { t = 1; s = clamp01((b - c) / a, 0, 1); }
    }
  }
  const double axp = ax + d1x * s, ayp = ay + d1y * s, azp = az + d1z * s;
  const double cxp = cx + d2x * t, cyp = cy + d2y * t, czp = cz + d2z * t;
  const double rx2 = axp - cxp, ry2 = ayp - cyp, rz2 = azp - czp;
  return Closest{s, t, hypot3(rx2, ry2, rz2), rx2, ry2, rz2};
}

// solveQuadratic01() cubic.ts:51-66
int solve_quadratic01(double a2, double a1, double a0, double* out) {
  const double eps = 1e-12;
  if (std::fabs(a2) < eps) {
    if (std::fabs(a1) < eps) return 0;
    const double r = -a0 / a1;
    if (r >= 0 && r <= 1) { out[0] = r; return 1; }
    return 0;
  }
  const double disc = a1 * a1 - 4 * a2 * a0;
  if (disc < 0) return 0;
  const double sd = std::sqrt(disc);
  int k = 0;
  const double rs[2] = {(-a1 - sd) / (2 * a2), (-a1 + sd) / (2 * a2)};
  for (double r : rs)
    if (r >= 0 && r <= 1) out[k++] = r;
  std::sort(out, out + k);
  return k;
}

// solveCubic01() cubic.ts:4-49
int solve_cubic01(double a3, double a2, double a1, double a0, double* res) {
  const double eps = 1e-12;
  if (std::fabs(a3) < eps) return solve_quadratic01(a2, a1, a0, res);
  const double A = a2 / a3, B = a1 / a3, C = a0 / a3;
  const double sqA = A * A;
  const double p = (3 * B - sqA) / 3;
  const double q = (2 * sqA * A - 9 * A * B + 27 * C) / 27;
  double roots[3];
  int nr = 0;
  const double disc = (q * q) / 4 + (p * p * p) / 27;
  const double shift = A / 3;
  if (disc > eps) {
    const double sd = std::sqrt(disc);
    const double u = orcm::cbrt(-q / 2 + sd);
    const double v = orcm::cbrt(-q / 2 - sd);
    roots[nr++] = u + v - shift;
  } else if (std::fabs(disc) <= eps) {
    const double u = orcm::cbrt(-q / 2);
    roots[nr++] = 2 * u - shift;
    roots[nr++] = -u - shift;
  } else {
    const double r = std::sqrt(-(p * p * p) / 27);
    const double phi = orcm::acos(jmin(1, jmax(-1, -q / (2 * r))));
    const double s = 2 * orcm::cbrt(r);
    roots[nr++] = s * orcm::cos(phi / 3) - shift;
    roots[nr++] = s * orcm::cos((phi + 2 * M_PI) / 3) - shift;
    roots[nr++] = s * orcm::cos((phi + 4 * M_PI) / 3) - shift;
  }
  double out[3];
  int no = 0;
  for (int i = 0; i < nr; i++) {
    double r = roots[i];
    if (!std::isfinite(r)) continue;
    for (int k = 0; k < 4; k++) {
      const double f = ((a3 * r + a2) * r + a1) * r + a0;
      const double df = (3 * a3 * r + 2 * a2) * r + a1;
      if (std::fabs(df) < 1e-18) break;
      r -= f / df;
    }
    if (r >= -1e-9 && r <= 1 + 1e-9) out[no++] = jmin(1, jmax(0, r));
  }
  std::sort(out, out + no);
  int k = 0;
  for (int i = 0; i < no; i++)
    if (i == 0 || out[i] - out[i - 1] > 1e-9) res[k++] = out[i];
  return k;
}

inline void cross_v(double* o, const double* u, const double* v) {
  o[0] = u[1] * v[2] - u[2] * v[1];
  o[1] = u[2] * v[0] - u[0] * v[2];
  o[2] = u[0] * v[1] - u[1] * v[0];
}
inline double dot_v(const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; }

// vtCCD() ccd-vt.ts:20-115
double vt_ccd(const double* x0, const double* x1, size_t p, size_t a, size_t b, size_t c, double thickness) {
  const double d0 = closest_vt(x0[p * 3], x0[p * 3 + 1], x0[p * 3 + 2], x0[a * 3], x0[a * 3 + 1], x0[a * 3 + 2],
                               x0[b * 3], x0[b * 3 + 1], x0[b * 3 + 2], x0[c * 3], x0[c * 3 + 1], x0[c * 3 + 2]).dist;
  if (d0 <= thickness) return 0;
  const size_t vs[4] = {p, a, b, c};
  double mx = 0, my = 0, mz = 0;
  for (size_t v : vs) {
    mx += x1[v * 3] - x0[v * 3];
    my += x1[v * 3 + 1] - x0[v * 3 + 1];
    mz += x1[v * 3 + 2] - x0[v * 3 + 2];
  }
  mx /= 4; my /= 4; mz /= 4;
  double dev = 0;
  for (size_t v : vs) {
    const double dx = x1[v * 3] - x0[v * 3] - mx;
    const double dy = x1[v * 3 + 1] - x0[v * 3 + 1] - my;
    const double dz = x1[v * 3 + 2] - x0[v * 3 + 2] - mz;
    const double m = jmax(jmax(std::fabs(dx), std::fabs(dy)), std::fabs(dz));
    if (m > dev) dev = m;
  }
  if (d0 - 4 * dev > thickness) return INF;
  double p0[3], dp[3], a0[3], da[3], b0[3], db[3], c0[3], dc[3];
  for (int i = 0; i < 3; i++) {
    p0[i] = x0[p * 3 + i]; dp[i] = x1[p * 3 + i] - x0[p * 3 + i];
    a0[i] = x0[a * 3 + i]; da[i] = x1[a * 3 + i] - x0[a * 3 + i];
    b0[i] = x0[b * 3 + i]; db[i] = x1[b * 3 + i] - x0[b * 3 + i];
    c0[i] = x0[c * 3 + i]; dc[i] = x1[c * 3 + i] - x0[c * 3 + i];
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
  auto at = [&](double t, double* pp, double* aa, double* bb, double* cc) {
    for (int i = 0; i < 3; i++) {
      pp[i] = p0[i] + dp[i] * t; aa[i] = a0[i] + da[i] * t;
      bb[i] = b0[i] + db[i] * t; cc[i] = c0[i] + dc[i] * t;
    }
  };
  for (int k = 0; k < nr; k++) {
    const double t = roots[k];
    double pp[3], aa[3], bb[3], cc[3];
    at(t, pp, aa, bb, cc);
    const Closest cp = closest_vt(pp[0], pp[1], pp[2], aa[0], aa[1], aa[2], bb[0], bb[1], bb[2], cc[0], cc[1], cc[2]);
    if (cp.dist <= thickness) return t;
  }
  if (g_fault == ORC_FAULT_CCD_NO_FALLBACK) return INF;
  double minD = INF, minT = 0;
  const int N = 16;
  for (int k = 0; k <= N; k++) {
    const double t = (double)k / N;
    double pp[3], aa[3], bb[3], cc[3];
    at(t, pp, aa, bb, cc);
    const double d = closest_vt(pp[0], pp[1], pp[2], aa[0], aa[1], aa[2], bb[0], bb[1], bb[2], cc[0], cc[1], cc[2]).dist;
    if (d < minD) { minD = d; minT = t; }
  }
  if (minD <= thickness) return minT;
  return INF;
}

// eeCCD() ccd-ee.ts:24-101
double ee_ccd(const double* x0, const double* x1, size_t a, size_t b, size_t c, size_t d, double thickness) {
  double a0[3], da[3], b0[3], db[3], c0[3], dc[3], d0[3], dd[3];
  for (int i = 0; i < 3; i++) {
    a0[i] = x0[a * 3 + i]; da[i] = x1[a * 3 + i] - x0[a * 3 + i];
    b0[i] = x0[b * 3 + i]; db[i] = x1[b * 3 + i] - x0[b * 3 + i];
    c0[i] = x0[c * 3 + i]; dc[i] = x1[c * 3 + i] - x0[c * 3 + i];
    d0[i] = x0[d * 3 + i]; dd[i] = x1[d * 3 + i] - x0[d * 3 + i];
  }
  auto q = [&](double t) {
    double A[3], B[3], C[3], D[3];
    for (int i = 0; i < 3; i++) {
      A[i] = a0[i] + da[i] * t; B[i] = b0[i] + db[i] * t;
      C[i] = c0[i] + dc[i] * t; D[i] = d0[i] + dd[i] * t;
    }
    return closest_ee(A[0], A[1], A[2], B[0], B[1], B[2], C[0], C[1], C[2], D[0], D[1], D[2]);
  };
  const double q0 = q(0).dist;
  if (q0 <= thickness) return 0;
  const size_t vs[4] = {a, b, c, d};
  double mx = 0, my = 0, mz = 0;
  for (size_t v : vs) {
    mx += x1[v * 3] - x0[v * 3];
    my += x1[v * 3 + 1] - x0[v * 3 + 1];
    mz += x1[v * 3 + 2] - x0[v * 3 + 2];
  }
  mx /= 4; my /= 4; mz /= 4;
  double dev = 0;
  for (size_t v : vs) {
    const double dx = x1[v * 3] - x0[v * 3] - mx;
    const double dy = x1[v * 3 + 1] - x0[v * 3 + 1] - my;
    const double dz = x1[v * 3 + 2] - x0[v * 3 + 2] - mz;
    const double m = jmax(jmax(std::fabs(dx), std::fabs(dy)), std::fabs(dz));
    if (m > dev) dev = m;
  }
  if (q0 - 4 * dev > thickness) return INF;
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
    const Closest cp = q(t);
    if (cp.dist <= thickness && cp.s >= -1e-6 && cp.s <= 1 + 1e-6 && cp.t >= -1e-6 && cp.t <= 1 + 1e-6) return t;
  }
  if (g_fault == ORC_FAULT_CCD_NO_FALLBACK) return INF;
  double minD = INF, minT = 0;
  const int N = 16;
  for (int k = 0; k <= N; k++) {
    const double t = (double)k / N;
    const double dd2 = q(t).dist;
    if (dd2 < minD) { minD = dd2; minT = t; }
  }
  if (minD <= thickness) return minT;
  return INF;
}

// barrier.ts:4-24
double barrier_value(double d, double dHat) {
  if (d >= dHat) return 0;
  const double dc = jmax(d, 1e-12);
  const double t = dc - dHat;
  return -t * t * orcm::log(dc / dHat);
}
double barrier_grad_scalar(double d, double dHat, double kappa, double w) {
  if (d >= dHat) return 0;
  const double dc = jmax(d, 1e-12);
  const double t = dc - dHat;
  const double db = -2 * t * orcm::log(dc / dHat) - (t * t) / dc;
  const double r = kappa * w * db;
  return g_fault == ORC_FAULT_BARRIER_SIGN ? -r : r;
}
double barrier_normal_force(double d, double dHat, double kappa, double w) {
  return jmax(0, -barrier_grad_scalar(d, dHat, kappa, w));
}

// coulombForce() friction.ts:14-23
void coulomb_force(double ux, double uy, double uz, double nx, double ny, double nz, double lambdaN, double mu,
                   double eps, double* f) {
  const double vn = ux * nx + uy * ny + uz * nz;
  const double tx = ux - vn * nx, ty = uy - vn * ny, tz = uz - vn * nz;
  const double mag = std::sqrt(tx * tx + ty * ty + tz * tz + eps * eps);
  const double s = (mu * jmax(0, lambdaN)) / mag;
  f[0] = -s * tx; f[1] = -s * ty; f[2] = -s * tz;
}

const int K_VT = 0, K_EE = 1, K_FLOOR = 2;

struct Cand {
  int kind, p, a, b, c, d;
  bool operator<(const Cand& o) const {
    if (kind != o.kind) return kind < o.kind;
    if (p != o.p) return p < o.p;
    if (a != o.a) return a < o.a;
    if (b != o.b) return b < o.b;
    if (c != o.c) return c < o.c;
    return d < o.d;
  }
};

// Normalised lagged key: vtKey / eeKey / `floor:${v}` (contact-assembly.ts:607-616, 210).
struct LKey {
  int64_t k[5];
  bool operator==(const LKey& o) const { return std::memcmp(k, o.k, sizeof(k)) == 0; }
};
struct LKeyHash {
  size_t operator()(const LKey& x) const {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < 5; i++) { h ^= (uint64_t)x.k[i]; h *= 1099511628211ull; }
    return (size_t)h;
  }
};
LKey vt_key(int p, int a, int b, int c) {
  int t[3] = {a, b, c};
  std::sort(t, t + 3);
  return LKey{{0, p, t[0], t[1], t[2]}};
}
LKey ee_key(int a, int b, int c, int d) {
  const int e1a = std::min(a, b), e1b = std::max(a, b), e2a = std::min(c, d), e2b = std::max(c, d);
  if (e1a < e2a || (e1a == e2a && e1b <= e2b)) return LKey{{1, e1a, e1b, e2a, e2b}};
  return LKey{{1, e2a, e2b, e1a, e1b}};
}
LKey floor_key(int v) { return LKey{{2, v, 0, 0, 0}}; }

struct Frozen {
  int kind, p, a, b, c, d;
  double w0, w1, w2, w3, nx, ny, nz;
  LKey key;
};
struct Lagged {
  double nx, ny, nz, lambdaN;
};

cs_contact_diag empty_diag() {
  cs_contact_diag d;
  std::memset(&d, 0, sizeof(d));
  d.min_distance = INF;
  d.min_toi = INF;
  return d;
}

// ContactSystem contact-assembly.ts:46-595
struct Contact {
  cs_contact_params params;
  std::vector<uint32_t> clothIdx;
  uint32_t triCount = 0;
  bool hasFloor = false;
  double floorY = 0;
  std::vector<float> staticPos;
  std::vector<uint32_t> staticIdx;
  std::vector<double> xStep;
  std::unordered_map<LKey, Lagged, LKeyHash> lagged;
  std::vector<Frozen> active;
  std::vector<Cand> candidates;
  cs_contact_diag diag = empty_diag();
  TriBvh bvh;

  bool has_static() const { return !staticIdx.empty(); }

  // beginStep() :75-81
  void begin_step(const std::vector<double>& x0) {
    xStep = x0;
    lagged.clear();
    active.clear();
    candidates.clear();
    diag = empty_diag();
  }

  void add_rejection() { diag.rejected_line_search_steps++; }

  // buildCandidates() :87-121
  std::vector<Cand> build_candidates(const double* x, double pad) {
    bvh.build(xStep.data(), x, clothIdx.data(), triCount, pad);
    const auto pairs = bvh.self_pairs(clothIdx.data());
    std::vector<Cand> out;
    out.reserve(pairs.size() * 15);
    static const int TE[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    for (const auto& pr : pairs) {
      const uint32_t tA = pr.first, tB = pr.second;
      const int A[3] = {(int)clothIdx[tA * 3], (int)clothIdx[tA * 3 + 1], (int)clothIdx[tA * 3 + 2]};
      const int B[3] = {(int)clothIdx[tB * 3], (int)clothIdx[tB * 3 + 1], (int)clothIdx[tB * 3 + 2]};
      // expandTriPair() self-collision.ts:23-41
      for (int p : A) out.push_back({K_VT, p, B[0], B[1], B[2], -1});
      for (int p : B) out.push_back({K_VT, p, A[0], A[1], A[2], -1});
      for (const auto& ea : TE)
        for (const auto& eb : TE) {
          const int a = A[ea[0]], b = A[ea[1]], c = B[eb[0]], d = B[eb[1]];
          if (a == c || a == d || b == c || b == d) continue;
          out.push_back({K_EE, -1, a, b, c, d});
        }
    }
    if (has_static()) {
      const size_t sTri = staticIdx.size() / 3;
      std::vector<Aabb> sBoxes(sTri);
      for (size_t t = 0; t < sTri; t++) sBoxes[t] = static_tri_aabb(t, pad);
      const auto clothTris = query_bvh_all(sBoxes);
      for (const auto& cs : clothTris) {
        const uint32_t ct = cs.first, st = cs.second;
        const int A[3] = {(int)clothIdx[ct * 3], (int)clothIdx[ct * 3 + 1], (int)clothIdx[ct * 3 + 2]};
        const int B[3] = {(int)staticIdx[st * 3], (int)staticIdx[st * 3 + 1], (int)staticIdx[st * 3 + 2]};
        for (int p : A) out.push_back({K_VT, p, -(B[0] + 2), -(B[1] + 2), -(B[2] + 2), -2});
        const int be[3][2] = {{B[0], B[1]}, {B[1], B[2]}, {B[2], B[0]}};
        const int ae[3][2] = {{A[0], A[1]}, {A[1], A[2]}, {A[2], A[0]}};
        for (const auto& e : ae)
          for (const auto& f : be) out.push_back({K_EE, -1, e[0], e[1], -(f[0] + 2), -(f[1] + 2)});
      }
    }
    std::sort(out.begin(), out.end());
    diag.candidate_pairs = out.size();
    return out;
  }

  // staticTriAabb() :597-605
  Aabb static_tri_aabb(size_t t, double pad) const {
    const uint32_t i0 = staticIdx[t * 3], i1 = staticIdx[t * 3 + 1], i2 = staticIdx[t * 3 + 2];
    auto mn3 = [](double a, double b, double c) { return jmin(jmin(a, b), c); };
    auto mx3 = [](double a, double b, double c) { return jmax(jmax(a, b), c); };
    const float* s = staticPos.data();
    if (g_fault == ORC_FAULT_STATIC_NO_PAD) pad = 0;
    return {mn3(s[i0 * 3], s[i1 * 3], s[i2 * 3]) - pad, mn3(s[i0 * 3 + 1], s[i1 * 3 + 1], s[i2 * 3 + 1]) - pad,
            mn3(s[i0 * 3 + 2], s[i1 * 3 + 2], s[i2 * 3 + 2]) - pad, mx3(s[i0 * 3], s[i1 * 3], s[i2 * 3]) + pad,
            mx3(s[i0 * 3 + 1], s[i1 * 3 + 1], s[i2 * 3 + 1]) + pad, mx3(s[i0 * 3 + 2], s[i1 * 3 + 2], s[i2 * 3 + 2]) + pad};
  }

  // queryBvhAll() :124-140
  std::vector<std::pair<uint32_t, uint32_t>> query_bvh_all(const std::vector<Aabb>& sBoxes) const {
    std::vector<std::pair<uint32_t, uint32_t>> out;
    if (bvh.root < 0) return out;
    std::vector<int> stack;
    for (size_t s = 0; s < sBoxes.size(); s++) {
      stack.assign(1, bvh.root);
      while (!stack.empty()) {
        const int nn = stack.back();
        stack.pop_back();
        const auto& nd = bvh.nodes[nn];
        if (!overlaps(nd.aabb, sBoxes[s])) continue;
        if (nd.tri >= 0) { out.push_back({(uint32_t)nd.tri, (uint32_t)s}); continue; }
        stack.push_back(nd.left);
        stack.push_back(nd.right);
      }
    }
    std::sort(out.begin(), out.end());
    return out;
  }

  void pos_of(const double* x, int v, double* o) const {
    if (v >= 0) { load3(x, (uint32_t)v, o); return; }
    const size_t i = (size_t)(-v - 2);
    o[0] = staticPos[i * 3]; o[1] = staticPos[i * 3 + 1]; o[2] = staticPos[i * 3 + 2];
  }

  // updateActiveSet() :164-220
  void update_active_set(const double* x, size_t n) {
    const double dHat = params.d_hat;
    candidates = build_candidates(x, dHat);
    std::vector<Frozen> act;
    double minD = INF;
    for (const Cand& q : candidates) {
      if (q.kind == K_VT) {
        double P[3], A[3], B[3], C[3];
        pos_of(x, q.p, P); pos_of(x, q.a, A); pos_of(x, q.b, B); pos_of(x, q.c, C);
        const Closest cp = closest_vt(P[0], P[1], P[2], A[0], A[1], A[2], B[0], B[1], B[2], C[0], C[1], C[2]);
        if (cp.dist < minD) minD = cp.dist;
        if (cp.dist >= dHat || cp.dist < 1e-12) continue;
        const double nx = cp.rx / cp.dist, ny = cp.ry / cp.dist, nz = cp.rz / cp.dist;
        const double w1 = cp.s, w2 = cp.t, w0 = 1 - cp.s - cp.t;
        act.push_back({K_VT, q.p, q.a, q.b, q.c, -1, w0, w1, w2, 0, nx, ny, nz, vt_key(q.p, q.a, q.b, q.c)});
      } else {
        double A[3], B[3], C[3], D[3];
        pos_of(x, q.a, A); pos_of(x, q.b, B); pos_of(x, q.c, C); pos_of(x, q.d, D);
        const Closest cp = closest_ee(A[0], A[1], A[2], B[0], B[1], B[2], C[0], C[1], C[2], D[0], D[1], D[2]);
        if (cp.dist < minD) minD = cp.dist;
        if (cp.dist >= dHat || cp.dist < 1e-12) continue;
        const double nx = cp.rx / cp.dist, ny = cp.ry / cp.dist, nz = cp.rz / cp.dist;
        act.push_back({K_EE, -1, q.a, q.b, q.c, q.d, 1 - cp.s, cp.s, -(1 - cp.t), -cp.t, nx, ny, nz,
                       ee_key(q.a, q.b, q.c, q.d)});
      }
    }
    if (hasFloor) {
      for (size_t v = 0; v < n; v++) {
        const double d = x[v * 3 + 1] - floorY;
        if (d < minD) minD = d;
        if (d >= dHat) continue;
        act.push_back({K_FLOOR, (int)v, -1, -1, -1, -1, 0, 0, 0, 0, 0, 1, 0, floor_key((int)v)});
      }
    }
    active = std::move(act);
    diag.active_pairs = active.size();
    uint64_t nvt = 0, nee = 0, nfl = 0;
    for (const Frozen& c : active) {
      if (c.kind == K_VT) nvt++;
      else if (c.kind == K_EE) nee++;
      else nfl++;
    }
    diag.vt_pairs = nvt; diag.ee_pairs = nee; diag.floor_pairs = nfl;
    if (minD < diag.min_distance) diag.min_distance = minD;
  }

  // frozenDist() :222-258
  double frozen_dist(const double* x, const Frozen& c) const {
    if (c.kind == K_FLOOR) return x[(size_t)c.p * 3 + 1] - floorY;
    if (c.kind == K_VT) {
      double ddx, ddy, ddz;
      if (c.a >= 0) {
        const size_t a = c.a, b = c.b, cc = c.c, p = c.p;
        const double qx = c.w0 * x[a * 3] + c.w1 * x[b * 3] + c.w2 * x[cc * 3];
        const double qy = c.w0 * x[a * 3 + 1] + c.w1 * x[b * 3 + 1] + c.w2 * x[cc * 3 + 1];
        const double qz = c.w0 * x[a * 3 + 2] + c.w1 * x[b * 3 + 2] + c.w2 * x[cc * 3 + 2];
        ddx = x[p * 3] - qx; ddy = x[p * 3 + 1] - qy; ddz = x[p * 3 + 2] - qz;
      } else {
        const float* s = staticPos.data();
        const size_t ia = -c.a - 2, ib = -c.b - 2, ic = -c.c - 2, p = c.p;
        const double sx = c.w0 * s[ia * 3] + c.w1 * s[ib * 3] + c.w2 * s[ic * 3];
        const double sy = c.w0 * s[ia * 3 + 1] + c.w1 * s[ib * 3 + 1] + c.w2 * s[ic * 3 + 1];
        const double sz = c.w0 * s[ia * 3 + 2] + c.w1 * s[ib * 3 + 2] + c.w2 * s[ic * 3 + 2];
        ddx = x[p * 3] - sx; ddy = x[p * 3 + 1] - sy; ddz = x[p * 3 + 2] - sz;
      }
      return c.nx * ddx + c.ny * ddy + c.nz * ddz;
    }
    double A[3], B[3], C[3], D[3];
    pos_of(x, c.a, A); pos_of(x, c.b, B); pos_of(x, c.c, C); pos_of(x, c.d, D);
    const double rx = c.w0 * A[0] + c.w1 * B[0] + c.w2 * C[0] + c.w3 * D[0];
    const double ry = c.w0 * A[1] + c.w1 * B[1] + c.w2 * C[1] + c.w3 * D[1];
    const double rz = c.w0 * A[2] + c.w1 * B[2] + c.w2 * C[2] + c.w3 * D[2];
    return c.nx * rx + c.ny * ry + c.nz * rz;
  }

  // addDistGrad() :288-315
  static void add_dist_grad(double* grad, const Frozen& c, double g) {
    if (c.kind == K_FLOOR) { grad[(size_t)c.p * 3 + 1] += g; return; }
    if (c.kind == K_VT) {
      grad[(size_t)c.p * 3] += g * c.nx;
      grad[(size_t)c.p * 3 + 1] += g * c.ny;
      grad[(size_t)c.p * 3 + 2] += g * c.nz;
      const double ws[3] = {c.w0, c.w1, c.w2};
      const int vs[3] = {c.a, c.b, c.c};
      for (int k = 0; k < 3; k++) {
        if (vs[k] < 0) continue;
        grad[(size_t)vs[k] * 3] += -g * ws[k] * c.nx;
        grad[(size_t)vs[k] * 3 + 1] += -g * ws[k] * c.ny;
        grad[(size_t)vs[k] * 3 + 2] += -g * ws[k] * c.nz;
      }
      return;
    }
    const double ws[4] = {c.w0, c.w1, c.w2, c.w3};
    const int vs[4] = {c.a, c.b, c.c, c.d};
    for (int k = 0; k < 4; k++) {
      if (vs[k] < 0) continue;
      grad[(size_t)vs[k] * 3] += g * ws[k] * c.nx;
      grad[(size_t)vs[k] * 3 + 1] += g * ws[k] * c.ny;
      grad[(size_t)vs[k] * 3 + 2] += g * ws[k] * c.nz;
    }
  }

  // addFrictionForce() :318-341
  static void add_friction_force(double* grad, const Frozen& c, double fx, double fy, double fz) {
    if (c.kind == K_FLOOR || c.kind == K_VT) {
      grad[(size_t)c.p * 3] += fx; grad[(size_t)c.p * 3 + 1] += fy; grad[(size_t)c.p * 3 + 2] += fz;
      if (c.kind == K_FLOOR) return;
      const double ws[3] = {c.w0, c.w1, c.w2};
      const int vs[3] = {c.a, c.b, c.c};
      for (int k = 0; k < 3; k++) {
        if (vs[k] < 0) continue;
        grad[(size_t)vs[k] * 3] += -fx * ws[k];
        grad[(size_t)vs[k] * 3 + 1] += -fy * ws[k];
        grad[(size_t)vs[k] * 3 + 2] += -fz * ws[k];
      }
      return;
    }
    const int vv[4] = {c.a, c.b, c.c, c.d};
    const double ww[4] = {c.w0, c.w1, -c.w2, -c.w3};
    for (int k = 0; k < 4; k++) {
      if (vv[k] < 0) continue;
      grad[(size_t)vv[k] * 3] += fx * ww[k];
      grad[(size_t)vv[k] * 3 + 1] += fy * ww[k];
      grad[(size_t)vv[k] * 3 + 2] += fz * ww[k];
    }
  }

  // relativeSlip() :344-367 (VT reads D(c.d), Q16).
  void relative_slip(const double* x, const Frozen& c, double* u) const {
    auto D = [&](int v, double* o) {
      if (v < 0) { o[0] = o[1] = o[2] = 0; return; }
      for (int i = 0; i < 3; i++) o[i] = x[(size_t)v * 3 + i] - xStep[(size_t)v * 3 + i];
    };
    if (c.kind == K_FLOOR) { D(c.p, u); return; }
    if (c.kind == K_VT) {
      double dp[3], da[3], db[3], dc[3];
      D(c.p, dp); D(c.a, da); D(c.b, db); D(c.d, dc);
      for (int i = 0; i < 3; i++) u[i] = dp[i] - (c.w0 * da[i] + c.w1 * db[i] + c.w2 * dc[i]);
      return;
    }
    double da[3], db[3], dc[3], dd[3];
    D(c.a, da); D(c.b, db); D(c.c, dc); D(c.d, dd);
    for (int i = 0; i < 3; i++) u[i] = c.w0 * da[i] + c.w1 * db[i] + c.w2 * dc[i] + c.w3 * dd[i];
  }

  // energyGrad() :262-285
  double energy_grad(const double* x, double* grad, size_t n3) {
    const double dHat = params.d_hat, kappa = params.kappa, mu = params.friction_mu, eps = params.friction_eps;
    std::fill(grad, grad + n3, 0.0);
    orcm::dd energy = orcm::dd_zero(), fwork = orcm::dd_zero();
    for (const Frozen& c : active) {
      const double d = jmax(frozen_dist(x, c), 1e-12);
      if (d >= dHat) continue;
      orcm::dd_add(energy, kappa * barrier_value(d, dHat));
      const double g = barrier_grad_scalar(d, dHat, kappa, 1);
      add_dist_grad(grad, c, g);
      auto it = lagged.find(c.key);
      const double nx = it != lagged.end() ? it->second.nx : c.nx;
      const double ny = it != lagged.end() ? it->second.ny : c.ny;
      const double nz = it != lagged.end() ? it->second.nz : c.nz;
      const double lambdaN = it != lagged.end() ? it->second.lambdaN : barrier_normal_force(d, dHat, kappa, 1);
      double u[3], f[3];
      relative_slip(x, c, u);
      coulomb_force(u[0], u[1], u[2], nx, ny, nz, lambdaN, mu, eps, f);
      add_friction_force(grad, c, -f[0], -f[1], -f[2]);
      orcm::dd_add(fwork, f[0] * u[0] + f[1] * u[1] + f[2] * u[2]);
    }
    diag.barrier_energy = orcm::dd_value(energy);
    diag.friction_work += orcm::dd_value(fwork);
    return diag.barrier_energy;
  }

  // distJvp() :394-414
  static double dist_jvp(const Frozen& c, const double* v) {
    if (c.kind == K_FLOOR) return v[(size_t)c.p * 3 + 1];
    if (c.kind == K_VT) {
      const size_t p = c.p;
      double j = c.nx * v[p * 3] + c.ny * v[p * 3 + 1] + c.nz * v[p * 3 + 2];
      const double ws[3] = {c.w0, c.w1, c.w2};
      const int vs[3] = {c.a, c.b, c.c};
      for (int k = 0; k < 3; k++) {
        if (vs[k] < 0) continue;
        const size_t q = vs[k];
        j -= ws[k] * (c.nx * v[q * 3] + c.ny * v[q * 3 + 1] + c.nz * v[q * 3 + 2]);
      }
      return j;
    }
    const double ws[4] = {c.w0, c.w1, c.w2, c.w3};
    const int vs[4] = {c.a, c.b, c.c, c.d};
    double j = 0;
    for (int k = 0; k < 4; k++) {
      if (vs[k] < 0) continue;
      const size_t q = vs[k];
      j += ws[k] * (c.nx * v[q * 3] + c.ny * v[q * 3 + 1] + c.nz * v[q * 3 + 2]);
    }
    return j;
  }

  // applyHvp() :371-392
  void apply_hvp(const double* x, const double* v, double* out, size_t n3) const {
    const double dHat = params.d_hat, kappa = params.kappa;
    std::fill(out, out + n3, 0.0);
    std::vector<double> dd(active.size(), 0.0);
    for (size_t i = 0; i < active.size(); i++) {
      const Frozen& c = active[i];
      const double d = jmax(frozen_dist(x, c), 1e-12);
      if (d >= dHat) { dd[i] = 0; continue; }
      const double L = orcm::log(d / dHat);
      const double u = d - dHat;
      const double b2 = -(2 * L + (4 * u) / d - (u * u) / (d * d));
      const double coeff = kappa * jmax(b2, 0);
      dd[i] = coeff * dist_jvp(c, v);
    }
    for (size_t i = 0; i < active.size(); i++) {
      if (dd[i] == 0) continue;
      add_dist_grad(out, active[i], dd[i]);
    }
  }

  // meritEnergy() :419-428
  double merit_energy(const double* x) const {
    const double dHat = params.d_hat, kappa = params.kappa;
    orcm::dd e = orcm::dd_zero();
    for (const Frozen& c : active) {
      const double d = jmax(frozen_dist(x, c), 1e-12);
      if (d >= dHat) continue;
      orcm::dd_add(e, kappa * barrier_value(d, dHat));
    }
    return orcm::dd_value(e);
  }

  // addDiagEstimate() :436-471
  void add_diag_estimate(double* dg, const double* x) const {
    const double dHat = params.d_hat, kappa = params.kappa;
    for (const Frozen& c : active) {
      const double d = jmax(frozen_dist(x, c), 1e-12);
      if (d >= dHat) continue;
      const double L = orcm::log(d / dHat);
      const double u = d - dHat;
      const double b2 = jmax(-(2 * L + (4 * u) / d - (u * u) / (d * d)), 0);
      const double k = kappa * b2;
      if (c.kind == K_FLOOR) { dg[(size_t)c.p * 3 + 1] += k; continue; }
      const double n2[3] = {c.nx * c.nx, c.ny * c.ny, c.nz * c.nz};
      if (c.kind == K_VT) {
        for (int kk = 0; kk < 3; kk++) dg[(size_t)c.p * 3 + kk] += k * n2[kk];
        const double ws[3] = {c.w0, c.w1, c.w2};
        const int vs[3] = {c.a, c.b, c.c};
        for (int j = 0; j < 3; j++) {
          if (vs[j] < 0) continue;
          for (int kk = 0; kk < 3; kk++) dg[(size_t)vs[j] * 3 + kk] += k * ws[j] * ws[j] * n2[kk];
        }
        continue;
      }
      const double ws[4] = {c.w0, c.w1, c.w2, c.w3};
      const int vs[4] = {c.a, c.b, c.c, c.d};
      for (int j = 0; j < 4; j++) {
        if (vs[j] < 0) continue;
        for (int kk = 0; kk < 3; kk++) dg[(size_t)vs[j] * 3 + kk] += k * ws[j] * ws[j] * n2[kk];
      }
    }
  }

  // commit() :475-485
  void commit(const double* x) {
    const double dHat = params.d_hat, kappa = params.kappa;
    for (const Frozen& c : active) {
      const double d = jmax(frozen_dist(x, c), 1e-12);
      if (d >= dHat) continue;
      lagged[c.key] = Lagged{c.nx, c.ny, c.nz, barrier_normal_force(d, dHat, kappa, 1)};
    }
  }

  double exact_dist(const double* x, const Cand& q) const {
    if (q.kind == K_VT) {
      double P[3], A[3], B[3], C[3];
      pos_of(x, q.p, P); pos_of(x, q.a, A); pos_of(x, q.b, B); pos_of(x, q.c, C);
      return closest_vt(P[0], P[1], P[2], A[0], A[1], A[2], B[0], B[1], B[2], C[0], C[1], C[2]).dist;
    }
    double A[3], B[3], C[3], D[3];
    pos_of(x, q.a, A); pos_of(x, q.b, B); pos_of(x, q.c, C); pos_of(x, q.d, D);
    return closest_ee(A[0], A[1], A[2], B[0], B[1], B[2], C[0], C[1], C[2], D[0], D[1], D[2]).dist;
  }

  // checkTrial() :489-553 (static CCD via extended arrays, :556-580).
  bool check_trial(const double* xT, size_t n, double* minDistOut, double* minToiOut) {
    const double dMin = params.d_min;
    const std::vector<Cand> cands = build_candidates(xT, dMin);
    double minDist = INF, minTOI = INF;
    auto ret = [&](bool v) {
      if (minDistOut) *minDistOut = minDist;
      if (minToiOut) *minToiOut = minTOI;
      return v;
    };
    for (const Cand& q : cands) {
      const double d = exact_dist(xT, q);
      if (d < minDist) minDist = d;
      if (!std::isfinite(d)) { diag.ccd_failures++; return ret(false); }
    }
    if (hasFloor) {
      for (size_t v = 0; v < n; v++) {
        const double d = xT[v * 3 + 1] - floorY;
        if (d < minDist) minDist = d;
      }
    }
    if (minDist <= dMin) return ret(false);
    std::vector<double> X0, X1;
    const size_t ns = staticPos.size() / 3;
    if (has_static()) {
      X0.assign(xStep.begin(), xStep.end());
      X1.assign(xT, xT + n * 3);
      for (size_t i = 0; i < ns * 3; i++) { X0.push_back(staticPos[i]); X1.push_back(staticPos[i]); }
    }
    auto remap = [&](int v) -> size_t { return v >= 0 ? (size_t)v : n + (size_t)(-v - 2); };
    for (const Cand& q : cands) {
      double toi = INF;
      if (q.kind == K_VT) {
        if (q.p < 0 && q.a < 0) continue;
        if (q.p < 0 || q.a < 0) {
          toi = vt_ccd(X0.data(), X1.data(), remap(q.p), remap(q.a), remap(q.b), remap(q.c), dMin);
        } else {
          toi = vt_ccd(xStep.data(), xT, q.p, q.a, q.b, q.c, dMin);
        }
      } else {
        if (q.a < 0 && q.b < 0 && q.c < 0 && q.d < 0) continue;
        if (q.a < 0 || q.b < 0 || q.c < 0 || q.d < 0) {
          toi = ee_ccd(X0.data(), X1.data(), remap(q.a), remap(q.b), remap(q.c), remap(q.d), dMin);
        } else {
          toi = ee_ccd(xStep.data(), xT, q.a, q.b, q.c, q.d, dMin);
        }
      }
      if (std::isnan(toi)) { diag.ccd_failures++; return ret(false); }
      if (toi < minTOI) minTOI = toi;
      if (toi > 0 && toi < 1 - 1e-9) return ret(false);
    }
    if (hasFloor) {
      for (size_t v = 0; v < n; v++) {
        const double y0 = xStep[v * 3 + 1] - floorY;
        const double yT = xT[v * 3 + 1] - floorY;
        if (y0 - std::fabs(yT - y0) > dMin) continue;
        if (y0 > dMin && yT <= dMin && y0 > yT) {
          const double toi = (y0 - dMin) / (y0 - yT);
          if (toi < minTOI) minTOI = toi;
          if (toi < 1 - 1e-9) return ret(false);
        }
      }
    }
    if (minDist < diag.min_distance) diag.min_distance = minDist;
    if (minTOI < diag.min_toi) diag.min_toi = minTOI;
    return ret(true);
  }
};

// ------------------------------------------------------------- solver

struct Solver {
  Mesh mesh;
  Mat mat;
  double gravity[3];
  std::vector<std::pair<uint32_t, std::array<double, 3>>> pinned;  // Map insertion order
  std::vector<double> positions, velocities;
  bool hasContact = false;
  Contact contact;
  cs_options opt;
  cs_step_stats last;
  bool hasLast = false;
  SchwarzDomains domains;
  std::vector<std::vector<uint32_t>> v2t;
  // Accepted Armijo trials of the last step: (energy before, energy after,
  // alpha, g.dx) per Newton iteration that moved (newton.ts:204-208).
  std::vector<std::array<double, 4>> armijo;

  long pin_index(uint32_t id) const {
    for (size_t i = 0; i < pinned.size(); i++)
      if (pinned[i].first == id) return (long)i;
    return -1;
  }
  // enforcePins() scene.ts:38-44
  void enforce_pins(double* x) const {
    for (const auto& pr : pinned) {
      x[(size_t)pr.first * 3] = pr.second[0];
      x[(size_t)pr.first * 3 + 1] = pr.second[1];
      x[(size_t)pr.first * 3 + 2] = pr.second[2];
    }
  }
};

// makePinFilter() scene.ts:46-55 (ids captured when the filter is made).
struct PinFilter {
  std::vector<uint32_t> ids;
  void operator()(double* v) const {
    for (uint32_t id : ids) { v[(size_t)id * 3] = 0; v[(size_t)id * 3 + 1] = 0; v[(size_t)id * 3 + 2] = 0; }
  }
};

double dot_self(const std::vector<double>& a) {
  orcm::dd s = orcm::dd_zero();
  for (double v : a) orcm::dd_add_prod(s, v, v);
  return orcm::dd_value(s);
}
double dot_cr(const double* a, const double* b, size_t n) {
  orcm::dd s = orcm::dd_zero();
  for (size_t i = 0; i < n; i++) orcm::dd_add_prod(s, a[i], b[i]);
  return orcm::dd_value(s);
}

// pcg() pcg.ts:11-95 (x0 = 0).  `precond` null => Jacobi on diag.
struct PcgResult {
  std::vector<double> x;
  int iters;
  double residual;
  bool converged;
};
template <typename Hvp, typename Pre>
PcgResult pcg(const std::vector<double>& b, Hvp hvp, const std::vector<double>& diag, int maxIters, double tol,
              const PinFilter& filter, Pre* precond) {
  const size_t n = b.size();
  std::vector<double> x(n, 0.0), r(b), z(n, 0.0), p(n, 0.0), Ap(n, 0.0);
  filter(r.data());
  auto apply_m = [&]() {
    if (precond) {
      (*precond)(r.data(), z.data());
    } else {
      for (size_t i = 0; i < n; i++) z[i] = diag[i] > 1e-12 ? r[i] / diag[i] : r[i];
    }
    filter(z.data());
  };
  apply_m();
  p = z;
  double rz = dot_cr(r.data(), z.data(), n);
  const double bNorm = std::sqrt(dot_self(b));
  const double absTol = jmax(tol * jmax(bNorm, 1e-12), 1e-12);
  double residual = std::sqrt(dot_self(r));
  if (residual < absTol) return {x, 0, residual, true};
  for (int k = 0; k < maxIters; k++) {
    filter(p.data());
    hvp(p.data(), Ap.data());
    filter(Ap.data());
    const double pAp = dot_cr(p.data(), Ap.data(), n);
    if (!(pAp > 1e-30)) return {x, k, residual, residual < absTol};
    const double alpha = rz / pAp;
    for (size_t i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    filter(x.data());
    filter(r.data());
    residual = std::sqrt(dot_self(r));
    if (residual < absTol) return {x, k + 1, residual, true};
    apply_m();
    const double rzNew = dot_cr(r.data(), z.data(), n);
    const double beta = g_fault == ORC_FAULT_PCG_BETA ? rz / rzNew : rzNew / rz;
    for (size_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    rz = rzNew;
  }
  return {x, maxIters, residual, false};
}

// killApproachVelocity() newton.ts:238-303
void kill_approach_velocity(Contact& contact, Solver& sc, const double* x) {
  contact.update_active_set(x, sc.mesh.n);
  double* v = sc.velocities.data();
  for (const Frozen& c : contact.active) {
    if (c.kind == K_FLOOR) {
      const double vy = v[(size_t)c.p * 3 + 1];
      if (vy < 0) v[(size_t)c.p * 3 + 1] = 0;
      continue;
    }
    if (c.kind == K_VT) {
      const size_t p = c.p;
      const double vp[3] = {v[p * 3], v[p * 3 + 1], v[p * 3 + 2]};
      double tx = 0, ty = 0, tz = 0;
      const double ws[3] = {c.w0, c.w1, c.w2};
      const int vs[3] = {c.a, c.b, c.c};
      for (int k = 0; k < 3; k++) {
        if (vs[k] < 0) continue;
        const size_t q = vs[k];
        tx += ws[k] * v[q * 3];
        ty += ws[k] * v[q * 3 + 1];
        tz += ws[k] * v[q * 3 + 2];
      }
      const double rel = (vp[0] - tx) * c.nx + (vp[1] - ty) * c.ny + (vp[2] - tz) * c.nz;
      if (rel < 0) {
        const double j = (g_fault == ORC_FAULT_KILL_VT_QUARTER ? -0.25 : -0.5) * rel;
        v[p * 3] += j * c.nx;
        v[p * 3 + 1] += j * c.ny;
        v[p * 3 + 2] += j * c.nz;
        for (int k = 0; k < 3; k++) {
          if (vs[k] < 0) continue;
          const size_t q = vs[k];
          v[q * 3] -= j * ws[k] * c.nx;
          v[q * 3 + 1] -= j * ws[k] * c.ny;
          v[q * 3 + 2] -= j * ws[k] * c.nz;
        }
      }
      continue;
    }
    auto avg = [&](int a, int b, double* o) {
      double xa[3] = {0, 0, 0}, xb[3] = {0, 0, 0};
      if (a >= 0) for (int i = 0; i < 3; i++) xa[i] = v[(size_t)a * 3 + i];
      if (b >= 0) for (int i = 0; i < 3; i++) xb[i] = v[(size_t)b * 3 + i];
      for (int i = 0; i < 3; i++) o[i] = (xa[i] + xb[i]) / 2;
    };
    double A[3], C[3];
    avg(c.a, c.b, A);
    avg(c.c, c.d, C);
    const double rel = (A[0] - C[0]) * c.nx + (A[1] - C[1]) * c.ny + (A[2] - C[2]) * c.nz;
    if (rel < 0) {
      const double j = -0.25 * rel;
      const int ab[2] = {c.a, c.b}, cd[2] = {c.c, c.d};
      for (int vv : ab) {
        if (vv < 0) continue;
        v[(size_t)vv * 3] += j * c.nx;
        v[(size_t)vv * 3 + 1] += j * c.ny;
        v[(size_t)vv * 3 + 2] += j * c.nz;
      }
      for (int vv : cd) {
        if (vv < 0) continue;
        v[(size_t)vv * 3] -= j * c.nx;
        v[(size_t)vv * 3 + 1] -= j * c.ny;
        v[(size_t)vv * 3 + 2] -= j * c.nz;
      }
    }
  }
}

struct NonFinite {};

// implicitStep() newton.ts:27-232 (+ the preconditioner hook pcg() exposes;
// factors are built once per Newton solve after the diagonal, as the device
// path does, gpu-newton.ts:1026-1048).
cs_step_stats implicit_step(Solver& sc, double h) {
  const int newtonIters = sc.opt.newton_iters;
  const int pcgMaxIters = sc.opt.pcg_max_iters;
  const double gradTol = sc.opt.grad_tol;
  const int lsIters = sc.opt.line_search_iters;
  const size_t n = sc.mesh.n, n3 = n * 3;
  const std::vector<float>& M = sc.mesh.masses;
  const std::vector<double> x0 = sc.positions;
  const std::vector<double>& v0 = sc.velocities;
  const double gx = sc.gravity[0], gy = sc.gravity[1], gz = sc.gravity[2];
  std::vector<double> yHat(n3);
  for (size_t i = 0; i < n; i++) {
    yHat[i * 3] = x0[i * 3] + h * v0[i * 3] + h * h * gx;
    yHat[i * 3 + 1] = x0[i * 3 + 1] + h * v0[i * 3 + 1] + h * h * gy;
    yHat[i * 3 + 2] = x0[i * 3 + 2] + h * v0[i * 3 + 2] + h * h * gz;
  }
  std::vector<double> x = yHat;
  sc.enforce_pins(x.data());
  sc.armijo.clear();
  PinFilter filter;
  for (const auto& pr : sc.pinned) filter.ids.push_back(pr.first);
  Contact* contact = sc.hasContact ? &sc.contact : nullptr;
  int lsTrials = 0;
  if (contact) {
    contact->begin_step(x0);
    if (!contact->check_trial(x.data(), n, nullptr, nullptr)) {
      double lo = 0, hi = 1;
      std::vector<double> probe(n3);
      for (int b = 0; b < 8; b++) {
        const double mid = 0.5 * (lo + hi);
        for (size_t i = 0; i < n3; i++) probe[i] = x0[i] + mid * (x[i] - x0[i]);
        sc.enforce_pins(probe.data());
        if (contact->check_trial(probe.data(), n, nullptr, nullptr)) lo = mid;
        else hi = mid;
        if (hi - lo < 1e-4) break;
      }
      for (size_t i = 0; i < n3; i++) x[i] = x0[i] + lo * (x[i] - x0[i]);
      sc.enforce_pins(x.data());
    }
  }
  auto totalEnergy = [&](const double* xx) {
    orcm::dd in = orcm::dd_zero();
    for (size_t i = 0; i < n; i++) {
      const double m = M[i];
      for (int k = 0; k < 3; k++) {
        const double d = xx[i * 3 + k] - yHat[i * 3 + k];
        orcm::dd_add(in, 0.5 * m * d * d / (h * h));
      }
    }
    double e = orcm::dd_value(in);
    e += internal_energy(sc.mesh, sc.mat, xx);
    if (contact) e += contact->merit_energy(xx);
    return e;
  };
  double energy = 0;
  {
    if (contact) contact->update_active_set(x.data(), n);
    energy = totalEnergy(x.data());
  }
  double gradNorm = INF, maxStrain = 0;
  int pcgItersTotal = 0, used = 0;
  bool converged = false;
  const double invH2 = 1 / (h * h);
  const double beta = jmax(jmax(sc.mat.c00, sc.mat.c11), sc.mat.g) * sc.mat.thick * 0.1 + 1e-6;
  std::vector<double> diag(n3);
  for (size_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) diag[i * 3 + k] = M[i] * invH2 + beta;
  std::vector<double> cHp(n3), evgrad(n3), ceg(n3), Hp(n3);
  std::vector<double> bjInv;
  std::vector<uint8_t> bjFlag;
  SchwarzFactors swf;
  for (int ni = 0; ni < newtonIters; ni++) {
    used++;
    double evEnergy;
    eval_internal(sc.mesh, sc.mat, x.data(), evgrad.data(), &evEnergy, &maxStrain);
    std::vector<double> g(n3);
    for (size_t i = 0; i < n3; i++) {
      const size_t vi = i / 3;
      g[i] = M[vi] * (x[i] - yHat[i]) * invH2 + evgrad[i];
    }
    if (contact) {
      contact->update_active_set(x.data(), n);
      contact->energy_grad(x.data(), ceg.data(), n3);
      for (size_t i = 0; i < n3; i++) g[i] += ceg[i];
      contact->add_diag_estimate(diag.data(), x.data());
    }
    filter(g.data());
    gradNorm = std::sqrt(dot_self(g));
    if (!std::isfinite(gradNorm)) throw NonFinite{};
    if (gradNorm < gradTol) { converged = true; break; }
    std::vector<double> b(n3);
    for (size_t i = 0; i < n3; i++) b[i] = -g[i];
    const bool analytic = sc.opt.hvp_mode == CS_HVP_ANALYTIC;
    auto hvpFull = [&](const double* p, double* out) {
      if (analytic) membrane_hvp_analytic(sc.mesh, sc.mat, x.data(), p, Hp.data());
      else membrane_hvp_fd(sc.mesh, sc.mat, x.data(), p, Hp.data());
      for (size_t i = 0; i < n3; i++) {
        const size_t vi = i / 3;
        out[i] = M[vi] * p[i] * invH2 + Hp[i];
      }
      if (contact) {
        contact->apply_hvp(x.data(), p, cHp.data(), n3);
        for (size_t i = 0; i < n3; i++) out[i] += cHp[i];
      }
    };
    PcgResult res;
    if (sc.opt.precond == CS_PRECOND_BLOCK_JACOBI) {
      build_block_factors(sc.mesh, sc.mat, x.data(), diag.data(), bjInv, bjFlag);
      auto pre = [&](const double* r, double* z) { apply_block_factors(bjInv, r, z, n); };
      res = pcg(b, hvpFull, diag, pcgMaxIters, 1e-3, filter, &pre);
    } else if (sc.opt.precond == CS_PRECOND_SCHWARZ) {
      swf = build_schwarz_factors(sc.mesh, sc.mat, x.data(), diag.data(), sc.domains, sc.v2t);
      auto pre = [&](const double* r, double* z) { apply_schwarz(swf, sc.domains, r, z, n); };
      res = pcg(b, hvpFull, diag, pcgMaxIters, 1e-3, filter, &pre);
    } else {
      struct NoPre { void operator()(const double*, double*) const {} };
      res = pcg(b, hvpFull, diag, pcgMaxIters, 1e-3, filter, (NoPre*)nullptr);
    }
    std::vector<double>& dx = res.x;
    pcgItersTotal += res.iters;
    double gtdx = dot_cr(g.data(), dx.data(), n3);
    if (!std::isfinite(gtdx) || gtdx >= 0) {
      for (size_t i = 0; i < n3; i++) dx[i] = -g[i] / diag[i];
      gtdx = dot_cr(g.data(), dx.data(), n3);
    }
    {
      double maxDx = 0;
      for (size_t i = 0; i < n3; i++) {
        const double m = std::fabs(dx[i]);
        if (m > maxDx) maxDx = m;
      }
      const double TRUST = 0.002;
      if (maxDx > TRUST) {
        const double s = TRUST / maxDx;
        for (size_t i = 0; i < n3; i++) dx[i] *= s;
        gtdx *= s;
      }
    }
    double alpha = 1;
    bool accepted = false;
    std::vector<double> xt(n3);
    for (int li = 0; li < lsIters; li++) {
      lsTrials++;
      for (size_t i = 0; i < n3; i++) xt[i] = x[i] + alpha * dx[i];
      sc.enforce_pins(xt.data());
      if (contact) {
        if (!contact->check_trial(xt.data(), n, nullptr, nullptr)) {
          contact->add_rejection();
          alpha *= 0.5;
          continue;
        }
      }
      const double eNew = totalEnergy(xt.data());
      if (!std::isfinite(eNew)) {
        if (contact) contact->add_rejection();
        alpha *= 0.5;
        continue;
      }
      if (eNew <= energy + 1e-4 * alpha * gtdx) {
        sc.armijo.push_back({energy, eNew, alpha, gtdx});
        x = xt;
        energy = eNew;
        accepted = true;
        if (contact) contact->commit(xt.data());
        break;
      }
      if (contact) contact->add_rejection();
      alpha *= 0.5;
    }
    if (!accepted) break;
    if (alpha < 1e-6) break;
  }
  const double damp = 1 - jmin(jmax(sc.mat.damp, 0), 0.1);
  for (size_t i = 0; i < n3; i++) sc.velocities[i] = ((x[i] - x0[i]) / h) * damp;
  if (contact) kill_approach_velocity(*contact, sc, x.data());
  sc.positions = x;
  sc.enforce_pins(sc.positions.data());
  cs_step_stats st;
  std::memset(&st, 0, sizeof(st));
  st.newton_iters = newtonIters;
  st.pcg_iters = pcgItersTotal;
  st.grad_norm = gradNorm;
  st.energy = energy;
  st.max_strain = maxStrain;
  st.converged = converged ? 1 : 0;
  st.has_contact = contact ? 1 : 0;
  st.contact = contact ? contact->diag : empty_diag();
  st.newton_used = used;
  st.line_search_trials = lsTrials;
  return st;
}

Mat mat_of(const cs_material& m) { return from_cs(m); }

}  // namespace

struct orc_solver {
  Solver s;
};

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }
void orc_set_fault(int fault) { g_fault = fault; }
double orc_js_hypot3(double a, double b, double c) { return hypot3(a, b, c); }

// buildGrid() mesh.ts:17-48
void orc_build_grid(int nx, int ny, double w, double h, float* pos, float* uv, uint32_t* idx) {
  for (int j = 0; j <= ny; j++) {
    for (int i = 0; i <= nx; i++) {
      const int id = j * (nx + 1) + i;
      const double x = ((double)i / nx) * w;
      const double z = ((double)j / ny) * h;
      pos[id * 3] = (float)x;
      pos[id * 3 + 1] = 0;
      pos[id * 3 + 2] = (float)z;
      uv[id * 2] = (float)x;
      uv[id * 2 + 1] = (float)z;
    }
  }
  size_t k = 0;
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      const uint32_t a = j * (nx + 1) + i, b = a + 1, c = a + (nx + 1), d = c + 1;
      idx[k++] = a; idx[k++] = c; idx[k++] = b;
      idx[k++] = b; idx[k++] = c; idx[k++] = d;
    }
  }
}

// preprocess() mesh.ts:50-97
int orc_preprocess(uint32_t n, uint32_t m, const float* pos, const float* uv, const uint32_t* indices,
                   double arealDensity, float* invDm, float* areas, float* masses, uint32_t* hv, double* rest,
                   double* len, double* asum, int max_hinges) {
  for (uint32_t i = 0; i < n; i++) masses[i] = 0;
  for (uint32_t t = 0; t < m; t++) {
    const uint32_t i0 = indices[t * 3], i1 = indices[t * 3 + 1], i2 = indices[t * 3 + 2];
    const double u0 = uv[i0 * 2], v0 = uv[i0 * 2 + 1];
    const double u1 = uv[i1 * 2], v1 = uv[i1 * 2 + 1];
    const double u2 = uv[i2 * 2], v2 = uv[i2 * 2 + 1];
    const double du1 = u1 - u0, dv1 = v1 - v0;
    const double du2 = u2 - u0, dv2 = v2 - v0;
    const double det = du1 * dv2 - du2 * dv1;
    if (std::fabs(det) < 1e-12) {
      g_err = "degenerate triangle " + std::to_string(t);
      return -1;
    }
    const double area = 0.5 * std::fabs(det);
    areas[t] = (float)area;
    invDm[t * 4] = (float)(dv2 / det);
    invDm[t * 4 + 1] = (float)(-du2 / det);
    invDm[t * 4 + 2] = (float)(-dv1 / det);
    invDm[t * 4 + 3] = (float)(du1 / det);
    const double tm = arealDensity * area / 3;
    masses[i0] = (float)((double)masses[i0] + tm);
    masses[i1] = (float)((double)masses[i1] + tm);
    masses[i2] = (float)((double)masses[i2] + tm);
  }
  std::vector<uint32_t> idx(indices, indices + (size_t)m * 3);
  std::vector<float> ar(areas, areas + m);
  const auto hinges = build_hinges(idx, pos, ar);
  if ((int)hinges.size() > max_hinges) {
    g_err = "hinge buffer too small";
    return -1;
  }
  for (size_t k = 0; k < hinges.size(); k++) {
    hv[k * 4] = hinges[k].v0; hv[k * 4 + 1] = hinges[k].v1; hv[k * 4 + 2] = hinges[k].v2; hv[k * 4 + 3] = hinges[k].v3;
    rest[k] = hinges[k].rest; len[k] = hinges[k].len; asum[k] = hinges[k].asum;
  }
  return (int)hinges.size();
}

double orc_dihedral(const double* x, uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3) {
  return dihedral(x, v0, v1, v2, v3);
}

void orc_tri_energy_grad(const double* x0, const double* x1, const double* x2, const double* inv, double area,
                         const cs_material* mat, double* energy, double* grad9, double* E3) {
  const TriGrad r = tri_energy_grad(x0, x1, x2, inv, area, mat_of(*mat));
  if (energy) *energy = r.energy;
  if (grad9) std::memcpy(grad9, r.grad, sizeof(r.grad));
  if (E3) { E3[0] = r.E00; E3[1] = r.E11; E3[2] = r.E01; }
}
void orc_tri_hvp(const double* x0, const double* x1, const double* x2, const double* v0, const double* v1,
                 const double* v2, const double* inv, double area, const cs_material* mat, double* out9) {
  tri_hvp(x0, x1, x2, v0, v1, v2, inv, area, mat_of(*mat), out9);
}
void orc_tri_hessian9(const double* x0, const double* x1, const double* x2, const double* inv, double area,
                      const cs_material* mat, double* h81) {
  tri_hessian9(x0, x1, x2, inv, area, mat_of(*mat), h81);
}
double orc_hinge_energy(const double* x, uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3, double rest,
                        double len, double asum, double k) {
  return hinge_energy(x, v0, v1, v2, v3, rest, len, asum, k);
}

void orc_closest_vt(const double* p, const double* a, const double* b, const double* c, double* o) {
  const Closest r = closest_vt(p[0], p[1], p[2], a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]);
  o[0] = r.s; o[1] = r.t; o[2] = r.dist; o[3] = r.rx; o[4] = r.ry; o[5] = r.rz;
}
void orc_closest_ee(const double* a, const double* b, const double* c, const double* d, double* o) {
  const Closest r = closest_ee(a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], d[0], d[1], d[2]);
  o[0] = r.s; o[1] = r.t; o[2] = r.dist; o[3] = r.rx; o[4] = r.ry; o[5] = r.rz;
}
int orc_solve_cubic01(double a3, double a2, double a1, double a0, double* roots3) {
  return solve_cubic01(a3, a2, a1, a0, roots3);
}
double orc_vt_ccd(const double* x0, const double* x1, int p, int a, int b, int c, double thickness) {
  return vt_ccd(x0, x1, p, a, b, c, thickness);
}
double orc_ee_ccd(const double* x0, const double* x1, int a, int b, int c, int d, double thickness) {
  return ee_ccd(x0, x1, a, b, c, d, thickness);
}
double orc_barrier_value(double d, double dhat) { return barrier_value(d, dhat); }
double orc_barrier_grad(double d, double dhat, double kappa, double w) { return barrier_grad_scalar(d, dhat, kappa, w); }
void orc_coulomb(const double* u, const double* nn, double lambda_n, double mu, double eps, double* f3) {
  coulomb_force(u[0], u[1], u[2], nn[0], nn[1], nn[2], lambda_n, mu, eps, f3);
}

int64_t orc_self_pairs(const double* x0, const double* x1, const uint32_t* idx, uint32_t m, double pad,
                       uint32_t* pairs, int64_t cap) {
  TriBvh bvh;
  bvh.build(x0, x1, idx, m, pad);
  const auto out = bvh.self_pairs(idx);
  for (int64_t k = 0; k < (int64_t)out.size() && k < cap; k++) {
    pairs[k * 2] = out[k].first;
    pairs[k * 2 + 1] = out[k].second;
  }
  return (int64_t)out.size();
}

orc_solver* orc_create(const cs_scene_desc* d, const cs_options* opt) {
  if (!d || !opt) { g_err = "null argument"; return nullptr; }
  auto* o = new orc_solver;
  Solver& s = o->s;
  Mesh& me = s.mesh;
  me.n = d->n_verts;
  me.m = d->n_tris;
  me.idx.assign(d->indices, d->indices + (size_t)me.m * 3);
  me.invDm.assign(d->inv_dm, d->inv_dm + (size_t)me.m * 4);
  me.areas.assign(d->areas, d->areas + me.m);
  me.masses.assign(d->masses, d->masses + me.n);
  if (d->rest_positions) {
    me.restPos.assign(d->rest_positions, d->rest_positions + (size_t)me.n * 3);
  } else {
    me.restPos.resize((size_t)me.n * 3);
    for (size_t i = 0; i < (size_t)me.n * 3; i++) me.restPos[i] = (float)d->positions[i];
  }
  for (uint32_t k = 0; k < d->n_hinges; k++) {
    me.hinges.push_back({d->hinge_verts[k * 4], d->hinge_verts[k * 4 + 1], d->hinge_verts[k * 4 + 2],
                         d->hinge_verts[k * 4 + 3], d->hinge_rest_angle[k], d->hinge_edge_len[k],
                         d->hinge_area_sum[k]});
  }
  s.mat = mat_of(d->material);
  for (int i = 0; i < 3; i++) s.gravity[i] = d->gravity[i];
  for (uint32_t k = 0; k < d->n_pins; k++) {
    s.pinned.push_back({d->pin_ids[k], {d->pin_pos[k * 3], d->pin_pos[k * 3 + 1], d->pin_pos[k * 3 + 2]}});
  }
  s.positions.assign(d->positions, d->positions + (size_t)me.n * 3);
  s.velocities.assign(d->velocities, d->velocities + (size_t)me.n * 3);
  s.hasContact = d->contact_enabled != 0;
  if (s.hasContact) {
    Contact& c = s.contact;
    c.params = d->contact;
    c.clothIdx = me.idx;
    c.triCount = me.m;
    c.hasFloor = d->has_floor != 0;
    c.floorY = d->floor_y;
    if (d->n_static_tris > 0) {
      c.staticPos.assign(d->static_pos, d->static_pos + (size_t)d->n_static_verts * 3);
      c.staticIdx.assign(d->static_idx, d->static_idx + (size_t)d->n_static_tris * 3);
    }
  }
  s.opt = *opt;
  if (s.opt.precond == CS_PRECOND_SCHWARZ) {
    s.domains = build_schwarz_domains(me, me.restPos.data(), s.opt.schwarz_target, s.opt.schwarz_min,
                                      s.opt.schwarz_max);
    s.v2t.assign(me.n, {});
    for (uint32_t t = 0; t < me.m; t++)
      for (int k = 0; k < 3; k++) s.v2t[me.idx[t * 3 + k]].push_back(t);
  }
  return o;
}

void orc_destroy(orc_solver* s) { delete s; }

int orc_step(orc_solver* o, double dt) {
  try {
    o->s.last = implicit_step(o->s, dt);
    o->s.hasLast = true;
    return CS_OK;
  } catch (const NonFinite&) {
    g_err = "non-finite gradient \xe2\x80\x94 blow up; reduce dt/stiffness";
    return CS_E_NONFINITE;
  }
}

// ContactSystem.setStaticMesh (contact-assembly.ts:71-73): the arrays replaced.
int orc_set_static_mesh(orc_solver* o, const float* pos, uint32_t ns, const uint32_t* idx, uint32_t ms) {
  Solver& s = o->s;
  if (!s.hasContact) { g_err = "scene has no contact system"; return CS_E_INVALID; }
  s.contact.staticPos.assign(pos, pos + (size_t)ns * 3);
  s.contact.staticIdx.assign(idx, idx + (size_t)ms * 3);
  return CS_OK;
}

int orc_pin_vertex(orc_solver* o, uint32_t id, const double* pos) {
  Solver& s = o->s;
  if (id >= s.mesh.n) { g_err = "vertex id out of range"; return CS_E_INVALID; }
  std::array<double, 3> p;
  if (pos) p = {pos[0], pos[1], pos[2]};
  else p = {s.positions[(size_t)id * 3], s.positions[(size_t)id * 3 + 1], s.positions[(size_t)id * 3 + 2]};
  const long k = s.pin_index(id);
  if (k >= 0) s.pinned[k].second = p;
  else s.pinned.push_back({id, p});
  s.enforce_pins(s.positions.data());
  return CS_OK;
}

int orc_unpin_vertex(orc_solver* o, uint32_t id) {
  Solver& s = o->s;
  const long k = s.pin_index(id);
  if (k >= 0) s.pinned.erase(s.pinned.begin() + k);
  return CS_OK;
}

void orc_set_material(orc_solver* o, const cs_material* mat) { o->s.mat = mat_of(*mat); }
void orc_get_positions(orc_solver* o, double* out) { std::copy(o->s.positions.begin(), o->s.positions.end(), out); }
void orc_get_velocities(orc_solver* o, double* out) { std::copy(o->s.velocities.begin(), o->s.velocities.end(), out); }
void orc_set_state(orc_solver* o, const double* x, const double* v) {
  const size_t n3 = (size_t)o->s.mesh.n * 3;
  if (x) o->s.positions.assign(x, x + n3);
  if (v) o->s.velocities.assign(v, v + n3);
}
void orc_last_stats(orc_solver* o, cs_step_stats* out) {
  if (o->s.hasLast) *out = o->s.last;
  else std::memset(out, 0, sizeof(*out));
}

void orc_eval_internal(orc_solver* o, const double* x, double* grad, double* energy, double* max_strain) {
  eval_internal(o->s.mesh, o->s.mat, x, grad, energy, max_strain);
}
double orc_internal_energy(orc_solver* o, const double* x) { return internal_energy(o->s.mesh, o->s.mat, x); }
void orc_membrane_hvp_fd(orc_solver* o, const double* x, const double* p, double* out) {
  membrane_hvp_fd(o->s.mesh, o->s.mat, x, p, out);
}
void orc_membrane_hvp_analytic(orc_solver* o, const double* x, const double* p, double* out) {
  membrane_hvp_analytic(o->s.mesh, o->s.mat, x, p, out);
}
void orc_membrane_blocks(orc_solver* o, const double* x, double* blocks) {
  membrane_blocks(o->s.mesh, o->s.mat, x, blocks);
}

// applyHvp (contact-assembly.ts:371-392) of the active set updateActiveSet
// builds at x after beginStep(x0); returns the active contact count.
int64_t orc_contact_hvp(orc_solver* o, const double* x0, const double* x, const double* p, double* out) {
  Solver& s = o->s;
  if (!s.hasContact) return -1;
  const size_t n3 = (size_t)s.mesh.n * 3;
  s.contact.begin_step(std::vector<double>(x0, x0 + n3));
  s.contact.update_active_set(x, s.mesh.n);
  s.contact.apply_hvp(x, p, out, n3);
  return (int64_t)s.contact.active.size();
}

// The contact gradient (energyAndGradient, contact-assembly.ts:262-367) at y
// of the active set updateActiveSet froze at x after beginStep(x0).
int64_t orc_contact_grad_frozen(orc_solver* o, const double* x0, const double* x, const double* y, double* grad) {
  Solver& s = o->s;
  if (!s.hasContact) return -1;
  const size_t n3 = (size_t)s.mesh.n * 3;
  s.contact.begin_step(std::vector<double>(x0, x0 + n3));
  s.contact.update_active_set(x, s.mesh.n);
  s.contact.energy_grad(y, grad, n3);
  return (int64_t)s.contact.active.size();
}

// buildCandidates (contact-assembly.ts:87-121) at x with the given pad after beginStep(x0).
int64_t orc_candidates(orc_solver* o, const double* x0, const double* x, int pad_is_dmin, int32_t* ids6, int64_t cap) {
  Solver& s = o->s;
  if (!s.hasContact) return -1;
  const size_t n3 = (size_t)s.mesh.n * 3;
  s.contact.begin_step(std::vector<double>(x0, x0 + n3));
  const auto c = s.contact.build_candidates(x, pad_is_dmin ? s.contact.params.d_min : s.contact.params.d_hat);
  for (int64_t k = 0; k < (int64_t)c.size() && k < cap; k++) {
    const int v[6] = {c[k].kind, c[k].p, c[k].a, c[k].b, c[k].c, c[k].d};
    for (int j = 0; j < 6; j++) ids6[k * 6 + j] = v[j];
  }
  return (int64_t)c.size();
}

int64_t orc_contact_probe(orc_solver* o, const double* x0, const double* x, double pad_is_dmin, int32_t* ids6,
                          double* vals7, int64_t cap, int64_t* n_candidates) {
  Solver& s = o->s;
  if (!s.hasContact) return -1;
  const size_t n3 = (size_t)s.mesh.n * 3;
  s.contact.begin_step(std::vector<double>(x0, x0 + n3));
  if (pad_is_dmin != 0) {
    const auto c = s.contact.build_candidates(x, s.contact.params.d_min);
    if (n_candidates) *n_candidates = (int64_t)c.size();
    for (int64_t k = 0; k < (int64_t)c.size() && k < cap; k++) {
      const int v[6] = {c[k].kind, c[k].p, c[k].a, c[k].b, c[k].c, c[k].d};
      for (int j = 0; j < 6; j++) ids6[k * 6 + j] = v[j];
    }
    return (int64_t)c.size();
  }
  s.contact.update_active_set(x, s.mesh.n);
  if (n_candidates) *n_candidates = (int64_t)s.contact.candidates.size();
  const auto& a = s.contact.active;
  for (int64_t k = 0; k < (int64_t)a.size() && k < cap; k++) {
    const int v[6] = {a[k].kind, a[k].p, a[k].a, a[k].b, a[k].c, a[k].d};
    for (int j = 0; j < 6; j++) ids6[k * 6 + j] = v[j];
    const double w[7] = {a[k].w0, a[k].w1, a[k].w2, a[k].w3, a[k].nx, a[k].ny, a[k].nz};
    for (int j = 0; j < 7; j++) vals7[k * 7 + j] = w[j];
  }
  return (int64_t)a.size();
}

int orc_check_trial(orc_solver* o, const double* x0, const double* xT, double* min_dist, double* min_toi) {
  Solver& s = o->s;
  if (!s.hasContact) return -1;
  const size_t n3 = (size_t)s.mesh.n * 3;
  s.contact.begin_step(std::vector<double>(x0, x0 + n3));
  return s.contact.check_trial(xT, s.mesh.n, min_dist, min_toi) ? 1 : 0;
}

int orc_schwarz_domains(orc_solver* o, int target, int minsz, int maxsz, uint32_t* domain_of, uint32_t* local_of) {
  Solver& s = o->s;
  const SchwarzDomains d = build_schwarz_domains(s.mesh, s.mesh.restPos.data(), target, minsz, maxsz);
  std::copy(d.domainOf.begin(), d.domainOf.end(), domain_of);
  std::copy(d.localOf.begin(), d.localOf.end(), local_of);
  return (int)d.members.size();
}

int orc_schwarz_apply(orc_solver* o, int target, int minsz, int maxsz, const double* x, const double* diag,
                      const double* r, double* z, uint8_t* flags) {
  Solver& s = o->s;
  const SchwarzDomains d = build_schwarz_domains(s.mesh, s.mesh.restPos.data(), target, minsz, maxsz);
  std::vector<std::vector<uint32_t>> v2t(s.mesh.n);
  for (uint32_t t = 0; t < s.mesh.m; t++)
    for (int k = 0; k < 3; k++) v2t[s.mesh.idx[t * 3 + k]].push_back(t);
  const SchwarzFactors f = build_schwarz_factors(s.mesh, s.mat, x, diag, d, v2t);
  apply_schwarz(f, d, r, z, s.mesh.n);
  std::copy(f.flag.begin(), f.flag.end(), flags);
  return (int)d.members.size();
}

// ---- two-level Schwarz probes (mas.ts, coarse-csr.ts) ----
namespace {
SchwarzDomains probe_domains(orc_solver* o, int target, int minsz, int maxsz) {
  return build_schwarz_domains(o->s.mesh, o->s.mesh.restPos.data(), target, minsz, maxsz);
}
}  // namespace

// buildCoarsePattern: off (D + 1) and up to cap block columns; returns nnz.
int64_t orc_coarse_pattern(orc_solver* o, int target, int minsz, int maxsz, uint32_t* off, uint32_t* col,
                           int64_t cap) {
  const SchwarzDomains d = probe_domains(o, target, minsz, maxsz);
  const CoarsePattern p = build_coarse_pattern(o->s.mesh, d);
  if (off) std::copy(p.off.begin(), p.off.end(), off);
  if (col)
    for (int64_t q = 0; q < (int64_t)p.col.size() && q < cap; q++) col[q] = p.col[q];
  return (int64_t)p.col.size();
}

// assembleCoarseRaw / assembleCoarseValues (9 nnz each, either may be null)
// and buildCoarseDiag (3D); returns D.
int orc_coarse_values(orc_solver* o, int target, int minsz, int maxsz, const double* x, const double* diag,
                      double* raw, double* sym, double* dc) {
  const SchwarzDomains d = probe_domains(o, target, minsz, maxsz);
  const CoarsePattern p = build_coarse_pattern(o->s.mesh, d);
  const std::vector<double> r = assemble_coarse_raw(o->s.mesh, o->s.mat, x, diag, d, p);
  if (raw) std::copy(r.begin(), r.end(), raw);
  if (sym) {
    const std::vector<double> v = symmetrize_coarse(p, r);
    std::copy(v.begin(), v.end(), sym);
  }
  if (dc) {
    const std::vector<double> c = build_coarse_diag(o->s.mesh, o->s.mat, x, diag, d);
    std::copy(c.begin(), c.end(), dc);
  }
  return (int)d.members.size();
}

// dir 0: out (3D) = R in (3n); dir 1: out (3n) += P in (3D).  Returns D.
int orc_coarse_transfer(orc_solver* o, int target, int minsz, int maxsz, int dir, const double* in, double* out) {
  const SchwarzDomains d = probe_domains(o, target, minsz, maxsz);
  if (dir == 0) restrict_coarse(in, d, out);
  else prolongate_add(in, d, out);
  return (int)d.members.size();
}

// buildCoarseExact: dense (3D)^2 Ac and its Cholesky factor (or the Jacobi
// fallback); returns the flag.
int orc_coarse_exact(orc_solver* o, int target, int minsz, int maxsz, const double* x, const double* diag,
                     double* ac, double* chol) {
  const SchwarzDomains d = probe_domains(o, target, minsz, maxsz);
  std::vector<double> a, c;
  const int flag = build_coarse_exact(o->s.mesh, o->s.mat, x, diag, d, a, c);
  if (ac) std::copy(a.begin(), a.end(), ac);
  if (chol) std::copy(c.begin(), c.end(), chol);
  return flag;
}

// Coarse solves of rhs (3D): mode 2 applyCoarseBlockJacobi, 3 coarsePcg of
// `iters` iterations at tolerance `tol` (returns its iteration count), 4 the
// exact Cholesky solve (returns the flag).
int orc_coarse_solve(orc_solver* o, int target, int minsz, int maxsz, const double* x, const double* diag, int mode,
                     int iters, double tol, const double* rhs, double* out, double* residual) {
  const SchwarzDomains d = probe_domains(o, target, minsz, maxsz);
  const CoarsePattern p = build_coarse_pattern(o->s.mesh, d);
  const std::vector<double> B = symmetrize_coarse(p, assemble_coarse_raw(o->s.mesh, o->s.mat, x, diag, d, p));
  const std::vector<double> dc = build_coarse_diag(o->s.mesh, o->s.mat, x, diag, d);
  if (mode == 2) {
    coarse_block_jacobi(p, B, dc, rhs, out);
    return 0;
  }
  if (mode == 3) return coarse_pcg(p, B, dc, rhs, iters, tol, out, residual);
  std::vector<double> a, c;
  const int flag = build_coarse_exact(o->s.mesh, o->s.mat, x, diag, d, a, c);
  solve_coarse_exact(c, flag, d.members.size() * 3, rhs, out);
  return flag;
}

// z = Schwarz1(r) + P C(R r) with C by `mode` (1 damped Jacobi, 2 block
// Jacobi, 3 coarse PCG of `iters` iterations); returns D.
int orc_two_level_apply(orc_solver* o, int target, int minsz, int maxsz, const double* x, const double* diag,
                        int mode, int iters, double omega, const double* r, double* z) {
  Solver& s = o->s;
  const SchwarzDomains d = probe_domains(o, target, minsz, maxsz);
  std::vector<std::vector<uint32_t>> v2t(s.mesh.n);
  for (uint32_t t = 0; t < s.mesh.m; t++)
    for (int k = 0; k < 3; k++) v2t[s.mesh.idx[t * 3 + k]].push_back(t);
  const SchwarzFactors f = build_schwarz_factors(s.mesh, s.mat, x, diag, d, v2t);
  const CoarsePattern p = build_coarse_pattern(s.mesh, d);
  const std::vector<double> B = symmetrize_coarse(p, assemble_coarse_raw(s.mesh, s.mat, x, diag, d, p));
  const std::vector<double> dc = build_coarse_diag(s.mesh, s.mat, x, diag, d);
  apply_two_level(f, d, p, B, dc, mode, iters, omega, r, z, s.mesh.n);
  return (int)d.members.size();
}

void orc_set_sum_mode(int faithful) { orcm::g_faithful_sums = faithful != 0; }

int64_t orc_armijo_trace(orc_solver* o, double* out4, int64_t cap) {
  const auto& a = o->s.armijo;
  for (int64_t k = 0; k < (int64_t)a.size() && k < cap; k++)
    for (int j = 0; j < 4; j++) out4[k * 4 + j] = a[k][j];
  return (int64_t)a.size();
}

// buildBlockFactors() block-jacobi.ts:30-87 at x with Jacobi diagonal `diag`.
void orc_block_factors(orc_solver* o, const double* x, const double* diag, double* inv9n, uint8_t* flags) {
  std::vector<double> inv;
  std::vector<uint8_t> flag;
  build_block_factors(o->s.mesh, o->s.mat, x, diag, inv, flag);
  std::copy(inv.begin(), inv.end(), inv9n);
  std::copy(flag.begin(), flag.end(), flags);
}

// applyBlockFactors() block-jacobi.ts:90-104.
void orc_apply_block_factors(const double* inv9n, const double* r, double* z, uint32_t n) {
  const std::vector<double> inv(inv9n, inv9n + (size_t)n * 9);
  apply_block_factors(inv, r, z, n);
}

// pcg() pcg.ts:11-95 with a dense operator out = A p (row-major ndof x ndof),
// the Jacobi preconditioner on `diag` and the pin filter of `pins` (vertex
// ids, xyz rows zeroed).  Returns converged; *iters / *residual as PcgResult.
int orc_pcg_dense(const double* A, uint32_t ndof, const double* b, const double* diag, const uint32_t* pins,
                  uint32_t npins, int max_iters, double tol, double* x, int* iters, double* residual) {
  PinFilter filter;
  for (uint32_t k = 0; k < npins; k++) filter.ids.push_back(pins[k]);
  auto hvp = [&](const double* p, double* out) {
    for (uint32_t i = 0; i < ndof; i++) {
      double acc = 0;
      for (uint32_t j = 0; j < ndof; j++) acc += A[(size_t)i * ndof + j] * p[j];
      out[i] = acc;
    }
  };
  struct NoPre { void operator()(const double*, double*) const {} };
  const std::vector<double> bb(b, b + ndof), dd(diag, diag + ndof);
  const PcgResult r = pcg(bb, hvp, dd, max_iters, tol, filter, (NoPre*)nullptr);
  std::copy(r.x.begin(), r.x.end(), x);
  if (iters) *iters = r.iters;
  if (residual) *residual = r.residual;
  return r.converged ? 1 : 0;
}

// addHingeGradient() bending.ts:19-42 on one hinge whose four vertices are
// x12 = (x_v0, x_v1, x_v2, x_v3); grad12 receives the FD gradient, the return
// value is hingeEnergy.
double orc_hinge_grad_fd(const double* x12, double rest, double len, double asum, double k, double* grad12) {
  std::vector<double> tmp(x12, x12 + 12);
  for (int i = 0; i < 12; i++) grad12[i] = 0;
  const Hinge h{0, 1, 2, 3, rest, len, asum};
  return add_hinge_gradient(tmp, grad12, h, k);
}

// queryBvhAll() contact-assembly.ts:124-140 with the static boxes of
// staticTriAabb() :597-605: (cloth triangle, static triangle) pairs, sorted.
int64_t orc_static_pairs(const double* x0, const double* x1, const uint32_t* idx, uint32_t m, const float* spos,
                         uint32_t ns, const uint32_t* sidx, uint32_t ms, double pad, uint32_t* pairs, int64_t cap) {
  Contact c;
  c.staticPos.assign(spos, spos + (size_t)ns * 3);
  c.staticIdx.assign(sidx, sidx + (size_t)ms * 3);
  c.bvh.build(x0, x1, idx, m, pad);
  std::vector<Aabb> boxes(ms);
  for (uint32_t t = 0; t < ms; t++) boxes[t] = c.static_tri_aabb(t, pad);
  const auto out = c.query_bvh_all(boxes);
  for (int64_t k = 0; k < (int64_t)out.size() && k < cap; k++) {
    pairs[k * 2] = out[k].first;
    pairs[k * 2 + 1] = out[k].second;
  }
  return (int64_t)out.size();
}

// killApproachVelocity() newton.ts:238-303: begin a step at x0, build the
// active set at x and filter the velocities v (3n, in place).
int orc_kill_velocity(orc_solver* o, const double* x0, const double* x, double* v) {
  Solver& s = o->s;
  if (!s.hasContact) return -1;
  const size_t n3 = (size_t)s.mesh.n * 3;
  s.contact.begin_step(std::vector<double>(x0, x0 + n3));
  s.velocities.assign(v, v + n3);
  kill_approach_velocity(s.contact, s, x);
  std::copy(s.velocities.begin(), s.velocities.end(), v);
  return (int)s.contact.active.size();
}

// commit() contact-assembly.ts:475-485: begin a step at x0, freeze the active
// set at x, commit at x and dump the lagged table: key (5 ints, the
// normalised vtKey/eeKey/floor key) and (nx, ny, nz, lambdaN) per entry, in
// no particular order.  Returns the entry count.
int64_t orc_commit_probe(orc_solver* o, const double* x0, const double* x, int64_t* keys5, double* vals4,
                         int64_t cap) {
  Solver& s = o->s;
  if (!s.hasContact) return -1;
  const size_t n3 = (size_t)s.mesh.n * 3;
  s.contact.begin_step(std::vector<double>(x0, x0 + n3));
  s.contact.update_active_set(x, s.mesh.n);
  s.contact.commit(x);
  int64_t k = 0;
  for (const auto& kv : s.contact.lagged) {
    if (k < cap) {
      for (int j = 0; j < 5; j++) keys5[k * 5 + j] = kv.first.k[j];
      vals4[k * 4] = kv.second.nx; vals4[k * 4 + 1] = kv.second.ny;
      vals4[k * 4 + 2] = kv.second.nz; vals4[k * 4 + 3] = kv.second.lambdaN;
    }
    k++;
  }
  return k;
}

}  // extern "C"
