// hessian.h — rows of the membrane triangle Hessian (triangleHessian9,
// membrane-blocks.ts:21-98) in the reference's loop nest and accumulation
// order, shared by the Schwarz local matrices and the coarse level.
#pragma once

#include "fem.h"

namespace cloth {

// Rows A = ci*3 + ra (ra = 0..2), all columns B, of triangleHessian9
// (membrane-blocks.ts:21-98), same loop nest and accumulation order.
// Rows ra0 .. ra1-1 of corner ci's three rows into h[(ra - ra0) * 9 + B];
// every entry is computed on its own, so any subset of rows gives the same
// values as the full block.
__device__ __forceinline__ void tri_hessian_rows_range(const double* x0, const double* x1, const double* x2,
                                                       const float* inv, double area, const DMat& mat, int ci,
                                                       int ra0, int ra1, double* h) {
  const double a = inv[0], b = inv[1], c = inv[2], d = inv[3];
  const double* X[3] = {x0, x1, x2};
  const double K[3][2] = {{-(a + c), -(b + d)}, {a, b}, {c, d}};
  double F[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 2; j++)
      for (int v = 0; v < 3; v++) F[i][j] += K[v][j] * X[v][i];
  double Cm[2][2] = {{0, 0}, {0, 0}};
  for (int l = 0; l < 2; l++)
    for (int m = 0; m < 2; m++)
      for (int i = 0; i < 3; i++) Cm[l][m] += F[i][l] * F[i][m];
  const double E00 = 0.5 * (Cm[0][0] - 1), E11 = 0.5 * (Cm[1][1] - 1), E01 = 0.5 * Cm[0][1];
  const double S00 = mat.c00 * E00 + mat.c01 * E11;
  const double S11 = mat.c11 * E11 + mat.c01 * E00;
  const double S01 = 2 * mat.g * E01;
  const double S[2][2] = {{S00, S01}, {S01, S11}};
  double D[2][2][2][2];
  for (int q = 0; q < 16; q++) (&D[0][0][0][0])[q] = 0;
  D[0][0][0][0] = mat.c00;
  D[1][1][1][1] = mat.c11;
  D[0][0][1][1] = mat.c01; D[1][1][0][0] = mat.c01;
  D[0][1][0][1] = mat.g; D[0][1][1][0] = mat.g; D[1][0][0][1] = mat.g; D[1][0][1][0] = mat.g;
  const double s = mat.thick * area;
  // M[A][i][j] = (i == A % 3) ? K[A / 3][j] : 0
  for (int ra = ra0; ra < ra1; ra++) {
    const int A = ci * 3 + ra, av = A / 3, ai = A % 3;
    for (int B = 0; B < 9; B++) {
      const int bv = B / 3, bi = B % 3;
      double sum = 0;
      for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 2; j++) {
          const double mA = (i == ai) ? K[av][j] : 0.0;
          if (mA == 0) continue;
          double dP = 0;
          for (int k = 0; k < 2; k++) {
            const double mBik = (i == bi) ? K[bv][k] : 0.0;
            dP += mBik * S[k][j];
            double dE = 0;
            for (int l = 0; l < 2; l++) {
              for (int m2 = 0; m2 < 2; m2++) {
                const double dd = D[k][j][l][m2];
                if (dd == 0) continue;
                double dElm = 0;
                for (int p = 0; p < 3; p++) {
                  const double mBpl = (p == bi) ? K[bv][l] : 0.0;
                  const double mBpm = (p == bi) ? K[bv][m2] : 0.0;
                  dElm += mBpl * F[p][m2] + F[p][l] * mBpm;
                }
                dE += dd * 0.5 * dElm;
              }
            }
            dP += F[i][k] * dE;
          }
          sum += mA * dP;
        }
      }
      h[(ra - ra0) * 9 + B] = s * sum;
    }
  }
}

__device__ __forceinline__ void tri_hessian_rows(const double* x0, const double* x1, const double* x2, const float* inv,
                                                 double area, const DMat& mat, int ci, double* h27) {
  tri_hessian_rows_range(x0, x1, x2, inv, area, mat, ci, 0, 3, h27);
}

}  // namespace cloth
