// pcg.hip — device-resident PCG of the Newton step (pcg.ts:11-95) in FP64.
//
// Per iteration three kernels, each ending in a grid reduction whose last
// block takes the scalar decision on the device (no host round trip):
//   k_hvp     Ap = hvpFull(p) (newton.ts:135-146), pins zeroed; pAp; alpha
//   k_update  x += alpha p, r -= alpha Ap; residual test; z = M^-1 r; beta
//   k_dir     p = z + beta p; |p| for the FD operator's step h
// A `done` latch turns the remaining launches of a converged solve into no-ops,
// which is exactly the reference's early return.
#include "elements.h"
#include "pcg.h"

namespace cloth {

// hvpFull at vertex v: M p / h^2 + H_membrane p (+ contact), unfiltered.
__device__ __forceinline__ void hvp_vertex(const PcgArgs& a, uint32_t v, const double* __restrict__ p, double h,
                                           double* out) {
  const DMesh& me = a.me;
  double acc0[3] = {0, 0, 0}, acc1[3] = {0, 0, 0};
  const uint32_t b = me.vt_off[v], e = me.vt_off[v + 1];
  for (uint32_t k = b; k < e; k++) {
    const uint32_t rec = me.vt_rec[k];
    const uint32_t t = rec >> 2;
    const int corner = rec & 3;
    const uint32_t i0 = me.idx[(size_t)t * 3], i1 = me.idx[(size_t)t * 3 + 1], i2 = me.idx[(size_t)t * 3 + 2];
    double x0[3], x1[3], x2[3], p0[3], p1[3], p2[3], gc[3];
    ld3(a.x, i0, x0); ld3(a.x, i1, x1); ld3(a.x, i2, x2);
    ld3(p, i0, p0); ld3(p, i1, p1); ld3(p, i2, p2);
    const float* inv = me.inv_dm + (size_t)t * 4;
    const double area = me.area[t];
    if (a.hvp_mode == HVP_FD) {
      // membraneHvp (fem.ts:108-130): gradients at x + h p and x - h p
      double xp0[3], xp1[3], xp2[3], xm0[3], xm1[3], xm2[3];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        xp0[i] = x0[i] + h * p0[i]; xm0[i] = x0[i] - h * p0[i];
        xp1[i] = x1[i] + h * p1[i]; xm1[i] = x1[i] - h * p1[i];
        xp2[i] = x2[i] + h * p2[i]; xm2[i] = x2[i] - h * p2[i];
      }
      const TriState sp = tri_state(xp0, xp1, xp2, inv);
      tri_grad_corner(sp, area, a.mat, corner, gc);
      acc0[0] += gc[0]; acc0[1] += gc[1]; acc0[2] += gc[2];
      const TriState sm = tri_state(xm0, xm1, xm2, inv);
      tri_grad_corner(sm, area, a.mat, corner, gc);
      acc1[0] += gc[0]; acc1[1] += gc[1]; acc1[2] += gc[2];
    } else {
      const TriState s = tri_state(x0, x1, x2, inv);
      tri_hvp_corner(s, p0, p1, p2, area, a.mat, corner, gc);
      acc0[0] += gc[0]; acc0[1] += gc[1]; acc0[2] += gc[2];
    }
  }
  double Hp[3];
  if (a.hvp_mode == HVP_FD) {
#pragma unroll
    for (int i = 0; i < 3; i++) Hp[i] = (acc0[i] - acc1[i]) / (2 * h);
  } else {
    Hp[0] = acc0[0]; Hp[1] = acc0[1]; Hp[2] = acc0[2];
  }
  const double m = me.mass[v];
#pragma unroll
  for (int i = 0; i < 3; i++) out[i] = m * p[(size_t)v * 3 + i] * a.inv_h2 + Hp[i];
  if (a.contact.act) {
    // applyHvp (contact-assembly.ts:371-392): out.fill(0) then addDistGrad per
    // active contact in list order, skipping dd == 0; then out += cHp.
    double c[3] = {0, 0, 0};
    const uint32_t cb = a.contact.v_off[v], ce = a.contact.v_off[v + 1];
    for (uint32_t k = cb; k < ce; k++) {
      const uint32_t ent = a.contact.v_ent[k];
      const uint32_t ci = ent >> 3;
      const int slot = ent & 7;
      const double g = a.contact.dd[ci];
      if (g == 0) continue;
      const Frozen& f = a.contact.act[ci];
      if (f.kind == 2) {
        c[1] += g;
      } else if (f.kind == 0) {
        if (slot == 0) {
          c[0] += g * f.nx; c[1] += g * f.ny; c[2] += g * f.nz;
        } else {
          const double w = slot == 1 ? f.w0 : slot == 2 ? f.w1 : f.w2;
          c[0] += -g * w * f.nx; c[1] += -g * w * f.ny; c[2] += -g * w * f.nz;
        }
      } else {
        const double w = slot == 0 ? f.w0 : slot == 1 ? f.w1 : slot == 2 ? f.w2 : f.w3;
        c[0] += g * w * f.nx; c[1] += g * w * f.ny; c[2] += g * w * f.nz;
      }
    }
    out[0] += c[0]; out[1] += c[1]; out[2] += c[2];
  }
}

// z = M^-1 r at vertex v (Jacobi with the diag > 1e-12 guard, or 3x3 block).
__device__ __forceinline__ void precond_vertex(const PcgArgs& a, uint32_t v, const double* r3, double* z3) {
  if (a.precond == PRE_BLOCK) {
    const double* f = a.binv + (size_t)v * 9;
    z3[0] = f[0] * r3[0] + f[1] * r3[1] + f[2] * r3[2];
    z3[1] = f[3] * r3[0] + f[4] * r3[1] + f[5] * r3[2];
    z3[2] = f[6] * r3[0] + f[7] * r3[1] + f[8] * r3[2];
  } else {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const double d = a.diag[(size_t)v * 3 + i];
      z3[i] = d > 1e-12 ? r3[i] / d : r3[i];
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_pcg_init(PcgArgs a) {
  jsm::dd rz = jsm::dd_zero(), bb = jsm::dd_zero(), rr = jsm::dd_zero(), pp = jsm::dd_zero();
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.me.n; v += gridDim.x * blockDim.x) {
    const bool pinned = a.me.pinmask[v] != 0;
    double r3[3], z3[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const size_t j = (size_t)v * 3 + i;
      const double bi = a.b[j];
      jsm::dd_add_prod(bb, bi, bi);
      r3[i] = pinned ? 0.0 : bi;
      a.sol[j] = 0.0;
      a.r[j] = r3[i];
    }
    if (a.precond != PRE_EXTERNAL) {
      precond_vertex(a, v, r3, z3);
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const size_t j = (size_t)v * 3 + i;
        if (pinned) z3[i] = 0.0;
        a.z[j] = z3[i];
        a.p[j] = z3[i];
        jsm::dd_add_prod(rz, r3[i], z3[i]);
        jsm::dd_add_prod(pp, z3[i], z3[i]);
      }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) jsm::dd_add_prod(rr, r3[i], r3[i]);
  }
  PcgScalars* s = a.s;
  jsm::dd t;
  if (grid_reduce_dd<kBlock>(bb, a.partials, &s->tickets[0], t)) {
    s->bnorm = jsm::sqrt_(jsm::dd_value(t));
    s->abs_tol = jsm::jmax(a.tol * jsm::jmax(s->bnorm, 1e-12), 1e-12);
  }
  if (grid_reduce_dd<kBlock>(rr, a.partials + 1024, &s->tickets[1], t)) s->residual = jsm::sqrt_(jsm::dd_value(t));
  if (a.precond != PRE_EXTERNAL) {
    if (grid_reduce_dd<kBlock>(rz, a.partials + 2048, &s->tickets[2], t)) s->rz = jsm::dd_value(t);
    if (grid_reduce_dd<kBlock>(pp, a.partials + 3072, &s->tickets[3], t))
      s->h_fd = 1e-8 * (1 + jsm::sqrt_(jsm::dd_value(t)));
  }
}

// Decide after init (separate tiny kernel: the init reductions finish in
// different last blocks).
__global__ void k_pcg_init_decide(PcgScalars* s) {
  s->iters = 0;
  s->done = (s->residual < s->abs_tol) ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_pcg_hvp(PcgArgs a, int k) {
  PcgScalars* s = a.s;
  if (s->done) return;
  const double h = s->h_fd;
  jsm::dd pap = jsm::dd_zero();
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.me.n; v += gridDim.x * blockDim.x) {
    double o[3];
    hvp_vertex(a, v, a.p, h, o);
    const bool pinned = a.me.pinmask[v] != 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const size_t j = (size_t)v * 3 + i;
      const double ap = pinned ? 0.0 : o[i];
      a.Ap[j] = ap;
      jsm::dd_add_prod(pap, a.p[j], ap);
    }
  }
  jsm::dd t;
  if (grid_reduce_dd<kBlock>(pap, a.partials, &s->tickets[4], t)) {
    const double pAp = jsm::dd_value(t);
    s->pAp = pAp;
    if (!(pAp > 1e-30)) {
      s->done = 1;
      s->iters = k;
    } else {
      s->alpha = s->rz / pAp;
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_pcg_update(PcgArgs a, int k) {
  PcgScalars* s = a.s;
  if (s->done) return;
  const double alpha = s->alpha;
  jsm::dd rr = jsm::dd_zero(), rz = jsm::dd_zero();
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.me.n; v += gridDim.x * blockDim.x) {
    const bool pinned = a.me.pinmask[v] != 0;
    double r3[3], z3[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const size_t j = (size_t)v * 3 + i;
      double xi = a.sol[j] + alpha * a.p[j];
      double ri = a.r[j] - alpha * a.Ap[j];
      if (pinned) { xi = 0.0; ri = 0.0; }
      a.sol[j] = xi;
      a.r[j] = ri;
      r3[i] = ri;
      jsm::dd_add_prod(rr, ri, ri);
    }
    if (a.precond != PRE_EXTERNAL) {
      precond_vertex(a, v, r3, z3);
#pragma unroll
      for (int i = 0; i < 3; i++) {
        if (pinned) z3[i] = 0.0;
        a.z[(size_t)v * 3 + i] = z3[i];
        jsm::dd_add_prod(rz, r3[i], z3[i]);
      }
    }
  }
  jsm::dd t;
  if (grid_reduce_dd<kBlock>(rr, a.partials, &s->tickets[5], t)) {
    s->residual = jsm::sqrt_(jsm::dd_value(t));
    if (s->residual < s->abs_tol) {
      s->done = 2;  // converged in this iteration (any non-zero value stops the solve)
      s->iters = k + 1;
    }
  }
  if (a.precond != PRE_EXTERNAL) {
    if (grid_reduce_dd<kBlock>(rz, a.partials + 1024, &s->tickets[6], t)) s->beta = jsm::dd_value(t);  // rzNew
  }
}

// External preconditioner path: z was written by another kernel; reduce rz.
__global__ void __launch_bounds__(kBlock) k_pcg_rz(PcgArgs a) {
  PcgScalars* s = a.s;
  if (s->done) return;
  jsm::dd rz = jsm::dd_zero();
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.me.n; v += gridDim.x * blockDim.x) {
    const bool pinned = a.me.pinmask[v] != 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const size_t j = (size_t)v * 3 + i;
      const double zi = pinned ? 0.0 : a.z[j];
      a.z[j] = zi;
      jsm::dd_add_prod(rz, a.r[j], zi);
    }
  }
  jsm::dd t;
  if (grid_reduce_dd<kBlock>(rz, a.partials + 1024, &s->tickets[6], t)) s->beta = jsm::dd_value(t);
}

__global__ void __launch_bounds__(kBlock) k_pcg_dir(PcgArgs a) {
  PcgScalars* s = a.s;
  if (s->done) return;
  // beta holds rzNew here: beta = rzNew / rz; rz = rzNew (pcg.ts:88-92)
  const double rzNew = s->beta;
  const double beta = rzNew / s->rz;
  jsm::dd pp = jsm::dd_zero();
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.me.n; v += gridDim.x * blockDim.x) {
    const bool pinned = a.me.pinmask[v] != 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const size_t j = (size_t)v * 3 + i;
      const double pi = pinned ? 0.0 : a.z[j] + beta * a.p[j];
      a.p[j] = pi;
      jsm::dd_add_prod(pp, pi, pi);
    }
  }
  jsm::dd t;
  if (grid_reduce_dd<kBlock>(pp, a.partials + 2048, &s->tickets[7], t)) {
    s->h_fd = 1e-8 * (1 + jsm::sqrt_(jsm::dd_value(t)));
    s->rz = rzNew;
  }
}

void pcg_init(hipStream_t st, const PcgArgs& a) {
  hipLaunchKernelGGL(k_pcg_init, dim3(red_grid(a.me.n)), dim3(kBlock), 0, st, a);
  CS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pcg_init_decide, dim3(1), dim3(1), 0, st, a.s);
  CS_LAUNCH_CHECK();
}

void pcg_iteration(hipStream_t st, const PcgArgs& a, int k) {
  pcg_iteration_hvp(st, a, k);
  pcg_iteration_rest(st, a, k);
}

void pcg_iteration_hvp(hipStream_t st, const PcgArgs& a, int k) {
  hipLaunchKernelGGL(k_pcg_hvp, dim3(red_grid(a.me.n)), dim3(kBlock), 0, st, a, k);
  CS_LAUNCH_CHECK();
}

void pcg_iteration_rest(hipStream_t st, const PcgArgs& a, int k) {
  const dim3 g(red_grid(a.me.n)), b(kBlock);
  hipLaunchKernelGGL(k_pcg_update, g, b, 0, st, a, k);
  hipLaunchKernelGGL(k_pcg_dir, g, b, 0, st, a);
  CS_LAUNCH_CHECK();
}

void pcg_iteration_split_a(hipStream_t st, const PcgArgs& a, int k) {
  const dim3 g(red_grid(a.me.n)), b(kBlock);
  hipLaunchKernelGGL(k_pcg_hvp, g, b, 0, st, a, k);
  hipLaunchKernelGGL(k_pcg_update, g, b, 0, st, a, k);
  CS_LAUNCH_CHECK();
}

void pcg_iteration_split_b(hipStream_t st, const PcgArgs& a, int k) {
  const dim3 g(red_grid(a.me.n)), b(kBlock);
  hipLaunchKernelGGL(k_pcg_rz, g, b, 0, st, a);
  hipLaunchKernelGGL(k_pcg_dir, g, b, 0, st, a);
  CS_LAUNCH_CHECK();
}

// ---- block-Jacobi factors ---------------------------------------------------

// Diagonal 3x3 block (corner vc) of triangleHessian9 (membrane-blocks.ts:21-98),
// same loop nest and accumulation order as the reference for those entries.
__device__ void tri_hessian_block(const double* x0, const double* x1, const double* x2, const float* inv,
                                  double area, const DMat& mat, int vc, double* hb) {
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
  for (int ra = 0; ra < 3; ra++) {
    for (int rb = 0; rb < 3; rb++) {
      const int A = vc * 3 + ra, B = vc * 3 + rb;
      const int av = A / 3, ai = A % 3, bv = B / 3, bi = B % 3;
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
      hb[ra * 3 + rb] = s * sum;
    }
  }
}

// buildBlockFactors (block-jacobi.ts:30-87), one thread per vertex.
__global__ void __launch_bounds__(kBlock) k_block_factors(DMesh me, DMat mat, const double* __restrict__ x,
                                                          const double* __restrict__ diag, double* binv,
                                                          uint8_t* flag) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= me.n) return;
  double bl[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, hb[9];
  for (uint32_t k = me.vt_off[v]; k < me.vt_off[v + 1]; k++) {
    const uint32_t rec = me.vt_rec[k];
    const uint32_t t = rec >> 2;
    const int corner = rec & 3;
    double x0[3], x1[3], x2[3];
    ld3(x, me.idx[(size_t)t * 3], x0);
    ld3(x, me.idx[(size_t)t * 3 + 1], x1);
    ld3(x, me.idx[(size_t)t * 3 + 2], x2);
    tri_hessian_block(x0, x1, x2, me.inv_dm + (size_t)t * 4, me.area[t], mat, corner, hb);
    for (int q = 0; q < 9; q++) bl[q] += hb[q];
  }
  const double d0 = diag[(size_t)v * 3], d1 = diag[(size_t)v * 3 + 1], d2 = diag[(size_t)v * 3 + 2];
  const double B00 = bl[0] + d0, B11 = bl[4] + d1, B22 = bl[8] + d2;
  const double B01 = 0.5 * (bl[1] + bl[3]);
  const double B02 = 0.5 * (bl[2] + bl[6]);
  const double B12 = 0.5 * (bl[5] + bl[7]);
  const double tr = jsm::fabs_(B00) + jsm::fabs_(B11) + jsm::fabs_(B22);
  const double tiny = 1e-20 * (1 + tr);
  bool ok = true;
  if (!(B00 > tiny)) ok = false;
  if (ok && !(B00 * B11 - B01 * B01 > tiny * (1 + jsm::fabs_(B00) + jsm::fabs_(B11)))) ok = false;
  const double det = B00 * (B11 * B22 - B12 * B12) - B01 * (B01 * B22 - B12 * B02) + B02 * (B01 * B12 - B11 * B02);
  if (ok && !(det > tiny * (1 + tr) * (1 + tr))) ok = false;
  double* o = binv + (size_t)v * 9;
  if (ok) {
    const double s = 1 / det;
    o[0] = (B11 * B22 - B12 * B12) * s;
    o[1] = (B02 * B12 - B01 * B22) * s;
    o[2] = (B01 * B12 - B11 * B02) * s;
    o[3] = (B02 * B12 - B01 * B22) * s;
    o[4] = (B00 * B22 - B02 * B02) * s;
    o[5] = (B01 * B02 - B00 * B12) * s;
    o[6] = (B01 * B12 - B11 * B02) * s;
    o[7] = (B01 * B02 - B00 * B12) * s;
    o[8] = (B00 * B11 - B01 * B01) * s;
    if (flag) flag[v] = 1;
  } else {
    const double i0 = d0 > 1e-12 ? 1 / d0 : 0;
    const double i1 = d1 > 1e-12 ? 1 / d1 : 0;
    const double i2 = d2 > 1e-12 ? 1 / d2 : 0;
    o[0] = i0; o[1] = 0; o[2] = 0;
    o[3] = 0; o[4] = i1; o[5] = 0;
    o[6] = 0; o[7] = 0; o[8] = i2;
    if (flag) flag[v] = 0;
  }
}

void launch_block_factors(hipStream_t st, const DMesh& me, const DMat& mat, const double* x, const double* diag,
                          double* binv, uint8_t* flag) {
  hipLaunchKernelGGL(k_block_factors, dim3(grid_for(me.n)), dim3(kBlock), 0, st, me, mat, x, diag, binv, flag);
  CS_LAUNCH_CHECK();
}

// ---- vector helpers ---------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_dot(const double* __restrict__ a, const double* __restrict__ b,
                                                size_t n, jsm::dd* partials, unsigned* ticket, double* out) {
  jsm::dd acc = jsm::dd_zero();
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    jsm::dd_add_prod(acc, a[i], b[i]);
  jsm::dd t;
  if (grid_reduce_dd<kBlock>(acc, partials, ticket, t)) *out = jsm::dd_value(t);
}

void launch_dot(hipStream_t st, const double* a, const double* b, size_t n, jsm::dd* partials, unsigned* ticket,
                double* out) {
  hipLaunchKernelGGL(k_dot, dim3(red_grid(n)), dim3(kBlock), 0, st, a, b, n, partials, ticket, out);
  CS_LAUNCH_CHECK();
}

__global__ void __launch_bounds__(kBlock) k_max_abs(const double* __restrict__ a, size_t n, double* partials,
                                                    unsigned* ticket, double* out) {
  double m = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double v = jsm::fabs_(a[i]);
    if (v > m) m = v;
  }
  double r;
  if (grid_reduce_ext<kBlock, GtMax>(m, partials, ticket, r)) *out = r > 0 ? r : 0.0;
}

void launch_max_abs(hipStream_t st, const double* a, size_t n, double* partials, unsigned* ticket, double* out) {
  hipLaunchKernelGGL(k_max_abs, dim3(red_grid(n)), dim3(kBlock), 0, st, a, n, partials, ticket, out);
  CS_LAUNCH_CHECK();
}

__global__ void k_fallback_dir(const double* __restrict__ g, const double* __restrict__ diag, double* dx, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) dx[i] = -g[i] / diag[i];
}
void launch_fallback_dir(hipStream_t st, const double* g, const double* diag, double* dx, size_t n) {
  hipLaunchKernelGGL(k_fallback_dir, dim3(grid_for(n)), dim3(kBlock), 0, st, g, diag, dx, n);
  CS_LAUNCH_CHECK();
}

__global__ void k_scale(double* v, double s, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] *= s;
}
void launch_scale(hipStream_t st, double* v, double s, size_t n) {
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(kBlock), 0, st, v, s, n);
  CS_LAUNCH_CHECK();
}

__global__ void k_trial(const double* __restrict__ x, const double* __restrict__ dx, double alpha, double* xt,
                        size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) xt[i] = x[i] + alpha * dx[i];
}
void launch_trial(hipStream_t st, const double* x, const double* dx, double alpha, double* xt, size_t n) {
  hipLaunchKernelGGL(k_trial, dim3(grid_for(n)), dim3(kBlock), 0, st, x, dx, alpha, xt, n);
  CS_LAUNCH_CHECK();
}

__global__ void k_neg(const double* __restrict__ g, double* b, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = -g[i];
}
void launch_neg(hipStream_t st, const double* g, double* b, size_t n) {
  hipLaunchKernelGGL(k_neg, dim3(grid_for(n)), dim3(kBlock), 0, st, g, b, n);
  CS_LAUNCH_CHECK();
}

}  // namespace cloth
