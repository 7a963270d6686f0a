// pcg.hip — device-resident PCG of the Newton step (pcg.ts:11-95) in FP64.
//
// Per iteration three kernels; each writes block partials of its dot
// products and the next kernel merges them (fixed order, every block the
// same bits), so no scalar ever makes a host round trip or needs a fence:
//   k_hvp     [merge |p|^2 -> FD step h] Ap = hvpFull(p) (newton.ts:135-146),
//             pins zeroed; partials of p.Ap
//   k_update  [merge p.Ap -> breakdown test, alpha] x += alpha p,
//             r -= alpha Ap, z = M^-1 r; partials of r.r and r.z
//   k_dir     [merge r.r -> residual test; r.z -> beta] p = z + beta p;
//             partials of p.p
// A `done` latch turns the remaining launches of a converged solve into
// no-ops, which is exactly the reference's early return.  Block 0 records
// decisions; every block takes them from the same merged values.
#include "elements.h"
#include "pcg.h"
#include "schwarz.h"

namespace cloth {

enum { PART_BB = 0, PART_RR = 1, PART_RZ = 2, PART_PP = 3, PART_PAP = 4 };

__device__ __forceinline__ jsm::dd* part(const PcgArgs& a, int which) {
  return a.partials + (size_t)which * kMaxRedBlocks;
}

// dd_c = kappa max(b'', 0) (J_c . p) of active contact c (distJvp,
// contact-assembly.ts:395-414, scaled as in applyHvp :376-386).
__device__ __forceinline__ double contact_dd(const DContactHvp& H, uint32_t ci, const double* __restrict__ p) {
  if (!H.inr[ci]) return 0.0;
  const Frozen& c = H.act[ci];
  double j;
  if (c.kind == 2) {
    j = p[(size_t)c.p * 3 + 1];
  } else if (c.kind == 0) {
    const size_t q = c.p;
    j = c.nx * p[q * 3] + c.ny * p[q * 3 + 1] + c.nz * p[q * 3 + 2];
    const double ws[3] = {c.w0, c.w1, c.w2};
    const int vs[3] = {c.a, c.b, c.c};
    for (int k = 0; k < 3; k++) {
      if (vs[k] < 0) continue;
      const size_t r = vs[k];
      j -= ws[k] * (c.nx * p[r * 3] + c.ny * p[r * 3 + 1] + c.nz * p[r * 3 + 2]);
    }
  } else {
    const double ws[4] = {c.w0, c.w1, c.w2, c.w3};
    const int vs[4] = {c.a, c.b, c.c, c.d};
    j = 0;
    for (int k = 0; k < 4; k++) {
      if (vs[k] < 0) continue;
      const size_t r = vs[k];
      j += ws[k] * (c.nx * p[r * 3] + c.ny * p[r * 3 + 1] + c.nz * p[r * 3 + 2]);
    }
  }
  return H.coeff[ci] * j;
}

// Contribution of one incident triangle corner to the membrane HVP at the
// vertex: analytic, or the two gradient corners of the FD operator.
__device__ __forceinline__ void hvp_corner(const PcgArgs& a, uint4 cr, const double* __restrict__ p, double h,
                                           double* g0, double* g1) {
  const DMesh& me = a.me;
  const uint32_t t = cr.w >> 2;
  const int corner = cr.w & 3;
  const uint32_t i0 = cr.x, i1 = cr.y, i2 = cr.z;
  double x0[3], x1[3], x2[3], p0[3], p1[3], p2[3];
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
    tri_grad_corner(sp, area, a.mat, corner, g0);
    const TriState sm = tri_state(xm0, xm1, xm2, inv);
    tri_grad_corner(sm, area, a.mat, corner, g1);
  } else {
    const TriState s = tri_state(x0, x1, x2, inv);
    tri_hvp_corner(s, p0, p1, p2, area, a.mat, corner, g0, a.psd != 0);
    g1[0] = g1[1] = g1[2] = 0;
  }
}

// One contact incidence entry's term at the vertex (applyHvp's addDistGrad).
__device__ __forceinline__ bool contact_term(const DContactHvp& H, uint32_t ent, const double* __restrict__ p,
                                             double* c) {
  const uint32_t ci = ent >> 3;
  const int slot = ent & 7;
  const double g = contact_dd(H, ci, p);
  if (g == 0) return false;
  const Frozen& f = H.act[ci];
  if (f.kind == 2) {
    c[0] = 0; c[1] = g; c[2] = 0;
  } else if (f.kind == 0) {
    if (slot == 0) {
      c[0] = g * f.nx; c[1] = g * f.ny; c[2] = g * f.nz;
    } else {
      const double w = slot == 1 ? f.w0 : slot == 2 ? f.w1 : f.w2;
      c[0] = -g * w * f.nx; c[1] = -g * w * f.ny; c[2] = -g * w * f.nz;
    }
  } else {
    const double w = slot == 0 ? f.w0 : slot == 1 ? f.w1 : slot == 2 ? f.w2 : f.w3;
    c[0] = g * w * f.nx; c[1] = g * w * f.ny; c[2] = g * w * f.nz;
  }
  return true;
}

constexpr int kLanesPerVertex = 8;
constexpr int kVertsPerBlock = kBlock / kLanesPerVertex;

// hvpFull at vertex v: M p / h^2 + H_membrane p (+ contact), unfiltered.
// The 8 lanes of a vertex group evaluate up to 8 incident corners (and
// contact entries) at once; lane 0 of the group adds them in list order, so
// the sums are the reference's element-order sums bit for bit.
__device__ __forceinline__ void hvp_vertex8(const PcgArgs& a, uint32_t v, bool valid, const double* __restrict__ p,
                                            double h, double* out) {
  const DMesh& me = a.me;
  const int sub = threadIdx.x % kLanesPerVertex;
  const int base = (threadIdx.x % kWave) - sub;
  double acc0[3] = {0, 0, 0}, acc1[3] = {0, 0, 0};
  const uint32_t b = valid ? me.vt_off[v] : 0, e = valid ? me.vt_off[v + 1] : 0;
  // rounds of 8 corners; the round count is uniform within the group
  for (uint32_t r0 = b; r0 < e; r0 += kLanesPerVertex) {
    double g0[3] = {0, 0, 0}, g1[3] = {0, 0, 0};
    const uint32_t q = r0 + sub;
    if (q < e) hvp_corner(a, me.vt_corner[q], p, h, g0, g1);
    const uint32_t cnt = min((uint32_t)kLanesPerVertex, e - r0);
    for (uint32_t j = 0; j < cnt; j++) {
#pragma unroll
      for (int i = 0; i < 3; i++) {
        acc0[i] += __shfl(g0[i], base + (int)j, kWave);
        acc1[i] += __shfl(g1[i], base + (int)j, kWave);
      }
    }
  }
  double Hp[3];
  if (a.hvp_mode == HVP_FD) {
#pragma unroll
    for (int i = 0; i < 3; i++) Hp[i] = (acc0[i] - acc1[i]) / (2 * h);
  } else {
    Hp[0] = acc0[0]; Hp[1] = acc0[1]; Hp[2] = acc0[2];
  }
  const double m = valid ? (double)me.mass[v] : 0.0;
#pragma unroll
  for (int i = 0; i < 3; i++) out[i] = m * (valid ? p[(size_t)v * 3 + i] : 0.0) * a.inv_h2 + Hp[i];
  if (a.contact.act && a.contact.nv && valid) {
    // applyHvp (contact-assembly.ts:371-392): cHp summed per vertex in list
    // order by k_contact_vsum, then out += cHp (vsum holds the sums of this
    // HVP only while the active set has incidence entries: nv > 0)
    out[0] += a.contact.vsum[(size_t)v * 3];
    out[1] += a.contact.vsum[(size_t)v * 3 + 1];
    out[2] += a.contact.vsum[(size_t)v * 3 + 2];
  }
}

// applyHvp's contact part (out.fill(0), then addDistGrad per active contact
// in list order, skipping dd == 0) in two passes: every incidence entry's
// term in parallel (a skipped entry stores +0, which leaves a sum that
// started at +0 unchanged), then one thread per contact vertex and component
// adds its entries in list order.
__global__ void __launch_bounds__(kBlock) k_contact_terms(DContactHvp H, const double* __restrict__ p,
                                                          const int32_t* done) {
  if (block_latched(done)) return;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= H.E) return;
  double t[3] = {0, 0, 0};
  if (!contact_term(H, H.v_ent[e], p, t)) t[0] = t[1] = t[2] = 0;
  H.term[(size_t)e * 3] = t[0];
  H.term[(size_t)e * 3 + 1] = t[1];
  H.term[(size_t)e * 3 + 2] = t[2];
}

__global__ void __launch_bounds__(kBlock) k_contact_vsum(DContactHvp H, const int32_t* done) {
  if (block_latched(done)) return;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= H.nv * 3) return;
  const uint32_t v = H.cverts[t / 3];
  const int q = t % 3;
  H.vsum[(size_t)v * 3 + q] = ordered_sum<3, 0>(H.term + q, H.v_off[v], H.v_off[v + 1], 0.0);
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
    if (a.precond != PRE_SCHWARZ) {
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
  block_partial_dd<kBlock>(bb, part(a, PART_BB));
  block_partial_dd<kBlock>(rr, part(a, PART_RR));
  if (a.precond != PRE_SCHWARZ) {
    block_partial_dd<kBlock>(rz, part(a, PART_RZ));
    block_partial_dd<kBlock>(pp, part(a, PART_PP));
  }
}

// External preconditioner (Schwarz): z was written by the apply; the pin
// filter on z (pcg.ts:46, :84), then p = z (init only), r.z and p.p partials.
__global__ void __launch_bounds__(kBlock) k_pcg_ext_z(PcgArgs a, int init) {
  if (!init && block_latched(&a.s->done)) return;
  jsm::dd rz = jsm::dd_zero(), pp = jsm::dd_zero();
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.me.n; v += gridDim.x * blockDim.x) {
    const bool pinned = a.me.pinmask[v] != 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const size_t j = (size_t)v * 3 + i;
      const double zi = pinned ? 0.0 : a.z[j];
      a.z[j] = zi;
      jsm::dd_add_prod(rz, a.r[j], zi);
      if (init) {
        a.p[j] = zi;
        jsm::dd_add_prod(pp, zi, zi);
      }
    }
  }
  block_partial_dd<kBlock>(rz, part(a, PART_RZ));
  if (init) block_partial_dd<kBlock>(pp, part(a, PART_PP));
}

// After init, one block: tolerance, initial residual and rz, early return.
__global__ void __launch_bounds__(kBlock) k_pcg_init_decide(PcgArgs a, unsigned count) {
  const double bb = jsm::dd_value(merge_partials_dd<kBlock>(part(a, PART_BB), count));
  const double rr = jsm::dd_value(merge_partials_dd<kBlock>(part(a, PART_RR), count));
  const double rz = jsm::dd_value(merge_partials_dd<kBlock>(part(a, PART_RZ), count));
  if (threadIdx.x == 0) {
    PcgScalars* s = a.s;
    s->bnorm = jsm::sqrt_(bb);
    s->abs_tol = jsm::jmax(a.tol * jsm::jmax(s->bnorm, 1e-12), 1e-12);
    s->residual = jsm::sqrt_(rr);
    s->rz[0] = rz;
    s->iters = 0;
    s->done = (s->residual < s->abs_tol) ? 2 : 0;  // 2 converged, 1 (k_pcg_update) breakdown
  }
}

__global__ void __launch_bounds__(kBlock) k_pcg_hvp(PcgArgs a, int k, unsigned pp_count) {
  PcgScalars* s = a.s;
  if (block_latched(&s->done)) return;
  double h = 0;
  if (a.hvp_mode == HVP_FD)
    h = 1e-8 * (1 + jsm::sqrt_(jsm::dd_value(merge_partials_dd<kBlock>(part(a, PART_PP), pp_count))));
  jsm::dd pap = jsm::dd_zero();
  const int sub = threadIdx.x % kLanesPerVertex;
  const uint32_t stride = gridDim.x * kVertsPerBlock;
  // the loop count is uniform per vertex group (all 8 lanes share v)
  for (uint32_t v0 = blockIdx.x * kVertsPerBlock; v0 < a.me.n; v0 += stride) {
    const uint32_t v = v0 + threadIdx.x / kLanesPerVertex;
    const bool valid = v < a.me.n;
    double o[3];
    hvp_vertex8(a, v, valid, a.p, h, o);
    if (valid && sub == 0) {
      const bool pinned = a.me.pinmask[v] != 0;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const size_t j = (size_t)v * 3 + i;
        const double ap = pinned ? 0.0 : o[i];
        a.Ap[j] = ap;
        jsm::dd_add_prod(pap, a.p[j], ap);
      }
    }
  }
  block_partial_dd<kBlock>(pap, part(a, PART_PAP));
}

// out = M p / h^2 + H_membrane(x) p (analytic), pins zeroed: the exact
// membrane-and-inertia part of the Hessian for the step adjoint.
__global__ void __launch_bounds__(kBlock) k_hvp64_analytic(PcgArgs a, const double* __restrict__ p,
                                                           double* __restrict__ out) {
  const int sub = threadIdx.x % kLanesPerVertex;
  const uint32_t stride = gridDim.x * kVertsPerBlock;
  for (uint32_t v0 = blockIdx.x * kVertsPerBlock; v0 < a.me.n; v0 += stride) {
    const uint32_t v = v0 + threadIdx.x / kLanesPerVertex;
    const bool valid = v < a.me.n;
    double o[3];
    hvp_vertex8(a, v, valid, p, 0.0, o);
    if (valid && sub == 0) {
      const bool pinned = a.me.pinmask[v] != 0;
#pragma unroll
      for (int i = 0; i < 3; i++) out[(size_t)v * 3 + i] = pinned ? 0.0 : o[i];
    }
  }
}

void launch_hvp64_analytic(hipStream_t st, const DMesh& me, const DMat& mat, double inv_h2, const double* x,
                           const double* p, double* out) {
  PcgArgs a{};
  a.me = me;
  a.mat = mat;
  a.inv_h2 = inv_h2;
  a.x = x;
  a.hvp_mode = HVP_ANALYTIC;
  a.contact.act = nullptr;
  const unsigned g = (unsigned)std::min<size_t>(kMaxHvpBlocks, ((size_t)me.n + kVertsPerBlock - 1) / kVertsPerBlock);
  hipLaunchKernelGGL(k_hvp64_analytic, dim3(std::max(1u, g)), dim3(kBlock), 0, st, a, p, out);
  CS_LAUNCH_CHECK();
}

__global__ void __launch_bounds__(kBlock) k_pcg_update(PcgArgs a, int k, unsigned count) {
  PcgScalars* s = a.s;
  if (block_latched(&s->done)) return;
  const double pAp = jsm::dd_value(merge_partials_dd<kBlock>(part(a, PART_PAP), count));
  if (!(pAp > 1e-30)) {
    // non-positive curvature: stop with the current x (pcg.ts:65-68)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      s->pAp = pAp;
      s->done = 1;
      s->iters = k;
    }
    return;
  }
  const double alpha = s->rz[k & 1] / pAp;
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
    if (a.precond != PRE_SCHWARZ) {
      precond_vertex(a, v, r3, z3);
#pragma unroll
      for (int i = 0; i < 3; i++) {
        if (pinned) z3[i] = 0.0;
        a.z[(size_t)v * 3 + i] = z3[i];
        jsm::dd_add_prod(rz, r3[i], z3[i]);
      }
    }
  }
  block_partial_dd<kBlock>(rr, part(a, PART_RR));
  if (a.precond != PRE_SCHWARZ) block_partial_dd<kBlock>(rz, part(a, PART_RZ));
}

__global__ void __launch_bounds__(kBlock) k_pcg_dir(PcgArgs a, int k) {
  PcgScalars* s = a.s;
  if (block_latched(&s->done)) return;
  const double residual = jsm::sqrt_(jsm::dd_value(merge_partials_dd<kBlock>(part(a, PART_RR), gridDim.x)));
  if (residual < s->abs_tol) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      s->residual = residual;
      s->done = 2;  // converged in this iteration
      s->iters = k + 1;
    }
    return;
  }
  const double rzNew = jsm::dd_value(merge_partials_dd<kBlock>(part(a, PART_RZ), gridDim.x));
  const double beta = rzNew / s->rz[k & 1];
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
  block_partial_dd<kBlock>(pp, part(a, PART_PP));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s->rz[(k + 1) & 1] = rzNew;
    s->residual = residual;
  }
}

unsigned pcg_grid(const PcgArgs& a) { return red_grid(a.me.n); }
unsigned pcg_hvp_grid(const PcgArgs& a) {
  return (unsigned)std::min<size_t>(kMaxHvpBlocks, ((size_t)a.me.n + kVertsPerBlock - 1) / kVertsPerBlock);
}

void pcg_init(hipStream_t st, const PcgArgs& a) {
  hipLaunchKernelGGL(k_pcg_init, dim3(pcg_grid(a)), dim3(kBlock), 0, st, a);
  CS_LAUNCH_CHECK();
  if (a.precond == PRE_SCHWARZ) {
    a.sw->apply64(a.r, a.z, nullptr);
    hipLaunchKernelGGL(k_pcg_ext_z, dim3(pcg_grid(a)), dim3(kBlock), 0, st, a, 1);
    CS_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_pcg_init_decide, dim3(1), dim3(kBlock), 0, st, a, pcg_grid(a));
  CS_LAUNCH_CHECK();
}

// Back-to-back HVP launches on the operands of the last solve (Ap and the
// p.Ap partials are overwritten; the latch is cleared for the run and set after).
double pcg_time_hvp(hipStream_t st, const PcgArgs& a, int reps) {
  if (reps < 1) return 0.0;
  CS_HIP(hipMemsetAsync(&a.s->done, 0, sizeof(int32_t), st));
  hipEvent_t e0, e1;
  CS_HIP(hipEventCreate(&e0));
  CS_HIP(hipEventCreate(&e1));
  const dim3 g(pcg_hvp_grid(a)), b(kBlock);
  hipLaunchKernelGGL(k_pcg_hvp, g, b, 0, st, a, 0, pcg_grid(a));
  CS_HIP(hipEventRecord(e0, st));
  for (int i = 0; i < reps; i++) hipLaunchKernelGGL(k_pcg_hvp, g, b, 0, st, a, 0, pcg_grid(a));
  CS_HIP(hipEventRecord(e1, st));
  CS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  CS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const int32_t one = 1;
  CS_HIP(hipMemcpyAsync(&a.s->done, &one, sizeof(one), hipMemcpyHostToDevice, st));
  CS_HIP(hipStreamSynchronize(st));
  return (double)ms / reps;
}

void pcg_iteration(hipStream_t st, const PcgArgs& a, int k) {
  pcg_iteration_hvp(st, a, k);
  pcg_iteration_rest(st, a, k);
}

// The HVP grid differs from the vector grid; its partial count is passed on.
void pcg_iteration_hvp(hipStream_t st, const PcgArgs& a, int k) {
  if (a.contact.act && a.contact.nv) {
    hipLaunchKernelGGL(k_contact_terms, dim3(grid_for(a.contact.E)), dim3(kBlock), 0, st, a.contact, a.p, &a.s->done);
    hipLaunchKernelGGL(k_contact_vsum, dim3(grid_for((size_t)a.contact.nv * 3)), dim3(kBlock), 0, st, a.contact,
                       &a.s->done);
  }
  hipLaunchKernelGGL(k_pcg_hvp, dim3(pcg_hvp_grid(a)), dim3(kBlock), 0, st, a, k, pcg_grid(a));
  CS_LAUNCH_CHECK();
}

void pcg_iteration_rest(hipStream_t st, const PcgArgs& a, int k) {
  const dim3 g(pcg_grid(a)), b(kBlock);
  hipLaunchKernelGGL(k_pcg_update, g, b, 0, st, a, k, pcg_hvp_grid(a));
  if (a.precond == PRE_SCHWARZ) {
    a.sw->apply64(a.r, a.z, &a.s->done);
    hipLaunchKernelGGL(k_pcg_ext_z, g, b, 0, st, a, 0);
  }
  hipLaunchKernelGGL(k_pcg_dir, g, b, 0, st, a, k);
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
                                                size_t n, jsm::dd* partials) {
  jsm::dd acc = jsm::dd_zero();
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    jsm::dd_add_prod(acc, a[i], b[i]);
  block_partial_dd<kBlock>(acc, partials);
}

void launch_dot(hipStream_t st, const double* a, const double* b, size_t n, jsm::dd* partials, double* out) {
  const unsigned grid = red_grid(n);
  hipLaunchKernelGGL(k_dot, dim3(grid), dim3(kBlock), 0, st, a, b, n, partials);
  CS_LAUNCH_CHECK();
  FinishJobs f{};
  f.j[0] = FinishJob{partials, grid, FIN_DD, 0, out};
  launch_finish(st, f, 1);
}

__global__ void __launch_bounds__(kBlock) k_max_abs(const double* __restrict__ a, size_t n, double* partials) {
  double m = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double v = jsm::fabs_(a[i]);
    if (v > m) m = v;
  }
  block_partial_ext<kBlock, GtMax>(m, partials);
}

void launch_max_abs(hipStream_t st, const double* a, size_t n, double* partials, double* out) {
  const unsigned grid = red_grid(n);
  hipLaunchKernelGGL(k_max_abs, dim3(grid), dim3(kBlock), 0, st, a, n, partials);
  CS_LAUNCH_CHECK();
  FinishJobs f{};
  f.j[0] = FinishJob{partials, grid, FIN_GTMAX, 1, out};
  launch_finish(st, f, 1);
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
