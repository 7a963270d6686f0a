// bend.hip — analytic hinge bending gradient and Hessian-vector product on
// the device (the per-hinge math is in bend.h).  Each hinge writes its 12
// terms (4 corners x xyz); each vertex then sums the terms of its hinges
// through the vertex -> (hinge << 2 | corner) CSR in ascending hinge order,
// so the per-vertex sums are deterministic and follow the reference's
// hinge-order scatter (bending.ts:29-39).
#include "bend.h"
#include "fem.h"

namespace cloth {

__device__ __forceinline__ void hinge_load(const DMesh& me, uint32_t h, uint32_t (&v)[4], double& rest, double& w) {
  const uint4 q = reinterpret_cast<const uint4*>(me.hv)[h];
  v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  rest = me.hrest[h];
  w = hinge_weight(me.hlen[h], me.hasum[h]);
}

template <bool kHvp>
__global__ void __launch_bounds__(kBlock) k_bend_terms(DMesh me, double k, const double* __restrict__ x,
                                                       const double* __restrict__ p, double* __restrict__ term) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= me.nh) return;
  uint32_t v[4];
  double rest, w;
  hinge_load(me, h, v, rest, w);
  double* out = term + (size_t)h * 12;
  if (!kHvp) {
    double X[4][3], G[4][3];
    for (int c = 0; c < 4; c++)
      for (int i = 0; i < 3; i++) X[c][i] = x[(size_t)v[c] * 3 + i];
    hinge_grad_t<double>(X, rest, w, k, G);
    for (int c = 0; c < 4; c++)
      for (int i = 0; i < 3; i++) out[c * 3 + i] = G[c][i];
  } else {
    Dual X[4][3], G[4][3];
    for (int c = 0; c < 4; c++)
      for (int i = 0; i < 3; i++) X[c][i] = mkd(x[(size_t)v[c] * 3 + i], p[(size_t)v[c] * 3 + i]);
    hinge_grad_t<Dual>(X, rest, w, k, G);
    for (int c = 0; c < 4; c++)
      for (int i = 0; i < 3; i++) out[c * 3 + i] = G[c][i].d;
  }
}

// out[3v + q] (+)= sum over v's hinges in ascending order; pinned rows are
// left as they are (accumulate) or zero.
__global__ void __launch_bounds__(kBlock) k_bend_vsum(DMesh me, const double* __restrict__ term,
                                                      double* __restrict__ out, int accumulate) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= me.n * 3) return;
  const uint32_t v = t / 3, q = t % 3;
  double s = 0;
  for (uint32_t r = me.vh_off[v]; r < me.vh_off[v + 1]; r++) {
    const uint32_t rec = me.vh_rec[r];
    s += term[(size_t)(rec >> 2) * 12 + (rec & 3) * 3 + q];
  }
  const bool pinned = me.pinmask && me.pinmask[v];
  if (accumulate) {
    if (!pinned) out[t] += s;
  } else {
    out[t] = pinned ? 0.0 : s;
  }
}

void launch_bend_grad(hipStream_t st, const DMesh& me, double k, const double* x, double* term, double* out,
                      bool accumulate) {
  if (me.nh)
    hipLaunchKernelGGL(k_bend_terms<false>, dim3(grid_for(me.nh)), dim3(kBlock), 0, st, me, k, x, nullptr, term);
  hipLaunchKernelGGL(k_bend_vsum, dim3(grid_for((size_t)me.n * 3)), dim3(kBlock), 0, st, me, term, out,
                     accumulate ? 1 : 0);
  CS_LAUNCH_CHECK();
}

void launch_bend_hvp(hipStream_t st, const DMesh& me, double k, const double* x, const double* p, double* term,
                     double* out, bool accumulate) {
  if (me.nh)
    hipLaunchKernelGGL(k_bend_terms<true>, dim3(grid_for(me.nh)), dim3(kBlock), 0, st, me, k, x, p, term);
  hipLaunchKernelGGL(k_bend_vsum, dim3(grid_for((size_t)me.n * 3)), dim3(kBlock), 0, st, me, term, out,
                     accumulate ? 1 : 0);
  CS_LAUNCH_CHECK();
}

// Independent hinges for component tests: X, P 12 doubles per hinge.
__global__ void __launch_bounds__(kBlock) k_probe_hinges(const double* __restrict__ X12,
                                                         const double* __restrict__ P12,
                                                         const double* __restrict__ rest,
                                                         const double* __restrict__ len,
                                                         const double* __restrict__ asum, double k, uint64_t count,
                                                         double* grad12, double* hvp12) {
  const uint64_t h = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (h >= count) return;
  const double w = hinge_weight(len[h], asum[h]);
  double X[4][3], G[4][3];
  Dual XD[4][3], GD[4][3];
  for (int c = 0; c < 4; c++)
    for (int i = 0; i < 3; i++) {
      X[c][i] = X12[h * 12 + c * 3 + i];
      XD[c][i] = mkd(X[c][i], P12[h * 12 + c * 3 + i]);
    }
  hinge_grad_t<double>(X, rest[h], w, k, G);
  hinge_grad_t<Dual>(XD, rest[h], w, k, GD);
  for (int c = 0; c < 4; c++)
    for (int i = 0; i < 3; i++) {
      grad12[h * 12 + c * 3 + i] = G[c][i];
      hvp12[h * 12 + c * 3 + i] = GD[c][i].d;
    }
}

void probe_hinges(const double* X12, const double* P12, const double* rest, const double* len, const double* asum,
                  double k, uint64_t count, double* grad12, double* hvp12) {
  hipStream_t st = nullptr;
  DevBuf<double> x, p, r, l, a, g, hv;
  x.upload(X12, count * 12, st);
  p.upload(P12, count * 12, st);
  r.upload(rest, count, st);
  l.upload(len, count, st);
  a.upload(asum, count, st);
  g.alloc(count * 12);
  hv.alloc(count * 12);
  if (count)
    hipLaunchKernelGGL(k_probe_hinges, dim3(grid_for(count)), dim3(kBlock), 0, st, x.ptr, p.ptr, r.ptr, l.ptr, a.ptr,
                       k, count, g.ptr, hv.ptr);
  CS_LAUNCH_CHECK();
  CS_HIP(hipMemcpy(grad12, g.ptr, count * 12 * sizeof(double), hipMemcpyDeviceToHost));
  CS_HIP(hipMemcpy(hvp12, hv.ptr, count * 12 * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace cloth
