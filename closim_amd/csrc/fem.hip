// fem.hip — Newton gradient, energies and strain of the cloth step (FP64).
#include "elements.h"

namespace cloth {

// One thread per vertex: evalInternal gradient at the vertex (membrane in
// triangle order, then bending in hinge order), plus inertia
// M (x - yHat) / h^2 (newton.ts:116-120), plus the contact gradient when
// present (:122-124), pins zeroed (:126); grid-reduces sum g^2 (:127).
__global__ void __launch_bounds__(kBlock) k_gradient(DMesh me, DMat mat, const double* __restrict__ x,
                                                     const double* __restrict__ yhat, double inv_h2,
                                                     const double* __restrict__ ceg, double* __restrict__ g,
                                                     jsm::dd* partials) {
  jsm::dd acc2 = jsm::dd_zero();
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < me.n; v += gridDim.x * blockDim.x) {
    double acc[3];
    internal_grad_vertex(me, mat, x, v, acc);
    const double m = me.mass[v];
    const bool pinned = me.pinmask[v] != 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const size_t i = (size_t)v * 3 + k;
      double gi = m * (x[i] - yhat[i]) * inv_h2 + acc[k];
      if (ceg) gi += ceg[i];
      if (pinned) gi = 0;
      g[i] = gi;
      jsm::dd_add_prod(acc2, gi, gi);
    }
  }
  block_partial_dd<kBlock>(acc2, partials);
}

void launch_gradient(hipStream_t st, const DMesh& me, const DMat& mat, const double* x, const double* yhat,
                     double inv_h2, const double* ceg, double* g, jsm::dd* partials, double* out_sumsq) {
  const unsigned grid = red_grid(me.n);
  hipLaunchKernelGGL(k_gradient, dim3(grid), dim3(kBlock), 0, st, me, mat, x, yhat, inv_h2, ceg, g, partials);
  CS_LAUNCH_CHECK();
  FinishJobs f{};
  f.j[0] = FinishJob{partials, grid, FIN_DD, 0, out_sumsq};
  launch_finish(st, f, 1);
}

// totalEnergy (newton.ts:76-88) minus the contact merit term: inertia sum
// over DOFs and internalEnergyOnly (fem.ts:56-75) over triangles then hinges,
// each returned correctly rounded.
__global__ void __launch_bounds__(kBlock) k_energy(DMesh me, DMat mat, const double* __restrict__ x,
                                                   const double* __restrict__ yhat, double h,
                                                   jsm::dd* partials) {
  jsm::dd ein = jsm::dd_zero(), eint = jsm::dd_zero();
  const size_t n3 = (size_t)me.n * 3;
  const size_t nh = mat.kb > 0 ? me.nh : 0;
  const size_t total = n3 + me.m + nh;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    if (i < n3) {
      const double m = me.mass[i / 3];
      const double d = x[i] - yhat[i];
      jsm::dd_add(ein, 0.5 * m * d * d / (h * h));
    } else if (i < n3 + me.m) {
      const size_t t = i - n3;
      double x0[3], x1[3], x2[3];
      ld3(x, me.idx[t * 3], x0);
      ld3(x, me.idx[t * 3 + 1], x1);
      ld3(x, me.idx[t * 3 + 2], x2);
      const TriState s = tri_state(x0, x1, x2, me.inv_dm + t * 4);
      jsm::dd_add(eint, tri_energy(s, me.area[t], mat));
    } else {
      const size_t hh = i - n3 - me.m;
      double P[4][3];
      for (int s = 0; s < 4; s++) ld3(x, me.hv[hh * 4 + s], P[s]);
      jsm::dd_add(eint, hinge_energy4(P, me.hrest[hh], me.hlen[hh], me.hasum[hh], mat.kb));
    }
  }
  block_partial_dd<kBlock>(ein, partials);
  block_partial_dd<kBlock>(eint, partials + kMaxRedBlocks);
}

void launch_energy(hipStream_t st, const DMesh& me, const DMat& mat, const double* x, const double* yhat,
                   double h, jsm::dd* partials, double* out2) {
  const size_t total = (size_t)me.n * 3 + me.m + me.nh;
  const unsigned grid = red_grid(total);
  hipLaunchKernelGGL(k_energy, dim3(grid), dim3(kBlock), 0, st, me, mat, x, yhat, h, partials);
  CS_LAUNCH_CHECK();
  FinishJobs f{};
  f.j[0] = FinishJob{partials, grid, FIN_DD, 0, out2};
  f.j[1] = FinishJob{partials + kMaxRedBlocks, grid, FIN_DD, 0, out2 + 1};
  launch_finish(st, f, 2);
}

// maxStrain (fem.ts:37): Math.max over |E00|, |E11|, |E01| of every triangle.
__global__ void __launch_bounds__(kBlock) k_max_strain(DMesh me, const double* __restrict__ x, double* partials) {
  double mx = 0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < me.m; t += gridDim.x * blockDim.x) {
    double x0[3], x1[3], x2[3];
    ld3(x, me.idx[(size_t)t * 3], x0);
    ld3(x, me.idx[(size_t)t * 3 + 1], x1);
    ld3(x, me.idx[(size_t)t * 3 + 2], x2);
    const TriState s = tri_state(x0, x1, x2, me.inv_dm + (size_t)t * 4);
    mx = jsm::jmax(jsm::jmax(jsm::jmax(mx, jsm::fabs_(s.E00)), jsm::fabs_(s.E11)), jsm::fabs_(s.E01));
  }
  block_partial_ext<kBlock, JsMax>(mx, partials);
}

void launch_max_strain(hipStream_t st, const DMesh& me, const double* x, double* partials, double* out) {
  const unsigned grid = red_grid(me.m);
  hipLaunchKernelGGL(k_max_strain, dim3(grid), dim3(kBlock), 0, st, me, x, partials);
  CS_LAUNCH_CHECK();
  FinishJobs f{};
  f.j[0] = FinishJob{partials, grid, FIN_JSMAX, 1, out};
  launch_finish(st, f, 1);
}

}  // namespace cloth
