// mixed.hip — FP32 inner PCG of the mixed-precision Newton step.
//
// Same algorithm as pcg.ts:11-95 (relative tolerance, breakdown test, pin
// filter on every vector) with the analytic membrane HVP
// (membrane-hvp.ts:18-108) evaluated in FP32 and FP32 reductions.  The
// per-vertex gather over the element-ordered CSR keeps the result
// deterministic run to run.
#include "fem.h"
#include "mixed.h"

namespace cloth {

#pragma clang fp contract(fast)

struct Frozen32 {
  int32_t kind, p, a, b, c, d;
  float w0, w1, w2, w3, nx, ny, nz, coeff;
};

struct MxScalars {
  float rz, pAp, alpha, rznew, residual, abs_tol;
  int32_t done, iters;
  unsigned tickets[8];
};

template <int BLOCK>
__device__ __forceinline__ bool grid_reduce_f(float v, float* partials, unsigned* ticket, float& out) {
  __shared__ float s[BLOCK / kWave];
  __shared__ bool s_last;
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0) s[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = 0;
    for (int w = 0; w < BLOCK / kWave; w++) b += s[w];
    partials[blockIdx.x] = b;
    __threadfence();
    s_last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  float acc = 0;
  for (unsigned i = threadIdx.x; i < gridDim.x; i += BLOCK) acc += partials[i];
  for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, kWave);
  __syncthreads();
  if (lane == 0) s[wid] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0;
    for (int w = 0; w < BLOCK / kWave; w++) t += s[w];
    *ticket = 0;
    out = t;
    return true;
  }
  return false;
}

struct MxArgs {
  DMesh32 me;
  float c00, c11, c01, g, thick, inv_h2;
  const float* x;
  const float* diag;
  const float* binv;  // null => Jacobi
  float* sol;
  float* r;
  float* z;
  float* p;
  float* Ap;
  float tol;
  MxScalars* s;
  float* partials;  // 4 * 1024
  // contacts
  uint32_t C;
  const Frozen32* act;
  float* dd;
  const uint32_t* v_off;
  const uint32_t* v_ent;
};

__device__ __forceinline__ void ld3f(const float* x, uint32_t v, float* o) {
  o[0] = x[(size_t)v * 3];
  o[1] = x[(size_t)v * 3 + 1];
  o[2] = x[(size_t)v * 3 + 2];
}

// Analytic membrane HVP corner contribution (membrane-hvp.ts:18-75) in FP32.
__device__ __forceinline__ void hvp_corner32(const MxArgs& a, uint32_t t, int corner, const float* p, float* out) {
  const uint32_t i0 = a.me.idx[(size_t)t * 3], i1 = a.me.idx[(size_t)t * 3 + 1], i2 = a.me.idx[(size_t)t * 3 + 2];
  float x0[3], x1[3], x2[3], v0[3], v1[3], v2[3];
  ld3f(a.x, i0, x0); ld3f(a.x, i1, x1); ld3f(a.x, i2, x2);
  ld3f(p, i0, v0); ld3f(p, i1, v1); ld3f(p, i2, v2);
  const float4 inv = *reinterpret_cast<const float4*>(a.me.inv_dm + (size_t)t * 4);
  const float ia = inv.x, ib = inv.y, ic = inv.z, id = inv.w;
  float F0[3], F1[3], dF0[3], dF1[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float e1 = x1[k] - x0[k], e2 = x2[k] - x0[k];
    const float w1 = v1[k] - v0[k], w2 = v2[k] - v0[k];
    F0[k] = ia * e1 + ic * e2;
    F1[k] = ib * e1 + id * e2;
    dF0[k] = ia * w1 + ic * w2;
    dF1[k] = ib * w1 + id * w2;
  }
  const float E00 = 0.5f * (F0[0] * F0[0] + F0[1] * F0[1] + F0[2] * F0[2] - 1.f);
  const float E11 = 0.5f * (F1[0] * F1[0] + F1[1] * F1[1] + F1[2] * F1[2] - 1.f);
  const float E01 = 0.5f * (F0[0] * F1[0] + F0[1] * F1[1] + F0[2] * F1[2]);
  const float dE00 = F0[0] * dF0[0] + F0[1] * dF0[1] + F0[2] * dF0[2];
  const float dE11 = F1[0] * dF1[0] + F1[1] * dF1[1] + F1[2] * dF1[2];
  const float dE01 = 0.5f * (dF0[0] * F1[0] + F0[0] * dF1[0] + dF0[1] * F1[1] + F0[1] * dF1[1] + dF0[2] * F1[2] +
                             F0[2] * dF1[2]);
  const float S00 = a.c00 * E00 + a.c01 * E11, S11 = a.c11 * E11 + a.c01 * E00, S01 = 2.f * a.g * E01;
  const float dS00 = a.c00 * dE00 + a.c01 * dE11, dS11 = a.c11 * dE11 + a.c01 * dE00, dS01 = 2.f * a.g * dE01;
  const float sc = a.thick * a.me.area[t];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float dP0 = dF0[k] * S00 + F0[k] * dS00 + dF1[k] * S01 + F1[k] * dS01;
    const float dP1 = dF0[k] * S01 + F0[k] * dS01 + dF1[k] * S11 + F1[k] * dS11;
    const float g1 = sc * (dP0 * ia + dP1 * ib);
    const float g2 = sc * (dP0 * ic + dP1 * id);
    out[k] = corner == 0 ? -(g1 + g2) : corner == 1 ? g1 : g2;
  }
}

__device__ __forceinline__ void precond32(const MxArgs& a, uint32_t v, const float* r3, float* z3) {
  if (a.binv) {
    const float* f = a.binv + (size_t)v * 9;
    z3[0] = f[0] * r3[0] + f[1] * r3[1] + f[2] * r3[2];
    z3[1] = f[3] * r3[0] + f[4] * r3[1] + f[5] * r3[2];
    z3[2] = f[6] * r3[0] + f[7] * r3[1] + f[8] * r3[2];
  } else {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const float d = a.diag[(size_t)v * 3 + i];
      z3[i] = d > 1e-12f ? r3[i] / d : r3[i];
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_mx_convert(uint32_t n, const double* __restrict__ x,
                                                       const double* __restrict__ b, const double* __restrict__ diag,
                                                       const double* __restrict__ binv, const uint8_t* pinmask,
                                                       float* x32, float* r32, float* d32, float* binv32, float* sol) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const bool pinned = pinmask[v] != 0;
  for (int i = 0; i < 3; i++) {
    const size_t j = (size_t)v * 3 + i;
    x32[j] = (float)x[j];
    r32[j] = pinned ? 0.f : (float)b[j];
    d32[j] = (float)diag[j];
    sol[j] = 0.f;
  }
  if (binv)
    for (int q = 0; q < 9; q++) binv32[(size_t)v * 9 + q] = (float)binv[(size_t)v * 9 + q];
}

__global__ void k_mx_contacts(const Frozen* __restrict__ act, const double* __restrict__ coeff, uint32_t C,
                              Frozen32* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  const Frozen f = act[i];
  Frozen32 o;
  o.kind = f.kind; o.p = f.p; o.a = f.a; o.b = f.b; o.c = f.c; o.d = f.d;
  o.w0 = (float)f.w0; o.w1 = (float)f.w1; o.w2 = (float)f.w2; o.w3 = (float)f.w3;
  o.nx = (float)f.nx; o.ny = (float)f.ny; o.nz = (float)f.nz;
  o.coeff = (float)coeff[i];
  out[i] = o;
}

__global__ void __launch_bounds__(kBlock) k_mx_init(MxArgs a) {
  float rz = 0, bb = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.me.n; v += gridDim.x * blockDim.x) {
    const bool pinned = a.me.pinmask[v] != 0;
    float r3[3], z3[3];
    ld3f(a.r, v, r3);
    precond32(a, v, r3, z3);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const size_t j = (size_t)v * 3 + i;
      if (pinned) z3[i] = 0.f;
      a.z[j] = z3[i];
      a.p[j] = z3[i];
      rz += r3[i] * z3[i];
      bb += r3[i] * r3[i];
    }
  }
  MxScalars* s = a.s;
  float t;
  if (grid_reduce_f<kBlock>(rz, a.partials, &s->tickets[0], t)) s->rz = t;
  if (grid_reduce_f<kBlock>(bb, a.partials + 1024, &s->tickets[1], t)) {
    s->residual = sqrtf(t);
    s->abs_tol = fmaxf(a.tol * fmaxf(s->residual, 1e-12f), 1e-12f);
  }
}

__global__ void k_mx_decide(MxScalars* s) {
  s->iters = 0;
  s->done = s->residual < s->abs_tol ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_mx_jv(MxArgs a) {
  if (a.s->done) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.C) return;
  const Frozen32 c = a.act[i];
  float j;
  if (c.kind == 2) {
    j = a.p[(size_t)c.p * 3 + 1];
  } else if (c.kind == 0) {
    float q[3];
    ld3f(a.p, c.p, q);
    j = c.nx * q[0] + c.ny * q[1] + c.nz * q[2];
    const float ws[3] = {c.w0, c.w1, c.w2};
    const int vs[3] = {c.a, c.b, c.c};
    for (int k = 0; k < 3; k++) {
      if (vs[k] < 0) continue;
      ld3f(a.p, vs[k], q);
      j -= ws[k] * (c.nx * q[0] + c.ny * q[1] + c.nz * q[2]);
    }
  } else {
    const float ws[4] = {c.w0, c.w1, c.w2, c.w3};
    const int vs[4] = {c.a, c.b, c.c, c.d};
    j = 0;
    for (int k = 0; k < 4; k++) {
      if (vs[k] < 0) continue;
      float q[3];
      ld3f(a.p, vs[k], q);
      j += ws[k] * (c.nx * q[0] + c.ny * q[1] + c.nz * q[2]);
    }
  }
  a.dd[i] = c.coeff * j;
}

__global__ void __launch_bounds__(kBlock) k_mx_hvp(MxArgs a, int k) {
  MxScalars* s = a.s;
  if (s->done) return;
  float pap = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.me.n; v += gridDim.x * blockDim.x) {
    float acc[3] = {0, 0, 0};
    const uint32_t e = a.me.vt_off[v + 1];
    for (uint32_t q = a.me.vt_off[v]; q < e; q++) {
      const uint32_t rec = a.me.vt_rec[q];
      float gc[3];
      hvp_corner32(a, rec >> 2, rec & 3, a.p, gc);
      acc[0] += gc[0]; acc[1] += gc[1]; acc[2] += gc[2];
    }
    float p3[3];
    ld3f(a.p, v, p3);
    const float mh = a.me.mass[v] * a.inv_h2;
    float o[3] = {mh * p3[0] + acc[0], mh * p3[1] + acc[1], mh * p3[2] + acc[2]};
    if (a.act) {
      for (uint32_t q = a.v_off[v]; q < a.v_off[v + 1]; q++) {
        const uint32_t ent = a.v_ent[q];
        const uint32_t ci = ent >> 3;
        const int slot = ent & 7;
        const float g = a.dd[ci];
        if (g == 0.f) continue;
        const Frozen32 c = a.act[ci];
        if (c.kind == 2) {
          o[1] += g;
        } else {
          float w;
          if (c.kind == 0) w = slot == 0 ? 1.f : -(slot == 1 ? c.w0 : slot == 2 ? c.w1 : c.w2);
          else w = slot == 0 ? c.w0 : slot == 1 ? c.w1 : slot == 2 ? c.w2 : c.w3;
          o[0] += g * w * c.nx; o[1] += g * w * c.ny; o[2] += g * w * c.nz;
        }
      }
    }
    const bool pinned = a.me.pinmask[v] != 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const float ap = pinned ? 0.f : o[i];
      a.Ap[(size_t)v * 3 + i] = ap;
      pap += p3[i] * ap;
    }
  }
  float t;
  if (grid_reduce_f<kBlock>(pap, a.partials, &s->tickets[2], t)) {
    s->pAp = t;
    if (!(t > 1e-30f)) {
      s->done = 1;
      s->iters = k;
    } else {
      s->alpha = s->rz / t;
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_mx_update(MxArgs a, int k) {
  MxScalars* s = a.s;
  if (s->done) return;
  const float alpha = s->alpha;
  float rr = 0, rz = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.me.n; v += gridDim.x * blockDim.x) {
    const bool pinned = a.me.pinmask[v] != 0;
    float r3[3], z3[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const size_t j = (size_t)v * 3 + i;
      float xi = a.sol[j] + alpha * a.p[j];
      float ri = a.r[j] - alpha * a.Ap[j];
      if (pinned) { xi = 0.f; ri = 0.f; }
      a.sol[j] = xi;
      a.r[j] = ri;
      r3[i] = ri;
      rr += ri * ri;
    }
    precond32(a, v, r3, z3);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      if (pinned) z3[i] = 0.f;
      a.z[(size_t)v * 3 + i] = z3[i];
      rz += r3[i] * z3[i];
    }
  }
  float t;
  if (grid_reduce_f<kBlock>(rr, a.partials, &s->tickets[3], t)) {
    s->residual = sqrtf(t);
    if (s->residual < s->abs_tol) {
      s->done = 2;
      s->iters = k + 1;
    }
  }
  if (grid_reduce_f<kBlock>(rz, a.partials + 1024, &s->tickets[4], t)) s->rznew = t;
}

__global__ void __launch_bounds__(kBlock) k_mx_dir(MxArgs a) {
  MxScalars* s = a.s;
  if (s->done) return;
  const float beta = s->rznew / s->rz;
  float dummy = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.me.n; v += gridDim.x * blockDim.x) {
    const bool pinned = a.me.pinmask[v] != 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const size_t j = (size_t)v * 3 + i;
      a.p[j] = pinned ? 0.f : a.z[j] + beta * a.p[j];
    }
  }
  float t;
  if (grid_reduce_f<kBlock>(dummy, a.partials + 2048, &s->tickets[5], t)) s->rz = s->rznew;
}

__global__ void k_mx_out(const float* sol, double* dx, size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n3) dx[i] = (double)sol[i];
}

struct MixedImpl {
  hipStream_t st;
  uint32_t n;
  DevBuf<float> x32, d32, binv32, sol, r, z, p, Ap, dd, partials;
  DevBuf<Frozen32> act32;
  DevBuf<MxScalars> s;
};

MixedSolve::MixedSolve(const cs_scene_desc& d, const cs_options&, hipStream_t st) : im_(new MixedImpl) {
  im_->st = st;
  im_->n = d.n_verts;
  const size_t n3 = (size_t)d.n_verts * 3;
  for (auto* b : {&im_->x32, &im_->d32, &im_->sol, &im_->r, &im_->z, &im_->p, &im_->Ap}) b->alloc(n3);
  im_->partials.alloc(4 * 1024);
  im_->s.alloc(1);
  CS_HIP(hipMemsetAsync(im_->s.ptr, 0, sizeof(MxScalars), st));
}

MixedSolve::~MixedSolve() { delete im_; }

int MixedSolve::solve_plain(const MixedInputs& in, void (*timer)(void*, int, bool), void* ctx) {
  MixedImpl& M = *im_;
  hipStream_t st = M.st;
  const uint32_t n = in.me.n;
  if (in.binv) M.binv32.alloc((size_t)n * 9);
  const uint32_t C = in.contact.act ? in.contact.C : 0;
  if (in.contact.act) {
    M.act32.alloc(C ? C : 1);
    M.dd.alloc(C ? C : 1);
    if (C)
      hipLaunchKernelGGL(k_mx_contacts, dim3(grid_for(C)), dim3(kBlock), 0, st, in.contact.act, in.contact.coeff, C,
                         M.act32.ptr);
  }
  hipLaunchKernelGGL(k_mx_convert, dim3(grid_for(n)), dim3(kBlock), 0, st, n, in.x, in.b, in.diag, in.binv,
                     in.me.pinmask, M.x32.ptr, M.r.ptr, M.d32.ptr, in.binv ? M.binv32.ptr : nullptr, M.sol.ptr);
  CS_LAUNCH_CHECK();
  MxArgs a;
  a.me = DMesh32{in.me.n, in.me.m, in.me.idx, in.me.inv_dm, in.me.area, in.me.mass, in.me.vt_off, in.me.vt_rec,
                 in.me.pinmask};
  a.c00 = (float)in.mat.c00; a.c11 = (float)in.mat.c11; a.c01 = (float)in.mat.c01; a.g = (float)in.mat.g;
  a.thick = (float)in.mat.thick;
  a.inv_h2 = (float)in.inv_h2;
  a.x = M.x32.ptr; a.diag = M.d32.ptr; a.binv = in.binv ? M.binv32.ptr : nullptr;
  a.sol = M.sol.ptr; a.r = M.r.ptr; a.z = M.z.ptr; a.p = M.p.ptr; a.Ap = M.Ap.ptr;
  a.tol = (float)in.tol;
  a.s = M.s.ptr;
  a.partials = M.partials.ptr;
  a.C = C;
  a.act = in.contact.act ? M.act32.ptr : nullptr;
  a.dd = in.contact.act ? M.dd.ptr : nullptr;
  a.v_off = in.contact.v_off;
  a.v_ent = in.contact.v_ent;
  const dim3 g(red_grid(n)), b(kBlock);
  hipLaunchKernelGGL(k_mx_init, g, b, 0, st, a);
  hipLaunchKernelGGL(k_mx_decide, dim3(1), dim3(1), 0, st, M.s.ptr);
  CS_LAUNCH_CHECK();
  const int chunk = 8;
  for (int k = 0; k < in.max_iters; k++) {
    if (C) hipLaunchKernelGGL(k_mx_jv, dim3(grid_for(C)), dim3(kBlock), 0, st, a);
    if (timer) timer(ctx, k, true);
    hipLaunchKernelGGL(k_mx_hvp, g, b, 0, st, a, k);
    if (timer) timer(ctx, k, false);
    hipLaunchKernelGGL(k_mx_update, g, b, 0, st, a, k);
    hipLaunchKernelGGL(k_mx_dir, g, b, 0, st, a);
    CS_LAUNCH_CHECK();
    if ((k + 1) % chunk == 0 || k + 1 == in.max_iters) {
      MxScalars hs;
      CS_HIP(hipMemcpyAsync(&hs, M.s.ptr, sizeof(hs), hipMemcpyDeviceToHost, st));
      CS_HIP(hipStreamSynchronize(st));
      if (hs.done) break;
    }
  }
  hipLaunchKernelGGL(k_mx_out, dim3(grid_for((size_t)n * 3)), dim3(kBlock), 0, st, M.sol.ptr, in.dx, (size_t)n * 3);
  CS_LAUNCH_CHECK();
  MxScalars hs;
  CS_HIP(hipMemcpyAsync(&hs, M.s.ptr, sizeof(hs), hipMemcpyDeviceToHost, st));
  CS_HIP(hipStreamSynchronize(st));
  return hs.done ? hs.iters : in.max_iters;
}

}  // namespace cloth
