// mixed.hip — FP32 inner PCG of the mixed-precision Newton step.
//
// Same algorithm as pcg.ts:11-95 (relative tolerance, breakdown test, pin
// filter on every vector) with the analytic membrane HVP
// (membrane-hvp.ts:18-108) evaluated in FP32 and FP32 reductions.  The
// per-vertex gather over the element-ordered CSR keeps the result
// deterministic run to run.
#include "elements.h"
#include "fem.h"
#include "mixed.h"
#include "schwarz.h"

namespace cloth {

#pragma clang fp contract(fast)

struct Frozen32 {
  int32_t kind, p, a, b, c, d;
  float w0, w1, w2, w3, nx, ny, nz, coeff;  // coeff = 0 outside dHat
};

struct MxScalars {
  float rz[2];  // r.z of iteration k in rz[k & 1]
  float pAp, residual, abs_tol;
  int32_t done, iters;
};

enum { MX_RZ = 0, MX_BB = 1, MX_RR = 2, MX_PAP = 3, MX_NPART = 3 + kMaxHvpBlocks / kMaxRedBlocks };

struct MxArgs {
  DMesh32 me;
  float c00, c11, c01, g, thick, inv_h2;
  const float* x;
  int psd;            // stress projected to PSD (stress_psd)
  const float* diag;
  const float* binv;  // null => Jacobi
  float* sol;
  float* r;
  float* z;
  float* p;
  float* Ap;
  float tol;
  int ext;  // z comes from the Schwarz apply between the kernels
  MxScalars* s;
  float* partials;  // MX_NPART * kMaxRedBlocks
  // contacts
  uint32_t C, E;
  const Frozen32* act;
  const uint32_t* v_off;
  const uint32_t* v_ent;
  const uint32_t* ent_vtx;  // [E] vertex of each incidence entry
  float* scan;              // [3E] per-vertex running sums of the entry terms within each 64-entry chunk
};

__device__ __forceinline__ float* mpart(const MxArgs& a, int which) {
  return a.partials + (size_t)which * kMaxRedBlocks;
}

// FP32 block sum (fixed tree), every thread gets the total.
__device__ __forceinline__ float block_sum_f(float v) {
  __shared__ float s[kBlock / kWave];
  __shared__ float s_res;
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0) s[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0;
    for (int w = 0; w < kBlock / kWave; w++) t += s[w];
    s_res = t;
  }
  __syncthreads();
  const float r = s_res;
  __syncthreads();
  return r;
}

__device__ __forceinline__ void block_partial_f(float v, float* partials, uint32_t bid) {
  const float b = block_sum_f(v);
  if (threadIdx.x == 0) partials[bid] = b;
}

__device__ __forceinline__ float merge_partials_f(const float* partials, unsigned count) {
  float acc = 0;
  for (unsigned i = threadIdx.x; i < count; i += kBlock) acc += partials[i];
  return block_sum_f(acc);
}

__device__ __forceinline__ void ld3f(const float* x, uint32_t v, float* o) {
  o[0] = x[(size_t)v * 3];
  o[1] = x[(size_t)v * 3 + 1];
  o[2] = x[(size_t)v * 3 + 2];
}

// Analytic membrane HVP corner contribution (membrane-hvp.ts:18-75) in FP32.
__device__ __forceinline__ void hvp_corner32(const MxArgs& a, uint4 cr, const float* p, float* out,
                                             float* pown = nullptr) {
  const uint32_t i0 = cr.x, i1 = cr.y, i2 = cr.z, t = cr.w >> 2;
  const int corner = cr.w & 3;
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
  float S00 = a.c00 * E00 + a.c01 * E11, S11 = a.c11 * E11 + a.c01 * E00, S01 = 2.f * a.g * E01;
  if (a.psd) stress_psd(S00, S11, S01);
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
  if (pown)
#pragma unroll
    for (int k = 0; k < 3; k++) pown[k] = corner == 0 ? v0[k] : corner == 1 ? v1[k] : v2[k];
}

// Frozen barrier HVP term of one incidence entry: coeff (J.p) times the
// entry's row of J^T (applyHvp, contact-assembly.ts:371-392).
__device__ __forceinline__ void contact_term32(const MxArgs& a, uint32_t ent, const float* p, float* o) {
  const Frozen32 c = a.act[ent >> 3];
  const int slot = ent & 7;
  o[0] = o[1] = o[2] = 0.f;
  if (c.coeff == 0.f) return;
  float j;
  float q[3];
  if (c.kind == 2) {
    j = p[(size_t)c.p * 3 + 1];
  } else if (c.kind == 0) {
    ld3f(p, c.p, q);
    j = c.nx * q[0] + c.ny * q[1] + c.nz * q[2];
    const float ws[3] = {c.w0, c.w1, c.w2};
    const int vs[3] = {c.a, c.b, c.c};
    for (int k = 0; k < 3; k++) {
      if (vs[k] < 0) continue;
      ld3f(p, vs[k], q);
      j -= ws[k] * (c.nx * q[0] + c.ny * q[1] + c.nz * q[2]);
    }
  } else {
    const float ws[4] = {c.w0, c.w1, c.w2, c.w3};
    const int vs[4] = {c.a, c.b, c.c, c.d};
    j = 0;
    for (int k = 0; k < 4; k++) {
      if (vs[k] < 0) continue;
      ld3f(p, vs[k], q);
      j += ws[k] * (c.nx * q[0] + c.ny * q[1] + c.nz * q[2]);
    }
  }
  const float g = c.coeff * j;
  if (c.kind == 2) {
    o[1] = g;
  } else {
    float w;
    if (c.kind == 0) w = slot == 0 ? 1.f : -(slot == 1 ? c.w0 : slot == 2 ? c.w1 : c.w2);
    else w = slot == 0 ? c.w0 : slot == 1 ? c.w1 : slot == 2 ? c.w2 : c.w3;
    o[0] = g * w * c.nx; o[1] = g * w * c.ny; o[2] = g * w * c.nz;
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

__global__ void k_mx_contacts(const Frozen* __restrict__ act, const uint8_t* __restrict__ inr,
                              const double* __restrict__ coeff, uint32_t C, Frozen32* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  const Frozen f = act[i];
  Frozen32 o;
  o.kind = f.kind; o.p = f.p; o.a = f.a; o.b = f.b; o.c = f.c; o.d = f.d;
  o.w0 = (float)f.w0; o.w1 = (float)f.w1; o.w2 = (float)f.w2; o.w3 = (float)f.w3;
  o.nx = (float)f.nx; o.ny = (float)f.ny; o.nz = (float)f.nz;
  o.coeff = inr[i] ? (float)coeff[i] : 0.f;
  out[i] = o;
}

__device__ __forceinline__ void k_mx_init_body(const MxArgs& a, uint32_t bid, uint32_t nb) {
  float rz = 0, bb = 0;
  for (uint32_t v = bid * blockDim.x + threadIdx.x; v < a.me.n; v += nb * blockDim.x) {
    const bool pinned = a.me.pinmask[v] != 0;
    float r3[3], z3[3];
    ld3f(a.r, v, r3);
    if (!a.ext) {
      precond32(a, v, r3, z3);
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const size_t j = (size_t)v * 3 + i;
        if (pinned) z3[i] = 0.f;
        a.z[j] = z3[i];
        a.p[j] = z3[i];
        rz += r3[i] * z3[i];
      }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) bb += r3[i] * r3[i];
  }
  if (!a.ext) block_partial_f(rz, mpart(a, MX_RZ), bid);
  block_partial_f(bb, mpart(a, MX_BB), bid);
}

__global__ void __launch_bounds__(kBlock) k_mx_init(MxArgs a) { k_mx_init_body(a, blockIdx.x, gridDim.x); }

// z from the Schwarz apply: pin filter, p = z (init), r.z partials.
__global__ void __launch_bounds__(kBlock) k_mx_ext_z(MxArgs a, int init) {
  if (!init && block_latched(&a.s->done)) return;
  const uint32_t bid = blockIdx.x;
  float rz = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < a.me.n; v += gridDim.x * blockDim.x) {
    const bool pinned = a.me.pinmask[v] != 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const size_t j = (size_t)v * 3 + i;
      const float zi = pinned ? 0.f : a.z[j];
      a.z[j] = zi;
      if (init) a.p[j] = zi;
      rz += a.r[j] * zi;
    }
  }
  block_partial_f(rz, mpart(a, MX_RZ), bid);
}

__device__ __forceinline__ void k_mx_decide_body(const MxArgs& a, unsigned count, uint32_t bid, uint32_t nb) {
  const float rz = merge_partials_f(mpart(a, MX_RZ), count);
  const float bb = merge_partials_f(mpart(a, MX_BB), count);
  if (threadIdx.x == 0) {
    MxScalars* s = a.s;
    s->rz[0] = rz;
    s->residual = sqrtf(bb);
    s->abs_tol = fmaxf(a.tol * fmaxf(s->residual, 1e-12f), 1e-12f);
    s->iters = 0;
    s->done = s->residual < s->abs_tol ? 2 : 0;  // 2 converged, 1 (update) breakdown
  }
}

__global__ void __launch_bounds__(kBlock) k_mx_decide(MxArgs a, unsigned count) {
  k_mx_decide_body(a, count, blockIdx.x, gridDim.x);
}

// Contact HVP terms of all incidence entries (one thread each) and, within
// each wave's 64 entries, their running sums per vertex (entries are sorted
// by vertex): the sum of a vertex's entries is then the last running sum in
// each 64-entry chunk its range touches.
__device__ __forceinline__ void k_mx_cscan_body(const MxArgs& a, uint32_t bid, uint32_t nb) {
  if (block_latched(&a.s->done)) return;
  const uint32_t e = bid * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x % kWave;
  float t[3] = {0.f, 0.f, 0.f};
  uint32_t u = 0xffffffffu;
  if (e < a.E) {
    contact_term32(a, a.v_ent[e], a.p, t);
    u = a.ent_vtx[e];
  }
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t uu = __shfl_up(u, d, kWave);
    float y[3];
    for (int i = 0; i < 3; i++) y[i] = __shfl_up(t[i], d, kWave);
    if (lane >= d && uu == u)
      for (int i = 0; i < 3; i++) t[i] += y[i];
  }
  if (e < a.E)
    for (int i = 0; i < 3; i++) a.scan[(size_t)e * 3 + i] = t[i];
}

__global__ void __launch_bounds__(kBlock) k_mx_cscan(MxArgs a) { k_mx_cscan_body(a, blockIdx.x, gridDim.x); }

constexpr int kMxLanes = 8;  // lanes per vertex in the HVP gather
constexpr int kMxVertsPerBlock = kBlock / kMxLanes;

// Ap = M p / h^2 + H_membrane p + contact, pins zeroed.  The 8 lanes of a
// vertex group evaluate the incident corners (and contact entries) in
// parallel; an xor-butterfly sums them (fixed order, deterministic).
__device__ __forceinline__ void k_mx_hvp_body(const MxArgs& a, int k, uint32_t bid, uint32_t nb) {
  if (block_latched(&a.s->done)) return;
  float pap = 0;
  const int sub = threadIdx.x % kMxLanes;
  const uint32_t stride = nb * kMxVertsPerBlock;
  for (uint32_t v0 = bid * kMxVertsPerBlock; v0 < a.me.n; v0 += stride) {
    const uint32_t v = v0 + threadIdx.x / kMxLanes;
    const bool valid = v < a.me.n;
    float acc[3] = {0.f, 0.f, 0.f};
    if (valid) {
      const uint32_t e = a.me.vt_off[v + 1];
      for (uint32_t q = a.me.vt_off[v] + sub; q < e; q += kMxLanes) {
        float gc[3];
        hvp_corner32(a, a.me.vt_corner[q], a.p, gc);
        acc[0] += gc[0]; acc[1] += gc[1]; acc[2] += gc[2];
      }
      if (a.scan) {
        const uint32_t cb = a.v_off[v], ce = a.v_off[v + 1];
        if (cb < ce)
          for (uint32_t k = cb / kWave + sub; k <= (ce - 1) / kWave; k += kMxLanes) {
            const size_t last = (size_t)min(ce, (k + 1) * (uint32_t)kWave) - 1;
            acc[0] += a.scan[last * 3];
            acc[1] += a.scan[last * 3 + 1];
            acc[2] += a.scan[last * 3 + 2];
          }
      }
    }
#pragma unroll
    for (int o = 1; o < kMxLanes; o <<= 1)
      for (int i = 0; i < 3; i++) acc[i] += __shfl_xor(acc[i], o, kWave);
    if (valid && sub == 0) {
      float p3[3];
      ld3f(a.p, v, p3);
      const float mh = a.me.mass[v] * a.inv_h2;
      const bool pinned = a.me.pinmask[v] != 0;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const float ap = pinned ? 0.f : mh * p3[i] + acc[i];
        a.Ap[(size_t)v * 3 + i] = ap;
        pap += p3[i] * ap;
      }
    }
  }
  block_partial_f(pap, mpart(a, MX_PAP), bid);
}

__global__ void __launch_bounds__(kBlock) k_mx_hvp(MxArgs a, int k) { k_mx_hvp_body(a, k, blockIdx.x, gridDim.x); }

// The same product over vertex-owner tiles: the work-group stages its local
// vertices' x and p in LDS, evaluates every tile triangle once (all three
// corners, membrane-hvp.ts:18-75 from x, no prepared records), then each
// owned vertex sums its corners in ascending triangle order, adds its contact
// sums and the mass term and writes Ap.  Tiles are in Morton order; each XCD
// walks a contiguous eighth of them.
__device__ __forceinline__ void k_mx_hvp_tile_body(const MxArgs& a, uint32_t bid, uint32_t nb) {
  if (__hip_atomic_load(&a.s->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
  __shared__ float xs[kTileLoc * 3], ps[kTileLoc * 3], out[kTileTris * 9];
  const DTiles& T = a.me.tl;
  const uint32_t G = nb;
  const uint32_t lb = (bid % kXcds) * (G / kXcds) + bid / kXcds;
  float pap = 0;
  for (uint32_t tile = lb; tile < T.count; tile += G) {
    const TileHdr h = T.hdr[tile];
    __syncthreads();  // the previous tile's LDS reads are done
    for (uint32_t l = threadIdx.x; l < h.n_loc; l += kBlock) {
      const uint32_t v = T.loc[h.loc + l];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        xs[l * 3 + k] = a.x[(size_t)v * 3 + k];
        ps[l * 3 + k] = a.p[(size_t)v * 3 + k];
      }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < h.n_tri; t += kBlock) {
      const uint2 li = T.lidx[h.tri + t];
      const float4 inv = T.inv[h.tri + t];
      const float sc = a.thick * T.area[h.tri + t];
      const uint32_t i0 = li.x & 0xffffu, i1 = li.x >> 16, i2 = li.y;
      const float ia = inv.x, ib = inv.y, ic = inv.z, id = inv.w;
      float F0[3], F1[3], dF0[3], dF1[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const float e1 = xs[i1 * 3 + k] - xs[i0 * 3 + k], e2 = xs[i2 * 3 + k] - xs[i0 * 3 + k];
        const float w1 = ps[i1 * 3 + k] - ps[i0 * 3 + k], w2 = ps[i2 * 3 + k] - ps[i0 * 3 + k];
        F0[k] = ia * e1 + ic * e2;
        F1[k] = ib * e1 + id * e2;
        dF0[k] = ia * w1 + ic * w2;
        dF1[k] = ib * w1 + id * w2;
      }
      const float E00 = 0.5f * (F0[0] * F0[0] + F0[1] * F0[1] + F0[2] * F0[2] - 1.f);
      const float E11 = 0.5f * (F1[0] * F1[0] + F1[1] * F1[1] + F1[2] * F1[2] - 1.f);
      const float E01 = 0.5f * (F0[0] * F1[0] + F0[1] * F1[1] + F0[2] * F1[2]);
      float S00 = a.c00 * E00 + a.c01 * E11, S11 = a.c11 * E11 + a.c01 * E00, S01 = 2.f * a.g * E01;
      if (a.psd) stress_psd(S00, S11, S01);
      const float dE00 = F0[0] * dF0[0] + F0[1] * dF0[1] + F0[2] * dF0[2];
      const float dE11 = F1[0] * dF1[0] + F1[1] * dF1[1] + F1[2] * dF1[2];
      const float dE01 = 0.5f * (dF0[0] * F1[0] + F0[0] * dF1[0] + dF0[1] * F1[1] + F0[1] * dF1[1] +
                                 dF0[2] * F1[2] + F0[2] * dF1[2]);
      const float dS00 = a.c00 * dE00 + a.c01 * dE11, dS11 = a.c11 * dE11 + a.c01 * dE00, dS01 = 2.f * a.g * dE01;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const float dP0 = dF0[k] * S00 + F0[k] * dS00 + dF1[k] * S01 + F1[k] * dS01;
        const float dP1 = dF0[k] * S01 + F0[k] * dS01 + dF1[k] * S11 + F1[k] * dS11;
        const float g1 = sc * (dP0 * ia + dP1 * ib);
        const float g2 = sc * (dP0 * ic + dP1 * id);
        out[t * 9 + k] = -(g1 + g2);
        out[t * 9 + 3 + k] = g1;
        out[t * 9 + 6 + k] = g2;
      }
    }
    __syncthreads();
    for (uint32_t o = threadIdx.x; o < h.n_own; o += kBlock) {
      const uint32_t v = T.loc[h.loc + o];
      float acc[3] = {0.f, 0.f, 0.f};
      const uint32_t e1 = T.off[h.off + o + 1];
      for (uint32_t e = T.off[h.off + o]; e < e1; e++) {
        const uint32_t q = T.ent[h.ent + e];
        const float* g = out + (q >> 2) * 9 + (q & 3) * 3;
        acc[0] += g[0];
        acc[1] += g[1];
        acc[2] += g[2];
      }
      float cs[3] = {0.f, 0.f, 0.f};
      if (a.scan) {
        const uint32_t cb = a.v_off[v], ce = a.v_off[v + 1];
        if (cb < ce)
          for (uint32_t k = cb / kWave; k <= (ce - 1) / kWave; k++) {
            const size_t e = (size_t)min(ce, (k + 1) * (uint32_t)kWave) - 1;
            cs[0] += a.scan[e * 3];
            cs[1] += a.scan[e * 3 + 1];
            cs[2] += a.scan[e * 3 + 2];
          }
      }
      const float mh = a.me.mass[v] * a.inv_h2;
      const bool pinned = a.me.pinmask[v] != 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const float pv = ps[o * 3 + k];
        const float ap = pinned ? 0.f : mh * pv + (acc[k] + cs[k]);
        a.Ap[(size_t)v * 3 + k] = ap;
        pap += pv * ap;
      }
    }
  }
  block_partial_f(pap, mpart(a, MX_PAP), bid);
}

__global__ void __launch_bounds__(kBlock) k_mx_hvp_tile(MxArgs a) {
  k_mx_hvp_tile_body(a, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void k_mx_update_body(const MxArgs& a, int k, unsigned pap_count, uint32_t bid, uint32_t nb) {
  MxScalars* s = a.s;
  if (block_latched(&s->done)) return;
  const float pAp = merge_partials_f(mpart(a, MX_PAP), pap_count);
  if (!(pAp > 1e-30f)) {
    if (bid == 0 && threadIdx.x == 0) {
      s->pAp = pAp;
      s->done = 1;
      s->iters = k;
    }
    return;
  }
  const float alpha = s->rz[k & 1] / pAp;
  float rr = 0, rz = 0;
  for (uint32_t v = bid * blockDim.x + threadIdx.x; v < a.me.n; v += nb * blockDim.x) {
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
    if (!a.ext) {
      precond32(a, v, r3, z3);
#pragma unroll
      for (int i = 0; i < 3; i++) {
        if (pinned) z3[i] = 0.f;
        a.z[(size_t)v * 3 + i] = z3[i];
        rz += r3[i] * z3[i];
      }
    }
  }
  block_partial_f(rr, mpart(a, MX_RR), bid);
  if (!a.ext) block_partial_f(rz, mpart(a, MX_RZ), bid);
}

__global__ void __launch_bounds__(kBlock) k_mx_update(MxArgs a, int k, unsigned pap_count) {
  k_mx_update_body(a, k, pap_count, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void k_mx_dir_body(const MxArgs& a, int k, uint32_t bid, uint32_t nb) {
  MxScalars* s = a.s;
  if (block_latched(&s->done)) return;
  const float residual = sqrtf(merge_partials_f(mpart(a, MX_RR), nb));
  if (residual < s->abs_tol) {
    if (bid == 0 && threadIdx.x == 0) {
      s->residual = residual;
      s->done = 2;
      s->iters = k + 1;
    }
    return;
  }
  const float rznew = merge_partials_f(mpart(a, MX_RZ), nb);
  const float beta = rznew / s->rz[k & 1];
  for (uint32_t v = bid * blockDim.x + threadIdx.x; v < a.me.n; v += nb * blockDim.x) {
    const bool pinned = a.me.pinmask[v] != 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const size_t j = (size_t)v * 3 + i;
      a.p[j] = pinned ? 0.f : a.z[j] + beta * a.p[j];
    }
  }
  if (bid == 0 && threadIdx.x == 0) {
    s->rz[(k + 1) & 1] = rznew;
    s->residual = residual;
  }
}

__global__ void __launch_bounds__(kBlock) k_mx_dir(MxArgs a, int k) { k_mx_dir_body(a, k, blockIdx.x, gridDim.x); }

__global__ void k_mx_out(const float* sol, double* dx, size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n3) dx[i] = (double)sol[i];
}

struct MixedImpl {
  hipStream_t st;
  uint32_t n;
  hipGraphExec_t exec = nullptr;  // the last captured chunk of kGraphIters iterations
  DevBuf<float> x32, d32, binv32, sol, r, z, p, Ap, partials, scan;
  DevBuf<Frozen32> act32;
  DevBuf<MxScalars> s;
  MxArgs last{};
  bool has_last = false;
};

static unsigned mx_grid(uint32_t n) { return red_grid(n); }
constexpr int kGraphIters = 8;  // PCG iterations per captured chunk (even)
static unsigned mx_hvp_grid(const DMesh32& me) {
  if (me.tl.count) {
    const size_t g = std::min<size_t>(kMaxHvpBlocks, me.tl.count);
    return (unsigned)((g + kXcds - 1) / kXcds * kXcds);  // a multiple of the XCD count
  }
  return (unsigned)std::min<size_t>(kMaxHvpBlocks, ((size_t)me.n + kMxVertsPerBlock - 1) / kMxVertsPerBlock);
}

static void launch_mx_hvp(const MxArgs& a, hipStream_t st) {
  const dim3 gh(mx_hvp_grid(a.me)), b(kBlock);
  if (a.me.tl.count) hipLaunchKernelGGL(k_mx_hvp_tile, gh, b, 0, st, a);
  else hipLaunchKernelGGL(k_mx_hvp, gh, b, 0, st, a, 0);
}

MixedSolve::MixedSolve(const cs_scene_desc& d, const cs_options&, hipStream_t st) : im_(new MixedImpl) {
  im_->st = st;
  im_->n = d.n_verts;
  const size_t n3 = (size_t)d.n_verts * 3;
  for (auto* b : {&im_->x32, &im_->d32, &im_->sol, &im_->r, &im_->z, &im_->p, &im_->Ap}) b->alloc(n3);
  im_->partials.alloc((size_t)MX_NPART * kMaxRedBlocks);
  im_->s.alloc(1);
  CS_HIP(hipMemsetAsync(im_->s.ptr, 0, sizeof(MxScalars), st));
}

const int32_t* MixedSolve::done_flag() const { return &im_->s.ptr->done; }

MixedSolve::~MixedSolve() {
  if (im_->exec) (void)hipGraphExecDestroy(im_->exec);
  delete im_;
}

// The per-solve set-up of solve_plain: FP32 copies of x, r = b, the diagonal
// and the block inverses, the FP32 contact records, and the kernel arguments.
static MxArgs prepare_args(MixedImpl& M, const MixedInputs& in, hipStream_t st) {
  const uint32_t n = in.me.n;
  if (in.binv) M.binv32.alloc((size_t)n * 9);
  const uint32_t C = in.contact.act ? in.contact.C : 0;
  const uint32_t E = C ? in.contact.E : 0;
  if (C) {
    M.act32.alloc(C);
    M.scan.alloc((size_t)std::max<uint32_t>(E, 1) * 3);
    hipLaunchKernelGGL(k_mx_contacts, dim3(grid_for(C)), dim3(kBlock), 0, st, in.contact.act, in.contact.inr,
                       in.contact.coeff, C, M.act32.ptr);
  }
  hipLaunchKernelGGL(k_mx_convert, dim3(grid_for(n)), dim3(kBlock), 0, st, n, in.x, in.b, in.diag, in.binv,
                     in.me.pinmask, M.x32.ptr, M.r.ptr, M.d32.ptr, in.binv ? M.binv32.ptr : nullptr, M.sol.ptr);
  CS_LAUNCH_CHECK();
  MxArgs a;
  a.me = DMesh32{in.me.n,         in.me.m,       in.me.idx,       in.me.inv_dm,   in.me.area,
                 in.me.mass,      in.me.vt_off,  in.me.vt_rec,    in.me.vt_corner, in.me.pinmask,
                 in.me.tl};
  a.c00 = (float)in.mat.c00; a.c11 = (float)in.mat.c11; a.c01 = (float)in.mat.c01; a.g = (float)in.mat.g;
  a.thick = (float)in.mat.thick;
  a.inv_h2 = (float)in.inv_h2;
  a.x = M.x32.ptr; a.diag = M.d32.ptr; a.binv = in.binv ? M.binv32.ptr : nullptr;
  a.psd = in.psd ? 1 : 0;
  a.sol = M.sol.ptr; a.r = M.r.ptr; a.z = M.z.ptr; a.p = M.p.ptr; a.Ap = M.Ap.ptr;
  a.tol = (float)in.tol;
  a.ext = in.sw ? 1 : 0;
  a.s = M.s.ptr;
  a.partials = M.partials.ptr;
  a.C = C;
  a.E = E;
  a.ent_vtx = in.contact.ent_vtx;
  a.scan = E ? M.scan.ptr : nullptr;
  a.act = C ? M.act32.ptr : nullptr;
  a.v_off = in.contact.v_off;
  a.v_ent = in.contact.v_ent;
  M.last = a;
  M.has_last = true;
  return a;
}

int MixedSolve::solve_plain(const MixedInputs& in, void (*timer)(void*, int, bool), void* ctx) {
  MixedImpl& M = *im_;
  hipStream_t st = M.st;
  const uint32_t n = in.me.n;
  const MxArgs a = prepare_args(M, in, st);
  const uint32_t E = a.E;
  const dim3 g(mx_grid(n)), b(kBlock);
  hipLaunchKernelGGL(k_mx_init, g, b, 0, st, a);
  if (in.sw) {
    in.sw->apply32(a.r, a.z, nullptr);
    hipLaunchKernelGGL(k_mx_ext_z, g, b, 0, st, a, 1);
  }
  hipLaunchKernelGGL(k_mx_decide, dim3(1), b, 0, st, a, mx_grid(n));
  CS_LAUNCH_CHECK();
  // one PCG iteration; k only selects the r.z slot (k & 1) and is the
  // iteration count the finishing kernel records (k + 1)
  auto iteration = [&](int k, bool timed) {
    if (E) hipLaunchKernelGGL(k_mx_cscan, dim3(grid_for(E)), b, 0, st, a);
    if (timed) timer(ctx, k, true);
    launch_mx_hvp(a, st);
    if (timed) timer(ctx, k, false);
    hipLaunchKernelGGL(k_mx_update, g, b, 0, st, a, k, mx_hvp_grid(a.me));
    if (in.sw) {
      in.sw->apply32(a.r, a.z, &M.s.ptr->done);
      hipLaunchKernelGGL(k_mx_ext_z, g, b, 0, st, a, 0);
    }
    hipLaunchKernelGGL(k_mx_dir, g, b, 0, st, a, k);
    CS_LAUNCH_CHECK();
  };
  auto finished = [&](MxScalars& hs) {
    CS_HIP(hipMemcpyAsync(&hs, M.s.ptr, sizeof(hs), hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
    return hs.done != 0;
  };
  const int chunk = kGraphIters;  // the host checks convergence every chunk
  int k0 = 0, iters = -1;
  if (in.graph && in.max_iters >= chunk) {
    // Capture kGraphIters iterations (k = 0 .. kGraphIters-1; kGraphIters is
    // even, so the r.z slots line up in every replay) and replay the chunk;
    // the executable graph of the previous solve is updated in place when the
    // launch sequence has the same shape.  Nothing in a chunk is a memset.
    hipGraph_t graph = nullptr;
    CS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < chunk; k++) iteration(k, false);
    CS_HIP(hipStreamEndCapture(st, &graph));
    bool updated = false;
    if (M.exec) {
      hipGraphNode_t err_node = nullptr;
      hipGraphExecUpdateResult res;
      updated = hipGraphExecUpdate(M.exec, graph, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
      if (!updated) {
        (void)hipGetLastError();
        CS_HIP(hipGraphExecDestroy(M.exec));
        M.exec = nullptr;
      }
    }
    if (!M.exec) CS_HIP(hipGraphInstantiate(&M.exec, graph, nullptr, nullptr, 0));
    CS_HIP(hipGraphDestroy(graph));
    for (; k0 + chunk <= in.max_iters; k0 += chunk) {
      CS_HIP(hipGraphLaunch(M.exec, st));
      MxScalars hs;
      if (finished(hs)) {
        iters = k0 + hs.iters;  // the chunk recorded its local count
        break;
      }
    }
  }
  if (iters < 0) {
    for (int k = k0; k < in.max_iters; k++) {
      iteration(k, timer != nullptr);
      MxScalars hs;
      if (((k + 1) % chunk == 0 || k + 1 == in.max_iters) && finished(hs)) {
        iters = hs.iters;
        break;
      }
    }
  }
  hipLaunchKernelGGL(k_mx_out, dim3(grid_for((size_t)n * 3)), dim3(kBlock), 0, st, M.sol.ptr, in.dx, (size_t)n * 3);
  CS_LAUNCH_CHECK();
  if (iters >= 0) return iters;
  MxScalars hs;
  return finished(hs) ? hs.iters : in.max_iters;
}

__global__ void k_mx_probe_in(const double* x, const double* p, float* x32, float* p32, size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n3) {
    x32[i] = (float)x[i];
    p32[i] = (float)p[i];
  }
}

// One HVP launch (the kernel the solve uses, no contact) at (x, p), for
// component parity tests; overwrites the solve's x, p and Ap buffers.
void MixedSolve::probe_hvp(const DMesh& me, const DMat& mat, double inv_h2, const double* x, const double* p,
                           double* out, const DContactHvp* contact) {
  MixedImpl& M = *im_;
  hipStream_t st = M.st;
  const size_t n3 = (size_t)me.n * 3;
  hipLaunchKernelGGL(k_mx_probe_in, dim3(grid_for(n3)), dim3(kBlock), 0, st, x, p, M.x32.ptr, M.p.ptr, n3);
  MxArgs a{};
  a.me = DMesh32{me.n,         me.m,        me.idx,       me.inv_dm,   me.area,
                 me.mass,      me.vt_off,   me.vt_rec,    me.vt_corner, me.pinmask,
                 me.tl};
  a.c00 = (float)mat.c00; a.c11 = (float)mat.c11; a.c01 = (float)mat.c01; a.g = (float)mat.g;
  a.thick = (float)mat.thick;
  a.inv_h2 = (float)inv_h2;
  a.x = M.x32.ptr; a.p = M.p.ptr; a.Ap = M.Ap.ptr;
  a.s = M.s.ptr;
  a.partials = M.partials.ptr;
  CS_HIP(hipMemsetAsync(&M.s.ptr->done, 0, sizeof(int32_t), st));
  const uint32_t C = contact && contact->act ? contact->C : 0;
  const uint32_t E = C ? contact->E : 0;
  if (C) {
    M.act32.alloc(C);
    M.scan.alloc((size_t)std::max<uint32_t>(E, 1) * 3);
    hipLaunchKernelGGL(k_mx_contacts, dim3(grid_for(C)), dim3(kBlock), 0, st, contact->act, contact->inr,
                       contact->coeff, C, M.act32.ptr);
    a.C = C;
    a.E = E;
    a.act = M.act32.ptr;
    a.ent_vtx = contact->ent_vtx;
    a.v_off = contact->v_off;
    a.v_ent = contact->v_ent;
    a.scan = E ? M.scan.ptr : nullptr;
    if (E) hipLaunchKernelGGL(k_mx_cscan, dim3(grid_for(E)), dim3(kBlock), 0, st, a);
  }
  launch_mx_hvp(a, st);
  hipLaunchKernelGGL(k_mx_out, dim3(grid_for(n3)), dim3(kBlock), 0, st, M.Ap.ptr, out, n3);
  CS_LAUNCH_CHECK();
  CS_HIP(hipStreamSynchronize(st));
}

// Back-to-back HVP launches on the operands of the last solve (inputs stay
// resident; Ap and the p.Ap partials are overwritten).
double MixedSolve::time_hvp(int reps) {
  MixedImpl& M = *im_;
  if (!M.has_last || reps < 1) return 0.0;
  hipStream_t st = M.st;
  CS_HIP(hipMemsetAsync(&M.s.ptr->done, 0, sizeof(int32_t), st));
  hipEvent_t e0, e1;
  CS_HIP(hipEventCreate(&e0));
  CS_HIP(hipEventCreate(&e1));
  launch_mx_hvp(M.last, st);
  CS_HIP(hipEventRecord(e0, st));
  for (int i = 0; i < reps; i++) launch_mx_hvp(M.last, st);
  CS_HIP(hipEventRecord(e1, st));
  CS_LAUNCH_CHECK();
  CS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  CS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const int32_t one = 1;
  CS_HIP(hipMemcpyAsync(&M.s.ptr->done, &one, sizeof(one), hipMemcpyHostToDevice, st));
  CS_HIP(hipStreamSynchronize(st));
  return (double)ms / reps;
}

// ---- packed multi-instance solve --------------------------------------------
// Block table entry: instance, local block, local grid, and a count the stage
// needs (the partial count a merge reads).
struct PkEntry {
  uint32_t inst, bid, nb, aux;
};

__device__ __forceinline__ bool pk_capped(const int* caps, uint32_t inst, int k) { return k >= caps[inst]; }

__global__ void __launch_bounds__(kBlock) k_pk_init(const MxArgs* A, const PkEntry* tab) {
  const PkEntry e = tab[blockIdx.x];
  k_mx_init_body(A[e.inst], e.bid, e.nb);
}

__global__ void __launch_bounds__(kBlock) k_pk_decide(const MxArgs* A, const PkEntry* tab) {
  const PkEntry e = tab[blockIdx.x];
  k_mx_decide_body(A[e.inst], e.aux, e.bid, e.nb);
}

__global__ void __launch_bounds__(kBlock) k_pk_cscan(const MxArgs* A, const PkEntry* tab, const int* caps, int k) {
  const PkEntry e = tab[blockIdx.x];
  if (pk_capped(caps, e.inst, k)) return;
  k_mx_cscan_body(A[e.inst], e.bid, e.nb);
}

__global__ void __launch_bounds__(kBlock) k_pk_hvp(const MxArgs* A, const PkEntry* tab, const int* caps, int k) {
  const PkEntry e = tab[blockIdx.x];
  if (pk_capped(caps, e.inst, k)) return;
  const MxArgs& a = A[e.inst];
  if (a.me.tl.count) k_mx_hvp_tile_body(a, e.bid, e.nb);
  else k_mx_hvp_body(a, 0, e.bid, e.nb);
}

__global__ void __launch_bounds__(kBlock) k_pk_update(const MxArgs* A, const PkEntry* tab, const int* caps, int k) {
  const PkEntry e = tab[blockIdx.x];
  if (pk_capped(caps, e.inst, k)) return;
  k_mx_update_body(A[e.inst], k, e.aux, e.bid, e.nb);
}

__global__ void __launch_bounds__(kBlock) k_pk_dir(const MxArgs* A, const PkEntry* tab, const int* caps, int k) {
  const PkEntry e = tab[blockIdx.x];
  if (pk_capped(caps, e.inst, k)) return;
  k_mx_dir_body(A[e.inst], k, e.bid, e.nb);
}

__global__ void k_pk_flags(const MxArgs* A, uint32_t count, int2* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = make_int2(A[i].s->done, A[i].s->iters);
}

PackedScratch::~PackedScratch() {
  if (args) (void)hipFree(args);
  if (tabs) (void)hipFree(tabs);
  if (caps) (void)hipFree(caps);
  if (flags) (void)hipFree(flags);
}

void MixedSolve::solve_packed(std::vector<PackedJob>& jobs, hipStream_t st, PackedScratch& ps) {
  const size_t ni = jobs.size();
  if (!ni) return;
  std::vector<MxArgs> A(ni);
  std::vector<int> caps(ni);
  int kmax = 0;
  for (size_t i = 0; i < ni; i++) {
    const MixedInputs& in = jobs[i].in;
    if (in.sw) throw std::invalid_argument("a packed solve takes Jacobi or block-Jacobi systems");
    A[i] = prepare_args(*jobs[i].solver->im_, in, st);
    caps[i] = in.max_iters;
    kmax = std::max(kmax, in.max_iters);
  }
  // block tables: vector stages, one block per instance, contact scans, HVP
  std::vector<PkEntry> vec, one, scan, hvp, upd;
  for (uint32_t i = 0; i < (uint32_t)ni; i++) {
    const uint32_t gv = mx_grid(A[i].me.n), gh = mx_hvp_grid(A[i].me), gs = A[i].E ? grid_for(A[i].E) : 0;
    for (uint32_t b = 0; b < gv; b++) {
      vec.push_back(PkEntry{i, b, gv, 0});
      upd.push_back(PkEntry{i, b, gv, gh});
    }
    one.push_back(PkEntry{i, 0, 1, gv});
    for (uint32_t b = 0; b < gs; b++) scan.push_back(PkEntry{i, b, gs, 0});
    for (uint32_t b = 0; b < gh; b++) hvp.push_back(PkEntry{i, b, gh, 0});
  }
  const size_t off_vec = 0, off_one = off_vec + vec.size(), off_scan = off_one + one.size(),
               off_hvp = off_scan + scan.size(), off_upd = off_hvp + hvp.size(), n_tab = off_upd + upd.size();
  std::vector<PkEntry> tab;
  tab.reserve(n_tab);
  for (auto* t : {&vec, &one, &scan, &hvp, &upd}) tab.insert(tab.end(), t->begin(), t->end());
  const size_t abytes = ni * sizeof(MxArgs), tbytes = n_tab * sizeof(PkEntry);
  if (abytes > ps.args_bytes) {
    if (ps.args) CS_HIP(hipFree(ps.args));
    CS_HIP(hipMalloc(&ps.args, abytes));
    ps.args_bytes = abytes;
  }
  if (tbytes > ps.tabs_bytes) {
    if (ps.tabs) CS_HIP(hipFree(ps.tabs));
    CS_HIP(hipMalloc(&ps.tabs, tbytes));
    ps.tabs_bytes = tbytes;
  }
  if (ni > ps.n_inst) {
    if (ps.caps) CS_HIP(hipFree(ps.caps));
    if (ps.flags) CS_HIP(hipFree(ps.flags));
    CS_HIP(hipMalloc(&ps.caps, ni * sizeof(int)));
    CS_HIP(hipMalloc(&ps.flags, ni * sizeof(int2)));
    ps.n_inst = ni;
  }
  CS_HIP(hipMemcpyAsync(ps.args, A.data(), abytes, hipMemcpyHostToDevice, st));
  CS_HIP(hipMemcpyAsync(ps.tabs, tab.data(), tbytes, hipMemcpyHostToDevice, st));
  CS_HIP(hipMemcpyAsync(ps.caps, caps.data(), ni * sizeof(int), hipMemcpyHostToDevice, st));
  const MxArgs* dA = static_cast<const MxArgs*>(ps.args);
  const PkEntry* T = static_cast<const PkEntry*>(ps.tabs);
  const dim3 b(kBlock);
  hipLaunchKernelGGL(k_pk_init, dim3((unsigned)vec.size()), b, 0, st, dA, T + off_vec);
  hipLaunchKernelGGL(k_pk_decide, dim3((unsigned)one.size()), b, 0, st, dA, T + off_one);
  CS_LAUNCH_CHECK();
  std::vector<int2> hf(ni);
  for (int k = 0; k < kmax; k++) {
    if (!scan.empty())
      hipLaunchKernelGGL(k_pk_cscan, dim3((unsigned)scan.size()), b, 0, st, dA, T + off_scan, ps.caps, k);
    hipLaunchKernelGGL(k_pk_hvp, dim3((unsigned)hvp.size()), b, 0, st, dA, T + off_hvp, ps.caps, k);
    hipLaunchKernelGGL(k_pk_update, dim3((unsigned)upd.size()), b, 0, st, dA, T + off_upd, ps.caps, k);
    hipLaunchKernelGGL(k_pk_dir, dim3((unsigned)vec.size()), b, 0, st, dA, T + off_vec, ps.caps, k);
    CS_LAUNCH_CHECK();
    if ((k + 1) % kGraphIters == 0 || k + 1 == kmax) {
      hipLaunchKernelGGL(k_pk_flags, dim3(grid_for(ni)), b, 0, st, dA, (uint32_t)ni, ps.flags);
      CS_HIP(hipMemcpyAsync(hf.data(), ps.flags, ni * sizeof(int2), hipMemcpyDeviceToHost, st));
      CS_HIP(hipStreamSynchronize(st));
      bool all = true;
      for (size_t i = 0; i < ni; i++) all = all && (hf[i].x != 0 || k + 1 >= caps[i]);
      if (all) break;
    }
  }
  for (size_t i = 0; i < ni; i++) {
    const uint32_t n3 = jobs[i].in.me.n * 3;
    hipLaunchKernelGGL(k_mx_out, dim3(grid_for(n3)), b, 0, st, A[i].sol, jobs[i].in.dx, (size_t)n3);
  }
  CS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pk_flags, dim3(grid_for(ni)), b, 0, st, dA, (uint32_t)ni, ps.flags);
  CS_HIP(hipMemcpyAsync(hf.data(), ps.flags, ni * sizeof(int2), hipMemcpyDeviceToHost, st));
  CS_HIP(hipStreamSynchronize(st));
  for (size_t i = 0; i < ni; i++) jobs[i].iters = hf[i].x ? hf[i].y : caps[i];
}

}  // namespace cloth
