// contact.hip — the reference ContactSystem (contact-assembly.ts:46-595) on the
// device: LBVH broadphase, frozen active multiset, barrier + lagged friction,
// contact HVP / Jacobi terms, trial validation (distances + CCD) and the
// restitution-0 velocity filter.
//
// Ordering: the reference sorts candidates by (kind, p, a, b, c, d)
// (contact-assembly.ts:118) and accumulates every per-vertex contact term in
// that order.  Here each prim gets a 64-bit key that sorts identically:
//   VT  (0, p, rank of the triangle's signed (a, b, c))
//   EE  (1, rank of directed edge (a, b), rank of directed edge (c, d))
//   floor (2, v)  (floor contacts follow the candidates, contact-assembly.ts:202-213)
// Ties are identical records.  The lagged friction map (keyed by vtKey/eeKey,
// :607-616) is a sorted table whose lookups take the last occurrence, which
// is the reference's Map.set last-write-wins.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "collide.h"
#include "contact.h"
#include "sort.h"

namespace cloth {

constexpr uint32_t LEAF = 0x80000000u;
constexpr uint32_t NONE = 0xffffffffu;
constexpr uint64_t SENT = ~0ull;

struct DGeom {
  uint32_t n, m, ns, ms;
  const uint32_t* idx;
  const uint32_t* sidx;
  const double* spos;
  const uint32_t* tri_rank;    // m + ms
  const uint32_t* dedge_rank;  // 3 (m + ms)
  const uint32_t* uedge;       // 3 (m + ms)
  int B;
  int has_floor;
  double floor_y, d_hat, kappa, d_min, mu, feps;
};

struct CScalars {
  double red[16];
  unsigned tickets[16];
  uint32_t pair_count, spair_count, active_count, entry_count, nonfinite, floor_fail;
  uint32_t kind_count[4];
  uint64_t nan_key, cross_key;
};

__device__ __forceinline__ void vpos(const DGeom& G, const double* x, int v, double* o) {
  if (v >= 0) {
    o[0] = x[(size_t)v * 3]; o[1] = x[(size_t)v * 3 + 1]; o[2] = x[(size_t)v * 3 + 2];
  } else {
    const size_t i = (size_t)(-v - 2);
    o[0] = G.spos[i * 3]; o[1] = G.spos[i * 3 + 1]; o[2] = G.spos[i * 3 + 2];
  }
}

__device__ __forceinline__ bool box_overlap(const double* a, const double* b) {
  return a[0] <= b[3] && a[3] >= b[0] && a[1] <= b[4] && a[4] >= b[1] && a[2] <= b[5] && a[5] >= b[2];
}

// ---------------------------------------------------------------- LBVH

// sweptTriAabb (aabb.ts:9-31) + box centre bounds for the Morton frame.
__global__ void __launch_bounds__(kBlock) k_tri_boxes(DGeom G, const double* __restrict__ x0,
                                                      const double* __restrict__ x1, double pad, double* boxes,
                                                      double* partials, unsigned* tickets, double* bounds) {
  double lo[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
  double hi[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < G.m; t += gridDim.x * blockDim.x) {
    double mn[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double mx[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    for (int c = 0; c < 3; c++) {
      const size_t v = G.idx[(size_t)t * 3 + c];
      const double* xs[2] = {x0, x1};
      for (int q = 0; q < 2; q++) {
        for (int k = 0; k < 3; k++) {
          const double pk = xs[q][v * 3 + k];
          if (pk < mn[k]) mn[k] = pk;
          if (pk > mx[k]) mx[k] = pk;
        }
      }
    }
    double* b = boxes + (size_t)t * 6;
    for (int k = 0; k < 3; k++) {
      b[k] = mn[k] - pad;
      b[3 + k] = mx[k] + pad;
      const double c = 0.5 * (b[k] + b[3 + k]);
      if (c < lo[k]) lo[k] = c;
      if (c > hi[k]) hi[k] = c;
    }
  }
  double r;
  for (int k = 0; k < 3; k++) {
    if (grid_reduce_ext<kBlock, LtMin>(lo[k], partials + k * 1024, &tickets[k], r)) bounds[k] = r;
    if (grid_reduce_ext<kBlock, GtMax>(hi[k], partials + (3 + k) * 1024, &tickets[3 + k], r)) bounds[3 + k] = r;
  }
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ void k_morton(uint32_t m, const double* __restrict__ boxes, const double* __restrict__ bounds,
                         uint64_t* keys, uint32_t* vals) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const double* b = boxes + (size_t)t * 6;
  uint32_t q[3];
  for (int k = 0; k < 3; k++) {
    const double ext = bounds[3 + k] - bounds[k];
    const double c = 0.5 * (b[k] + b[3 + k]);
    double u = ext > 0 ? (c - bounds[k]) / ext : 0.0;
    if (!(u >= 0)) u = 0;
    if (u > 1) u = 1;
    q[k] = (uint32_t)fmin(1023.0, u * 1024.0);
  }
  keys[t] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
  vals[t] = t;
}

__device__ __forceinline__ int lcp(const uint64_t* codes, int m, int i, int j) {
  if (j < 0 || j >= m) return -1;
  const uint32_t ci = (uint32_t)codes[i], cj = (uint32_t)codes[j];
  if (ci == cj) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(ci ^ cj);
}

// Karras 2012 hierarchy over sorted Morton codes: internal node i covers the
// sorted leaf range [first, last]; children carry LEAF for leaves.
__global__ void k_karras(int m, const uint64_t* __restrict__ codes, uint32_t* left, uint32_t* right,
                         uint32_t* first, uint32_t* last, uint32_t* parent_int, uint32_t* parent_leaf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m - 1) return;
  const int d = (lcp(codes, m, i, i + 1) - lcp(codes, m, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = lcp(codes, m, i, i - d);
  int lmax = 2;
  while (lcp(codes, m, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (lcp(codes, m, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = lcp(codes, m, i, j);
  int s = 0, div = 2, t;
  do {
    t = (l + div - 1) / div;
    if (lcp(codes, m, i, i + (s + t) * d) > dnode) s += t;
    div *= 2;
  } while (t > 1);
  const int gamma = i + s * d + (d < 0 ? d : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const uint32_t L = (lo == gamma) ? (LEAF | (uint32_t)gamma) : (uint32_t)gamma;
  const uint32_t R = (hi == gamma + 1) ? (LEAF | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
  left[i] = L;
  right[i] = R;
  first[i] = lo;
  last[i] = hi;
  if (L & LEAF) parent_leaf[L & ~LEAF] = i; else parent_int[L] = i;
  if (R & LEAF) parent_leaf[R & ~LEAF] = i; else parent_int[R] = i;
}

__device__ __forceinline__ const double* child_box(uint32_t c, const uint32_t* sorted_tri, const double* boxes,
                                                   const double* nodebox) {
  return (c & LEAF) ? boxes + (size_t)sorted_tri[c & ~LEAF] * 6 : nodebox + (size_t)c * 6;
}

// Bottom-up refit: the second child to arrive at a node writes its union.
__global__ void k_refit(int m, const uint32_t* __restrict__ sorted_tri, const double* __restrict__ boxes,
                        const uint32_t* __restrict__ left, const uint32_t* __restrict__ right,
                        const uint32_t* __restrict__ parent_int, const uint32_t* __restrict__ parent_leaf,
                        double* nodebox, unsigned* flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  uint32_t node = parent_leaf[i];
  while (node != NONE) {
    __threadfence();
    if (atomicAdd(&flags[node], 1u) == 0) return;
    __threadfence();
    const double* a = child_box(left[node], sorted_tri, boxes, nodebox);
    const double* b = child_box(right[node], sorted_tri, boxes, nodebox);
    double* o = nodebox + (size_t)node * 6;
    for (int k = 0; k < 3; k++) {
      o[k] = jsm::jmin(a[k], b[k]);
      o[3 + k] = jsm::jmax(a[3 + k], b[3 + k]);
    }
    node = parent_int[node];
  }
}

constexpr int kStack = 64;

// selfPairs (bvh.ts:79-124): every unordered pair of overlapping triangles
// sharing no vertex, found once from the leaf with the smaller sorted index.
__global__ void __launch_bounds__(kBlock) k_traverse_self(DGeom G, const uint32_t* __restrict__ sorted_tri,
                                                          const double* __restrict__ boxes,
                                                          const uint32_t* __restrict__ left,
                                                          const uint32_t* __restrict__ right,
                                                          const uint32_t* __restrict__ last,
                                                          const double* __restrict__ nodebox, uint2* pairs,
                                                          uint32_t cap, CScalars* cs) {
  const int m = (int)G.m;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || m < 2) return;
  const uint32_t ti = sorted_tri[i];
  double Q[6];
  for (int k = 0; k < 6; k++) Q[k] = boxes[(size_t)ti * 6 + k];
  const uint32_t A0 = G.idx[(size_t)ti * 3], A1 = G.idx[(size_t)ti * 3 + 1], A2 = G.idx[(size_t)ti * 3 + 2];
  uint32_t stack[kStack];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const uint32_t node = stack[--sp];
    const uint32_t ch[2] = {left[node], right[node]};
    for (int c = 0; c < 2; c++) {
      const uint32_t cc = ch[c];
      if (cc & LEAF) {
        const int j = (int)(cc & ~LEAF);
        if (j <= i) continue;
        const uint32_t tj = sorted_tri[j];
        if (!box_overlap(Q, boxes + (size_t)tj * 6)) continue;
        const uint32_t B0 = G.idx[(size_t)tj * 3], B1 = G.idx[(size_t)tj * 3 + 1], B2 = G.idx[(size_t)tj * 3 + 2];
        if (A0 == B0 || A0 == B1 || A0 == B2 || A1 == B0 || A1 == B1 || A1 == B2 || A2 == B0 || A2 == B1 ||
            A2 == B2)
          continue;
        const uint32_t k = atomicAdd(&cs->pair_count, 1u);
        if (k < cap) pairs[k] = make_uint2(ti < tj ? ti : tj, ti < tj ? tj : ti);
      } else {
        if ((int)last[cc] <= i) continue;
        if (!box_overlap(Q, nodebox + (size_t)cc * 6)) continue;
        if (sp < kStack) stack[sp++] = cc;
        else atomicOr(&cs->nonfinite, 2u);  // traversal stack overflow
      }
    }
  }
}

// queryBvhAll (contact-assembly.ts:124-140): cloth triangles overlapping each
// padded static triangle box (staticTriAabb :597-605).
__global__ void __launch_bounds__(kBlock) k_traverse_static(DGeom G, double pad, const uint32_t* __restrict__ sorted_tri,
                                                            const double* __restrict__ boxes,
                                                            const uint32_t* __restrict__ left,
                                                            const uint32_t* __restrict__ right,
                                                            const double* __restrict__ nodebox, uint2* pairs,
                                                            uint32_t cap, CScalars* cs) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= G.ms) return;
  double Q[6];
  {
    double P[3][3];
    for (int c = 0; c < 3; c++) vpos(G, nullptr, -(int)G.sidx[(size_t)s * 3 + c] - 2, P[c]);
    for (int k = 0; k < 3; k++) {
      Q[k] = jsm::jmin(jsm::jmin(P[0][k], P[1][k]), P[2][k]) - pad;
      Q[3 + k] = jsm::jmax(jsm::jmax(P[0][k], P[1][k]), P[2][k]) + pad;
    }
  }
  if (G.m == 1) {
    if (box_overlap(boxes, Q)) {
      const uint32_t k = atomicAdd(&cs->spair_count, 1u);
      if (k < cap) pairs[k] = make_uint2(0, s);
    }
    return;
  }
  uint32_t stack[kStack];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const uint32_t node = stack[--sp];
    const uint32_t ch[2] = {left[node], right[node]};
    for (int c = 0; c < 2; c++) {
      const uint32_t cc = ch[c];
      const double* bx = child_box(cc, sorted_tri, boxes, nodebox);
      if (!box_overlap(bx, Q)) continue;
      if (cc & LEAF) {
        const uint32_t k = atomicAdd(&cs->spair_count, 1u);
        if (k < cap) pairs[k] = make_uint2(sorted_tri[cc & ~LEAF], s);
      } else {
        if (sp < kStack) stack[sp++] = cc;
        else atomicOr(&cs->nonfinite, 2u);
      }
    }
  }
}

// ------------------------------------------------------------- prims

// One candidate primitive (expandTriPair self-collision.ts:23-41 for cloth
// pairs; :105-114 for cloth/static pairs).
struct Prim {
  int kind, p, a, b, c, d;
  uint64_t key, nkey;
};

__device__ __forceinline__ uint64_t rkey(const DGeom& G, uint64_t kind, uint64_t f1, uint64_t f2) {
  return (kind << (2 * G.B)) | (f1 << G.B) | f2;
}

__device__ __forceinline__ void prim_of(const DGeom& G, const uint2* pairs, uint32_t np, const uint2* spairs,
                                        uint64_t tid, Prim& q) {
  const int TE[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  if (tid < (uint64_t)np * 15) {
    const uint2 pr = pairs[tid / 15];
    const int k = (int)(tid % 15);
    const uint32_t tA = pr.x, tB = pr.y;
    const int A[3] = {(int)G.idx[(size_t)tA * 3], (int)G.idx[(size_t)tA * 3 + 1], (int)G.idx[(size_t)tA * 3 + 2]};
    const int B[3] = {(int)G.idx[(size_t)tB * 3], (int)G.idx[(size_t)tB * 3 + 1], (int)G.idx[(size_t)tB * 3 + 2]};
    if (k < 6) {
      const bool fromA = k < 3;
      const int* P = fromA ? A : B;
      const int* T = fromA ? B : A;
      const uint32_t tri = fromA ? tB : tA;
      q.kind = 0; q.p = P[k % 3]; q.a = T[0]; q.b = T[1]; q.c = T[2]; q.d = -1;
      q.key = rkey(G, 0, (uint64_t)q.p, G.tri_rank[tri]);
      q.nkey = rkey(G, 0, (uint64_t)q.p, tri);
    } else {
      const int e = k - 6, ea = e / 3, eb = e % 3;
      q.kind = 1; q.p = -1;
      q.a = A[TE[ea][0]]; q.b = A[TE[ea][1]]; q.c = B[TE[eb][0]]; q.d = B[TE[eb][1]];
      const uint32_t de1 = tA * 3 + ea, de2 = tB * 3 + eb;
      q.key = rkey(G, 1, G.dedge_rank[de1], G.dedge_rank[de2]);
      const uint32_t u1 = G.uedge[de1], u2 = G.uedge[de2];
      q.nkey = rkey(G, 1, u1 < u2 ? u1 : u2, u1 < u2 ? u2 : u1);
    }
  } else {
    const uint64_t t2 = tid - (uint64_t)np * 15;
    const uint2 pr = spairs[t2 / 12];
    const int k = (int)(t2 % 12);
    const uint32_t ct = pr.x, st = pr.y;
    const int A[3] = {(int)G.idx[(size_t)ct * 3], (int)G.idx[(size_t)ct * 3 + 1], (int)G.idx[(size_t)ct * 3 + 2]};
    const int B[3] = {-(int)G.sidx[(size_t)st * 3] - 2, -(int)G.sidx[(size_t)st * 3 + 1] - 2,
                      -(int)G.sidx[(size_t)st * 3 + 2] - 2};
    const uint32_t stri = G.m + st;
    if (k < 3) {
      q.kind = 0; q.p = A[k]; q.a = B[0]; q.b = B[1]; q.c = B[2]; q.d = -2;
      q.key = rkey(G, 0, (uint64_t)q.p, G.tri_rank[stri]);
      q.nkey = rkey(G, 0, (uint64_t)q.p, stri);
    } else {
      const int e = k - 3, ea = e / 3, eb = e % 3;
      q.kind = 1; q.p = -1;
      q.a = A[TE[ea][0]]; q.b = A[TE[ea][1]]; q.c = B[TE[eb][0]]; q.d = B[TE[eb][1]];
      const uint32_t de1 = ct * 3 + ea, de2 = stri * 3 + eb;
      q.key = rkey(G, 1, G.dedge_rank[de1], G.dedge_rank[de2]);
      const uint32_t u1 = G.uedge[de1], u2 = G.uedge[de2];
      q.nkey = rkey(G, 1, u1 < u2 ? u1 : u2, u1 < u2 ? u2 : u1);
    }
  }
}

__device__ __forceinline__ Closest prim_closest(const DGeom& G, const double* x, const Prim& q) {
  double P0[3], P1[3], P2[3], P3[3];
  if (q.kind == 0) {
    vpos(G, x, q.p, P0); vpos(G, x, q.a, P1); vpos(G, x, q.b, P2); vpos(G, x, q.c, P3);
    return closest_vt(P0, P1, P2, P3);
  }
  vpos(G, x, q.a, P0); vpos(G, x, q.b, P1); vpos(G, x, q.c, P2); vpos(G, x, q.d, P3);
  return closest_ee(P0, P1, P2, P3);
}

// updateActiveSet (contact-assembly.ts:164-199): closest point per candidate,
// active when 1e-12 <= d < dHat; min distance over all candidates.
__global__ void __launch_bounds__(kBlock) k_prims_active(DGeom G, const double* __restrict__ x,
                                                         const uint2* __restrict__ pairs, uint32_t np,
                                                         const uint2* __restrict__ spairs, uint32_t nsp,
                                                         Frozen* rec, uint64_t* keys, uint64_t* nkeys, uint32_t cap,
                                                         double* partials, CScalars* cs) {
  const uint64_t total = (uint64_t)np * 15 + (uint64_t)nsp * 12;
  double minD = __builtin_inf();
  for (uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; tid < total;
       tid += (uint64_t)gridDim.x * blockDim.x) {
    Prim q;
    prim_of(G, pairs, np, spairs, tid, q);
    const Closest cp = prim_closest(G, x, q);
    if (cp.dist < minD) minD = cp.dist;
    if (cp.dist >= G.d_hat || cp.dist < 1e-12) continue;
    Frozen f;
    // frozen VT records carry d = -1 even for static candidates (contact-assembly.ts:181)
    f.kind = q.kind; f.p = q.p; f.a = q.a; f.b = q.b; f.c = q.c; f.d = q.kind == 0 ? -1 : q.d;
    f.nx = cp.rx / cp.dist; f.ny = cp.ry / cp.dist; f.nz = cp.rz / cp.dist;
    if (q.kind == 0) {
      f.w1 = cp.s; f.w2 = cp.t; f.w0 = 1 - cp.s - cp.t; f.w3 = 0;
    } else {
      f.w0 = 1 - cp.s; f.w1 = cp.s; f.w2 = -(1 - cp.t); f.w3 = -cp.t;
    }
    const uint32_t k = atomicAdd(&cs->active_count, 1u);
    if (k < cap) {
      rec[k] = f;
      keys[k] = q.key;
      nkeys[k] = q.nkey;
    }
  }
  double r;
  if (grid_reduce_ext<kBlock, LtMin>(minD, partials, &cs->tickets[0], r)) cs->red[0] = r;
}

// Floor actives (contact-assembly.ts:202-213), appended after the candidates.
__global__ void __launch_bounds__(kBlock) k_floor_active(DGeom G, const double* __restrict__ x, Frozen* rec,
                                                         uint64_t* keys, uint64_t* nkeys, uint32_t cap,
                                                         double* partials, CScalars* cs) {
  double minD = __builtin_inf();
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < G.n; v += gridDim.x * blockDim.x) {
    const double d = x[(size_t)v * 3 + 1] - G.floor_y;
    if (d < minD) minD = d;
    if (d >= G.d_hat) continue;
    Frozen f;
    f.kind = 2; f.p = (int)v; f.a = f.b = f.c = f.d = -1;
    f.w0 = f.w1 = f.w2 = f.w3 = 0;
    f.nx = 0; f.ny = 1; f.nz = 0;
    const uint32_t k = atomicAdd(&cs->active_count, 1u);
    if (k < cap) {
      rec[k] = f;
      keys[k] = rkey(G, 2, 0, v);
      nkeys[k] = rkey(G, 2, 0, v);
    }
  }
  double r;
  if (grid_reduce_ext<kBlock, LtMin>(minD, partials, &cs->tickets[1], r)) cs->red[1] = r;
}

__global__ void k_count_kinds(const Frozen* __restrict__ act, uint32_t C, CScalars* cs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < C) atomicAdd(&cs->kind_count[act[i].kind & 3], 1u);
}

__global__ void k_iota(uint32_t* v, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

__global__ void k_gather_active(const uint32_t* __restrict__ perm, const Frozen* __restrict__ rec,
                                const uint64_t* __restrict__ nkeys, uint32_t C, Frozen* act, uint64_t* anke) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  const uint32_t j = perm[i];
  act[i] = rec[j];
  anke[i] = nkeys[j];
}

// Vertex -> (contact << 3 | slot) incidence, one key per cloth stencil vertex.
__global__ void k_emit_entries(const Frozen* __restrict__ act, uint32_t C, int cshift, uint64_t* keys,
                               uint32_t* vals, CScalars* cs) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const Frozen f = act[c];
  int vs[4];
  if (f.kind == 2) {
    vs[0] = f.p; vs[1] = vs[2] = vs[3] = -1;
  } else if (f.kind == 0) {
    vs[0] = f.p; vs[1] = f.a; vs[2] = f.b; vs[3] = f.c;
  } else {
    vs[0] = f.a; vs[1] = f.b; vs[2] = f.c; vs[3] = f.d;
  }
  uint32_t valid = 0;
  for (int s = 0; s < 4; s++) {
    const size_t o = (size_t)c * 4 + s;
    if (vs[s] >= 0) {
      keys[o] = ((uint64_t)vs[s] << cshift) | ((uint64_t)c << 3) | (uint64_t)s;
      valid++;
    } else {
      keys[o] = SENT;
    }
    vals[o] = 0;
  }
  atomicAdd(&cs->entry_count, valid);
}

__global__ void k_entry_bounds(const uint64_t* __restrict__ keys, uint32_t cnt, int cshift, uint32_t n,
                               uint32_t* v_off, uint32_t* v_ent) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const uint64_t k = keys[i];
  v_ent[i] = (uint32_t)(k & ((1ull << cshift) - 1));
  const int64_t v = (int64_t)(k >> cshift);
  const int64_t prev = i ? (int64_t)(keys[i - 1] >> cshift) : -1;
  for (int64_t u = prev + 1; u <= v; u++) v_off[u] = i;
  if (i == cnt - 1)
    for (int64_t u = v + 1; u <= (int64_t)n; u++) v_off[u] = cnt;
}

// ---------------------------------------------------------- per contact

// frozenDist (contact-assembly.ts:222-258)
__device__ __forceinline__ double frozen_dist(const DGeom& G, const double* x, const Frozen& c) {
  if (c.kind == 2) return x[(size_t)c.p * 3 + 1] - G.floor_y;
  if (c.kind == 0) {
    double dd[3];
    const size_t p = c.p;
    if (c.a >= 0) {
      const size_t a = c.a, b = c.b, cc = c.c;
      for (int k = 0; k < 3; k++) {
        const double q = c.w0 * x[a * 3 + k] + c.w1 * x[b * 3 + k] + c.w2 * x[cc * 3 + k];
        dd[k] = x[p * 3 + k] - q;
      }
    } else {
      const size_t ia = -c.a - 2, ib = -c.b - 2, ic = -c.c - 2;
      for (int k = 0; k < 3; k++) {
        const double s = c.w0 * G.spos[ia * 3 + k] + c.w1 * G.spos[ib * 3 + k] + c.w2 * G.spos[ic * 3 + k];
        dd[k] = x[p * 3 + k] - s;
      }
    }
    return c.nx * dd[0] + c.ny * dd[1] + c.nz * dd[2];
  }
  double A[3], B[3], Cc[3], D[3];
  vpos(G, x, c.a, A); vpos(G, x, c.b, B); vpos(G, x, c.c, Cc); vpos(G, x, c.d, D);
  const double rx = c.w0 * A[0] + c.w1 * B[0] + c.w2 * Cc[0] + c.w3 * D[0];
  const double ry = c.w0 * A[1] + c.w1 * B[1] + c.w2 * Cc[1] + c.w3 * D[1];
  const double rz = c.w0 * A[2] + c.w1 * B[2] + c.w2 * Cc[2] + c.w3 * D[2];
  return c.nx * rx + c.ny * ry + c.nz * rz;
}

// relativeSlip (contact-assembly.ts:344-367); VT reads D(c.d) as the reference.
__device__ __forceinline__ void relative_slip(const double* x, const double* xs, const Frozen& c, double* u) {
  auto D = [&](int v, double* o) {
    if (v < 0) { o[0] = o[1] = o[2] = 0; return; }
    for (int i = 0; i < 3; i++) o[i] = x[(size_t)v * 3 + i] - xs[(size_t)v * 3 + i];
  };
  if (c.kind == 2) { D(c.p, u); return; }
  if (c.kind == 0) {
    double dp[3], da[3], db[3], dc[3];
    D(c.p, dp); D(c.a, da); D(c.b, db); D(c.d, dc);
    for (int i = 0; i < 3; i++) u[i] = dp[i] - (c.w0 * da[i] + c.w1 * db[i] + c.w2 * dc[i]);
    return;
  }
  double da[3], db[3], dc[3], dd[3];
  D(c.a, da); D(c.b, db); D(c.c, dc); D(c.d, dd);
  for (int i = 0; i < 3; i++) u[i] = c.w0 * da[i] + c.w1 * db[i] + c.w2 * dc[i] + c.w3 * dd[i];
}

// Last table entry with `key` (Map.set last-write-wins), or -1.
__device__ __forceinline__ int64_t lag_find(const uint64_t* keys, uint32_t L, uint64_t key) {
  uint32_t lo = 0, hi = L;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (keys[mid] <= key) lo = mid + 1;
    else hi = mid;
  }
  if (lo == 0 || keys[lo - 1] != key) return -1;
  return (int64_t)lo - 1;
}

struct CEval {
  uint8_t* inr;
  double* g;
  double* fr;
  double* coeff;
};

// energyGrad (contact-assembly.ts:262-285) per contact + the curvature
// coefficient applyHvp / addDiagEstimate use (:376-386, :443-448).
__global__ void __launch_bounds__(kBlock) k_contact_eval(DGeom G, const double* __restrict__ x,
                                                         const double* __restrict__ xstep,
                                                         const Frozen* __restrict__ act,
                                                         const uint64_t* __restrict__ anke, uint32_t C,
                                                         const uint64_t* __restrict__ lag_keys,
                                                         const uint32_t* __restrict__ lag_idx, uint32_t L,
                                                         const double* __restrict__ lag_vals, CEval ev,
                                                         jsm::dd* partials, CScalars* cs) {
  jsm::dd energy = jsm::dd_zero(), fwork = jsm::dd_zero();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < C; i += gridDim.x * blockDim.x) {
    const Frozen c = act[i];
    const double d = jsm::jmax(frozen_dist(G, x, c), 1e-12);
    if (d >= G.d_hat) {
      ev.inr[i] = 0;
      ev.g[i] = 0;
      ev.coeff[i] = 0;
      ev.fr[(size_t)i * 3] = ev.fr[(size_t)i * 3 + 1] = ev.fr[(size_t)i * 3 + 2] = 0;
      continue;
    }
    ev.inr[i] = 1;
    jsm::dd_add(energy, G.kappa * barrier_value(d, G.d_hat));
    ev.g[i] = barrier_grad_scalar(d, G.d_hat, G.kappa, 1);
    double nx = c.nx, ny = c.ny, nz = c.nz, lam;
    const int64_t li = lag_find(lag_keys, L, anke[i]);
    if (li >= 0) {
      const double* lv = lag_vals + (size_t)lag_idx[li] * 4;
      nx = lv[0]; ny = lv[1]; nz = lv[2]; lam = lv[3];
    } else {
      lam = barrier_normal_force(d, G.d_hat, G.kappa, 1);
    }
    double u[3], f[3];
    relative_slip(x, xstep, c, u);
    coulomb_force(u, nx, ny, nz, lam, G.mu, G.feps, f);
    ev.fr[(size_t)i * 3] = f[0];
    ev.fr[(size_t)i * 3 + 1] = f[1];
    ev.fr[(size_t)i * 3 + 2] = f[2];
    jsm::dd_add(fwork, f[0] * u[0] + f[1] * u[1] + f[2] * u[2]);
    const double Lg = jsm::log(d / G.d_hat);
    const double uu = d - G.d_hat;
    const double b2 = -(2 * Lg + (4 * uu) / d - (uu * uu) / (d * d));
    ev.coeff[i] = G.kappa * jsm::jmax(b2, 0);
  }
  jsm::dd t;
  if (grid_reduce_dd<kBlock>(energy, partials, &cs->tickets[2], t)) cs->red[2] = jsm::dd_value(t);
  if (grid_reduce_dd<kBlock>(fwork, partials + 1024, &cs->tickets[3], t)) cs->red[3] = jsm::dd_value(t);
}

// Contact gradient per vertex: addDistGrad then addFrictionForce for each
// active contact in list order (contact-assembly.ts:272-279, 288-341).
__global__ void __launch_bounds__(kBlock) k_contact_grad(uint32_t n, const Frozen* __restrict__ act,
                                                         const uint32_t* __restrict__ v_off,
                                                         const uint32_t* __restrict__ v_ent, CEval ev,
                                                         double* __restrict__ ceg) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  double acc[3] = {0, 0, 0};
  for (uint32_t k = v_off[v]; k < v_off[v + 1]; k++) {
    const uint32_t ent = v_ent[k];
    const uint32_t ci = ent >> 3;
    const int slot = ent & 7;
    if (!ev.inr[ci]) continue;
    const Frozen c = act[ci];
    const double g = ev.g[ci];
    const double n3[3] = {c.nx, c.ny, c.nz};
    const double fp[3] = {-ev.fr[(size_t)ci * 3], -ev.fr[(size_t)ci * 3 + 1], -ev.fr[(size_t)ci * 3 + 2]};
    if (c.kind == 2) {
      acc[1] += g;
      acc[0] += fp[0]; acc[1] += fp[1]; acc[2] += fp[2];
    } else if (c.kind == 0) {
      if (slot == 0) {
        for (int q = 0; q < 3; q++) acc[q] += g * n3[q];
        for (int q = 0; q < 3; q++) acc[q] += fp[q];
      } else {
        const double w = slot == 1 ? c.w0 : slot == 2 ? c.w1 : c.w2;
        for (int q = 0; q < 3; q++) acc[q] += -g * w * n3[q];
        for (int q = 0; q < 3; q++) acc[q] += -fp[q] * w;
      }
    } else {
      const double w = slot == 0 ? c.w0 : slot == 1 ? c.w1 : slot == 2 ? c.w2 : c.w3;
      const double ww = slot == 0 ? c.w0 : slot == 1 ? c.w1 : slot == 2 ? -c.w2 : -c.w3;
      for (int q = 0; q < 3; q++) acc[q] += g * w * n3[q];
      for (int q = 0; q < 3; q++) acc[q] += fp[q] * ww;
    }
  }
  ceg[(size_t)v * 3] = acc[0];
  ceg[(size_t)v * 3 + 1] = acc[1];
  ceg[(size_t)v * 3 + 2] = acc[2];
}

// addDiagEstimate (contact-assembly.ts:436-471): accumulate into diag.
__global__ void __launch_bounds__(kBlock) k_contact_diag(uint32_t n, const Frozen* __restrict__ act,
                                                         const uint32_t* __restrict__ v_off,
                                                         const uint32_t* __restrict__ v_ent, CEval ev,
                                                         double* __restrict__ diag) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  double dg[3] = {diag[(size_t)v * 3], diag[(size_t)v * 3 + 1], diag[(size_t)v * 3 + 2]};
  for (uint32_t k = v_off[v]; k < v_off[v + 1]; k++) {
    const uint32_t ent = v_ent[k];
    const uint32_t ci = ent >> 3;
    const int slot = ent & 7;
    if (!ev.inr[ci]) continue;
    const Frozen c = act[ci];
    const double kk = ev.coeff[ci];
    if (c.kind == 2) {
      dg[1] += kk;
      continue;
    }
    const double n2[3] = {c.nx * c.nx, c.ny * c.ny, c.nz * c.nz};
    if (c.kind == 0 && slot == 0) {
      for (int q = 0; q < 3; q++) dg[q] += kk * n2[q];
    } else {
      const double w = c.kind == 0 ? (slot == 1 ? c.w0 : slot == 2 ? c.w1 : c.w2)
                                   : (slot == 0 ? c.w0 : slot == 1 ? c.w1 : slot == 2 ? c.w2 : c.w3);
      for (int q = 0; q < 3; q++) dg[q] += kk * w * w * n2[q];
    }
  }
  diag[(size_t)v * 3] = dg[0];
  diag[(size_t)v * 3 + 1] = dg[1];
  diag[(size_t)v * 3 + 2] = dg[2];
}

// distJvp (contact-assembly.ts:394-414) times the frozen curvature.
__global__ void __launch_bounds__(kBlock) k_contact_jv(const Frozen* __restrict__ act, uint32_t C, CEval ev,
                                                       const double* __restrict__ p, double* __restrict__ dd,
                                                       const int32_t* done) {
  if (done && *done) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  if (!ev.inr[i]) { dd[i] = 0; return; }
  const Frozen c = act[i];
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
  dd[i] = ev.coeff[i] * j;
}

// meritEnergy (contact-assembly.ts:419-428)
__global__ void __launch_bounds__(kBlock) k_merit(DGeom G, const double* __restrict__ x,
                                                  const Frozen* __restrict__ act, uint32_t C, jsm::dd* partials,
                                                  CScalars* cs) {
  jsm::dd e = jsm::dd_zero();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < C; i += gridDim.x * blockDim.x) {
    const double d = jsm::jmax(frozen_dist(G, x, act[i]), 1e-12);
    if (d >= G.d_hat) continue;
    jsm::dd_add(e, G.kappa * barrier_value(d, G.d_hat));
  }
  jsm::dd t;
  if (grid_reduce_dd<kBlock>(e, partials, &cs->tickets[4], t)) cs->red[4] = jsm::dd_value(t);
}

// commit (contact-assembly.ts:475-485): one table entry per in-range contact
// (sentinel otherwise), values appended at `base`.
__global__ void k_commit_emit(DGeom G, const double* __restrict__ x, const Frozen* __restrict__ act,
                              const uint64_t* __restrict__ anke, uint32_t C, uint32_t L, uint32_t base,
                              uint64_t* lag_keys, uint32_t* lag_idx, double* lag_vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  const Frozen c = act[i];
  const double d = jsm::jmax(frozen_dist(G, x, c), 1e-12);
  const uint32_t slot = base + i;
  if (d >= G.d_hat) {
    lag_keys[L + i] = SENT;
  } else {
    lag_keys[L + i] = anke[i];
    double* lv = lag_vals + (size_t)slot * 4;
    lv[0] = c.nx; lv[1] = c.ny; lv[2] = c.nz;
    lv[3] = barrier_normal_force(d, G.d_hat, G.kappa, 1);
  }
  lag_idx[L + i] = slot;
}

__global__ void k_count_below(const uint64_t* keys, uint32_t n, uint32_t* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool here = keys[i] != SENT;
  const bool next = (i + 1 < n) ? keys[i + 1] != SENT : false;
  if (here && !next) *out = i + 1;
  if (i == 0 && !here) *out = 0;
}

// killApproachVelocity (newton.ts:238-303): the reference sweeps the active
// list once in order.  A contact may run as soon as every earlier contact that
// shares one of its vertices has run (it is then at the head of each of its
// vertices' incidence lists), so rounds of ready contacts reproduce the
// sequential result exactly.  One work-group; every round runs at least the
// earliest pending contact, so the loop ends within C rounds.
__global__ void __launch_bounds__(1024) k_kill_approach(const Frozen* __restrict__ act, uint32_t C,
                                                        const uint32_t* __restrict__ v_ent, uint32_t* heads,
                                                        uint8_t* state, double* v) {
  __shared__ uint32_t s_left;
  for (uint32_t round = 0; round <= C; round++) {
    if (threadIdx.x == 0) s_left = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < C; i += blockDim.x) {
      if (state[i] == 2) continue;
      const Frozen c = act[i];
      int vs[4];
      if (c.kind == 2) { vs[0] = c.p; vs[1] = vs[2] = vs[3] = -1; }
      else if (c.kind == 0) { vs[0] = c.p; vs[1] = c.a; vs[2] = c.b; vs[3] = c.c; }
      else { vs[0] = c.a; vs[1] = c.b; vs[2] = c.c; vs[3] = c.d; }
      bool ready = true;
      for (int s = 0; s < 4 && ready; s++)
        if (vs[s] >= 0 && (v_ent[heads[vs[s]]] >> 3) != i) ready = false;
      state[i] = ready ? 1 : 0;
      atomicAdd(&s_left, 1u);
    }
    __syncthreads();
    if (s_left == 0) break;
    for (uint32_t i = threadIdx.x; i < C; i += blockDim.x) {
      if (state[i] != 1) continue;
      const Frozen c = act[i];
      if (c.kind == 2) {
        const double vy = v[(size_t)c.p * 3 + 1];
        if (vy < 0) v[(size_t)c.p * 3 + 1] = 0;
        heads[c.p]++;
      } else if (c.kind == 0) {
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
          const double j = -0.5 * rel;
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
        heads[p]++;
        for (int k = 0; k < 3; k++)
          if (vs[k] >= 0) heads[vs[k]]++;
      } else {
        double A[3] = {0, 0, 0}, B[3] = {0, 0, 0}, Cc[3] = {0, 0, 0}, D[3] = {0, 0, 0};
        for (int q = 0; q < 3; q++) {
          if (c.a >= 0) A[q] = v[(size_t)c.a * 3 + q];
          if (c.b >= 0) B[q] = v[(size_t)c.b * 3 + q];
          if (c.c >= 0) Cc[q] = v[(size_t)c.c * 3 + q];
          if (c.d >= 0) D[q] = v[(size_t)c.d * 3 + q];
        }
        double ab[3], cd[3];
        for (int q = 0; q < 3; q++) {
          ab[q] = (A[q] + B[q]) / 2;
          cd[q] = (Cc[q] + D[q]) / 2;
        }
        const double rel = (ab[0] - cd[0]) * c.nx + (ab[1] - cd[1]) * c.ny + (ab[2] - cd[2]) * c.nz;
        if (rel < 0) {
          const double j = -0.25 * rel;
          const int plus[2] = {c.a, c.b}, minus[2] = {c.c, c.d};
          for (int vv : plus) {
            if (vv < 0) continue;
            v[(size_t)vv * 3] += j * c.nx;
            v[(size_t)vv * 3 + 1] += j * c.ny;
            v[(size_t)vv * 3 + 2] += j * c.nz;
          }
          for (int vv : minus) {
            if (vv < 0) continue;
            v[(size_t)vv * 3] -= j * c.nx;
            v[(size_t)vv * 3 + 1] -= j * c.ny;
            v[(size_t)vv * 3 + 2] -= j * c.nz;
          }
        }
        const int all[4] = {c.a, c.b, c.c, c.d};
        for (int vv : all)
          if (vv >= 0) heads[vv]++;
      }
      state[i] = 2;
    }
    __syncthreads();
  }
}

// checkTrial prims (contact-assembly.ts:497-535): exact distance at the trial
// and CCD from step start; the first failure in candidate order decides
// whether a CCD failure is counted (nan_key vs cross_key).
__global__ void __launch_bounds__(kBlock) k_prims_trial(DGeom G, const double* __restrict__ xs,
                                                        const double* __restrict__ xt,
                                                        const uint2* __restrict__ pairs, uint32_t np,
                                                        const uint2* __restrict__ spairs, uint32_t nsp,
                                                        double* partials, CScalars* cs) {
  const uint64_t total = (uint64_t)np * 15 + (uint64_t)nsp * 12;
  double minD = __builtin_inf(), minT = __builtin_inf();
  for (uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; tid < total;
       tid += (uint64_t)gridDim.x * blockDim.x) {
    Prim q;
    prim_of(G, pairs, np, spairs, tid, q);
    const double d = prim_closest(G, xt, q).dist;
    if (d < minD) minD = d;
    if (!jsm::isfinite_(d)) atomicOr(&cs->nonfinite, 1u);
    Seg4 S;
    const int ids[4] = {q.kind == 0 ? q.p : q.a, q.kind == 0 ? q.a : q.b, q.kind == 0 ? q.b : q.c,
                        q.kind == 0 ? q.c : q.d};
    for (int s = 0; s < 4; s++) {
      vpos(G, xs, ids[s], S.x0[s]);
      vpos(G, xt, ids[s], S.x1[s]);
    }
    const double toi = q.kind == 0 ? vt_ccd(S, G.d_min) : ee_ccd(S, G.d_min);
    if (jsm::isnan_(toi)) {
      atomicMin((unsigned long long*)&cs->nan_key, (unsigned long long)q.key);
    } else {
      if (toi < minT) minT = toi;
      if (toi > 0 && toi < 1 - 1e-9) atomicMin((unsigned long long*)&cs->cross_key, (unsigned long long)q.key);
    }
  }
  double r;
  if (grid_reduce_ext<kBlock, LtMin>(minD, partials, &cs->tickets[5], r)) cs->red[5] = r;
  if (grid_reduce_ext<kBlock, LtMin>(minT, partials + 1024, &cs->tickets[6], r)) cs->red[6] = r;
}

// checkTrial floor terms (contact-assembly.ts:503-509, 537-549).
__global__ void __launch_bounds__(kBlock) k_floor_trial(DGeom G, const double* __restrict__ xs,
                                                        const double* __restrict__ xt, double* partials,
                                                        CScalars* cs) {
  double minD = __builtin_inf(), minT = __builtin_inf();
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < G.n; v += gridDim.x * blockDim.x) {
    const double d = xt[(size_t)v * 3 + 1] - G.floor_y;
    if (d < minD) minD = d;
    const double y0 = xs[(size_t)v * 3 + 1] - G.floor_y;
    const double yT = d;
    if (y0 - jsm::fabs_(yT - y0) > G.d_min) continue;
    if (y0 > G.d_min && yT <= G.d_min && y0 > yT) {
      const double toi = (y0 - G.d_min) / (y0 - yT);
      if (toi < minT) minT = toi;
      if (toi < 1 - 1e-9) atomicOr(&cs->floor_fail, 1u);
    }
  }
  double r;
  if (grid_reduce_ext<kBlock, LtMin>(minD, partials, &cs->tickets[7], r)) cs->red[7] = r;
  if (grid_reduce_ext<kBlock, LtMin>(minT, partials + 1024, &cs->tickets[8], r)) cs->red[8] = r;
}

// ====================================================================== host

static int bits_for(uint64_t v) {
  int b = 1;
  while (b < 63 && (1ull << b) <= v) b++;
  return b;
}

struct ContactImpl {
  hipStream_t st;
  cs_contact_params prm;
  uint32_t n, m, ns, ms;
  bool has_floor;
  double floor_y;
  int B;
  DevBuf<uint32_t> idx, sidx, tri_rank, dedge_rank, uedge;
  DevBuf<double> spos;
  // LBVH
  DevBuf<double> boxes, nodebox, bounds, dpart;
  DevBuf<uint64_t> mkeys, mkeys_t;
  DevBuf<uint32_t> mvals, mvals_t, left, right, first, last, parent_int, parent_leaf;
  DevBuf<unsigned> flags;
  const uint32_t* sorted_tri = nullptr;
  // candidates
  DevBuf<uint2> pairs, spairs;
  uint32_t pair_cap = 0, spair_cap = 0, np = 0, nsp = 0;
  // active set
  DevBuf<Frozen> rec, act;
  DevBuf<uint64_t> akeys, akeys_t, anke_raw, anke;
  DevBuf<uint32_t> avals, avals_t;
  uint32_t act_cap = 0, C = 0;
  // incidence
  DevBuf<uint64_t> ekeys, ekeys_t;
  DevBuf<uint32_t> evals, evals_t, v_off, v_ent;
  // per-contact evaluation
  DevBuf<uint8_t> inr, kstate;
  DevBuf<double> cg, fr, coeff, dd;
  DevBuf<uint32_t> heads;
  // lagged friction table
  DevBuf<uint64_t> lag_keys, lag_keys_t;
  DevBuf<uint32_t> lag_idx, lag_idx_t, lag_count;
  DevBuf<double> lag_vals;
  uint32_t lag_L = 0, lag_vbase = 0;
  // scalars
  DevBuf<CScalars> cs;
  CScalars* hcs = nullptr;
  DevBuf<jsm::dd> ddpart;
  SortScratch sort;
  DevBuf<uint32_t> sort_hist;
  const double* xstep = nullptr;
  cs_contact_diag diag;

  DGeom geom() const {
    DGeom G;
    G.n = n; G.m = m; G.ns = ns; G.ms = ms;
    G.idx = idx.ptr; G.sidx = sidx.ptr; G.spos = spos.ptr;
    G.tri_rank = tri_rank.ptr; G.dedge_rank = dedge_rank.ptr; G.uedge = uedge.ptr;
    G.B = B;
    G.has_floor = has_floor ? 1 : 0;
    G.floor_y = floor_y;
    G.d_hat = prm.d_hat; G.kappa = prm.kappa; G.d_min = prm.d_min; G.mu = prm.friction_mu;
    G.feps = prm.friction_eps;
    return G;
  }

  void reset_diag() {
    std::memset(&diag, 0, sizeof(diag));
    diag.min_distance = INFINITY;
    diag.min_toi = INFINITY;
  }

  void sync_cs() {
    CS_HIP(hipMemcpyAsync(hcs, cs.ptr, sizeof(CScalars), hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
  }

  void ensure_sort(size_t n_items) {
    const size_t e = radix_hist_entries(n_items);
    if (sort.hist_cap < e) {
      sort_hist.alloc(e);
      sort.hist = sort_hist.ptr;
      sort.hist_cap = e;
    }
  }

  ContactImpl(const cs_scene_desc& d, hipStream_t s) : st(s) {
    prm = d.contact;
    n = d.n_verts;
    m = d.n_tris;
    ns = d.n_static_tris ? d.n_static_verts : 0;
    ms = d.n_static_tris;
    has_floor = d.has_floor != 0;
    floor_y = d.floor_y;
    reset_diag();
    for (size_t k = 0; k < (size_t)ms * 3; k++)
      if (d.static_idx[k] >= ns) throw std::invalid_argument("static triangle index out of range");
    idx.upload(d.indices, (size_t)m * 3, st);
    if (ms) {
      sidx.upload(d.static_idx, (size_t)ms * 3, st);
      std::vector<double> sp(d.static_pos, d.static_pos + (size_t)ns * 3);
      spos.upload(sp.data(), sp.size(), st);
    } else {
      sidx.alloc(3);
      spos.alloc(3);
    }
    // signed vertex tuples of cloth and static triangles (static ids -(i+2))
    const size_t T = (size_t)m + ms;
    std::vector<std::array<int64_t, 3>> tup(T);
    for (size_t t = 0; t < m; t++)
      for (int c = 0; c < 3; c++) tup[t][c] = d.indices[t * 3 + c];
    for (size_t t = 0; t < ms; t++)
      for (int c = 0; c < 3; c++) tup[m + t][c] = -(int64_t)d.static_idx[t * 3 + c] - 2;
    std::vector<uint32_t> ord(T);
    std::iota(ord.begin(), ord.end(), 0u);
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return tup[a] < tup[b]; });
    std::vector<uint32_t> trank(T);
    for (size_t r = 0; r < T; r++) trank[ord[r]] = (uint32_t)r;
    // directed edges (a, b) for edge e of each triangle: (0,1), (1,2), (2,0)
    const int TE[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    std::vector<std::pair<int64_t, int64_t>> de(T * 3);
    for (size_t t = 0; t < T; t++)
      for (int e = 0; e < 3; e++) de[t * 3 + e] = {tup[t][TE[e][0]], tup[t][TE[e][1]]};
    std::vector<uint32_t> eo(T * 3);
    std::iota(eo.begin(), eo.end(), 0u);
    std::stable_sort(eo.begin(), eo.end(), [&](uint32_t a, uint32_t b) { return de[a] < de[b]; });
    std::vector<uint32_t> drank(T * 3), urank(T * 3);
    uint32_t r = 0;
    for (size_t k = 0; k < eo.size(); k++) {
      if (k && de[eo[k]] != de[eo[k - 1]]) r++;
      drank[eo[k]] = r;
    }
    auto und = [&](uint32_t i) {
      const auto& pr = de[i];
      return std::make_pair(std::min(pr.first, pr.second), std::max(pr.first, pr.second));
    };
    std::stable_sort(eo.begin(), eo.end(), [&](uint32_t a, uint32_t b) { return und(a) < und(b); });
    r = 0;
    for (size_t k = 0; k < eo.size(); k++) {
      if (k && und(eo[k]) != und(eo[k - 1])) r++;
      urank[eo[k]] = r;
    }
    tri_rank.upload(trank.data(), T, st);
    dedge_rank.upload(drank.data(), T * 3, st);
    uedge.upload(urank.data(), T * 3, st);
    B = bits_for(std::max<uint64_t>({(uint64_t)n, (uint64_t)T * 3}));
    if (2 * B + 2 > 64) throw std::invalid_argument("mesh too large for 64-bit contact keys");
    // LBVH buffers
    boxes.alloc((size_t)m * 6);
    nodebox.alloc((size_t)std::max<uint32_t>(m, 1) * 6);
    bounds.alloc(6);
    dpart.alloc(9 * 1024);
    mkeys.alloc(m); mkeys_t.alloc(m); mvals.alloc(m); mvals_t.alloc(m);
    for (auto* b : {&left, &right, &first, &last, &parent_int, &parent_leaf}) b->alloc(m);
    flags.alloc(m);
    cs.alloc(1);
    ddpart.alloc(2 * 1024);
    CS_HIP(hipHostMalloc(&hcs, sizeof(CScalars)));
    std::memset(hcs, 0, sizeof(CScalars));
    CS_HIP(hipMemsetAsync(cs.ptr, 0, sizeof(CScalars), st));
    v_off.alloc(n + 1);
    heads.alloc(n);
    lag_count.alloc(1);
    ensure_sort(m);
    grow_pairs(std::max<uint32_t>(1024, m * 4), std::max<uint32_t>(256, ms * 8));
    grow_active(std::max<uint32_t>(4096, n));
    CS_HIP(hipStreamSynchronize(st));
  }

  ~ContactImpl() {
    if (hcs) (void)hipHostFree(hcs);
  }

  void grow_pairs(uint32_t pc, uint32_t sc) {
    if (pc > pair_cap) { pairs.alloc(pc); pair_cap = pc; }
    if (sc > spair_cap) { spairs.alloc(sc); spair_cap = sc; }
  }

  void grow_active(uint32_t cap) {
    if (cap <= act_cap) return;
    rec.alloc(cap); act.alloc(cap);
    akeys.alloc(cap); akeys_t.alloc(cap); anke_raw.alloc(cap); anke.alloc(cap);
    avals.alloc(cap); avals_t.alloc(cap);
    ekeys.alloc((size_t)cap * 4); ekeys_t.alloc((size_t)cap * 4);
    evals.alloc((size_t)cap * 4); evals_t.alloc((size_t)cap * 4); v_ent.alloc((size_t)cap * 4);
    inr.alloc(cap); kstate.alloc(cap);
    cg.alloc(cap); fr.alloc((size_t)cap * 3); coeff.alloc(cap); dd.alloc(cap);
    act_cap = cap;
    ensure_sort((size_t)cap * 4);
  }

  void clear_counters() {
    CS_HIP(hipMemsetAsync(cs.ptr, 0, sizeof(CScalars), st));
  }

  // buildCandidates (contact-assembly.ts:87-121) minus the expansion into
  // prims, which the prim kernels do on the fly.
  void build_candidates(const double* x, double pad) {
    const DGeom G = geom();
    clear_counters();
    CS_HIP(hipMemsetAsync(&cs.ptr->nan_key, 0xff, 2 * sizeof(uint64_t), st));
    hipLaunchKernelGGL(k_tri_boxes, dim3(red_grid(m)), dim3(kBlock), 0, st, G, xstep, x, pad, boxes.ptr, dpart.ptr,
                       cs.ptr->tickets, bounds.ptr);
    hipLaunchKernelGGL(k_morton, dim3(grid_for(m)), dim3(kBlock), 0, st, m, boxes.ptr, bounds.ptr, mkeys.ptr,
                       mvals.ptr);
    CS_LAUNCH_CHECK();
    const bool in_tmp = radix_sort_pairs(st, mkeys.ptr, mvals.ptr, mkeys_t.ptr, mvals_t.ptr, m, 0, 30, sort);
    const uint64_t* codes = in_tmp ? mkeys_t.ptr : mkeys.ptr;
    sorted_tri = in_tmp ? mvals_t.ptr : mvals.ptr;
    if (m > 1) {
      CS_HIP(hipMemsetAsync(parent_int.ptr, 0xff, (size_t)m * sizeof(uint32_t), st));
      CS_HIP(hipMemsetAsync(flags.ptr, 0, (size_t)m * sizeof(unsigned), st));
      hipLaunchKernelGGL(k_karras, dim3(grid_for(m - 1)), dim3(kBlock), 0, st, (int)m, codes, left.ptr, right.ptr,
                         first.ptr, last.ptr, parent_int.ptr, parent_leaf.ptr);
      hipLaunchKernelGGL(k_refit, dim3(grid_for(m)), dim3(kBlock), 0, st, (int)m, sorted_tri, boxes.ptr, left.ptr,
                         right.ptr, parent_int.ptr, parent_leaf.ptr, nodebox.ptr, flags.ptr);
      CS_LAUNCH_CHECK();
    }
    for (int attempt = 0; attempt < 3; attempt++) {
      CS_HIP(hipMemsetAsync(&cs.ptr->pair_count, 0, 2 * sizeof(uint32_t), st));
      if (m > 1)
        hipLaunchKernelGGL(k_traverse_self, dim3(grid_for(m)), dim3(kBlock), 0, st, G, sorted_tri, boxes.ptr,
                           left.ptr, right.ptr, last.ptr, nodebox.ptr, pairs.ptr, pair_cap, cs.ptr);
      if (ms)
        hipLaunchKernelGGL(k_traverse_static, dim3(grid_for(ms)), dim3(kBlock), 0, st, G, pad, sorted_tri,
                           boxes.ptr, left.ptr, right.ptr, nodebox.ptr, spairs.ptr, spair_cap, cs.ptr);
      CS_LAUNCH_CHECK();
      sync_cs();
      if (hcs->nonfinite & 2u) throw OverflowError("LBVH traversal stack overflow");
      np = hcs->pair_count;
      nsp = hcs->spair_count;
      if (np <= pair_cap && nsp <= spair_cap) break;
      grow_pairs(std::max(pair_cap, (uint32_t)(np * 1.25) + 1024), std::max(spair_cap, (uint32_t)(nsp * 1.25) + 256));
      if (attempt == 2) throw OverflowError("candidate pair buffer could not be grown");
    }
    diag.candidate_pairs = (uint64_t)np * 15 + (uint64_t)nsp * 12;
  }

  void update_active_set(const double* x) {
    build_candidates(x, prm.d_hat);
    const DGeom G = geom();
    const uint64_t total = (uint64_t)np * 15 + (uint64_t)nsp * 12;
    for (int attempt = 0; attempt < 3; attempt++) {
      CS_HIP(hipMemsetAsync(&cs.ptr->active_count, 0, sizeof(uint32_t), st));
      hipLaunchKernelGGL(k_prims_active, dim3(red_grid(total)), dim3(kBlock), 0, st, G, x, pairs.ptr, np, spairs.ptr,
                         nsp, rec.ptr, akeys.ptr, anke_raw.ptr, act_cap, dpart.ptr, cs.ptr);
      if (has_floor)
        hipLaunchKernelGGL(k_floor_active, dim3(red_grid(n)), dim3(kBlock), 0, st, G, x, rec.ptr, akeys.ptr,
                           anke_raw.ptr, act_cap, dpart.ptr + 1024, cs.ptr);
      CS_LAUNCH_CHECK();
      sync_cs();
      if (hcs->active_count <= act_cap) break;
      grow_active((uint32_t)(hcs->active_count * 1.25) + 1024);
      if (attempt == 2) throw OverflowError("active contact buffer could not be grown");
    }
    C = hcs->active_count;
    double minD = total ? hcs->red[0] : INFINITY;
    if (has_floor && hcs->red[1] < minD) minD = hcs->red[1];
    // sort the frozen multiset into the reference's candidate order
    if (C) {
      hipLaunchKernelGGL(k_iota, dim3(grid_for(C)), dim3(kBlock), 0, st, avals.ptr, C);
      const bool in_tmp =
          radix_sort_pairs(st, akeys.ptr, avals.ptr, akeys_t.ptr, avals_t.ptr, C, 0, 2 * B + 2, sort);
      const uint32_t* perm = in_tmp ? avals_t.ptr : avals.ptr;
      hipLaunchKernelGGL(k_gather_active, dim3(grid_for(C)), dim3(kBlock), 0, st, perm, rec.ptr, anke_raw.ptr, C,
                         act.ptr, anke.ptr);
      CS_LAUNCH_CHECK();
    }
    CS_HIP(hipMemsetAsync(cs.ptr->kind_count, 0, sizeof(hcs->kind_count), st));
    if (C) hipLaunchKernelGGL(k_count_kinds, dim3(grid_for(C)), dim3(kBlock), 0, st, act.ptr, C, cs.ptr);
    build_incidence();  // syncs the scalars
    // diagnostics (contact-assembly.ts:215-219)
    diag.active_pairs = C;
    diag.vt_pairs = hcs->kind_count[0];
    diag.ee_pairs = hcs->kind_count[1];
    diag.floor_pairs = hcs->kind_count[2];
    if (minD < diag.min_distance) diag.min_distance = minD;
  }

  int cshift = 3;
  void build_incidence() {
    CS_HIP(hipMemsetAsync(v_off.ptr, 0, (size_t)(n + 1) * sizeof(uint32_t), st));
    if (!C) {
      sync_cs();
      return;
    }
    const int cb = bits_for(C);
    cshift = cb + 3;
    const int vb = bits_for(n) + 1;
    CS_HIP(hipMemsetAsync(&cs.ptr->entry_count, 0, sizeof(uint32_t), st));
    hipLaunchKernelGGL(k_emit_entries, dim3(grid_for(C)), dim3(kBlock), 0, st, act.ptr, C, cshift, ekeys.ptr,
                       evals.ptr, cs.ptr);
    CS_LAUNCH_CHECK();
    const size_t E = (size_t)C * 4;
    const bool in_tmp = radix_sort_pairs(st, ekeys.ptr, evals.ptr, ekeys_t.ptr, evals_t.ptr, E, 0, cshift + vb, sort);
    const uint64_t* sk = in_tmp ? ekeys_t.ptr : ekeys.ptr;
    sync_cs();
    const uint32_t cnt = hcs->entry_count;
    if (cnt)
      hipLaunchKernelGGL(k_entry_bounds, dim3(grid_for(cnt)), dim3(kBlock), 0, st, sk, cnt, cshift, n, v_off.ptr,
                         v_ent.ptr);
    CS_LAUNCH_CHECK();
  }

  CEval cev() { return CEval{inr.ptr, cg.ptr, fr.ptr, coeff.ptr}; }

  void energy_grad(const double* x, double* ceg) {
    const DGeom G = geom();
    if (C) {
      hipLaunchKernelGGL(k_contact_eval, dim3(red_grid(C)), dim3(kBlock), 0, st, G, x, xstep, act.ptr, anke.ptr, C,
                         lag_keys.ptr, lag_idx.ptr, lag_L, lag_vals.ptr, cev(), ddpart.ptr, cs.ptr);
      CS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_contact_grad, dim3(grid_for(n)), dim3(kBlock), 0, st, n, act.ptr, v_off.ptr, v_ent.ptr, cev(),
                       ceg);
    CS_LAUNCH_CHECK();
    if (C) {
      sync_cs();
      diag.barrier_energy = hcs->red[2];
      diag.friction_work += hcs->red[3];
    } else {
      diag.barrier_energy = 0;
    }
  }

  void add_diag_estimate(double* dg) {
    if (!C) return;
    hipLaunchKernelGGL(k_contact_diag, dim3(grid_for(n)), dim3(kBlock), 0, st, n, act.ptr, v_off.ptr, v_ent.ptr, cev(),
                       dg);
    CS_LAUNCH_CHECK();
  }

  double merit_energy(const double* x) {
    if (!C) return 0;
    hipLaunchKernelGGL(k_merit, dim3(red_grid(C)), dim3(kBlock), 0, st, geom(), x, act.ptr, C, ddpart.ptr, cs.ptr);
    CS_LAUNCH_CHECK();
    sync_cs();
    return hcs->red[4];
  }

  void commit(const double* x) {
    if (!C) return;
    const uint32_t need = lag_L + C;
    if (lag_keys.count < need) {
      // grow, preserving the current table
      DevBuf<uint64_t> nk;
      DevBuf<uint32_t> ni;
      nk.alloc((size_t)need * 2);
      ni.alloc((size_t)need * 2);
      if (lag_L) {
        CS_HIP(hipMemcpyAsync(nk.ptr, lag_keys.ptr, (size_t)lag_L * 8, hipMemcpyDeviceToDevice, st));
        CS_HIP(hipMemcpyAsync(ni.ptr, lag_idx.ptr, (size_t)lag_L * 4, hipMemcpyDeviceToDevice, st));
      }
      CS_HIP(hipStreamSynchronize(st));
      std::swap(lag_keys.ptr, nk.ptr); std::swap(lag_keys.count, nk.count);
      std::swap(lag_idx.ptr, ni.ptr); std::swap(lag_idx.count, ni.count);
      lag_keys_t.alloc(lag_keys.count);
      lag_idx_t.alloc(lag_idx.count);
    }
    if (lag_vals.count < (size_t)(lag_vbase + C) * 4) {
      DevBuf<double> nv;
      nv.alloc((size_t)(lag_vbase + C) * 8);
      if (lag_vbase)
        CS_HIP(hipMemcpyAsync(nv.ptr, lag_vals.ptr, (size_t)lag_vbase * 4 * sizeof(double), hipMemcpyDeviceToDevice, st));
      CS_HIP(hipStreamSynchronize(st));
      std::swap(lag_vals.ptr, nv.ptr); std::swap(lag_vals.count, nv.count);
    }
    hipLaunchKernelGGL(k_commit_emit, dim3(grid_for(C)), dim3(kBlock), 0, st, geom(), x, act.ptr, anke.ptr, C, lag_L,
                       lag_vbase, lag_keys.ptr, lag_idx.ptr, lag_vals.ptr);
    CS_LAUNCH_CHECK();
    const uint32_t total = lag_L + C;
    ensure_sort(total);
    const bool in_tmp =
        radix_sort_pairs(st, lag_keys.ptr, lag_idx.ptr, lag_keys_t.ptr, lag_idx_t.ptr, total, 0, 2 * B + 2, sort);
    if (in_tmp) {
      std::swap(lag_keys.ptr, lag_keys_t.ptr);
      std::swap(lag_idx.ptr, lag_idx_t.ptr);
    }
    hipLaunchKernelGGL(k_count_below, dim3(grid_for(total)), dim3(kBlock), 0, st, lag_keys.ptr, total, lag_count.ptr);
    CS_LAUNCH_CHECK();
    uint32_t L;
    CS_HIP(hipMemcpyAsync(&L, lag_count.ptr, 4, hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
    lag_L = L;
    lag_vbase += C;
  }

  void kill_approach_velocity(const double* x, double* v) {
    update_active_set(x);
    if (!C) return;
    CS_HIP(hipMemcpyAsync(heads.ptr, v_off.ptr, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    CS_HIP(hipMemsetAsync(kstate.ptr, 0, C, st));
    hipLaunchKernelGGL(k_kill_approach, dim3(1), dim3(1024), 0, st, act.ptr, C, v_ent.ptr, heads.ptr, kstate.ptr, v);
    CS_LAUNCH_CHECK();
  }

  bool check_trial(const double* xt) {
    build_candidates(xt, prm.d_min);
    const DGeom G = geom();
    const uint64_t total = (uint64_t)np * 15 + (uint64_t)nsp * 12;
    if (total)
      hipLaunchKernelGGL(k_prims_trial, dim3(red_grid(total)), dim3(kBlock), 0, st, G, xstep, xt, pairs.ptr, np,
                         spairs.ptr, nsp, dpart.ptr, cs.ptr);
    if (has_floor)
      hipLaunchKernelGGL(k_floor_trial, dim3(red_grid(n)), dim3(kBlock), 0, st, G, xstep, xt, dpart.ptr + 2048, cs.ptr);
    CS_LAUNCH_CHECK();
    sync_cs();
    double minDist = total ? hcs->red[5] : INFINITY;
    double minTOI = total ? hcs->red[6] : INFINITY;
    if (hcs->nonfinite & 1u) {
      diag.ccd_failures++;
      return false;
    }
    if (has_floor && hcs->red[7] < minDist) minDist = hcs->red[7];
    if (minDist <= prm.d_min) return false;
    if (hcs->nan_key != SENT && hcs->nan_key < hcs->cross_key) {
      diag.ccd_failures++;
      return false;
    }
    if (hcs->cross_key != SENT) return false;
    if (has_floor) {
      if (hcs->red[8] < minTOI) minTOI = hcs->red[8];
      if (hcs->floor_fail) return false;
    }
    if (minDist < diag.min_distance) diag.min_distance = minDist;
    if (minTOI < diag.min_toi) diag.min_toi = minTOI;
    return true;
  }
};

// ---------------------------------------------------------------- facade

ContactSystem::ContactSystem(const cs_scene_desc& d, hipStream_t st) : im_(new ContactImpl(d, st)) {}
ContactSystem::~ContactSystem() { delete im_; }
void ContactSystem::begin_step(const double* x0) {
  im_->xstep = x0;
  im_->lag_L = 0;
  im_->lag_vbase = 0;
  im_->C = 0;
  im_->reset_diag();
}
bool ContactSystem::check_trial(const double* xt) { return im_->check_trial(xt); }
void ContactSystem::update_active_set(const double* x) { im_->update_active_set(x); }
void ContactSystem::energy_grad(const double* x, double* ceg) { im_->energy_grad(x, ceg); }
void ContactSystem::add_diag_estimate(double* diag, const double*) { im_->add_diag_estimate(diag); }
double ContactSystem::merit_energy(const double* x) { return im_->merit_energy(x); }
void ContactSystem::add_rejection() { im_->diag.rejected_line_search_steps++; }
void ContactSystem::commit(const double* x) { im_->commit(x); }
void ContactSystem::kill_approach_velocity(const double* x, double* v) { im_->kill_approach_velocity(x, v); }
cs_contact_diag ContactSystem::diagnostics() const { return im_->diag; }
uint64_t ContactSystem::active_count() const { return im_->C; }

uint64_t ContactSystem::candidate_count() const { return im_->diag.candidate_pairs; }

uint64_t ContactSystem::copy_active(int32_t* ids6, double* vals7, uint64_t cap) {
  const uint64_t C = im_->C;
  if (C) {
    std::vector<Frozen> tmp(C);
    CS_HIP(hipMemcpyAsync(tmp.data(), im_->act.ptr, C * sizeof(Frozen), hipMemcpyDeviceToHost, im_->st));
    CS_HIP(hipStreamSynchronize(im_->st));
    for (uint64_t k = 0; k < C && k < cap; k++) {
      const Frozen& f = tmp[k];
      const int32_t v[6] = {f.kind, f.p, f.a, f.b, f.c, f.d};
      const double w[7] = {f.w0, f.w1, f.w2, f.w3, f.nx, f.ny, f.nz};
      for (int j = 0; j < 6; j++) ids6[k * 6 + j] = v[j];
      for (int j = 0; j < 7; j++) vals7[k * 7 + j] = w[j];
    }
  }
  return C;
}

__global__ void k_probe_ccd(const double* x0, const double* x1, const int32_t* kinds, const int32_t* ids4,
                            uint64_t count, double thickness, double* toi) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= count) return;
  Seg4 S;
  for (int s = 0; s < 4; s++) {
    const size_t v = ids4[i * 4 + s];
    for (int k = 0; k < 3; k++) {
      S.x0[s][k] = x0[v * 3 + k];
      S.x1[s][k] = x1[v * 3 + k];
    }
  }
  toi[i] = kinds[i] == 0 ? vt_ccd(S, thickness) : ee_ccd(S, thickness);
}

__global__ void k_probe_closest(const double* x0, const int32_t* kinds, const int32_t* ids4, uint64_t count,
                                double* out3) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= count) return;
  double P[4][3];
  for (int s = 0; s < 4; s++) {
    const size_t v = ids4[i * 4 + s];
    for (int k = 0; k < 3; k++) P[s][k] = x0[v * 3 + k];
  }
  const Closest c = kinds[i] == 0 ? closest_vt(P[0], P[1], P[2], P[3]) : closest_ee(P[0], P[1], P[2], P[3]);
  out3[i * 3] = c.s;
  out3[i * 3 + 1] = c.t;
  out3[i * 3 + 2] = c.dist;
}

void probe_ccd(hipStream_t st, const double* x0, const double* x1, uint32_t n, const int32_t* kinds,
               const int32_t* ids4, uint64_t count, double thickness, double* toi) {
  for (uint64_t k = 0; k < count * 4; k++)
    if (ids4[k] < 0 || (uint32_t)ids4[k] >= n) throw std::invalid_argument("probe vertex id out of range");
  DevBuf<double> a, b, t;
  DevBuf<int32_t> kd, id;
  a.upload(x0, (size_t)n * 3, st);
  b.upload(x1, (size_t)n * 3, st);
  kd.upload(kinds, count, st);
  id.upload(ids4, count * 4, st);
  t.alloc(count);
  hipLaunchKernelGGL(k_probe_ccd, dim3(grid_for(count)), dim3(kBlock), 0, st, a.ptr, b.ptr, kd.ptr, id.ptr, count,
                     thickness, t.ptr);
  CS_LAUNCH_CHECK();
  CS_HIP(hipMemcpyAsync(toi, t.ptr, count * sizeof(double), hipMemcpyDeviceToHost, st));
  CS_HIP(hipStreamSynchronize(st));
}

void probe_closest(hipStream_t st, const double* x0, uint32_t n, const int32_t* kinds, const int32_t* ids4,
                   uint64_t count, double* out3) {
  for (uint64_t k = 0; k < count * 4; k++)
    if (ids4[k] < 0 || (uint32_t)ids4[k] >= n) throw std::invalid_argument("probe vertex id out of range");
  DevBuf<double> a, o;
  DevBuf<int32_t> kd, id;
  a.upload(x0, (size_t)n * 3, st);
  kd.upload(kinds, count, st);
  id.upload(ids4, count * 4, st);
  o.alloc(count * 3);
  hipLaunchKernelGGL(k_probe_closest, dim3(grid_for(count)), dim3(kBlock), 0, st, a.ptr, kd.ptr, id.ptr, count, o.ptr);
  CS_LAUNCH_CHECK();
  CS_HIP(hipMemcpyAsync(out3, o.ptr, count * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
  CS_HIP(hipStreamSynchronize(st));
}

DContactHvp ContactSystem::hvp_hook() {
  DContactHvp h;
  h.C = im_->C;
  h.act = im_->act.ptr;
  h.coeff = im_->coeff.ptr;
  h.dd = im_->dd.ptr;
  h.v_off = im_->v_off.ptr;
  h.v_ent = im_->v_ent.ptr;
  return h;
}

void ContactSystem::launch_hvp_coeffs(const double* p, const int32_t* done) {
  if (!im_->C) return;
  hipLaunchKernelGGL(k_contact_jv, dim3(grid_for(im_->C)), dim3(kBlock), 0, im_->st, im_->act.ptr, im_->C, im_->cev(),
                     p, im_->dd.ptr, done);
  CS_LAUNCH_CHECK();
}

}  // namespace cloth
