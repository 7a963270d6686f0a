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
  uint64_t mink[16];   // okey images of LtMin reductions (all ones = no value)
  uint32_t pair_count, spair_count, active_count, entry_count, nonfinite, floor_fail, nverts, kill_rounds;
  uint32_t tail_count;
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

__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long u = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// ---------------------------------------------------------------- LBVH
//
// The candidate set of selfPairs (bvh.ts:79-124) is defined by the inclusive
// swept-AABB predicate and the shared-vertex exclusion alone, so any
// hierarchy whose inner boxes contain their leaves finds it exactly.  Here the
// hierarchy is implicit: leaves are the triangles in 30-bit Morton order and
// level l node k covers leaves [k 8^l, (k+1) 8^l).  Inner boxes are exact
// double unions (NaN-ignoring), leaf tests use the reference's own box values.
// A refit is one fused kernel (leaf boxes + three levels by wave shuffles) and
// one small block for the levels above; the Morton order is reused until the
// owner asks for a re-sort, so most builds skip the sort.

constexpr int kFan = 8;
constexpr int kMaxLevels = 9;   // stack bound 7 L + 1 <= 64 entries per wave
constexpr int kTravStack = 64;

struct Leaf {     // 64 B
  double lo[3], hi[3];
  int32_t v[3];   // cloth vertex ids of the triangle
  uint32_t tri;   // triangle id
};

struct Levels {   // level l (1..L) holds cnt[l] inner boxes at nodes + off[l] * 6
  int L;
  uint32_t cnt[kMaxLevels + 1];
  uint32_t off[kMaxLevels + 1];
  uint32_t span[kMaxLevels + 1];  // leaves under one level-l node (8^l)
};

static Levels make_levels(uint32_t m, size_t* total_nodes) {
  Levels LV;
  std::memset(&LV, 0, sizeof(LV));
  LV.cnt[0] = m;
  LV.span[0] = 1;
  uint32_t off = 0;
  int L = 0;
  uint32_t c = m;
  while (c > 1) {
    if (L == kMaxLevels) throw std::invalid_argument("mesh too large for the 8-ary LBVH (more than 8^9 triangles)");
    c = (c + kFan - 1) / kFan;
    L++;
    LV.cnt[L] = c;
    LV.off[L] = off;
    LV.span[L] = LV.span[L - 1] * kFan;
    off += c;
  }
  LV.L = L;
  *total_nodes = off;
  return LV;
}

// Order-preserving u64 image of a double (for atomic min / max of non-NaN values).
__device__ __forceinline__ uint64_t okey(double d) {
  const uint64_t u = (uint64_t)__double_as_longlong(d);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double okey_dec(uint64_t k) {
  const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)u);
}

// Block-wide `if (v < best) best = v` (NaN never taken) folded into a global
// okey slot with one atomic per block.  Every thread of the block must call it.
__device__ __forceinline__ void block_min_to(double v, uint64_t* slot) {
  __shared__ double s_min[kBlock / kWave];
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const double o = __shfl_down(v, off, kWave);
    if (o < v) v = o;
  }
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0) s_min[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = s_min[0];
    for (int w = 1; w < kBlock / kWave; w++)
      if (s_min[w] < r) r = s_min[w];
    if (r < __builtin_inf()) atomicMin((unsigned long long*)slot, (unsigned long long)okey(r));
  }
  __syncthreads();
}

__device__ __forceinline__ bool box_overlap6(const double* a, const double* lo, const double* hi) {
  return a[0] <= hi[0] && a[3] >= lo[0] && a[1] <= hi[1] && a[4] >= lo[1] && a[2] <= hi[2] && a[5] >= lo[2];
}

__device__ __forceinline__ uint32_t rank_in(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

// Wave-aggregated allocation: the lanes with `want` get consecutive slots
// from one atomic per wave.  Every lane of the wave must call it.
__device__ __forceinline__ uint32_t wave_alloc(bool want, uint32_t* counter) {
  const uint64_t mask = __ballot(want);
  if (!mask) return 0;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if ((int)(threadIdx.x % kWave) == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, leader, kWave);
  return base + rank_in(mask);
}

// Per-wave staging of found pairs in LDS; one global atomic per flush.
constexpr int kPairStage = 256;
struct PairStage {
  uint2* buf;  // kPairStage entries of this wave
  int count;   // wave-uniform
  __device__ __forceinline__ void push(bool hit, uint2 v, uint32_t* counter, uint2* out, uint32_t cap) {
    const uint64_t mask = __ballot(hit);
    if (!mask) return;
    if (hit) buf[count + rank_in(mask)] = v;
    count += __popcll(mask);
    if (count > kPairStage - kWave) flush(counter, out, cap);
  }
  __device__ __forceinline__ void flush(uint32_t* counter, uint2* out, uint32_t cap) {
    if (count == 0) return;
    __builtin_amdgcn_wave_barrier();
    const int lane = threadIdx.x % kWave;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(counter, (uint32_t)count);
    base = __shfl(base, 0, kWave);
    for (int t = lane; t < count; t += kWave)
      if (base + t < cap) out[base + t] = buf[t];
    __builtin_amdgcn_wave_barrier();
    count = 0;
  }
};

// sweptTriAabb (aabb.ts:9-31): min / max over the three vertices at x0 and
// x1 in the reference's visiting order, then -/+ pad.  With x2 the box also
// covers a third configuration (segment supersets, never a reference box).
__device__ __forceinline__ void swept_box(const DGeom& G, const double* x0, const double* x1, uint32_t t, double pad,
                                          double* lo, double* hi, int32_t* v, const double* x2 = nullptr) {
  double mn[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
  double mx[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
  const int nq = x2 ? 3 : 2;
  for (int c = 0; c < 3; c++) {
    const uint32_t vi = G.idx[(size_t)t * 3 + c];
    v[c] = (int32_t)vi;
    const double* xs[3] = {x0, x1, x2};
    for (int q = 0; q < nq; q++) {
      for (int k = 0; k < 3; k++) {
        const double pk = xs[q][(size_t)vi * 3 + k];
        if (pk < mn[k]) mn[k] = pk;
        if (pk > mx[k]) mx[k] = pk;
      }
    }
  }
  for (int k = 0; k < 3; k++) {
    lo[k] = mn[k] - pad;
    hi[k] = mx[k] + pad;
  }
}

// Morton frame: bounds of the vertex positions (NaN ignored), one atomic per
// block and axis.  bkeys = {lo x,y,z (min), hi x,y,z (max)} as okey images.
__global__ void __launch_bounds__(kBlock) k_vert_bounds(uint32_t n, const double* __restrict__ x, uint64_t* bkeys) {
  double lo[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
  double hi[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    for (int k = 0; k < 3; k++) {
      const double p = x[(size_t)v * 3 + k];
      lo[k] = fmin(lo[k], p);
      hi[k] = fmax(hi[k], p);
    }
  __shared__ double s[6][kBlock / kWave];
  for (int k = 0; k < 3; k++)
    for (int off = kWave / 2; off > 0; off >>= 1) {
      lo[k] = fmin(lo[k], __shfl_down(lo[k], off, kWave));
      hi[k] = fmax(hi[k], __shfl_down(hi[k], off, kWave));
    }
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0)
    for (int k = 0; k < 3; k++) { s[k][wid] = lo[k]; s[3 + k][wid] = hi[k]; }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    double r = s[k][0];
    for (int w = 1; w < kBlock / kWave; w++) r = k < 3 ? fmin(r, s[k][w]) : fmax(r, s[k][w]);
    if (r == r && (k < 3 ? r < __builtin_inf() : r > -__builtin_inf())) {
      if (k < 3) atomicMin((unsigned long long*)&bkeys[k], (unsigned long long)okey(r));
      else atomicMax((unsigned long long*)&bkeys[k], (unsigned long long)okey(r));
    }
  }
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// 30-bit Morton code of each triangle centroid in the cube spanned by the
// largest extent of the vertex bounds (equal cells on every axis, so a thin
// sheet's jitter across its plane does not take over the high bits).
__global__ void k_morton(DGeom G, const double* __restrict__ x, const uint64_t* __restrict__ bkeys, uint64_t* keys,
                         uint32_t* vals) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.m) return;
  double ext = 0;
  for (int k = 0; k < 3; k++) {
    const double e = okey_dec(bkeys[3 + k]) - okey_dec(bkeys[k]);
    if (e > ext) ext = e;
  }
  uint32_t q[3];
  for (int k = 0; k < 3; k++) {
    const double lo = okey_dec(bkeys[k]);
    double c = 0;
    for (int j = 0; j < 3; j++) c += x[(size_t)G.idx[(size_t)t * 3 + j] * 3 + k];
    double u = ext > 0 ? (c * (1.0 / 3.0) - lo) / ext : 0.0;
    if (!(u >= 0)) u = 0;
    if (u > 1) u = 1;
    q[k] = (uint32_t)fmin(1023.0, u * 1024.0);
  }
  keys[t] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
  vals[t] = t;
}

__device__ __forceinline__ void store_box(double* dst, const double* lo, const double* hi) {
  for (int k = 0; k < 3; k++) { dst[k] = lo[k]; dst[3 + k] = hi[k]; }
}

// Refit: one wave per 512 sorted leaves.  Lane L owns leaves 8L..8L+7 of its
// wave's range: writes their records and the level-1 box; levels 2 and 3
// come from xor-shuffles over 8 and 64 lanes.
__global__ void __launch_bounds__(kBlock) k_refit8(DGeom G, const double* __restrict__ x0,
                                                   const double* __restrict__ x1, const double* __restrict__ x2,
                                                   double pad, const uint32_t* __restrict__ order, Leaf* leaves,
                                                   double* nodes, Levels LV) {
  const uint32_t lane = threadIdx.x % kWave;
  const uint32_t wbase = (blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) * (kWave * kFan);
  double lo[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
  double hi[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
  for (int c = 0; c < kFan; c++) {
    const uint32_t j = wbase + lane * kFan + c;
    if (j >= G.m) break;
    Leaf R;
    const uint32_t t = order[j];
    swept_box(G, x0, x1, t, pad, R.lo, R.hi, R.v, x2);
    R.tri = t;
    leaves[j] = R;
    for (int k = 0; k < 3; k++) {
      lo[k] = fmin(lo[k], R.lo[k]);
      hi[k] = fmax(hi[k], R.hi[k]);
    }
  }
  const uint32_t n1 = wbase / kFan + lane;
  if (LV.L >= 1 && n1 < LV.cnt[1]) store_box(nodes + ((size_t)LV.off[1] + n1) * 6, lo, hi);
  for (int o = 1; o < kFan; o <<= 1)
    for (int k = 0; k < 3; k++) {
      lo[k] = fmin(lo[k], __shfl_xor(lo[k], o, kWave));
      hi[k] = fmax(hi[k], __shfl_xor(hi[k], o, kWave));
    }
  const uint32_t n2 = wbase / (kFan * kFan) + lane / kFan;
  if (LV.L >= 2 && lane % kFan == 0 && n2 < LV.cnt[2]) store_box(nodes + ((size_t)LV.off[2] + n2) * 6, lo, hi);
  for (int o = kFan; o < kWave; o <<= 1)
    for (int k = 0; k < 3; k++) {
      lo[k] = fmin(lo[k], __shfl_xor(lo[k], o, kWave));
      hi[k] = fmax(hi[k], __shfl_xor(hi[k], o, kWave));
    }
  const uint32_t n3 = wbase / (kFan * kFan * kFan);
  if (LV.L >= 3 && lane == 0 && n3 < LV.cnt[3]) store_box(nodes + ((size_t)LV.off[3] + n3) * 6, lo, hi);
}

// Levels 4..L in one block (level 3 has m/512 nodes).
__global__ void __launch_bounds__(1024) k_refit_top(double* nodes, Levels LV) {
  for (int l = 4; l <= LV.L; l++) {
    for (uint32_t k = threadIdx.x; k < LV.cnt[l]; k += blockDim.x) {
      double lo[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
      double hi[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
      const uint32_t c0 = k * kFan, c1 = min(c0 + kFan, LV.cnt[l - 1]);
      for (uint32_t c = c0; c < c1; c++) {
        const double* b = nodes + ((size_t)LV.off[l - 1] + c) * 6;
        for (int q = 0; q < 3; q++) {
          lo[q] = fmin(lo[q], b[q]);
          hi[q] = fmax(hi[q], b[3 + q]);
        }
      }
      store_box(nodes + ((size_t)LV.off[l] + k) * 6, lo, hi);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ uint32_t stk_enc(int l, uint32_t k) { return ((uint32_t)l << 28) | k; }

// selfPairs (bvh.ts:79-124): every unordered pair of overlapping triangles
// sharing no vertex, found once from the leaf with the smaller sorted index.
// Wave-packet traversal: a wave queries its 64 consecutive leaves together;
// a node is entered when any lane's query box overlaps it and a lane can
// still pair with a later leaf under it.  The node stack is wave-uniform in LDS.
__global__ void __launch_bounds__(kBlock) k_traverse_self(const Leaf* __restrict__ leaves,
                                                          const double* __restrict__ nodes, Levels LV, uint2* pairs,
                                                          uint32_t cap, CScalars* cs) {
  __shared__ uint32_t stk[kBlock / kWave][kTravStack];
  __shared__ uint2 stage[kBlock / kWave][kPairStage];
  const int wid = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const uint32_t m = LV.cnt[0];
  PairStage ps{stage[wid], 0};
  const uint32_t i = (blockIdx.x * (kBlock / kWave) + wid) * kWave + threadIdx.x % kWave;
  const bool act = i < m;
  double Q[6];
  int32_t A[3] = {-1, -1, -1};
  uint32_t ti = 0;
  if (act) {
    const Leaf& R = leaves[i];
    for (int k = 0; k < 3; k++) { Q[k] = R.lo[k]; Q[3 + k] = R.hi[k]; A[k] = R.v[k]; }
    ti = R.tri;
  } else {
    for (int k = 0; k < 6; k++) Q[k] = __builtin_nan("");
  }
  if (LV.L == 0) return;
  uint32_t* S = stk[wid];
  int sp = 0;
  S[sp++] = stk_enc(LV.L, 0);
  bool overflow = false;
  while (sp > 0) {
    const uint32_t e = __builtin_amdgcn_readfirstlane(S[--sp]);
    const int cl = (int)(e >> 28) - 1;
    const uint32_t k = e & 0x0fffffffu;
    const uint32_t c0 = k * kFan, c1 = min(c0 + kFan, LV.cnt[cl]);
    // lanes 0..7 fetch the (up to 8) children with one vector load; each is
    // then broadcast to the packet with readlane
    const int nch = (int)(c1 - c0);
    if (cl == 0) {
      Leaf mine;
      if (lane < nch) mine = leaves[c0 + lane];
      for (int q = 0; q < nch; q++) {
        double lo[3], hi[3];
        for (int k = 0; k < 3; k++) { lo[k] = readlane_d(mine.lo[k], q); hi[k] = readlane_d(mine.hi[k], q); }
        const int32_t v0 = __builtin_amdgcn_readlane(mine.v[0], q), v1 = __builtin_amdgcn_readlane(mine.v[1], q),
                      v2 = __builtin_amdgcn_readlane(mine.v[2], q);
        const uint32_t tj = (uint32_t)__builtin_amdgcn_readlane((int)mine.tri, q);
        const uint32_t j = c0 + q;
        const bool hit = act && j > i && box_overlap6(Q, lo, hi) &&
                         !(A[0] == v0 || A[0] == v1 || A[0] == v2 || A[1] == v0 || A[1] == v1 || A[1] == v2 ||
                           A[2] == v0 || A[2] == v1 || A[2] == v2);
        ps.push(hit, make_uint2(ti < tj ? ti : tj, ti < tj ? tj : ti), &cs->pair_count, pairs, cap);
      }
    } else {
      double mb[6];
      if (lane < nch) {
        const double* b = nodes + ((size_t)LV.off[cl] + c0 + lane) * 6;
        for (int k = 0; k < 6; k++) mb[k] = b[k];
      }
      for (int q = nch - 1; q >= 0; q--) {
        double lo[3], hi[3];
        for (int k = 0; k < 3; k++) { lo[k] = readlane_d(mb[k], q); hi[k] = readlane_d(mb[3 + k], q); }
        const uint32_t c = c0 + q;
        const uint64_t last = min((uint64_t)(c + 1) * LV.span[cl], (uint64_t)m) - 1;
        const bool hit = act && last > i && box_overlap6(Q, lo, hi);
        if (__any(hit)) {
          if (sp < kTravStack) S[sp++] = stk_enc(cl, c);
          else overflow = true;
        }
      }
    }
  }
  ps.flush(&cs->pair_count, pairs, cap);
  if (overflow && (threadIdx.x % kWave) == 0) atomicOr(&cs->nonfinite, 2u);
}

// queryBvhAll (contact-assembly.ts:124-140): cloth triangles whose swept box
// overlaps a static triangle box padded by `pad` (staticTriAabb :597-605).
// Queries are the cloth leaves (wave packets as above) against a fixed
// hierarchy over the static triangles holding their unpadded boxes; the
// padded leaf box is formed exactly as the reference forms it, and inner
// boxes are widened by pad, which keeps them containing the padded leaves.
__global__ void __launch_bounds__(kBlock) k_traverse_static(const Leaf* __restrict__ leaves, uint32_t m,
                                                            const Leaf* __restrict__ sleaves,
                                                            const double* __restrict__ snodes, Levels SL, double pad,
                                                            uint2* pairs, uint32_t cap, CScalars* cs) {
  __shared__ uint32_t stk[kBlock / kWave][kTravStack];
  __shared__ uint2 stage[kBlock / kWave][kPairStage];
  const int wid = threadIdx.x / kWave;
  const uint32_t i = (blockIdx.x * (kBlock / kWave) + wid) * kWave + threadIdx.x % kWave;
  const bool act = i < m;
  PairStage ps{stage[wid], 0};
  double Q[6];
  uint32_t ti = 0;
  if (act) {
    const Leaf& R = leaves[i];
    for (int k = 0; k < 3; k++) { Q[k] = R.lo[k]; Q[3 + k] = R.hi[k]; }
    ti = R.tri;
  } else {
    for (int k = 0; k < 6; k++) Q[k] = __builtin_nan("");
  }
  uint32_t* S = stk[wid];
  int sp = 0;
  // the root is entered like any node: stack entries name the node whose
  // children are tested; a single static triangle is tested directly.
  const uint32_t ms = SL.cnt[0];
  bool overflow = false;
  const int lane = threadIdx.x % kWave;
  // leaves [c0, c1) (at most 8): one vector load, readlane broadcast
  auto test_leaves = [&](uint32_t c0, uint32_t c1) {
    const int nch = (int)(c1 - c0);
    Leaf mine;
    if (lane < nch) mine = sleaves[c0 + lane];
    for (int q = 0; q < nch; q++) {
      double lo[3], hi[3];
      for (int k = 0; k < 3; k++) {
        lo[k] = readlane_d(mine.lo[k], q) - pad;
        hi[k] = readlane_d(mine.hi[k], q) + pad;
      }
      const uint32_t st_id = (uint32_t)__builtin_amdgcn_readlane((int)mine.tri, q);
      ps.push(act && box_overlap6(Q, lo, hi), make_uint2(ti, st_id), &cs->spair_count, pairs, cap);
    }
  };
  if (SL.L == 0) {
    if (ms) test_leaves(0, 1);
    ps.flush(&cs->spair_count, pairs, cap);
    return;
  }
  S[sp++] = stk_enc(SL.L, 0);
  while (sp > 0) {
    const uint32_t e = __builtin_amdgcn_readfirstlane(S[--sp]);
    const int cl = (int)(e >> 28) - 1;
    const uint32_t k = e & 0x0fffffffu;
    const uint32_t c0 = k * kFan, c1 = min(c0 + kFan, SL.cnt[cl]);
    if (cl == 0) {
      test_leaves(c0, c1);
    } else {
      const int nch = (int)(c1 - c0);
      double mb[6];
      if (lane < nch) {
        const double* b = snodes + ((size_t)SL.off[cl] + c0 + lane) * 6;
        for (int k = 0; k < 6; k++) mb[k] = b[k];
      }
      for (int q = nch - 1; q >= 0; q--) {
        double lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
          lo[k] = readlane_d(mb[k], q) - pad;
          hi[k] = readlane_d(mb[3 + k], q) + pad;
        }
        const bool hit = act && box_overlap6(Q, lo, hi);
        if (__any(hit)) {
          if (sp < kTravStack) S[sp++] = stk_enc(cl, c0 + q);
          else overflow = true;
        }
      }
    }
  }
  ps.flush(&cs->spair_count, pairs, cap);
  if (overflow && (threadIdx.x % kWave) == 0) atomicOr(&cs->nonfinite, 2u);
}

// ------------------------------------------------------------- prims

// One candidate primitive (expandTriPair self-collision.ts:23-41 for cloth
// pairs; :105-114 for cloth/static pairs).
struct Prim {
  int kind, p, a, b, c, d;
  uint32_t r1, r2;  // VT: triangle id; EE: the two directed edge slots (tri * 3 + e)
};

__device__ __forceinline__ uint64_t rkey(const DGeom& G, uint64_t kind, uint64_t f1, uint64_t f2) {
  return (kind << (2 * G.B)) | (f1 << G.B) | f2;
}

// Sort key (candidate order) and lagged-friction key of a prim.
__device__ __forceinline__ uint64_t prim_key(const DGeom& G, const Prim& q) {
  if (q.kind == 0) return rkey(G, 0, (uint64_t)q.p, G.tri_rank[q.r1]);
  return rkey(G, 1, G.dedge_rank[q.r1], G.dedge_rank[q.r2]);
}
__device__ __forceinline__ uint64_t prim_nkey(const DGeom& G, const Prim& q) {
  if (q.kind == 0) return rkey(G, 0, (uint64_t)q.p, q.r1);
  const uint32_t u1 = G.uedge[q.r1], u2 = G.uedge[q.r2];
  return rkey(G, 1, u1 < u2 ? u1 : u2, u1 < u2 ? u2 : u1);
}

// ------------------------------------------------------------ pair kernels
//
// One thread per triangle pair: the pair's six corners are loaded once, at
// both configurations, and its 15 (cloth/cloth) or 12 (cloth/static) prims
// (the expansion of expandTriPair, self-collision.ts:23-41 and :105-114) are
// evaluated from registers, instead of one thread per prim gathering its four
// corners.  With kFilter the pairs are the step-segment superset and each is
// first tested with the broadphase's padded swept-box predicate (the same
// arithmetic as swept_box + box_overlap6), which selects exactly the
// candidate pairs of a traversal at these positions.

struct PairCorners {
  int v[6];               // slots 0-2: cloth triangle A, 3-5: triangle B (static corners as -id-2)
  double x0[6][3], x1[6][3];
};

__device__ __forceinline__ void load_pair(const DGeom& G, const double* __restrict__ xa, const double* __restrict__ xb,
                                          uint2 pr, bool stat, PairCorners& P) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const int v = (int)G.idx[(size_t)pr.x * 3 + c];
    P.v[c] = v;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      P.x0[c][k] = xa[(size_t)v * 3 + k];
      P.x1[c][k] = xb[(size_t)v * 3 + k];
    }
  }
  // triangle B: cloth corners from xa / xb, or static corners (fixed)
  const uint32_t* bidx = stat ? G.sidx : G.idx;
  const double* ba = stat ? G.spos : xa;
  const double* bb = stat ? G.spos : xb;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const uint32_t v = bidx[(size_t)pr.y * 3 + c];
    P.v[3 + c] = stat ? -(int)v - 2 : (int)v;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      P.x0[3 + c][k] = ba[(size_t)v * 3 + k];
      P.x1[3 + c][k] = bb[(size_t)v * 3 + k];
    }
  }
}

// swept_box over corner slots base..base+2 (same visiting order)
__device__ __forceinline__ void corner_box(const PairCorners& P, int base, double pad, double* lo, double* hi) {
  double mn[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
  double mx[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
#pragma unroll
  for (int c = 0; c < 3; c++) {
#pragma unroll
    for (int q = 0; q < 2; q++) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double pk = q == 0 ? P.x0[base + c][k] : P.x1[base + c][k];
        if (pk < mn[k]) mn[k] = pk;
        if (pk > mx[k]) mx[k] = pk;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = mn[k] - pad;
    hi[k] = mx[k] + pad;
  }
}

// Candidate pairs of one build from the superset of its segment: a trial
// or iterate x on the segment (x_a, x_b) has every swept box (x_step -> x)
// inside the segment box (x_step, x_a, x_b), so the superset holds every
// pair of the build, and the reference's own predicate (exact swept boxes,
// inclusive overlap; shared-vertex exclusion holds already) selects them.
__device__ __forceinline__ bool pair_keep(const PairCorners& P, bool stat, double pad, const double* __restrict__ sbox,
                                          uint32_t st) {
  double alo[3], ahi[3];
  corner_box(P, 0, pad, alo, ahi);
  const double A[6] = {alo[0], alo[1], alo[2], ahi[0], ahi[1], ahi[2]};
  if (!stat) {
    double blo[3], bhi[3];
    corner_box(P, 3, pad, blo, bhi);
    return box_overlap6(A, blo, bhi);
  }
  const double* b = sbox + (size_t)st * 6;
  const double slo[3] = {b[0] - pad, b[1] - pad, b[2] - pad};
  const double shi[3] = {b[3] + pad, b[4] + pad, b[5] + pad};
  return box_overlap6(A, slo, shi);
}

// Corner ids by slot from scalar copies: a select over values, never an
// index into the pair's arrays (which would move them to scratch memory).
struct Ids6 {
  int s0, s1, s2, s3, s4, s5;
  __device__ __forceinline__ int at(int s) const {
    return s == 0 ? s0 : s == 1 ? s1 : s == 2 ? s2 : s == 3 ? s3 : s == 4 ? s4 : s5;
  }
};

__device__ __forceinline__ Ids6 pair_ids(const PairCorners& P) {
  return Ids6{P.v[0], P.v[1], P.v[2], P.v[3], P.v[4], P.v[5]};
}

// The prim (kind, corner slots, reference ids) of prim k of a pair: cloth
// pairs k < 6 VT (A's corners against B, then B's against A), 6..14 EE;
// static pairs k < 3 VT (cloth corners against the static triangle), 3..11
// EE.  The set matches expandTriPair; the order inside a pair is free since
// every consumer reduces with min or sorts by key.
__device__ __forceinline__ void pair_prim(const DGeom& G, const Ids6 V, uint2 pr, bool stat, int k, Prim& q) {
  const int nvt = stat ? 3 : 6;
  if (k < nvt) {
    // VT: corner k against the other triangle (k < 3: A's corner vs B)
    const int tb = k < 3 ? 3 : 0;
    q.kind = 0; q.p = V.at(k); q.a = V.at(tb); q.b = V.at(tb + 1); q.c = V.at(tb + 2);
    q.d = stat ? -2 : -1;
    q.r1 = stat ? G.m + pr.y : (k < 3 ? pr.y : pr.x);
    q.r2 = 0;
  } else {
    // EE: edge ea of A against edge eb of B, edges (0,1), (1,2), (2,0)
    const int e = k - nvt, ea = e / 3, eb = e % 3;
    q.kind = 1; q.p = -1;
    q.a = V.at(ea); q.b = V.at(ea == 2 ? 0 : ea + 1);
    q.c = V.at(3 + eb); q.d = V.at(3 + (eb == 2 ? 0 : eb + 1));
    q.r1 = pr.x * 3 + ea;
    q.r2 = (stat ? G.m + pr.y : pr.y) * 3 + eb;
  }
}

// Corners of prim k of a pair copied out of the pair's registers.  k is the
// same for every lane of a wave (a loop counter), so the switch is a scalar
// branch and every register index is a constant.
template <int A0, int A1, int A2, int A3, bool kBoth>
__device__ __forceinline__ void take4(const PairCorners& P, Seg4& S) {
  constexpr int sl[4] = {A0, A1, A2, A3};
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      S.x1[q][k] = P.x1[sl[q]][k];
      if (kBoth) S.x0[q][k] = P.x0[sl[q]][k];
    }
}

template <bool kBoth>
__device__ __forceinline__ void prim_seg(const PairCorners& P, bool stat, int k, Seg4& S) {
  if (!stat) {
    switch (k) {
      case 0: take4<0, 3, 4, 5, kBoth>(P, S); break;
      case 1: take4<1, 3, 4, 5, kBoth>(P, S); break;
      case 2: take4<2, 3, 4, 5, kBoth>(P, S); break;
      case 3: take4<3, 0, 1, 2, kBoth>(P, S); break;
      case 4: take4<4, 0, 1, 2, kBoth>(P, S); break;
      case 5: take4<5, 0, 1, 2, kBoth>(P, S); break;
      case 6: take4<0, 1, 3, 4, kBoth>(P, S); break;
      case 7: take4<0, 1, 4, 5, kBoth>(P, S); break;
      case 8: take4<0, 1, 5, 3, kBoth>(P, S); break;
      case 9: take4<1, 2, 3, 4, kBoth>(P, S); break;
      case 10: take4<1, 2, 4, 5, kBoth>(P, S); break;
      case 11: take4<1, 2, 5, 3, kBoth>(P, S); break;
      case 12: take4<2, 0, 3, 4, kBoth>(P, S); break;
      case 13: take4<2, 0, 4, 5, kBoth>(P, S); break;
      default: take4<2, 0, 5, 3, kBoth>(P, S); break;
    }
  } else {
    switch (k) {
      case 0: take4<0, 3, 4, 5, kBoth>(P, S); break;
      case 1: take4<1, 3, 4, 5, kBoth>(P, S); break;
      case 2: take4<2, 3, 4, 5, kBoth>(P, S); break;
      case 3: take4<0, 1, 3, 4, kBoth>(P, S); break;
      case 4: take4<0, 1, 4, 5, kBoth>(P, S); break;
      case 5: take4<0, 1, 5, 3, kBoth>(P, S); break;
      case 6: take4<1, 2, 3, 4, kBoth>(P, S); break;
      case 7: take4<1, 2, 4, 5, kBoth>(P, S); break;
      case 8: take4<1, 2, 5, 3, kBoth>(P, S); break;
      case 9: take4<2, 0, 3, 4, kBoth>(P, S); break;
      case 10: take4<2, 0, 4, 5, kBoth>(P, S); break;
      default: take4<2, 0, 5, 3, kBoth>(P, S); break;
    }
  }
}

// Swept boxes of the two triangles (unpadded) further apart than the
// thickness bound of the prim prefilter, all corners finite: every prim of
// the pair passes that prefilter too, since a prim's two parts lie inside the
// two triangles' boxes and the rounded gap is monotone in the boxes.
__device__ __forceinline__ bool pair_separated(const PairCorners& P, double d_min) {
  double gap2 = 0;
  bool all_finite = true;
#pragma unroll
  for (int kk = 0; kk < 3; kk++) {
    double lo1 = __builtin_inf(), hi1 = -__builtin_inf(), lo2 = __builtin_inf(), hi2 = -__builtin_inf();
#pragma unroll
    for (int s = 0; s < 6; s++) {
      const double a0 = P.x0[s][kk], a1 = P.x1[s][kk];
      all_finite = all_finite && jsm::isfinite_(a0) && jsm::isfinite_(a1);
      const double l = fmin(a0, a1), h = fmax(a0, a1);
      if (s < 3) { lo1 = fmin(lo1, l); hi1 = fmax(hi1, h); }
      else { lo2 = fmin(lo2, l); hi2 = fmax(hi2, h); }
    }
    const double g = fmax(fmax(lo2 - hi1, lo1 - hi2), 0.0);
    gap2 += g * g;
  }
  const double thr = d_min * (1 + 1e-6);
  return all_finite && gap2 > thr * thr;
}

// Swept boxes of a prim's two parts (VT: the point / the triangle; EE: edge
// / edge) further apart than the thickness, with a 1e-6 relative margin, all
// corners finite: no configuration on the segment x0 -> x1 comes within the
// thickness, and the reference's CCD returns Infinity on every branch (not
// resting, no validated root, no sample within the thickness).
__device__ __forceinline__ bool prim_separated(const Seg4& S, int kind, double d_min) {
  const int split = kind == 0 ? 1 : 2;
  double gap2 = 0;
  bool all_finite = true;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double lo1 = __builtin_inf(), hi1 = -__builtin_inf(), lo2 = __builtin_inf(), hi2 = -__builtin_inf();
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const double a0 = S.x0[s][k], a1 = S.x1[s][k];
      all_finite = all_finite && jsm::isfinite_(a0) && jsm::isfinite_(a1);
      const double l = fmin(a0, a1), h = fmax(a0, a1);
      if (s < split) { lo1 = fmin(lo1, l); hi1 = fmax(hi1, h); }
      else { lo2 = fmin(lo2, l); hi2 = fmax(hi2, h); }
    }
    const double g = fmax(fmax(lo2 - hi1, lo1 - hi2), 0.0);
    gap2 += g * g;
  }
  const double thr = d_min * (1 + 1e-6);
  return all_finite && gap2 > thr * thr;
}

// A CCD result of checkTrial (contact-assembly.ts:515-533): NaN marks a CCD
// failure, a TOI in (0, 1 - 1e-9) a crossing; the smallest key of each kind
// decides which came first in candidate order.
__device__ __forceinline__ void trial_toi(const DGeom& G, const Prim& q, double toi, double& minT, CScalars* cs) {
  if (jsm::isnan_(toi)) {
    atomicMin((unsigned long long*)&cs->nan_key, (unsigned long long)prim_key(G, q));
  } else {
    if (toi < minT) minT = toi;
    if (toi > 0 && toi < 1 - 1e-9) atomicMin((unsigned long long*)&cs->cross_key, (unsigned long long)prim_key(G, q));
  }
}

// updateActiveSet (contact-assembly.ts:164-199) per pair: closest point per
// prim at x, active when 1e-12 <= d < dHat (frozen record, sort key and
// lagged-friction key emitted wave-aggregated); min distance over all
// candidates.  With kFilter the pair lists are the superset, filtered with
// the padded swept boxes from the step start to x.
template <bool kFilter>
__global__ void __launch_bounds__(kBlock) k_active_pairs(DGeom G, const double* __restrict__ xs,
                                                         const double* __restrict__ x, double pad,
                                                         const uint2* __restrict__ pairs, uint32_t np,
                                                         const uint2* __restrict__ spairs, uint32_t nsp,
                                                         const double* __restrict__ sbox, Frozen* rec, uint64_t* keys,
                                                         uint64_t* nkeys, uint32_t cap, CScalars* cs) {
  const uint32_t total = np + nsp;
  double minD = __builtin_inf();
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    const bool stat = i >= np;
    bool keep = false;
    PairCorners P;
    uint2 pr = make_uint2(0, 0);
    if (i < total) {
      pr = stat ? spairs[i - np] : pairs[i];
      load_pair(G, xs, x, pr, stat, P);
      keep = kFilter ? pair_keep(P, stat, pad, sbox, pr.y) : true;
    }
    if (kFilter) {
      const uint64_t mc = __ballot(keep && !stat), ms = __ballot(keep && stat);
      if ((threadIdx.x % kWave) == 0) {
        if (mc) atomicAdd(&cs->pair_count, (uint32_t)__popcll(mc));
        if (ms) atomicAdd(&cs->spair_count, (uint32_t)__popcll(ms));
      }
    }
    const int nk = stat ? 12 : 15;
    // uniform trip count so that the wave-aggregated emission sees every lane
    for (int k = 0; k < 15; k++) {
      bool emit = false;
      Prim q;
      Frozen f;
      if (keep && k < nk) {
        pair_prim(G, pair_ids(P), pr, stat, k, q);
        Seg4 S;
        prim_seg<false>(P, stat, k, S);
        Closest cp = q.kind == 0 ? closest_vt<false>(S.x1[0], S.x1[1], S.x1[2], S.x1[3])
                                 : closest_ee<false>(S.x1[0], S.x1[1], S.x1[2], S.x1[3]);
        // the exact distance matters only below dHat or below this thread's
        // minimum; |r|^2 above both (with a 1e-9 margin) rules both out
        const double r2 = norm2(cp);
        const double lim = minD < G.d_hat ? G.d_hat : minD;
        const bool far = G.d_hat >= 1e-100 && lim <= 1e100 && r2 <= 1e200 && r2 > lim * lim * (1 + 1e-9);
        cp.dist = far ? __builtin_inf() : jsm::hypot3(cp.rx, cp.ry, cp.rz);
        if (cp.dist < minD) minD = cp.dist;
        if (!(cp.dist >= G.d_hat || cp.dist < 1e-12)) {
          emit = true;
          // frozen VT records carry d = -1 even for static candidates (contact-assembly.ts:181)
          f.kind = q.kind; f.p = q.p; f.a = q.a; f.b = q.b; f.c = q.c; f.d = q.kind == 0 ? -1 : q.d;
          f.nx = cp.rx / cp.dist; f.ny = cp.ry / cp.dist; f.nz = cp.rz / cp.dist;
          if (q.kind == 0) {
            f.w1 = cp.s; f.w2 = cp.t; f.w0 = 1 - cp.s - cp.t; f.w3 = 0;
          } else {
            f.w0 = 1 - cp.s; f.w1 = cp.s; f.w2 = -(1 - cp.t); f.w3 = -cp.t;
          }
        }
      }
      const uint32_t slot = wave_alloc(emit, &cs->active_count);
      if (emit && slot < cap) {
        rec[slot] = f;
        keys[slot] = prim_key(G, q);
        nkeys[slot] = prim_nkey(G, q);
      }
    }
  }
  block_min_to(minD, &cs->mink[0]);
}

// Floor actives (contact-assembly.ts:202-213), appended after the candidates.
__global__ void __launch_bounds__(kBlock) k_floor_active(DGeom G, const double* __restrict__ x, Frozen* rec,
                                                         uint64_t* keys, uint64_t* nkeys, uint32_t cap,
                                                         double* partials, CScalars* cs) {
  double minD = __builtin_inf();
  for (uint32_t base = blockIdx.x * blockDim.x; base < G.n; base += gridDim.x * blockDim.x) {
    const uint32_t v = base + threadIdx.x;
    bool emit = false;
    if (v < G.n) {
      const double d = x[(size_t)v * 3 + 1] - G.floor_y;
      if (d < minD) minD = d;
      emit = !(d >= G.d_hat);
    }
    const uint32_t k = wave_alloc(emit, &cs->active_count);
    if (emit && k < cap) {
      Frozen f;
      f.kind = 2; f.p = (int)v; f.a = f.b = f.c = f.d = -1;
      f.w0 = f.w1 = f.w2 = f.w3 = 0;
      f.nx = 0; f.ny = 1; f.nz = 0;
      rec[k] = f;
      keys[k] = rkey(G, 2, 0, v);
      nkeys[k] = rkey(G, 2, 0, v);
    }
  }
  block_min_to(minD, &cs->mink[1]);
}

__global__ void __launch_bounds__(kBlock) k_count_kinds(const Frozen* __restrict__ act, uint32_t C, CScalars* cs) {
  __shared__ uint32_t s_cnt[4];
  if (threadIdx.x < 4) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < C; i += gridDim.x * blockDim.x)
    atomicAdd(&s_cnt[act[i].kind & 3], 1u);
  __syncthreads();
  if (threadIdx.x < 4 && s_cnt[threadIdx.x]) atomicAdd(&cs->kind_count[threadIdx.x], s_cnt[threadIdx.x]);
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
__device__ __forceinline__ void emit_entry(const Frozen* __restrict__ act, uint32_t c, int cshift, uint64_t* keys,
                                           uint32_t* vals, uint32_t* s_valid) {
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
  atomicAdd(s_valid, valid);
}

__global__ void __launch_bounds__(kBlock) k_emit_entries(const Frozen* __restrict__ act, uint32_t C, int cshift,
                                                         uint64_t* keys, uint32_t* vals, CScalars* cs) {
  __shared__ uint32_t s_valid;
  if (threadIdx.x == 0) s_valid = 0;
  __syncthreads();
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C) emit_entry(act, c, cshift, keys, vals, &s_valid);
  __syncthreads();
  if (threadIdx.x == 0 && s_valid) atomicAdd(&cs->entry_count, s_valid);
}

// v_ent[i] = low bits of the sorted incidence keys; v_off[u] = first entry
// of vertex u (lower bound of u in the sorted vertex field), u in [0, n].
__global__ void k_entry_ent(const uint64_t* __restrict__ keys, uint32_t cnt, int cshift, uint32_t* v_ent,
                            uint32_t* ent_vtx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  v_ent[i] = (uint32_t)(keys[i] & ((1ull << cshift) - 1));
  ent_vtx[i] = (uint32_t)(keys[i] >> cshift);
}

// Vertices that own incidence entries (order irrelevant).
__global__ void __launch_bounds__(kBlock) k_contact_verts(const uint64_t* __restrict__ keys, uint32_t cnt, int cshift,
                                                          uint32_t* cverts, CScalars* cs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool head = false;
  uint32_t u = 0;
  if (i < cnt) {
    u = (uint32_t)(keys[i] >> cshift);
    head = i == 0 || (uint32_t)(keys[i - 1] >> cshift) != u;
  }
  const uint32_t k = wave_alloc(head, &cs->nverts);
  if (head) cverts[k] = u;
}

__global__ void k_entry_offsets(const uint64_t* __restrict__ keys, uint32_t cnt, int cshift, uint32_t n,
                                uint32_t* v_off) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u > n) return;
  uint32_t lo = 0, hi = cnt;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if ((keys[mid] >> cshift) < (uint64_t)u) lo = mid + 1;
    else hi = mid;
  }
  v_off[u] = lo;
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
  block_partial_dd<kBlock>(energy, partials);
  block_partial_dd<kBlock>(fwork, partials + kMaxRedBlocks);
}

// Contact gradient per vertex: addDistGrad then addFrictionForce for each
// active contact in list order (contact-assembly.ts:272-279, 288-341), in
// two passes: every incidence entry's two additions (distance-gradient term,
// then friction term) in parallel, then one thread per contact vertex and
// component adds them in list order.  Components an entry does not touch
// carry +0, which leaves a sum that started at +0 unchanged.
__global__ void __launch_bounds__(kBlock) k_grad_terms(uint32_t E, const Frozen* __restrict__ act,
                                                       const uint32_t* __restrict__ v_ent, CEval ev,
                                                       double* __restrict__ gterm) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= E) return;
  double ta[3] = {0, 0, 0}, tb[3] = {0, 0, 0};
  const uint32_t ent = v_ent[k];
  const uint32_t ci = ent >> 3;
  const int slot = ent & 7;
  if (ev.inr[ci]) {
    const Frozen c = act[ci];
    const double g = ev.g[ci];
    const double n3[3] = {c.nx, c.ny, c.nz};
    const double fp[3] = {-ev.fr[(size_t)ci * 3], -ev.fr[(size_t)ci * 3 + 1], -ev.fr[(size_t)ci * 3 + 2]};
    if (c.kind == 2) {
      ta[1] = g;
      for (int q = 0; q < 3; q++) tb[q] = fp[q];
    } else if (c.kind == 0) {
      if (slot == 0) {
        for (int q = 0; q < 3; q++) ta[q] = g * n3[q];
        for (int q = 0; q < 3; q++) tb[q] = fp[q];
      } else {
        const double w = slot == 1 ? c.w0 : slot == 2 ? c.w1 : c.w2;
        for (int q = 0; q < 3; q++) ta[q] = -g * w * n3[q];
        for (int q = 0; q < 3; q++) tb[q] = -fp[q] * w;
      }
    } else {
      const double w = slot == 0 ? c.w0 : slot == 1 ? c.w1 : slot == 2 ? c.w2 : c.w3;
      const double ww = slot == 0 ? c.w0 : slot == 1 ? c.w1 : slot == 2 ? -c.w2 : -c.w3;
      for (int q = 0; q < 3; q++) ta[q] = g * w * n3[q];
      for (int q = 0; q < 3; q++) tb[q] = fp[q] * ww;
    }
  }
  for (int q = 0; q < 3; q++) {
    gterm[(size_t)k * 6 + q] = ta[q];
    gterm[(size_t)k * 6 + 3 + q] = tb[q];
  }
}

__global__ void __launch_bounds__(kBlock) k_grad_vsum(uint32_t nv, const uint32_t* __restrict__ cverts,
                                                      const uint32_t* __restrict__ v_off,
                                                      const double* __restrict__ gterm, double* __restrict__ ceg) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nv * 3) return;
  const uint32_t v = cverts[t / 3];
  const int q = t % 3;
  ceg[(size_t)v * 3 + q] = ordered_sum<6, 3>(gterm + q, v_off[v], v_off[v + 1], 0.0);
}

// Adjoint of the lagged friction with respect to the step's start positions.
// The friction term -f(u) that energyGrad adds (contact-assembly.ts:272-279,
// distributed by addFrictionForce :318-341 with weights wW) depends on the
// start positions x_step only through the slip u = sum_k wS_k (x - x_step)_k
// (relativeSlip :344-367: the VT slip leaves out the triangle's third
// vertex, the EE slip adds all four), the lagged normal and normal force
// held.  With f = -c t / |t~| (coulombForce, friction.ts:14-23), t = P u,
// P = I - n n^T, |t~| = sqrt(|t|^2 + eps^2), c = mu max(0, lambdaN):
//   -df/du = K = c (P / |t~| - t t^T / |t~|^3)   (symmetric),
// so for the adjoint lambda of the step's root
//   dL/dx_step[v] += wS_v K (sum_k wW_k lambda_k)      (vertex v of the stencil).
// Pass 1 forms y = K (W^T lambda) per contact inside dHat; pass 2 the
// per-entry terms wS y; pass 3 adds them per vertex in list order.
__global__ void __launch_bounds__(kBlock) k_fric_adj_contact(DGeom G, const double* __restrict__ x,
                                                             const double* __restrict__ xstep,
                                                             const Frozen* __restrict__ act,
                                                             const uint64_t* __restrict__ anke, uint32_t C,
                                                             const uint64_t* __restrict__ lag_keys,
                                                             const uint32_t* __restrict__ lag_idx, uint32_t L,
                                                             const double* __restrict__ lag_vals,
                                                             const double* __restrict__ lam, double* __restrict__ y) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  const Frozen c = act[i];
  double out[3] = {0, 0, 0};
  const double d = jsm::jmax(frozen_dist(G, x, c), 1e-12);
  if (d < G.d_hat) {
    double nx = c.nx, ny = c.ny, nz = c.nz, lamN;
    const int64_t li = lag_find(lag_keys, L, anke[i]);
    if (li >= 0) {
      const double* lv = lag_vals + (size_t)lag_idx[li] * 4;
      nx = lv[0]; ny = lv[1]; nz = lv[2]; lamN = lv[3];
    } else {
      lamN = barrier_normal_force(d, G.d_hat, G.kappa, 1);
    }
    const double cc = G.mu * jsm::jmax(0, lamN);
    if (cc != 0) {
      double u[3];
      relative_slip(x, xstep, c, u);
      const double vn = u[0] * nx + u[1] * ny + u[2] * nz;
      const double t[3] = {u[0] - vn * nx, u[1] - vn * ny, u[2] - vn * nz};
      const double mag = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2] + G.feps * G.feps);
      // w = W^T lambda with addFrictionForce's weights (static vertices skipped)
      double w[3] = {0, 0, 0};
      auto acc = [&](int v, double wt) {
        if (v < 0) return;
        for (int q = 0; q < 3; q++) w[q] += wt * lam[(size_t)v * 3 + q];
      };
      if (c.kind == 2) acc(c.p, 1.0);
      else if (c.kind == 0) {
        acc(c.p, 1.0); acc(c.a, -c.w0); acc(c.b, -c.w1); acc(c.c, -c.w2);
      } else {
        acc(c.a, c.w0); acc(c.b, c.w1); acc(c.c, -c.w2); acc(c.d, -c.w3);
      }
      const double wn = w[0] * nx + w[1] * ny + w[2] * nz;
      const double pw[3] = {w[0] - wn * nx, w[1] - wn * ny, w[2] - wn * nz};
      const double tw = t[0] * w[0] + t[1] * w[1] + t[2] * w[2];
      for (int q = 0; q < 3; q++) out[q] = cc * (pw[q] / mag - t[q] * tw / (mag * mag * mag));
    }
  }
  for (int q = 0; q < 3; q++) y[(size_t)i * 3 + q] = out[q];
}

__global__ void __launch_bounds__(kBlock) k_fric_adj_terms(uint32_t E, const Frozen* __restrict__ act,
                                                           const uint32_t* __restrict__ v_ent,
                                                           const double* __restrict__ y, double* __restrict__ term) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= E) return;
  const uint32_t ent = v_ent[k];
  const uint32_t ci = ent >> 3;
  const int slot = ent & 7;
  const Frozen c = act[ci];
  // relativeSlip's weight of this stencil slot
  double ws;
  if (c.kind == 2) ws = 1.0;
  else if (c.kind == 0) ws = slot == 0 ? 1.0 : slot == 1 ? -c.w0 : slot == 2 ? -c.w1 : 0.0;
  else ws = slot == 0 ? c.w0 : slot == 1 ? c.w1 : slot == 2 ? c.w2 : c.w3;
  for (int q = 0; q < 3; q++) term[(size_t)k * 3 + q] = ws * y[(size_t)ci * 3 + q];
}

__global__ void __launch_bounds__(kBlock) k_fric_adj_vsum(uint32_t nv, const uint32_t* __restrict__ cverts,
                                                          const uint32_t* __restrict__ v_off,
                                                          const double* __restrict__ term, double* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nv * 3) return;
  const uint32_t v = cverts[t / 3];
  const int q = t % 3;
  out[(size_t)v * 3 + q] += ordered_sum<3, 0>(term + q, v_off[v], v_off[v + 1], 0.0);
}

// addDiagEstimate (contact-assembly.ts:436-471): accumulate into diag, the
// same two passes (a floor contact touches only the y component; +0 added to
// a positive diagonal leaves it unchanged).
__global__ void __launch_bounds__(kBlock) k_diag_terms(uint32_t E, const Frozen* __restrict__ act,
                                                       const uint32_t* __restrict__ v_ent, CEval ev,
                                                       double* __restrict__ dterm) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= E) return;
  double t[3] = {0, 0, 0};
  const uint32_t ent = v_ent[k];
  const uint32_t ci = ent >> 3;
  const int slot = ent & 7;
  if (ev.inr[ci]) {
    const Frozen c = act[ci];
    const double kk = ev.coeff[ci];
    if (c.kind == 2) {
      t[1] = kk;
    } else {
      const double n2[3] = {c.nx * c.nx, c.ny * c.ny, c.nz * c.nz};
      if (c.kind == 0 && slot == 0) {
        for (int q = 0; q < 3; q++) t[q] = kk * n2[q];
      } else {
        const double w = c.kind == 0 ? (slot == 1 ? c.w0 : slot == 2 ? c.w1 : c.w2)
                                     : (slot == 0 ? c.w0 : slot == 1 ? c.w1 : slot == 2 ? c.w2 : c.w3);
        for (int q = 0; q < 3; q++) t[q] = kk * w * w * n2[q];
      }
    }
  }
  for (int q = 0; q < 3; q++) dterm[(size_t)k * 3 + q] = t[q];
}

__global__ void __launch_bounds__(kBlock) k_diag_vsum(uint32_t nv, const uint32_t* __restrict__ cverts,
                                                      const uint32_t* __restrict__ v_off,
                                                      const double* __restrict__ dterm, double* __restrict__ diag) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nv * 3) return;
  const uint32_t v = cverts[t / 3];
  const int q = t % 3;
  const size_t o = (size_t)v * 3 + q;
  diag[o] = ordered_sum<3, 0>(dterm + q, v_off[v], v_off[v + 1], diag[o]);
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
  block_partial_dd<kBlock>(e, partials);
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
// list once in order.  A contact depends only on the earlier contacts that
// share one of its cloth vertices, so the sweep is a DAG: each contact's
// successors are the next contacts of its vertices' incidence lists (which
// are in list order) and it may run once all its predecessors have.  Ready
// contacts share no vertex (of two contacts on one vertex the later waits for
// the earlier), so any set of ready contacts can run in parallel and the
// result is the sequential sweep's, bit for bit.
//
// The sweep is latency-bound (its critical path is a large part of C), so it
// runs in one work-group with the touched vertices' velocities and the
// predecessor counters in LDS: a thread that finishes a contact decrements
// its successors' counters and goes on at once with one that became ready
// (the others start in the next round).

constexpr uint32_t kNoSucc = 0xffffffffu;

// One contact of the sweep, in list order: normal, barycentric weights, the
// cloth vertices (global id and LDS slot; -1 / 0xffff for none or static),
// and the next contact of each vertex.
struct KillRec {
  double nx, ny, nz, w0, w1, w2;
  int32_t vid[4];
  uint16_t slot[4];
  uint32_t succ[4];
  uint32_t kind;
  uint32_t indeg;  // contacts that precede this one on its vertices
};

__global__ void k_kill_slots(const uint32_t* __restrict__ cverts, uint32_t nv, uint32_t* vslot) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nv) vslot[cverts[k]] = k;
}

__global__ void k_kill_rec(const Frozen* __restrict__ act, uint32_t C, const uint32_t* __restrict__ vslot,
                           KillRec* rec) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  const Frozen c = act[i];
  KillRec r;
  r.nx = c.nx; r.ny = c.ny; r.nz = c.nz;
  r.w0 = c.w0; r.w1 = c.w1; r.w2 = c.w2;
  const int vs[4] = {c.kind == 2 ? c.p : c.kind == 0 ? c.p : c.a, c.kind == 2 ? -1 : c.kind == 0 ? c.a : c.b,
                     c.kind == 2 ? -1 : c.kind == 0 ? c.b : c.c, c.kind == 2 ? -1 : c.kind == 0 ? c.c : c.d};
  for (int s = 0; s < 4; s++) {
    r.vid[s] = vs[s] >= 0 ? vs[s] : -1;
    r.slot[s] = vs[s] >= 0 ? (uint16_t)vslot[vs[s]] : (uint16_t)0xffff;
    r.succ[s] = kNoSucc;
  }
  r.kind = (uint32_t)c.kind;
  r.indeg = 0;
  rec[i] = r;
}

// Per incidence entry (vertex u, contact i, stencil slot s): the successor of
// i on u and one predecessor count if i is not u's first contact.
__global__ void k_kill_links(const uint32_t* __restrict__ v_off, const uint32_t* __restrict__ v_ent,
                             const uint32_t* __restrict__ ent_vtx, uint32_t E, KillRec* rec) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= E) return;
  const uint32_t u = ent_vtx[h];
  const uint32_t e = v_ent[h];
  const uint32_t i = e >> 3, s = e & 7;
  if (h + 1 < v_off[u + 1]) rec[i].succ[s] = v_ent[h + 1] >> 3;
  if (h > v_off[u]) atomicAdd(&rec[i].indeg, 1u);
}

// The reference's per-contact update (newton.ts:245-301) on velocities reached
// through V(vid, slot), in its operation order.
template <class VelAt>
__device__ __forceinline__ void kill_contact(const KillRec& c, VelAt V) {
  if (c.kind == 2) {
    double* vp = V(c.vid[0], c.slot[0]);
    if (vp[1] < 0) vp[1] = 0;
  } else if (c.kind == 0) {
    double* vp = V(c.vid[0], c.slot[0]);
    const double p0 = vp[0], p1 = vp[1], p2 = vp[2];
    double tx = 0, ty = 0, tz = 0;
    const double ws[3] = {c.w0, c.w1, c.w2};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      if (c.vid[k + 1] < 0) continue;
      const double* q = V(c.vid[k + 1], c.slot[k + 1]);
      tx += ws[k] * q[0];
      ty += ws[k] * q[1];
      tz += ws[k] * q[2];
    }
    const double rel = (p0 - tx) * c.nx + (p1 - ty) * c.ny + (p2 - tz) * c.nz;
    if (rel < 0) {
      const double j = -0.5 * rel;
      vp[0] += j * c.nx;
      vp[1] += j * c.ny;
      vp[2] += j * c.nz;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (c.vid[k + 1] < 0) continue;
        double* q = V(c.vid[k + 1], c.slot[k + 1]);
        q[0] -= j * ws[k] * c.nx;
        q[1] -= j * ws[k] * c.ny;
        q[2] -= j * ws[k] * c.nz;
      }
    }
  } else {
    double A[3] = {0, 0, 0}, B[3] = {0, 0, 0}, Cc[3] = {0, 0, 0}, D[3] = {0, 0, 0};
    double* pa = c.vid[0] >= 0 ? V(c.vid[0], c.slot[0]) : nullptr;
    double* pb = c.vid[1] >= 0 ? V(c.vid[1], c.slot[1]) : nullptr;
    double* pc = c.vid[2] >= 0 ? V(c.vid[2], c.slot[2]) : nullptr;
    double* pd = c.vid[3] >= 0 ? V(c.vid[3], c.slot[3]) : nullptr;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      if (pa) A[q] = pa[q];
      if (pb) B[q] = pb[q];
      if (pc) Cc[q] = pc[q];
      if (pd) D[q] = pd[q];
    }
    double ab[3], cd[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      ab[q] = (A[q] + B[q]) / 2;
      cd[q] = (Cc[q] + D[q]) / 2;
    }
    const double rel = (ab[0] - cd[0]) * c.nx + (ab[1] - cd[1]) * c.ny + (ab[2] - cd[2]) * c.nz;
    if (rel < 0) {
      const double j = -0.25 * rel;
      double* plus[2] = {pa, pb};
      double* minus[2] = {pc, pd};
#pragma unroll
      for (int k = 0; k < 2; k++) {
        if (!plus[k]) continue;
        plus[k][0] += j * c.nx;
        plus[k][1] += j * c.ny;
        plus[k][2] += j * c.nz;
      }
#pragma unroll
      for (int k = 0; k < 2; k++) {
        if (!minus[k]) continue;
        minus[k][0] -= j * c.nx;
        minus[k][1] -= j * c.ny;
        minus[k][2] -= j * c.nz;
      }
    }
  }
}

constexpr int kKillBlock = 256;

// The sweep.  kLds: velocities of the nv touched vertices (slot order) and
// byte-packed predecessor counters in LDS, sized by the host; otherwise the
// velocities are updated in place in global memory and the counters live in
// `gcnt`.  listA / listB (LDS, nv entries each: a frontier is vertex-disjoint)
// hold the round's ready contacts.
template <bool kLds>
__global__ void __launch_bounds__(kKillBlock) k_kill_run(const KillRec* __restrict__ rec, uint32_t C,
                                                         const uint32_t* __restrict__ cverts, uint32_t nv, double* v,
                                                         uint32_t* gcnt, uint32_t* glist, uint32_t* rounds_out) {
  extern __shared__ __align__(16) unsigned char kill_lds[];
  __shared__ uint32_t nA, nB;
  double* sv = reinterpret_cast<double*>(kill_lds);                              // kLds: nv * 3
  uint32_t* cnt = kLds ? reinterpret_cast<uint32_t*>(sv + (size_t)nv * 3) : gcnt;  // kLds: ceil(C / 4) words
  uint32_t* listA = kLds ? cnt + (C + 3) / 4 : glist;
  uint32_t* listB = listA + nv;
  auto vel = [&](int vid, uint32_t slot) -> double* {
    return kLds ? sv + (size_t)slot * 3 : v + (size_t)vid * 3;
  };
  // decrement contact j's counter; true when it reached zero.  The LDS sweep
  // orders its velocity writes before the decrements and the reads of a
  // released contact after them with LDS-only fences (lgkmcnt), so that the
  // prefetched record load stays in flight; the global sweep needs full
  // work-group acquire/release.
  auto release = [&](uint32_t j) -> bool {
    if (kLds) {
      const uint32_t sh = 8 * (j & 3);
      const uint32_t old = __hip_atomic_fetch_sub(&cnt[j >> 2], 1u << sh, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP);
      return ((old >> sh) & 0xff) == 1;
    }
    return __hip_atomic_fetch_sub(&cnt[j], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 1;
  };
  if (threadIdx.x == 0) nA = 0;
  {
    // pull the records into this XCD's L2 (one dword per 128-byte line):
    // the chains below then wait on L2 hits instead of memory
    const uint32_t* words = reinterpret_cast<const uint32_t*>(rec);
    const size_t lines = ((size_t)C * sizeof(KillRec) + 127) / 128;
    uint32_t acc = 0;
    for (size_t l = threadIdx.x; l < lines; l += blockDim.x) acc ^= words[l * 32];
    asm volatile("" ::"v"(acc));
  }
  if (kLds) {
    for (uint32_t k = threadIdx.x; k < nv; k += blockDim.x) {
      const size_t u = cverts[k];
      sv[(size_t)k * 3] = v[u * 3];
      sv[(size_t)k * 3 + 1] = v[u * 3 + 1];
      sv[(size_t)k * 3 + 2] = v[u * 3 + 2];
    }
    for (uint32_t w = threadIdx.x; w < (C + 3) / 4; w += blockDim.x) {
      uint32_t word = 0;
      for (uint32_t b = 0; b < 4; b++)
        if (w * 4 + b < C) word |= (rec[w * 4 + b].indeg & 0xff) << (8 * b);
      cnt[w] = word;
    }
  } else {
    for (uint32_t i = threadIdx.x; i < C; i += blockDim.x) cnt[i] = rec[i].indeg;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < C; i += blockDim.x)
    if (rec[i].indeg == 0) listA[atomicAdd(&nA, 1u)] = i;
  __syncthreads();
  uint32_t round = 0;
  while (nA > 0) {
    round++;
    if (threadIdx.x == 0) nB = 0;
    __syncthreads();
    const uint32_t na = nA;
    for (uint32_t k = threadIdx.x; k < na; k += blockDim.x) {
      const uint32_t first = listA[k];
      KillRec c = rec[first];
      // guard: a chain never revisits a contact, so it ends within C steps
      for (uint32_t guard = 0; guard < C; guard++) {
        // the chain goes on with one of the successors: fetch them all
        // while this contact runs (an absent one reads a valid record)
        const uint32_t self = first;
        KillRec pre[4];
#pragma unroll
        for (int s = 0; s < 4; s++) pre[s] = rec[c.succ[s] != kNoSucc ? c.succ[s] : self];
        kill_contact(c, vel);
        if (kLds) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        int take = -1;
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const uint32_t j = c.succ[s];
          if (j == kNoSucc) continue;
          // a contact that follows this one on two of its vertices is
          // released twice
          if (release(j)) {
            if (take < 0) take = s;
            else listB[atomicAdd(&nB, 1u)] = j;
          }
        }
        if (take < 0) break;
        if (kLds) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        if (take == 0) c = pre[0];
        else if (take == 1) c = pre[1];
        else if (take == 2) c = pre[2];
        else c = pre[3];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) nA = nB;
    __syncthreads();
    const uint32_t nb = nA;
    for (uint32_t k = threadIdx.x; k < nb; k += blockDim.x) listA[k] = listB[k];
    __syncthreads();
  }
  if (kLds) {
    for (uint32_t k = threadIdx.x; k < nv; k += blockDim.x) {
      const size_t u = cverts[k];
      v[u * 3] = sv[(size_t)k * 3];
      v[u * 3 + 1] = sv[(size_t)k * 3 + 1];
      v[u * 3 + 2] = sv[(size_t)k * 3 + 2];
    }
  }
  if (threadIdx.x == 0) *rounds_out = round;
}

// checkTrial prims (contact-assembly.ts:497-535) per pair; pairs [0, np)
// are cloth/cloth, then nsp cloth/static.  Per prim: exact distance at the
// trial, the swept-box prefilter, then the head of the reference's CCD; the
// few prims that need the cubic and sampling tail are queued for
// k_trial_ccd.  With kFilter the lists are the superset and the kept pair
// counts go to cs->pair_count / spair_count.
template <bool kFilter>
__global__ void __launch_bounds__(kBlock) k_trial_pairs(DGeom G, const double* __restrict__ xs,
                                                        const double* __restrict__ xt, double pad,
                                                        const uint2* __restrict__ pairs, uint32_t np,
                                                        const uint2* __restrict__ spairs, uint32_t nsp,
                                                        const double* __restrict__ sbox, uint2* tail, CScalars* cs) {
  const uint32_t total = np + nsp;
  double minD = __builtin_inf(), minT = __builtin_inf();
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    const bool stat = i >= np;
    bool keep = false, sep = true;
    PairCorners P;
    uint2 pr = make_uint2(0, 0);
    if (i < total) {
      pr = stat ? spairs[i - np] : pairs[i];
      load_pair(G, xs, xt, pr, stat, P);
      keep = kFilter ? pair_keep(P, stat, pad, sbox, pr.y) : true;
      sep = pair_separated(P, G.d_min);
    }
    if (kFilter) {
      const uint64_t mc = __ballot(keep && !stat), ms = __ballot(keep && stat);
      if ((threadIdx.x % kWave) == 0) {
        if (mc) atomicAdd(&cs->pair_count, (uint32_t)__popcll(mc));
        if (ms) atomicAdd(&cs->spair_count, (uint32_t)__popcll(ms));
      }
    }
    const int nk = stat ? 12 : 15;
    // uniform trip count: the switch in prim_seg and the tail queue's
    // wave-aggregated allocation see every lane
    for (int k = 0; k < 15; k++) {
      bool queue = false;
      if (keep && k < nk) {
        Seg4 S;
        prim_seg<true>(P, stat, k, S);
        const int kind = k < (stat ? 3 : 6) ? 0 : 1;
        const Closest c1 = kind == 0 ? closest_vt<false>(S.x1[0], S.x1[1], S.x1[2], S.x1[3])
                                     : closest_ee<false>(S.x1[0], S.x1[1], S.x1[2], S.x1[3]);
        // the distance only matters when it can lower this thread's minimum
        // (or is not finite): |r|^2 above minD^2 (1 + 1e-9) rules that out
        const double r2 = norm2(c1);
        if (!(minD >= 1e-100 && minD <= 1e100 && r2 <= 1e200 && r2 > minD * minD * (1 + 1e-9))) {
          const double d = jsm::hypot3(c1.rx, c1.ry, c1.rz);
          if (d < minD) minD = d;
          if (!jsm::isfinite_(d)) atomicOr(&cs->nonfinite, 1u);
        }
        if (!sep && !prim_separated(S, kind, G.d_min)) {
          double toi;
          if (ccd_head_fast(S, kind, G.d_min, toi)) {
            if (toi < minT) minT = toi;  // 0 or Infinity: no crossing, no failure
          } else {
            queue = true;
          }
        }
      }
      const uint32_t slot = wave_alloc(queue, &cs->tail_count);
      if (queue) tail[slot] = make_uint2(i, (uint32_t)k);
    }
  }
  block_min_to(minD, &cs->mink[5]);
  block_min_to(minT, &cs->mink[6]);
}

// The CCD tail (cubic roots, sampled fallback) of the prims k_trial_pairs
// queued; the queue length is read on the device.
__global__ void __launch_bounds__(kBlock) k_trial_ccd(DGeom G, const double* __restrict__ xs,
                                                      const double* __restrict__ xt,
                                                      const uint2* __restrict__ pairs, uint32_t np,
                                                      const uint2* __restrict__ spairs,
                                                      const uint2* __restrict__ tail, CScalars* cs) {
  const uint32_t cnt = cs->tail_count;
  double minT = __builtin_inf();
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x) {
    const uint2 it = tail[j];
    const bool stat = it.x >= np;
    const uint2 pr = stat ? spairs[it.x - np] : pairs[it.x];
    PairCorners P;
    for (int c = 0; c < 3; c++) {
      P.v[c] = (int)G.idx[(size_t)pr.x * 3 + c];
      P.v[3 + c] = stat ? -(int)G.sidx[(size_t)pr.y * 3 + c] - 2 : (int)G.idx[(size_t)pr.y * 3 + c];
    }
    Prim q;
    pair_prim(G, pair_ids(P), pr, stat, (int)it.y, q);
    const int ids[4] = {q.kind == 0 ? q.p : q.a, q.kind == 0 ? q.a : q.b, q.kind == 0 ? q.b : q.c,
                        q.kind == 0 ? q.c : q.d};
    Seg4 S;
    for (int s = 0; s < 4; s++) {
      vpos(G, xs, ids[s], S.x0[s]);
      vpos(G, xt, ids[s], S.x1[s]);
    }
    const double toi = q.kind == 0 ? vt_ccd(S, G.d_min) : ee_ccd(S, G.d_min);
    trial_toi(G, q, toi, minT, cs);
  }
  block_min_to(minT, &cs->mink[6]);
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
  block_min_to(minD, &cs->mink[7]);
  block_min_to(minT, &cs->mink[8]);
}

// ====================================================================== host

static int bits_for(uint64_t v) {
  int b = 1;
  while (b < 63 && (1ull << b) <= v) b++;
  return b;
}

// Dynamic LDS available to the LDS velocity-filter sweep: the device's
// opt-in per-work-group maximum (raised for that kernel), less its static LDS.
static size_t kill_lds_limit() {
  int dev = 0, optin = 0, base = 0;
  CS_HIP(hipGetDevice(&dev));
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess) optin = 0;
  CS_HIP(hipDeviceGetAttribute(&base, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  size_t lim = (size_t)std::max(optin, base);
  if (lim > (size_t)base &&
      hipFuncSetAttribute(reinterpret_cast<const void*>(&k_kill_run<true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)(lim - 64)) != hipSuccess)
    lim = (size_t)base;
  return lim > 64 ? lim - 64 : 0;
}

// Grid-size autotuner for a grid-stride kernel whose results do not depend on
// its grid (every reduction is an atomic min or count, every emission is
// sorted or consumed by an order-free pass).  The first calls run the
// candidate grid caps in turn, timed with HIP events on the stream; once each
// has `kSamples` timings the cap with the smallest median is kept for the
// life of the solver.  The timings are read after the caller's own
// synchronisation, so tuning adds no host waits.
struct GridTuner {
  static constexpr int kSamples = 3;
  const char* name;
  std::vector<unsigned> caps{512, 1024, 2048, 4096, 8192};
  std::vector<std::vector<float>> ms;
  int chosen = -1, cur = -1, turn = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  explicit GridTuner(const char* n) : name(n), ms(caps.size()) {}
  GridTuner(const char* n, std::vector<unsigned> c) : name(n), caps(std::move(c)), ms(caps.size()) {}
  ~GridTuner() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  // grid for `work` items of one thread each; brackets the launch when tuning
  unsigned grid(size_t work, hipStream_t st) {
    const size_t need = std::max<size_t>(1, (work + kBlock - 1) / kBlock);
    if (chosen >= 0) return (unsigned)std::min<size_t>(caps[chosen], need);
    if (!e0) {
      CS_HIP(hipEventCreate(&e0));
      CS_HIP(hipEventCreate(&e1));
    }
    cur = turn;
    turn = (turn + 1) % (int)caps.size();
    CS_HIP(hipEventRecord(e0, st));
    return (unsigned)std::min<size_t>(caps[cur], need);
  }
  void after_launch(hipStream_t st) {
    if (cur >= 0) CS_HIP(hipEventRecord(e1, st));
  }
  // after the stream has been synchronised
  void collect() {
    if (cur < 0) return;
    float t = 0;
    CS_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[cur].push_back(t);
    cur = -1;
    bool done = true;
    for (auto& v : ms) done = done && (int)v.size() >= kSamples;
    if (!done) return;
    float best = 0;
    for (size_t c = 0; c < caps.size(); c++) {
      std::vector<float> v = ms[c];
      std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
      if (chosen < 0 || v[v.size() / 2] < best) {
        chosen = (int)c;
        best = v[v.size() / 2];
      }
    }
  }
  std::string describe() const {
    return std::string(name) + "=" + (chosen >= 0 ? std::to_string(caps[chosen]) : std::string("tuning"));
  }
};

struct ContactImpl {
  GridTuner tune_trial{"k_trial_pairs"}, tune_active{"k_active_pairs"};
  // the CCD tail's length is only known on the device: fixed grids
  GridTuner tune_ccd{"k_trial_ccd", {128, 256, 512, 1024, 2048}};
  hipStream_t st;
  cs_contact_params prm;
  uint32_t n, m, ns, ms;
  bool has_floor;
  double floor_y;
  int B;
  DevBuf<uint32_t> idx, sidx, tri_rank, dedge_rank, uedge;
  DevBuf<double> spos;
  // LBVH over the cloth triangles (Morton order kept until a re-sort) and a
  // fixed one over the static triangles
  Levels LV, SL;
  DevBuf<Leaf> leaves, sleaves;
  DevBuf<double> nodes, snodes, dpart;
  DevBuf<uint64_t> mkeys, mkeys_t, bkeys;
  DevBuf<uint32_t> mvals, mvals_t;
  const uint32_t* order = nullptr;
  bool need_sort = true;
  DevBuf<double> sbox;  // unpadded static triangle boxes by triangle id
  // candidate superset of the current segment (see set_segment)
  DevBuf<uint2> pairs_sup, spairs_sup;
  uint32_t sup_cap = 0, ssup_cap = 0, np_sup = 0, nsp_sup = 0;
  bool sup_valid = false;
  // candidates
  DevBuf<uint2> pairs, spairs;
  DevBuf<uint2> tailq;  // (pair, prim) items for the CCD tail
  size_t tail_cap = 0;
  uint32_t pair_cap = 0, spair_cap = 0, np = 0, nsp = 0;
  // active set
  DevBuf<Frozen> rec, act;
  DevBuf<uint64_t> akeys, akeys_t, anke_raw, anke;
  DevBuf<uint32_t> avals, avals_t;
  uint32_t act_cap = 0, C = 0;
  // incidence
  DevBuf<uint64_t> ekeys, ekeys_t;
  DevBuf<uint32_t> evals, evals_t, v_off, v_ent, ent_vtx;
  // per-contact evaluation
  DevBuf<uint8_t> inr;
  DevBuf<KillRec> krec;           // velocity-filter sweep records
  DevBuf<uint32_t> kcnt, klist;    // counters / frontier lists of the global-memory sweep
  size_t kill_lds_max = 0;         // dynamic LDS the LDS sweep may use
  DevBuf<double> cg, fr, coeff, dd, hvsum, hterm;
  DevBuf<uint32_t> heads, cverts;
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

  double mink_value(int k) const {
    const uint64_t key = hcs->mink[k];
    if (key == ~0ull) return INFINITY;
    const uint64_t u = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
    double d;
    std::memcpy(&d, &u, 8);
    return d;
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
      h_sidx.assign(d.static_idx, d.static_idx + (size_t)ms * 3);
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
    size_t nn = 0;
    LV = make_levels(m, &nn);
    leaves.alloc(std::max<uint32_t>(m, 1));
    nodes.alloc(std::max<size_t>(nn, 1) * 6);
    bkeys.alloc(6);
    dpart.alloc(9 * 1024);
    mkeys.alloc(m); mkeys_t.alloc(m); mvals.alloc(m); mvals_t.alloc(m);
    build_static_tree(d);
    cs.alloc(1);
    ddpart.alloc(2 * kMaxRedBlocks);
    CS_HIP(hipHostMalloc(&hcs, sizeof(CScalars)));
    std::memset(hcs, 0, sizeof(CScalars));
    CS_HIP(hipMemsetAsync(cs.ptr, 0, sizeof(CScalars), st));
    v_off.alloc(n + 1);
    heads.alloc(n);
    klist.alloc((size_t)2 * std::max<uint32_t>(n, 1));
    kill_lds_max = kill_lds_limit();
    cverts.alloc(std::max<uint32_t>(n, 1));
    hvsum.alloc((size_t)std::max<uint32_t>(n, 1) * 3);
    hvsum.zero(st);
    lag_count.alloc(1);
    ensure_sort(m);
    grow_pairs(std::max<uint32_t>(1024, m * 4), std::max<uint32_t>(256, ms * 8));
    sup_cap = std::max<uint32_t>(1024, m * 8);
    ssup_cap = std::max<uint32_t>(256, ms * 16);
    pairs_sup.alloc(sup_cap);
    spairs_sup.alloc(ssup_cap);
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
    ent_vtx.alloc((size_t)cap * 4); hterm.alloc((size_t)cap * 24);
    inr.alloc(cap); krec.alloc(cap); kcnt.alloc(cap);
    cg.alloc(cap); fr.alloc((size_t)cap * 3); coeff.alloc(cap); dd.alloc(cap);
    act_cap = cap;
    ensure_sort((size_t)cap * 4);
  }

  void clear_counters() {
    CS_HIP(hipMemsetAsync(cs.ptr, 0, sizeof(CScalars), st));
    CS_HIP(hipMemsetAsync(cs.ptr->mink, 0xff, sizeof(hcs->mink), st));
  }

  // Fixed hierarchy over the static triangles: unpadded exact boxes
  // (Math.min / Math.max of the three corners, staticTriAabb :597-605) in
  // Morton order of their centroids, inner boxes as NaN-ignoring unions.
  std::vector<uint32_t> h_sidx;  // the static triangles (their topology is fixed for the solver's life)
  void build_static_tree(const cs_scene_desc& d) { build_static_tree(d.static_pos, d.static_idx); }
  void build_static_tree(const float* static_pos, const uint32_t* static_idx) {
    struct {
      const float* static_pos;
      const uint32_t* static_idx;
    } d{static_pos, static_idx};
    size_t nn = 0;
    SL = make_levels(ms, &nn);
    std::vector<Leaf> sl(ms);
    std::vector<double> cen((size_t)ms * 3);
    double blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t t = 0; t < ms; t++) {
      Leaf& R = sl[t];
      for (int k = 0; k < 3; k++) {
        const double a = d.static_pos[(size_t)d.static_idx[(size_t)t * 3] * 3 + k];
        const double b = d.static_pos[(size_t)d.static_idx[(size_t)t * 3 + 1] * 3 + k];
        const double c = d.static_pos[(size_t)d.static_idx[(size_t)t * 3 + 2] * 3 + k];
        R.lo[k] = jsm::jmin(jsm::jmin(a, b), c);
        R.hi[k] = jsm::jmax(jsm::jmax(a, b), c);
        cen[(size_t)t * 3 + k] = 0.5 * (R.lo[k] + R.hi[k]);
        blo[k] = std::fmin(blo[k], cen[(size_t)t * 3 + k]);
        bhi[k] = std::fmax(bhi[k], cen[(size_t)t * 3 + k]);
      }
      for (int c = 0; c < 3; c++) R.v[c] = -(int32_t)d.static_idx[(size_t)t * 3 + c] - 2;
      R.tri = t;
    }
    auto expand = [](uint32_t v) {
      uint32_t r = 0;
      for (int b = 0; b < 10; b++) r |= ((v >> b) & 1u) << (3 * b);
      return r;
    };
    std::vector<uint64_t> code(ms);
    double ext = 0;
    for (int k = 0; k < 3; k++) ext = std::fmax(ext, bhi[k] - blo[k]);
    for (uint32_t t = 0; t < ms; t++) {
      uint32_t q[3];
      for (int k = 0; k < 3; k++) {
        double u = ext > 0 ? (cen[(size_t)t * 3 + k] - blo[k]) / ext : 0.0;
        if (!(u >= 0)) u = 0;
        if (u > 1) u = 1;
        q[k] = (uint32_t)std::fmin(1023.0, u * 1024.0);
      }
      code[t] = ((uint64_t)(expand(q[0]) << 2 | expand(q[1]) << 1 | expand(q[2])) << 32) | t;
    }
    std::sort(code.begin(), code.end());
    std::vector<Leaf> sorted(ms);
    for (uint32_t j = 0; j < ms; j++) sorted[j] = sl[(uint32_t)code[j]];
    std::vector<double> nb(std::max<size_t>(nn, 1) * 6);
    for (int l = 1; l <= SL.L; l++)
      for (uint32_t k = 0; k < SL.cnt[l]; k++) {
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        const uint32_t c0 = k * kFan, c1 = std::min(c0 + kFan, SL.cnt[l - 1]);
        for (uint32_t c = c0; c < c1; c++) {
          const double* b = l == 1 ? nullptr : &nb[((size_t)SL.off[l - 1] + c) * 6];
          for (int q = 0; q < 3; q++) {
            lo[q] = std::fmin(lo[q], l == 1 ? sorted[c].lo[q] : b[q]);
            hi[q] = std::fmax(hi[q], l == 1 ? sorted[c].hi[q] : b[3 + q]);
          }
        }
        for (int q = 0; q < 3; q++) {
          nb[((size_t)SL.off[l] + k) * 6 + q] = lo[q];
          nb[((size_t)SL.off[l] + k) * 6 + 3 + q] = hi[q];
        }
      }
    if (ms) sleaves.upload(sorted.data(), ms, st);
    else sleaves.alloc(1);
    {
      std::vector<double> sb(std::max<size_t>((size_t)ms * 6, 6));
      for (uint32_t t = 0; t < ms; t++)
        for (int k = 0; k < 3; k++) {
          sb[(size_t)t * 6 + k] = sl[t].lo[k];
          sb[(size_t)t * 6 + 3 + k] = sl[t].hi[k];
        }
      sbox.upload(sb.data(), sb.size(), st);
    }
    snodes.upload(nb.data(), nb.size(), st);
  }

  // Morton order of the cloth triangles at x (radix sort of 30-bit codes).
  void sort_leaves(const double* x) {
    const DGeom G = geom();
    CS_HIP(hipMemsetAsync(bkeys.ptr, 0xff, 3 * sizeof(uint64_t), st));
    CS_HIP(hipMemsetAsync(bkeys.ptr + 3, 0, 3 * sizeof(uint64_t), st));
    hipLaunchKernelGGL(k_vert_bounds, dim3(red_grid(n)), dim3(kBlock), 0, st, n, x, bkeys.ptr);
    hipLaunchKernelGGL(k_morton, dim3(grid_for(m)), dim3(kBlock), 0, st, G, x, bkeys.ptr, mkeys.ptr, mvals.ptr);
    CS_LAUNCH_CHECK();
    const bool in_tmp = radix_sort_pairs(st, mkeys.ptr, mvals.ptr, mkeys_t.ptr, mvals_t.ptr, m, 0, 30, sort);
    order = in_tmp ? mvals_t.ptr : mvals.ptr;
    need_sort = false;
  }

  // LBVH refit over (x_step, xa[, xb]) and both traversals into the given
  // pair buffers (grown and re-run on overflow).
  void traverse(const double* xa, const double* xb, double pad, DevBuf<uint2>& P, uint32_t& pcap,
                DevBuf<uint2>& SP, uint32_t& spcap, uint32_t& np_out, uint32_t& nsp_out) {
    const DGeom G = geom();
    const unsigned waves8 = (unsigned)((m + kWave * kFan - 1) / (kWave * kFan));
    hipLaunchKernelGGL(k_refit8, dim3((waves8 + 3) / 4), dim3(kBlock), 0, st, G, xstep, xa, xb, pad, order,
                       leaves.ptr, nodes.ptr, LV);
    if (LV.L > 3) hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(1024), 0, st, nodes.ptr, LV);
    CS_LAUNCH_CHECK();
    for (int attempt = 0; attempt < 3; attempt++) {
      CS_HIP(hipMemsetAsync(&cs.ptr->pair_count, 0, 2 * sizeof(uint32_t), st));
      if (m > 1)
        hipLaunchKernelGGL(k_traverse_self, dim3(grid_for(m)), dim3(kBlock), 0, st, leaves.ptr, nodes.ptr, LV, P.ptr,
                           pcap, cs.ptr);
      if (ms)
        hipLaunchKernelGGL(k_traverse_static, dim3(grid_for(m)), dim3(kBlock), 0, st, leaves.ptr, m, sleaves.ptr,
                           snodes.ptr, SL, pad, SP.ptr, spcap, cs.ptr);
      CS_LAUNCH_CHECK();
      sync_cs();
      if (hcs->nonfinite & 2u) throw OverflowError("LBVH traversal stack overflow");
      np_out = hcs->pair_count;
      nsp_out = hcs->spair_count;
      if (np_out <= pcap && nsp_out <= spcap) return;
      if (np_out > pcap) { pcap = (uint32_t)(np_out * 1.25) + 1024; P.alloc(pcap); }
      if (nsp_out > spcap) { spcap = (uint32_t)(nsp_out * 1.25) + 256; SP.alloc(spcap); }
    }
    throw OverflowError("candidate pair buffer could not be grown");
  }

  // The segment (xa -> xb) that the next builds lie on: line-search trials
  // x + alpha dx (alpha in (0, 1]) and the accepted iterate, or the predictor
  // bisection probes.  Every such build's swept boxes lie inside the boxes
  // over (x_step, xa, xb) padded by max(dHat, dMin), so one traversal gives a
  // superset of all their candidate pairs.
  void set_segment(const double* xa, const double* xb) {
    sup_valid = false;
    if (m == 0) return;
    if (need_sort) sort_leaves(xb);
    const double pad = std::max(prm.d_hat, prm.d_min);
    traverse(xa, xb, pad, pairs_sup, sup_cap, spairs_sup, ssup_cap, np_sup, nsp_sup);
    sup_valid = true;
  }

  // buildCandidates (contact-assembly.ts:87-121) minus the expansion into
  // prims, which the prim kernels do on the fly.
  // Candidate pairs of buildCandidates (contact-assembly.ts:87-121) at x
  // (swept from the step start, padded by `pad`) minus the expansion into
  // prims, which the pair kernels do.  Returns true when the pairs are the
  // step-segment superset, which the pair kernels then filter with the exact
  // predicate; otherwise pairs / spairs hold the traversal's exact list.
  struct PairSrc {
    const uint2* p;
    uint32_t np;
    const uint2* s;
    uint32_t nsp;
    bool filter;
  };

  PairSrc candidate_source(const double* x, double pad) {
    clear_counters();
    CS_HIP(hipMemsetAsync(&cs.ptr->nan_key, 0xff, 2 * sizeof(uint64_t), st));
    if (m == 0) {
      np = 0;
      nsp = 0;
      return PairSrc{nullptr, 0, nullptr, 0, false};
    }
    if (sup_valid) return PairSrc{pairs_sup.ptr, np_sup, spairs_sup.ptr, nsp_sup, true};
    if (need_sort) sort_leaves(x);
    traverse(x, nullptr, pad, pairs, pair_cap, spairs, spair_cap, np, nsp);
    return PairSrc{pairs.ptr, np, spairs.ptr, nsp, false};
  }

  // The exact candidate counts after a pair kernel ran (scalars synced).
  void candidate_counts(const PairSrc& src) {
    if (src.filter) {
      np = hcs->pair_count;
      nsp = hcs->spair_count;
    }
    diag.candidate_pairs = (uint64_t)np * 15 + (uint64_t)nsp * 12;
  }

  void update_active_set(const double* x) {
    const PairSrc src = candidate_source(x, prm.d_hat);
    const DGeom G = geom();
    const uint32_t total = src.np + src.nsp;
    for (int attempt = 0; attempt < 3; attempt++) {
      CS_HIP(hipMemsetAsync(&cs.ptr->pair_count, 0, 3 * sizeof(uint32_t), st));  // pair, spair, active counts
      if (total) {
        const dim3 ag(tune_active.grid(total, st));
        if (src.filter)
          hipLaunchKernelGGL(k_active_pairs<true>, ag, dim3(kBlock), 0, st, G, xstep, x, prm.d_hat, src.p, src.np,
                             src.s, src.nsp, sbox.ptr, rec.ptr, akeys.ptr, anke_raw.ptr, act_cap, cs.ptr);
        else
          hipLaunchKernelGGL(k_active_pairs<false>, ag, dim3(kBlock), 0, st, G, xstep, x, prm.d_hat, src.p, src.np,
                             src.s, src.nsp, sbox.ptr, rec.ptr, akeys.ptr, anke_raw.ptr, act_cap, cs.ptr);
        tune_active.after_launch(st);
      }
      if (has_floor)
        hipLaunchKernelGGL(k_floor_active, dim3(red_grid(n)), dim3(kBlock), 0, st, G, x, rec.ptr, akeys.ptr,
                           anke_raw.ptr, act_cap, dpart.ptr + 1024, cs.ptr);
      CS_LAUNCH_CHECK();
      sync_cs();
      if (total) tune_active.collect();
      if (hcs->active_count <= act_cap) break;
      grow_active((uint32_t)(hcs->active_count * 1.25) + 1024);
      if (attempt == 2) throw OverflowError("active contact buffer could not be grown");
    }
    candidate_counts(src);
    C = hcs->active_count;
    double minD = mink_value(0);
    if (has_floor && mink_value(1) < minD) minD = mink_value(1);
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
    if (C) hipLaunchKernelGGL(k_count_kinds, dim3(red_grid(C)), dim3(kBlock), 0, st, act.ptr, C, cs.ptr);
    build_incidence();  // syncs the scalars
    // diagnostics (contact-assembly.ts:215-219)
    diag.active_pairs = C;
    diag.vt_pairs = hcs->kind_count[0];
    diag.ee_pairs = hcs->kind_count[1];
    diag.floor_pairs = hcs->kind_count[2];
    if (minD < diag.min_distance) diag.min_distance = minD;
  }

  int cshift = 3;
  uint32_t E = 0;   // vertex -> contact incidence entries
  uint32_t nv = 0;  // vertices with entries
  void build_incidence() {
    CS_HIP(hipMemsetAsync(v_off.ptr, 0, (size_t)(n + 1) * sizeof(uint32_t), st));
    E = 0;
    nv = 0;  // no contact vertex until the entries are rebuilt below
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
    const size_t nent = (size_t)C * 4;
    const bool in_tmp = radix_sort_pairs(st, ekeys.ptr, evals.ptr, ekeys_t.ptr, evals_t.ptr, nent, 0, cshift + vb, sort);
    const uint64_t* sk = in_tmp ? ekeys_t.ptr : ekeys.ptr;
    sync_cs();
    const uint32_t cnt = hcs->entry_count;
    E = cnt;
    nv = 0;
    if (cnt) {
      hipLaunchKernelGGL(k_entry_ent, dim3(grid_for(cnt)), dim3(kBlock), 0, st, sk, cnt, cshift, v_ent.ptr,
                         ent_vtx.ptr);
      hipLaunchKernelGGL(k_entry_offsets, dim3(grid_for((size_t)n + 1)), dim3(kBlock), 0, st, sk, cnt, cshift, n,
                         v_off.ptr);
      CS_HIP(hipMemsetAsync(&cs.ptr->nverts, 0, sizeof(uint32_t), st));
      hipLaunchKernelGGL(k_contact_verts, dim3(grid_for(cnt)), dim3(kBlock), 0, st, sk, cnt, cshift, cverts.ptr,
                         cs.ptr);
      CS_HIP(hipMemsetAsync(hvsum.ptr, 0, (size_t)n * 3 * sizeof(double), st));
      CS_LAUNCH_CHECK();
      sync_cs();
      nv = hcs->nverts;
    }
    CS_LAUNCH_CHECK();
  }

  CEval cev() { return CEval{inr.ptr, cg.ptr, fr.ptr, coeff.ptr}; }

  void energy_grad(const double* x, double* ceg, bool energies = true) {
    const DGeom G = geom();
    if (C) {
      hipLaunchKernelGGL(k_contact_eval, dim3(red_grid(C)), dim3(kBlock), 0, st, G, x, xstep, act.ptr, anke.ptr, C,
                         lag_keys.ptr, lag_idx.ptr, lag_L, lag_vals.ptr, cev(), ddpart.ptr, cs.ptr);
      CS_LAUNCH_CHECK();
      if (energies) {
        FinishJobs f{};
        f.j[0] = FinishJob{ddpart.ptr, red_grid(C), FIN_DD, 0, &cs.ptr->red[2]};
        f.j[1] = FinishJob{ddpart.ptr + kMaxRedBlocks, red_grid(C), FIN_DD, 0, &cs.ptr->red[3]};
        launch_finish(st, f, 2);
      }
    }
    CS_HIP(hipMemsetAsync(ceg, 0, (size_t)n * 3 * sizeof(double), st));
    if (E) {
      hipLaunchKernelGGL(k_grad_terms, dim3(grid_for(E)), dim3(kBlock), 0, st, E, act.ptr, v_ent.ptr, cev(),
                         hterm.ptr);
      hipLaunchKernelGGL(k_grad_vsum, dim3(grid_for((size_t)nv * 3)), dim3(kBlock), 0, st, nv, cverts.ptr, v_off.ptr,
                         hterm.ptr, ceg);
    }
    CS_LAUNCH_CHECK();
    if (!energies) return;
    if (C) {
      sync_cs();
      diag.barrier_energy = hcs->red[2];
      diag.friction_work += hcs->red[3];
    } else {
      diag.barrier_energy = 0;
    }
  }

  // setStaticMesh (contact-assembly.ts:71-73) between steps: new positions of
  // the same static triangles (a moving rigid collider, avatar-collider.ts:
  // 205-230 samples its mesh per tick and attaches it this way).
  void set_static_mesh(const float* pos, uint32_t nverts, const uint32_t* sidx_h, uint32_t ntris) {
    if (ntris != ms || (ms && nverts != ns)) throw UnsupportedError("the static mesh's vertex and triangle counts are fixed at creation");
    for (size_t k = 0; k < (size_t)ms * 3; k++)
      if (sidx_h[k] != h_sidx[k]) throw UnsupportedError("the static mesh's triangles are fixed at creation");
    if (!ms) return;
    std::vector<double> sp(pos, pos + (size_t)ns * 3);
    CS_HIP(hipStreamSynchronize(st));
    spos.upload(sp.data(), sp.size(), st);
    build_static_tree(pos, sidx_h);
    CS_HIP(hipStreamSynchronize(st));
  }

  DevBuf<double> fadj_y, fadj_term;
  void friction_xstep_adjoint(const double* x, const double* lam, double* out) {
    if (!C || !E || !(prm.friction_mu > 0)) return;
    fadj_y.alloc((size_t)C * 3);
    fadj_term.alloc((size_t)E * 3);
    hipLaunchKernelGGL(k_fric_adj_contact, dim3(grid_for(C)), dim3(kBlock), 0, st, geom(), x, xstep, act.ptr,
                       anke.ptr, C, lag_keys.ptr, lag_idx.ptr, lag_L, lag_vals.ptr, lam, fadj_y.ptr);
    hipLaunchKernelGGL(k_fric_adj_terms, dim3(grid_for(E)), dim3(kBlock), 0, st, E, act.ptr, v_ent.ptr, fadj_y.ptr,
                       fadj_term.ptr);
    hipLaunchKernelGGL(k_fric_adj_vsum, dim3(grid_for((size_t)nv * 3)), dim3(kBlock), 0, st, nv, cverts.ptr,
                       v_off.ptr, fadj_term.ptr, out);
    CS_LAUNCH_CHECK();
  }

  void add_diag_estimate(double* dg) {
    if (!C) return;
    if (E) {
      hipLaunchKernelGGL(k_diag_terms, dim3(grid_for(E)), dim3(kBlock), 0, st, E, act.ptr, v_ent.ptr, cev(),
                         hterm.ptr);
      hipLaunchKernelGGL(k_diag_vsum, dim3(grid_for((size_t)nv * 3)), dim3(kBlock), 0, st, nv, cverts.ptr, v_off.ptr,
                         hterm.ptr, dg);
    }
    CS_LAUNCH_CHECK();
  }

  double merit_energy(const double* x) {
    if (!C) return 0;
    hipLaunchKernelGGL(k_merit, dim3(red_grid(C)), dim3(kBlock), 0, st, geom(), x, act.ptr, C, ddpart.ptr, cs.ptr);
    CS_LAUNCH_CHECK();
    FinishJobs f{};
    f.j[0] = FinishJob{ddpart.ptr, red_grid(C), FIN_DD, 0, &cs.ptr->red[4]};
    launch_finish(st, f, 1);
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
    // heads[] (n entries) serves as the vertex -> LDS slot map here
    hipLaunchKernelGGL(k_kill_slots, dim3(grid_for(nv)), dim3(kBlock), 0, st, cverts.ptr, nv, heads.ptr);
    hipLaunchKernelGGL(k_kill_rec, dim3(grid_for(C)), dim3(kBlock), 0, st, act.ptr, C, heads.ptr, krec.ptr);
    hipLaunchKernelGGL(k_kill_links, dim3(grid_for(E)), dim3(kBlock), 0, st, v_off.ptr, v_ent.ptr, ent_vtx.ptr, E,
                       krec.ptr);
    const size_t lds = (size_t)nv * 3 * sizeof(double) + (size_t)((C + 3) / 4) * 4 + (size_t)nv * 2 * 4;
    if (lds <= kill_lds_max)
      hipLaunchKernelGGL(k_kill_run<true>, dim3(1), dim3(kKillBlock), lds, st, krec.ptr, C, cverts.ptr, nv, v,
                         kcnt.ptr, klist.ptr, &cs.ptr->kill_rounds);
    else
      hipLaunchKernelGGL(k_kill_run<false>, dim3(1), dim3(kKillBlock), 0, st, krec.ptr, C, cverts.ptr, nv, v,
                         kcnt.ptr, klist.ptr, &cs.ptr->kill_rounds);
    CS_LAUNCH_CHECK();
    sync_cs();
    diag.filter_rounds = hcs->kill_rounds;
    CS_LAUNCH_CHECK();
  }

  // ---- trial validation queued for a batched line search ----------------
  // trial_enqueue() launches one trial's validation and keeps its counters in
  // slot `slot` without reading them; trial_fetch() reads every slot with one
  // synchronisation; trial_result() then judges a slot exactly as
  // check_trial() judges its own counters (diagnostics included), so a caller
  // that judges the slots in order up to its first accepted trial reproduces
  // the sequential line search.
  DevBuf<CScalars> cs_slots;
  std::vector<CScalars> hslots;
  std::vector<PairSrc> slot_src;
  void trial_enqueue(const double* xt, int slot) {
    if ((size_t)slot >= cs_slots.count) {
      cs_slots.alloc((size_t)slot + 8);
      hslots.resize((size_t)slot + 8);
      slot_src.resize((size_t)slot + 8);
    }
    slot_src[slot] = launch_trial_kernels(xt);
    CS_HIP(hipMemcpyAsync(cs_slots.ptr + slot, cs.ptr, sizeof(CScalars), hipMemcpyDeviceToDevice, st));
  }
  void trial_fetch(int count) {
    if (count <= 0) return;
    CS_HIP(hipMemcpyAsync(hslots.data(), cs_slots.ptr, (size_t)count * sizeof(CScalars), hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
    tune_trial.collect();
    tune_ccd.collect();
  }
  bool trial_result(int slot) { return judge_trial(hslots[slot], slot_src[slot]); }

  PairSrc launch_trial_kernels(const double* xt) {
    const PairSrc src = candidate_source(xt, prm.d_min);
    const DGeom G = geom();
    const uint32_t total = src.np + src.nsp;
    if (total) {
      if ((size_t)total * 15 > tail_cap) {
        tail_cap = (size_t)total * 15;
        tailq.alloc(tail_cap);
      }
      const dim3 tg(tune_trial.grid(total, st));
      if (src.filter)
        hipLaunchKernelGGL(k_trial_pairs<true>, tg, dim3(kBlock), 0, st, G, xstep, xt, prm.d_min, src.p, src.np,
                           src.s, src.nsp, sbox.ptr, tailq.ptr, cs.ptr);
      else
        hipLaunchKernelGGL(k_trial_pairs<false>, tg, dim3(kBlock), 0, st, G, xstep, xt, prm.d_min, src.p, src.np,
                           src.s, src.nsp, sbox.ptr, tailq.ptr, cs.ptr);
      tune_trial.after_launch(st);
      const dim3 cg(tune_ccd.grid((size_t)1 << 40, st));
      hipLaunchKernelGGL(k_trial_ccd, cg, dim3(kBlock), 0, st, G, xstep, xt, src.p, src.np, src.s, tailq.ptr,
                         cs.ptr);
      tune_ccd.after_launch(st);
    }
    if (has_floor)
      hipLaunchKernelGGL(k_floor_trial, dim3(red_grid(n)), dim3(kBlock), 0, st, G, xstep, xt, dpart.ptr + 2048, cs.ptr);
    CS_LAUNCH_CHECK();
    return src;
  }

  static double mink_of(const CScalars& h, int k) {
    const uint64_t key = h.mink[k];
    if (key == ~0ull) return INFINITY;
    const uint64_t u = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
    double d;
    std::memcpy(&d, &u, 8);
    return d;
  }

  // checkTrial's verdict (contact-assembly.ts:489-553) from a trial's counters
  bool judge_trial(const CScalars& h, const PairSrc& src) {
    if (src.filter) {
      np = h.pair_count;
      nsp = h.spair_count;
    }
    diag.candidate_pairs = (uint64_t)np * 15 + (uint64_t)nsp * 12;
    double minDist = mink_of(h, 5);
    double minTOI = mink_of(h, 6);
    if (h.nonfinite & 1u) {
      diag.ccd_failures++;
      return false;
    }
    if (has_floor && mink_of(h, 7) < minDist) minDist = mink_of(h, 7);
    if (minDist <= prm.d_min) return false;
    if (h.nan_key != SENT && h.nan_key < h.cross_key) {
      diag.ccd_failures++;
      return false;
    }
    if (h.cross_key != SENT) return false;
    if (has_floor) {
      if (mink_of(h, 8) < minTOI) minTOI = mink_of(h, 8);
      if (h.floor_fail) return false;
    }
    if (minDist < diag.min_distance) diag.min_distance = minDist;
    if (minTOI < diag.min_toi) diag.min_toi = minTOI;
    return true;
  }

  // The cloth/cloth and cloth/static triangle pairs of buildCandidates at x
  // (swept from the step start, padded by `pad`) from a fresh traversal, in
  // the traversal's order.
  void probe_pairs(const double* x, double pad, std::vector<uint32_t>& self2, std::vector<uint32_t>& stat2) {
    sup_valid = false;
    const PairSrc src = candidate_source(x, pad);
    self2.resize((size_t)src.np * 2);
    stat2.resize((size_t)src.nsp * 2);
    if (src.np) CS_HIP(hipMemcpyAsync(self2.data(), src.p, (size_t)src.np * 8, hipMemcpyDeviceToHost, st));
    if (src.nsp) CS_HIP(hipMemcpyAsync(stat2.data(), src.s, (size_t)src.nsp * 8, hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
  }

  // meritEnergy (contact-assembly.ts:419-428) at x into the device double `out`
  void merit_enqueue(const double* x, double* out) {
    if (!C) {
      CS_HIP(hipMemsetAsync(out, 0, sizeof(double), st));
      return;
    }
    hipLaunchKernelGGL(k_merit, dim3(red_grid(C)), dim3(kBlock), 0, st, geom(), x, act.ptr, C, ddpart.ptr, cs.ptr);
    CS_LAUNCH_CHECK();
    FinishJobs f{};
    f.j[0] = FinishJob{ddpart.ptr, red_grid(C), FIN_DD, 0, out};
    launch_finish(st, f, 1);
  }

  bool check_trial(const double* xt) {
    const PairSrc src = launch_trial_kernels(xt);
    sync_cs();
    if (src.np + src.nsp) {
      tune_trial.collect();
      tune_ccd.collect();
    }
    return judge_trial(*hcs, src);
  }
};

// ---------------------------------------------------------------- facade

ContactSystem::ContactSystem(const cs_scene_desc& d, hipStream_t st) : im_(new ContactImpl(d, st)) {}
ContactSystem::~ContactSystem() { delete im_; }
void ContactSystem::begin_step(const double* x0) {
  im_->xstep = x0;
  im_->need_sort = true;
  im_->sup_valid = false;
  im_->lag_L = 0;
  im_->lag_vbase = 0;
  im_->C = 0;
  im_->reset_diag();
}
bool ContactSystem::check_trial(const double* xt) { return im_->check_trial(xt); }
void ContactSystem::set_segment(const double* xa, const double* xb) { im_->set_segment(xa, xb); }
void ContactSystem::update_active_set(const double* x) { im_->update_active_set(x); }
void ContactSystem::energy_grad(const double* x, double* ceg) { im_->energy_grad(x, ceg); }
void ContactSystem::grad_only(const double* x, double* ceg) { im_->energy_grad(x, ceg, false); }
void ContactSystem::add_diag_estimate(double* diag, const double*) { im_->add_diag_estimate(diag); }
double ContactSystem::merit_energy(const double* x) { return im_->merit_energy(x); }
void ContactSystem::probe_pairs(const double* x, double pad, std::vector<uint32_t>& self2,
                                std::vector<uint32_t>& stat2) {
  im_->probe_pairs(x, pad, self2, stat2);
}
double ContactSystem::pad(bool d_min) const { return d_min ? im_->prm.d_min : im_->prm.d_hat; }
void ContactSystem::trial_enqueue(const double* xt, int slot) { im_->trial_enqueue(xt, slot); }
void ContactSystem::trial_fetch(int count) { im_->trial_fetch(count); }
bool ContactSystem::trial_result(int slot) { return im_->trial_result(slot); }
void ContactSystem::merit_enqueue(const double* x, double* out) { im_->merit_enqueue(x, out); }
void ContactSystem::set_static_mesh(const float* pos, uint32_t nverts, const uint32_t* idx, uint32_t ntris) {
  im_->set_static_mesh(pos, nverts, idx, ntris);
}
void ContactSystem::friction_xstep_adjoint(const double* x, const double* lam, double* out) {
  im_->friction_xstep_adjoint(x, lam, out);
}
void ContactSystem::add_rejection() { im_->diag.rejected_line_search_steps++; }
void ContactSystem::commit(const double* x) { im_->commit(x); }
void ContactSystem::kill_approach_velocity(const double* x, double* v) { im_->kill_approach_velocity(x, v); }
cs_contact_diag ContactSystem::diagnostics() const { return im_->diag; }
uint64_t ContactSystem::active_count() const { return im_->C; }

uint64_t ContactSystem::candidate_count() const { return im_->diag.candidate_pairs; }
std::string ContactSystem::tuned() const {
  return im_->tune_trial.describe() + " " + im_->tune_ccd.describe() + " " + im_->tune_active.describe();
}
uint64_t ContactSystem::incidence_count() const { return im_->E; }
uint64_t ContactSystem::incidence_vertices() const { return im_->nv; }

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

// Batched CCD over device arrays: the ids are checked in the kernel (an id
// outside [0, n) gives NaN and sets *bad).
__global__ void __launch_bounds__(kBlock) k_ccd_queries(const double* __restrict__ x0, const double* __restrict__ x1,
                                                        uint32_t n, const int32_t* __restrict__ kinds,
                                                        const int32_t* __restrict__ ids4, uint64_t count,
                                                        double thickness, double* toi, uint32_t* bad) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int4 id = reinterpret_cast<const int4*>(ids4)[i];
  const int32_t v4[4] = {id.x, id.y, id.z, id.w};
  for (int s = 0; s < 4; s++)
    if (v4[s] < 0 || (uint32_t)v4[s] >= n) {
      toi[i] = __builtin_nan("");
      atomicOr(bad, 1u);
      return;
    }
  Seg4 S;
  for (int s = 0; s < 4; s++)
    for (int k = 0; k < 3; k++) {
      S.x0[s][k] = x0[(size_t)v4[s] * 3 + k];
      S.x1[s][k] = x1[(size_t)v4[s] * 3 + k];
    }
  toi[i] = kinds[i] == 0 ? vt_ccd(S, thickness) : ee_ccd(S, thickness);
}

void ccd_queries(const double* x0, const double* x1, uint32_t n, const int32_t* kinds, const int32_t* ids4,
                 uint64_t count, double thickness, double* toi, double* kernel_ms) {
  hipStream_t st = nullptr;
  DevBuf<uint32_t> bad;
  bad.alloc(1);
  CS_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(uint32_t), st));
  hipEvent_t e[2] = {nullptr, nullptr};
  if (kernel_ms)
    for (auto& x : e) CS_HIP(hipEventCreate(&x));
  struct G {
    hipEvent_t* e;
    ~G() {
      for (int k = 0; k < 2; k++)
        if (e[k]) (void)hipEventDestroy(e[k]);
    }
  } guard{e};
  if (kernel_ms) CS_HIP(hipEventRecord(e[0], st));
  if (count)
    hipLaunchKernelGGL(k_ccd_queries, dim3(grid_for(count)), dim3(kBlock), 0, st, x0, x1, n, kinds, ids4, count,
                       thickness, toi, bad.ptr);
  CS_LAUNCH_CHECK();
  if (kernel_ms) CS_HIP(hipEventRecord(e[1], st));
  uint32_t hb = 0;
  CS_HIP(hipMemcpy(&hb, bad.ptr, sizeof(hb), hipMemcpyDeviceToHost));
  if (hb) throw std::invalid_argument("primitive vertex id out of range");
  if (kernel_ms) {
    float ms = 0;
    CS_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
    *kernel_ms = ms;
  }
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
  h.E = im_->E;
  h.nv = im_->nv;
  h.ent_vtx = im_->ent_vtx.ptr;
  h.term = im_->hterm.ptr;
  h.cverts = im_->cverts.ptr;
  h.vsum = im_->hvsum.ptr;
  h.act = im_->act.ptr;
  h.inr = im_->inr.ptr;
  h.coeff = im_->coeff.ptr;
  h.v_off = im_->v_off.ptr;
  h.v_ent = im_->v_ent.ptr;
  return h;
}

}  // namespace cloth
