// lbvh.hip — standalone self-contact broadphase over device arrays.
//
// Computes the candidate set of TriBvh.build + selfPairs (bvh.ts:19-124) with
// the adjacency exclusion of buildAdjacencyExclusions (bvh.ts:133-151): every
// triangle pair tA < tB whose swept boxes sweptTriAabb(x0, x1, pad)
// (aabb.ts:9-33) overlap under the inclusive test of `overlaps` (aabb.ts:35-39)
// and that share no vertex, sorted by (tA, tB).  The set does not depend on
// the hierarchy, so the build is a linear BVH shaped for the GPU rather than
// the reference's recursive median split:
//
//   1. swept FP64 box per triangle; bounds of the box centres (block minima
//      folded into order-preserving u64 atomics);
//   2. 30-bit Morton code of each centre in the cube of the largest extent;
//   3. LSD radix sort of (code, triangle) with LDS ranking (sort.hip);
//   4. leaf records in Morton order (FP64 box, corner ids, triangle);
//   5. Karras hierarchy (one thread per inner node: direction, range end and
//      split by binary searches on the common-prefix length of the sorted
//      codes, ties broken by index), with parent links and, per split
//      position, the escape link (the right child of the node split there);
//   6. bottom-up refit: each leaf climbs its parents, the second arrival at a
//      node (agent-scope acquire/release counter) forms its box, so every
//      inner node is written once; inner boxes are FP32 rounded outward
//      (they only need to contain their leaves), leaf boxes stay FP64;
//   7. stackless traversal, one thread per Morton-ordered leaf i: descend into
//      a node when its box overlaps the query and it holds a leaf after i,
//      otherwise follow the escape link of its last leaf; a leaf j > i that
//      passes the exact FP64 test and shares no vertex is a pair.  Each
//      unordered pair is met once, from its smaller Morton index.  A counting
//      pass, an exclusive scan over triangle ids and an emitting pass place
//      every pair in its tA segment; a last pass sorts each segment by tB, so
//      the output is the reference's sorted list whatever the atomic order.
//
// Inner boxes are conservative (outward rounding; a NaN coordinate is ignored
// by the FP32 min/max, and a leaf with a NaN coordinate fails the exact test
// anyway), so the result is exactly the set of the FP64 predicate.
#include <mutex>

#include "../../include/cloth_hip.h"
#include "lbvh.h"
#include "sort.h"

namespace cloth {

namespace {

struct LbLeaf {  // 64 B
  double lo[3];
  double hi[3];
  uint32_t v[3];
  uint32_t tri;
};

struct LbNode {  // 32 B: two float4 loads
  float lo[3];
  uint32_t left;  // child node id (kLeafBit | j for leaf j)
  float hi[3];
  uint32_t last;  // last Morton index of the node's leaves
};

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLocalBit = 0x80000000u;  // in a parent link: see k_lb_build
constexpr uint32_t kEnd = 0xffffffffu;
constexpr int kScanBlock = 1024;

// misc words
struct LbMisc {
  uint64_t bounds[6];  // okey images: min x,y,z then max x,y,z of the box centres
  uint32_t error;      // bit 0: index out of range; bit 1: traversal or refit bound hit
  uint32_t total;      // pair count (saturates at 0xffffffff)
  uint32_t pad_[2];
};

constexpr size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets of the scratch regions for m triangles.  keys_tmp / vals_tmp
// are free once the sort is done: the refit counters reuse vals_tmp and the
// per-triangle counts and offsets reuse keys_tmp.
struct Layout {
  size_t leaves, nodes, iparent, lparent, esc, keys, keys_tmp, vals, vals_tmp, hist, bsum, misc, total;
};

constexpr Layout layout_for(uint64_t m) {
  Layout L{};
  size_t o = 0;
  const uint64_t mi = m > 1 ? m - 1 : 1;
  L.leaves = o; o += al256(64 * (m ? m : 1));
  L.nodes = o; o += al256(32 * mi);
  L.iparent = o; o += al256(4 * mi);
  L.lparent = o; o += al256(4 * (m ? m : 1));
  L.esc = o; o += al256(4 * mi);
  L.keys = o; o += al256(8 * (m ? m : 1));
  L.keys_tmp = o; o += al256(8 * (m ? m : 1));
  L.vals = o; o += al256(4 * (m ? m : 1));
  L.vals_tmp = o; o += al256(4 * (m ? m : 1));
  L.hist = o; o += al256(4 * 256 * ((m + 4095) / 4096 + 1));
  L.bsum = o; o += al256(4 * ((m + kScanBlock - 1) / kScanBlock + 1));
  L.misc = o; o += al256(sizeof(LbMisc));
  L.total = o;
  return L;
}

__device__ __forceinline__ uint64_t ord_key(double d) {
  const uint64_t u = (uint64_t)__double_as_longlong(d);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double ord_dec(uint64_t k) {
  const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)u);
}

// sweptTriAabb (aabb.ts:9-33): vertices i0, i1, i2, each at x0 then x1,
// `<` / `>` updates from +-Infinity (a NaN coordinate never enters), then pad.
// False when a corner index is out of range (the box is then all NaN).
__device__ __forceinline__ bool swept_box(const double* __restrict__ x0, const double* __restrict__ x1,
                                          const uint32_t* __restrict__ idx, uint32_t n, uint32_t t, double pad,
                                          double lo[3], double hi[3], uint32_t v[3]) {
  for (int k = 0; k < 3; k++) v[k] = idx[(size_t)t * 3 + k];
  if (v[0] >= n || v[1] >= n || v[2] >= n) {
    for (int k = 0; k < 3; k++) lo[k] = hi[k] = __builtin_nan("");
    return false;
  }
  for (int k = 0; k < 3; k++) {
    lo[k] = __builtin_inf();
    hi[k] = -__builtin_inf();
  }
  for (int c = 0; c < 3; c++)
    for (int s = 0; s < 2; s++) {
      const double* p = (s ? x1 : x0) + (size_t)v[c] * 3;
      for (int k = 0; k < 3; k++) {
        const double q = p[k];
        if (q < lo[k]) lo[k] = q;
        if (q > hi[k]) hi[k] = q;
      }
    }
  for (int k = 0; k < 3; k++) {
    lo[k] = lo[k] - pad;
    hi[k] = hi[k] + pad;
  }
  return true;
}

__global__ void k_lb_zero(uint32_t* a, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) a[i] = 0;
}

__global__ void k_lb_reset(LbMisc* misc) {
  if (threadIdx.x < 3) {
    misc->bounds[threadIdx.x] = ~0ull;
    misc->bounds[3 + threadIdx.x] = 0ull;
  }
  if (threadIdx.x == 0) {
    misc->error = 0;
    misc->total = 0;
  }
}

// Bounds of the finite box centres, one atomic per block and axis.
__global__ void __launch_bounds__(kBlock) k_lb_bounds(const double* __restrict__ x0, const double* __restrict__ x1,
                                                      const uint32_t* __restrict__ idx, uint32_t n, uint32_t m,
                                                      double pad, LbMisc* misc) {
  double lo[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
  double hi[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
  bool bad = false;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x) {
    double a[3], b[3];
    uint32_t v[3];
    if (!swept_box(x0, x1, idx, n, t, pad, a, b, v)) {
      bad = true;
      continue;
    }
    for (int k = 0; k < 3; k++) {
      const double c = 0.5 * (a[k] + b[k]);
      if (c < lo[k]) lo[k] = c;
      if (c > hi[k]) hi[k] = c;
    }
  }
  for (int k = 0; k < 3; k++)
    for (int off = kWave / 2; off > 0; off >>= 1) {
      const double ol = __shfl_down(lo[k], off, kWave), oh = __shfl_down(hi[k], off, kWave);
      if (ol < lo[k]) lo[k] = ol;
      if (oh > hi[k]) hi[k] = oh;
    }
  __shared__ double s[6][kBlock / kWave];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0)
    for (int k = 0; k < 3; k++) {
      s[k][wid] = lo[k];
      s[3 + k][wid] = hi[k];
    }
  if (__any(bad) && lane == 0) atomicOr(&misc->error, 1u);
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    double r = s[k][0];
    for (int w = 1; w < kBlock / kWave; w++) r = k < 3 ? (s[k][w] < r ? s[k][w] : r) : (s[k][w] > r ? s[k][w] : r);
    if (r == r && r > -__builtin_inf() && r < __builtin_inf()) {
      if (k < 3) atomicMin((unsigned long long*)&misc->bounds[k], (unsigned long long)ord_key(r));
      else atomicMax((unsigned long long*)&misc->bounds[k], (unsigned long long)ord_key(r));
    }
  }
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// 30-bit Morton code of each box centre in the cube spanned by the largest
// extent of the centre bounds (10 bits per axis, x highest).
__global__ void __launch_bounds__(kBlock) k_lb_morton(const double* __restrict__ x0, const double* __restrict__ x1,
                                                      const uint32_t* __restrict__ idx, uint32_t n, uint32_t m,
                                                      double pad, const LbMisc* __restrict__ misc, uint64_t* keys,
                                                      uint32_t* vals) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  double lo3[3], ext = 0;
  const bool any = misc->bounds[0] != ~0ull;
  for (int k = 0; k < 3; k++) {
    lo3[k] = any ? ord_dec(misc->bounds[k]) : 0.0;
    const double e = any ? ord_dec(misc->bounds[3 + k]) - lo3[k] : 0.0;
    if (e > ext) ext = e;
  }
  double a[3], b[3];
  uint32_t v[3];
  uint32_t q[3] = {0, 0, 0};
  if (swept_box(x0, x1, idx, n, t, pad, a, b, v))
    for (int k = 0; k < 3; k++) {
      double u = ext > 0 ? (0.5 * (a[k] + b[k]) - lo3[k]) / ext : 0.0;
      if (!(u >= 0)) u = 0;  // NaN and below
      if (u > 1) u = 1;
      q[k] = (uint32_t)fmin(1023.0, u * 1024.0);
    }
  keys[t] = ((uint64_t)spread10(q[0]) << 2) | ((uint64_t)spread10(q[1]) << 1) | (uint64_t)spread10(q[2]);
  vals[t] = t;
}

// Leaf records in Morton order; the refit counters are cleared here.
__global__ void __launch_bounds__(kBlock) k_lb_leaves(const double* __restrict__ x0, const double* __restrict__ x1,
                                                      const uint32_t* __restrict__ idx, uint32_t n, uint32_t m,
                                                      double pad, const uint32_t* __restrict__ order, LbLeaf* leaves,
                                                      uint32_t* flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  LbLeaf R;
  const uint32_t t = order[i];
  swept_box(x0, x1, idx, n, t, pad, R.lo, R.hi, R.v);
  R.tri = t;
  leaves[i] = R;
  if (i + 1 < m) flags[i] = 0;
}

// Common-prefix length of sorted codes i and j (Karras 2012, with ties broken
// by index); -1 outside [0, m).
__device__ __forceinline__ int lb_delta(const uint64_t* __restrict__ keys, uint32_t m, int64_t i, int64_t j) {
  if (j < 0 || j >= (int64_t)m) return -1;
  const uint32_t a = (uint32_t)keys[i], b = (uint32_t)keys[j];
  if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
  return __clz(a ^ b);
}

// Karras hierarchy: inner node i of m - 1 covers the Morton range [first,
// last] that has i as one end; its split gamma sends [first, gamma] left and
// [gamma + 1, last] right (a child that is a single index is that leaf).
__global__ void __launch_bounds__(kBlock) k_lb_build(const uint64_t* __restrict__ keys, uint32_t m, LbNode* nodes,
                                                     uint32_t* iparent, uint32_t* lparent, uint32_t* esc) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= m) return;
  const int64_t ii = i;
  const int dl = lb_delta(keys, m, ii, ii + 1), dr = lb_delta(keys, m, ii, ii - 1);
  const int d = dl > dr ? 1 : -1;
  const int dmin = lb_delta(keys, m, ii, ii - d);
  // range end: exponential then binary search
  int64_t lmax = 2;
  while (lb_delta(keys, m, ii, ii + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (lb_delta(keys, m, ii, ii + (l + t) * d) > dmin) l += t;
  const int64_t j = ii + l * d;
  const int dnode = lb_delta(keys, m, ii, j);
  // split: binary search for the last index sharing more than dnode bits with i
  int64_t s = 0;
  for (int64_t div = 2;; div *= 2) {
    const int64_t t = (l + div - 1) / div;
    if (lb_delta(keys, m, ii, ii + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  const uint32_t gamma = (uint32_t)(ii + s * d + (d < 0 ? -1 : 0));
  const uint32_t first = (uint32_t)(ii < j ? ii : j), last = (uint32_t)(ii < j ? j : ii);
  const uint32_t left = first == gamma ? (kLeafBit | gamma) : gamma;
  const uint32_t right = last == gamma + 1 ? (kLeafBit | (gamma + 1)) : gamma + 1;
  nodes[i].left = left;
  nodes[i].last = last;
  // the parent links carry kLocalBit when node i's leaves all belong to one
  // refit work-group (so both of its climbers run in that work-group)
  const uint32_t link = i | ((first / kBlock == last / kBlock) ? kLocalBit : 0u);
  if (left & kLeafBit) lparent[gamma] = link;
  else iparent[gamma] = link;
  if (right & kLeafBit) lparent[gamma + 1] = link;
  else iparent[gamma + 1] = link;
  esc[gamma] = right;
}

template <int kScope>
__device__ __forceinline__ void child_box(const LbLeaf* __restrict__ leaves, const LbNode* nodes, uint32_t c,
                                          float lo[3], float hi[3]) {
  if (c & kLeafBit) {
    const LbLeaf& L = leaves[c & ~kLeafBit];
    for (int k = 0; k < 3; k++) {
      lo[k] = __double2float_rd(L.lo[k]);
      hi[k] = __double2float_ru(L.hi[k]);
    }
  } else {
    for (int k = 0; k < 3; k++) {
      lo[k] = __hip_atomic_load(&nodes[c].lo[k], __ATOMIC_RELAXED, kScope);
      hi[k] = __hip_atomic_load(&nodes[c].hi[k], __ATOMIC_RELAXED, kScope);
    }
  }
}

// Node p's box from its children's, once both have arrived: the first
// arrival at p's counter stops, the second (which sees the sibling's box
// through the counter's release / acquire) writes the union.  Returns whether
// this thread continues upwards.
template <int kScope>
__device__ __forceinline__ bool refit_node(const LbLeaf* __restrict__ leaves, LbNode* nodes, uint32_t p,
                                           uint32_t* flags) {
  if (__hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_ACQ_REL, kScope) == 0) return false;
  const uint32_t left = nodes[p].left, last = nodes[p].last;
  const uint32_t g = left & ~kLeafBit;
  const uint32_t right = last == g + 1 ? (kLeafBit | (g + 1)) : g + 1;
  float a[3], b[3], c[3], d[3];
  child_box<kScope>(leaves, nodes, left, a, b);
  child_box<kScope>(leaves, nodes, right, c, d);
  for (int k = 0; k < 3; k++) {
    __hip_atomic_store(&nodes[p].lo[k], fminf(a[k], c[k]), __ATOMIC_RELAXED, kScope);
    __hip_atomic_store(&nodes[p].hi[k], fmaxf(b[k], d[k]), __ATOMIC_RELAXED, kScope);
  }
  return true;
}

// Bottom-up refit: leaf j climbs from its parent.  A node whose leaves all
// belong to this work-group synchronises at work-group scope; a node spanning
// work-groups at agent scope, whose release also publishes the boxes this
// thread wrote below it.  Depth is at most 30 + log2(m) + 1.
__global__ void __launch_bounds__(kBlock) k_lb_refit(const LbLeaf* __restrict__ leaves, uint32_t m, LbNode* nodes,
                                                     const uint32_t* __restrict__ iparent,
                                                     const uint32_t* __restrict__ lparent, uint32_t* flags,
                                                     LbMisc* misc) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m || m < 2) return;
  uint32_t link = lparent[j];
  for (int depth = 0;; depth++) {
    const uint32_t p = link & ~kLocalBit;
    if (depth >= 128 || p + 1 >= m) {
      atomicOr(&misc->error, 2u);
      return;
    }
    const bool up = (link & kLocalBit) ? refit_node<__HIP_MEMORY_SCOPE_WORKGROUP>(leaves, nodes, p, flags)
                                       : refit_node<__HIP_MEMORY_SCOPE_AGENT>(leaves, nodes, p, flags);
    if (!up || p == 0) return;
    link = iparent[p];
  }
}

__device__ __forceinline__ bool overlap64(const double* a, const double* b) {
  // overlaps (aabb.ts:35-39), a = {lo[3], hi[3]}
  return a[0] <= b[3] && a[3] >= b[0] && a[1] <= b[4] && a[4] >= b[1] && a[2] <= b[5] && a[5] >= b[2];
}

// Traversal from Morton leaf i; kEmit = false counts pairs per tA, true
// writes them to their tA segments.  Every step moves forward in the
// hierarchy's pre-order, so at most 2m - 1 nodes are visited.
template <bool kEmit>
__global__ void __launch_bounds__(kBlock) k_lb_query(const LbLeaf* __restrict__ leaves,
                                                     const LbNode* __restrict__ nodes,
                                                     const uint32_t* __restrict__ esc, uint32_t m, uint32_t* cnt,
                                                     const uint32_t* __restrict__ off, uint2* out, LbMisc* misc) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const LbLeaf Q = leaves[i];
  const double q[6] = {Q.lo[0], Q.lo[1], Q.lo[2], Q.hi[0], Q.hi[1], Q.hi[2]};
  float qf[6];
  for (int k = 0; k < 3; k++) {
    qf[k] = __double2float_rd(Q.lo[k]);
    qf[3 + k] = __double2float_ru(Q.hi[k]);
  }
  uint32_t node = m == 1 ? kLeafBit : 0u;
  const uint64_t bound = 2 * (uint64_t)m;
  for (uint64_t steps = 0; node != kEnd; steps++) {
    if (steps > bound) {
      atomicOr(&misc->error, 2u);
      return;
    }
    uint32_t last;
    if (node & kLeafBit) {
      const uint32_t j = node & ~kLeafBit;
      if (j >= m) {
        atomicOr(&misc->error, 2u);
        return;
      }
      last = j;
      if (j > i) {
        const LbLeaf& R = leaves[j];
        const double r[6] = {R.lo[0], R.lo[1], R.lo[2], R.hi[0], R.hi[1], R.hi[2]};
        const bool share = Q.v[0] == R.v[0] || Q.v[0] == R.v[1] || Q.v[0] == R.v[2] || Q.v[1] == R.v[0] ||
                           Q.v[1] == R.v[1] || Q.v[1] == R.v[2] || Q.v[2] == R.v[0] || Q.v[2] == R.v[1] ||
                           Q.v[2] == R.v[2];
        if (!share && overlap64(q, r)) {
          const uint32_t ta = Q.tri < R.tri ? Q.tri : R.tri, tb = Q.tri < R.tri ? R.tri : Q.tri;
          const uint32_t k = atomicAdd(&cnt[ta], 1u);
          if (kEmit) out[(size_t)off[ta] + k] = make_uint2(ta, tb);
        }
      }
    } else {
      if (node + 1 >= m) {
        atomicOr(&misc->error, 2u);
        return;
      }
      const float4 w0 = *reinterpret_cast<const float4*>(&nodes[node].lo[0]);
      const float4 w1 = *reinterpret_cast<const float4*>(&nodes[node].hi[0]);
      last = __float_as_uint(w1.w);
      if (last > i && qf[0] <= w1.x && qf[3] >= w0.x && qf[1] <= w1.y && qf[4] >= w0.y && qf[2] <= w1.z &&
          qf[5] >= w0.z) {
        node = __float_as_uint(w0.w);
        continue;
      }
    }
    node = last + 1 >= m ? kEnd : esc[last];
  }
}

// Exclusive scan of cnt[0..m) into off, total into misc->total; cnt is
// cleared for reuse as the emit cursor.  Three phases over 1024-entry blocks.
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t* s_warp) {
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t u = __shfl_up(v, o, kWave);
    if (lane >= o) v += u;
  }
  if (lane == kWave - 1) s_warp[wid] = v;
  __syncthreads();
  if (wid == 0) {
    uint32_t w = lane < kScanBlock / kWave ? s_warp[lane] : 0;
    for (int o = 1; o < kScanBlock / kWave; o <<= 1) {
      const uint32_t u = __shfl_up(w, o, kWave);
      if (lane >= o) w += u;
    }
    if (lane < kScanBlock / kWave) s_warp[lane] = w;
  }
  __syncthreads();
  const uint32_t r = v + (wid ? s_warp[wid - 1] : 0);
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kScanBlock) k_lb_scan_sums(const uint32_t* __restrict__ cnt, uint32_t m,
                                                             uint32_t* bsum) {
  __shared__ uint32_t s_warp[kScanBlock / kWave];
  const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
  const uint32_t v = i < m ? cnt[i] : 0;
  const uint32_t incl = block_incl_scan(v, s_warp);
  if (threadIdx.x == kScanBlock - 1) bsum[blockIdx.x] = incl;
}

__global__ void __launch_bounds__(kScanBlock) k_lb_scan_top(uint32_t* bsum, uint32_t nb, LbMisc* misc) {
  __shared__ uint32_t s_warp[kScanBlock / kWave];
  uint64_t carry = 0;
  for (uint32_t base = 0; base < nb; base += kScanBlock) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < nb ? bsum[i] : 0;
    const uint32_t incl = block_incl_scan(v, s_warp);
    if (i < nb) bsum[i] = (uint32_t)(carry + incl - v);
    __shared__ uint32_t s_last;
    if (threadIdx.x == kScanBlock - 1) s_last = incl;
    __syncthreads();
    carry += s_last;
    __syncthreads();
  }
  if (threadIdx.x == 0) misc->total = carry > 0xffffffffull ? 0xffffffffu : (uint32_t)carry;
}

__global__ void __launch_bounds__(kScanBlock) k_lb_scan_apply(uint32_t* cnt, uint32_t m,
                                                              const uint32_t* __restrict__ bsum, uint32_t* off) {
  __shared__ uint32_t s_warp[kScanBlock / kWave];
  const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
  const uint32_t v = i < m ? cnt[i] : 0;
  const uint32_t incl = block_incl_scan(v, s_warp);
  if (i < m) {
    off[i] = bsum[blockIdx.x] + incl - v;
    cnt[i] = 0;
  }
}

// Each tA segment sorted by tB (insertion sort; segments are short).
__global__ void __launch_bounds__(kBlock) k_lb_seg_sort(const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ cnt, uint32_t m, uint2* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const size_t b = off[t], e = b + cnt[t];
  for (size_t a = b + 1; a < e; a++) {
    const uint2 key = out[a];
    size_t c = a;
    while (c > b && out[c - 1].y > key.y) {
      out[c] = out[c - 1];
      c--;
    }
    out[c] = key;
  }
}

struct Ptrs {
  LbLeaf* leaves;
  LbNode* nodes;
  uint32_t *iparent, *lparent, *esc, *vals, *vals_tmp, *hist, *bsum, *cnt, *off;
  uint64_t *keys, *keys_tmp;
  LbMisc* misc;
};

Ptrs carve(unsigned char* base, uint32_t m) {
  const Layout L = layout_for(m);
  Ptrs P;
  P.leaves = reinterpret_cast<LbLeaf*>(base + L.leaves);
  P.nodes = reinterpret_cast<LbNode*>(base + L.nodes);
  P.iparent = reinterpret_cast<uint32_t*>(base + L.iparent);
  P.lparent = reinterpret_cast<uint32_t*>(base + L.lparent);
  P.esc = reinterpret_cast<uint32_t*>(base + L.esc);
  P.keys = reinterpret_cast<uint64_t*>(base + L.keys);
  P.keys_tmp = reinterpret_cast<uint64_t*>(base + L.keys_tmp);
  P.vals = reinterpret_cast<uint32_t*>(base + L.vals);
  P.vals_tmp = reinterpret_cast<uint32_t*>(base + L.vals_tmp);
  P.hist = reinterpret_cast<uint32_t*>(base + L.hist);
  P.bsum = reinterpret_cast<uint32_t*>(base + L.bsum);
  P.misc = reinterpret_cast<LbMisc*>(base + L.misc);
  // after the sort: per-triangle counts and offsets in keys_tmp
  P.cnt = reinterpret_cast<uint32_t*>(P.keys_tmp);
  P.off = P.cnt + m;
  return P;
}

// The whole query on `st`.  Returns false (and *count) when the pairs do not
// fit `cap`; throws on an out-of-range index or a failed bound.
bool run_self_pairs(hipStream_t st, unsigned char* base, const double* x0, const double* x1, uint32_t n,
                    const uint32_t* idx, uint32_t m, double pad, uint32_t* pairs, uint64_t cap, uint64_t* count,
                    hipEvent_t* ev) {
  Ptrs P = carve(base, m);
  if (ev) CS_HIP(hipEventRecord(ev[0], st));
  hipLaunchKernelGGL(k_lb_reset, dim3(1), dim3(64), 0, st, P.misc);
  const unsigned gb = std::min<unsigned>(grid_for(m), 1024);
  hipLaunchKernelGGL(k_lb_bounds, dim3(gb), dim3(kBlock), 0, st, x0, x1, idx, n, m, pad, P.misc);
  hipLaunchKernelGGL(k_lb_morton, dim3(grid_for(m)), dim3(kBlock), 0, st, x0, x1, idx, n, m, pad, P.misc, P.keys,
                     P.vals);
  CS_LAUNCH_CHECK();
  SortScratch ss;
  ss.hist = P.hist;
  ss.hist_cap = radix_hist_entries(m);
  const bool in_tmp = radix_sort_pairs(st, P.keys, P.vals, P.keys_tmp, P.vals_tmp, m, 0, 30, ss);
  const uint64_t* keys = in_tmp ? P.keys_tmp : P.keys;
  const uint32_t* order = in_tmp ? P.vals_tmp : P.vals;
  if (in_tmp) throw HipError("LBVH: unexpected sort parity");  // 30 bits = 4 passes, result in place
  hipLaunchKernelGGL(k_lb_leaves, dim3(grid_for(m)), dim3(kBlock), 0, st, x0, x1, idx, n, m, pad, order, P.leaves,
                     P.vals_tmp);
  hipLaunchKernelGGL(k_lb_build, dim3(grid_for(m)), dim3(kBlock), 0, st, keys, m, P.nodes, P.iparent, P.lparent,
                     P.esc);
  hipLaunchKernelGGL(k_lb_refit, dim3(grid_for(m)), dim3(kBlock), 0, st, P.leaves, m, P.nodes, P.iparent,
                     P.lparent, P.vals_tmp, P.misc);
  CS_LAUNCH_CHECK();
  if (ev) CS_HIP(hipEventRecord(ev[1], st));
  // counts per triangle (keys_tmp is free now)
  hipLaunchKernelGGL(k_lb_zero, dim3(grid_for(m)), dim3(kBlock), 0, st, P.cnt, m);
  hipLaunchKernelGGL(k_lb_query<false>, dim3(grid_for(m)), dim3(kBlock), 0, st, P.leaves, P.nodes, P.esc, m, P.cnt,
                     nullptr, nullptr, P.misc);
  const unsigned nb = (unsigned)((m + kScanBlock - 1) / kScanBlock);
  hipLaunchKernelGGL(k_lb_scan_sums, dim3(nb), dim3(kScanBlock), 0, st, P.cnt, m, P.bsum);
  hipLaunchKernelGGL(k_lb_scan_top, dim3(1), dim3(kScanBlock), 0, st, P.bsum, nb, P.misc);
  hipLaunchKernelGGL(k_lb_scan_apply, dim3(nb), dim3(kScanBlock), 0, st, P.cnt, m, P.bsum, P.off);
  CS_LAUNCH_CHECK();
  LbMisc hm;
  CS_HIP(hipMemcpyAsync(&hm, P.misc, sizeof(hm), hipMemcpyDeviceToHost, st));
  CS_HIP(hipStreamSynchronize(st));
  if (hm.error & 1u) throw std::invalid_argument("triangle index out of range");
  if (hm.error & 2u) throw HipError("LBVH: hierarchy bound exceeded");
  *count = hm.total;
  if (hm.total == 0xffffffffu || hm.total > cap) {
    if (ev) {
      CS_HIP(hipEventRecord(ev[2], st));
      CS_HIP(hipEventSynchronize(ev[2]));
    }
    return false;
  }
  uint2* out = reinterpret_cast<uint2*>(pairs);
  hipLaunchKernelGGL(k_lb_query<true>, dim3(grid_for(m)), dim3(kBlock), 0, st, P.leaves, P.nodes, P.esc, m, P.cnt,
                     P.off, out, P.misc);
  hipLaunchKernelGGL(k_lb_seg_sort, dim3(grid_for(m)), dim3(kBlock), 0, st, P.off, P.cnt, m, out);
  CS_LAUNCH_CHECK();
  if (ev) CS_HIP(hipEventRecord(ev[2], st));
  CS_HIP(hipStreamSynchronize(st));
  return true;
}

constexpr size_t kNoAllocBytes = layout_for(kLbvhMaxTris).total;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Scratch of the allocation-free entry point: one copy per GPU (each device
// holds its own instance at its own address), for up to kLbvhMaxTris
// triangles.  Host code never passes this name to a kernel: it passes the
// device address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

size_t lbvh_scratch_bytes(uint32_t m) { return layout_for(m).total; }
size_t lbvh_noalloc_bytes() { return kNoAllocBytes; }

bool lbvh_self_pairs(const double* x0, const double* x1, uint32_t n, const uint32_t* idx, uint32_t m, double pad,
                     uint32_t* pairs, uint64_t cap, uint64_t* count, double* build_ms, double* query_ms) {
  *count = 0;
  if (m == 0) return true;
  hipStream_t st = nullptr;
  DevBuf<unsigned char> buf;
  buf.alloc(layout_for(m).total);
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  const bool timed = build_ms || query_ms;
  if (timed)
    for (auto& e : ev) CS_HIP(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int k = 0; k < 3; k++)
        if (e[k]) (void)hipEventDestroy(e[k]);
    }
  } guard{ev};
  const bool ok = run_self_pairs(st, buf.ptr, x0, x1, n, idx, m, pad, pairs, cap, count, timed ? ev : nullptr);
  if (timed) {
    float a = 0, b = 0;
    CS_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
    CS_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
    if (build_ms) *build_ms = a;
    if (query_ms) *query_ms = b;
  }
  return ok;
}

bool lbvh_self_pairs_noalloc(int device, const double* x0, const double* x1, uint32_t n, const uint32_t* idx,
                             uint32_t m, double pad, uint32_t* pairs, uint64_t cap, uint64_t* count) {
  *count = 0;
  if (m > kLbvhMaxTris)
    throw std::invalid_argument("n_tris exceeds the allocation-free scratch (at most " +
                                std::to_string(kLbvhMaxTris) + " triangles)");
  if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device ordinal out of range");
  void* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
      // the array's device address on the current device (each GPU holds its own copy)
      CS_HIP(hipGetSymbolAddress(&g_scratch_ptr[device], HIP_SYMBOL(scratch)));
    }
    ptr = g_scratch_ptr[device];
  }
  if (m == 0) return true;
  return run_self_pairs(nullptr, static_cast<unsigned char*>(ptr), x0, x1, n, idx, m, pad, pairs, cap, count,
                        nullptr);
}

}  // namespace cloth
