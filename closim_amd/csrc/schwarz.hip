// schwarz.hip — one-level additive Schwarz (schwarz.ts:44-309) on the device.
//
// Domains: built once on the host exactly as buildSchwarzDomains does
// (Morton-ordered seeds over the rest positions, BFS growth along mesh
// connectivity to the target size, merge of sub-minimum domains into the
// smallest adjacent one, ids compacted in creation order, members sorted).
// Factors (per Newton iteration): one work-group per domain assembles the
// dense local Hessian row by row (each thread owns the rows of one member
// vertex, so every entry receives its triangle contributions in ascending
// triangle order, as in the reference) and embeds the Jacobi diagonal; a
// second kernel symmetrises while loading the lower triangle into LDS and
// runs the reference's left-looking Cholesky there (every sum in the
// reference's order, rows below the pivot in parallel), then stores the
// factor packed by columns (FP64, and FP32 for the mixed solve).  Apply: one
// wave per domain, each column read contiguously.  The FP64 apply keeps the
// reference's operation order (forward substitution column by column, which
// subtracts in the same order as the row loop; back substitution row by row,
// products by the wave, subtractions in order by one lane); the FP32 apply
// substitutes forward by columns and backward by wave dot products.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "elements.h"
#include "schwarz.h"
#include "hessian.h"

namespace cloth {

struct SchwarzImpl {
  hipStream_t st;
  uint32_t n = 0, D = 0, dmax = 0, DW = 0;
  std::vector<uint32_t> h_domain_of, h_local_of;
  DevBuf<uint32_t> members, size, domain_of, local_of;
  DevBuf<double> A;      // assembled local Hessians, D * DW * DW, row stride dw of each domain
  DevBuf<float> A32;     // the same in FP32 for the mixed solve's factor (domains of at most 12 blocks)
  DevBuf<double> L;      // factors, D * T (T = DW (DW + 1) / 2), packed lower triangle by columns
  DevBuf<float> L32;     // FP32 copy of L for the mixed-precision apply (reciprocal diagonal)
  DevBuf<float> L32T;    // the same packed by rows, for its back substitution
  DevBuf<float> T32;     // FP32 factor as 16x16 tiles (domains of at most 12 blocks), see k_sw_apply32t
  uint32_t nbt = 0;      // 16-row blocks of the widest domain
  DevBuf<double> W;      // factorisation workspace when the triangle does not fit in LDS
  DevBuf<uint8_t> flag;  // 1 = Cholesky, 0 = Jacobi fallback
  size_t lds_max = 0;    // dynamic LDS the factor kernel may use
  std::unique_ptr<Coarse> coarse;  // two-level: the coarse level
};

// ---------------------------------------------------------------- host

// morton3 (schwarz.ts:27-36): 10 bits per axis, x in the top bit of each triple.
static uint32_t sw_morton(double x, double y, double z) {
  auto q = [](double v) -> int {
    // Math.max(0, Math.min(1023, Math.floor(v * 1024))); NaN becomes 0 in
    // the bit operations that follow
    const double f = jsm::jmax(0.0, jsm::jmin(1023.0, std::floor(v * 1024)));
    return f == f ? (int)f : 0;
  };
  const int xi = q(x), yi = q(y), zi = q(z);
  uint32_t m = 0;
  for (int b = 9; b >= 0; b--)
    m = (m << 3) | (uint32_t)((((xi >> b) & 1) << 2) | (((yi >> b) & 1) << 1) | ((zi >> b) & 1));
  return m;
}

static void build_domains(const cs_scene_desc& d, int target, int min_size, std::vector<std::vector<uint32_t>>& out,
                          std::vector<uint32_t>& domain_of, std::vector<uint32_t>& local_of) {
  const uint32_t n = d.n_verts, m = d.n_tris;
  const uint32_t* idx = d.indices;
  std::vector<std::vector<uint32_t>> adj(n);
  for (uint32_t t = 0; t < m; t++) {
    const uint32_t a = idx[t * 3], b = idx[t * 3 + 1], c = idx[t * 3 + 2];
    adj[a].push_back(b); adj[a].push_back(c);
    adj[b].push_back(a); adj[b].push_back(c);
    adj[c].push_back(a); adj[c].push_back(b);
  }
  for (auto& l : adj) {
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
  }
  auto rest = [&](uint32_t v, int k) -> double {
    return d.rest_positions ? (double)d.rest_positions[(size_t)v * 3 + k] : d.positions[(size_t)v * 3 + k];
  };
  // bounds with Math.min / Math.max (NaN propagates, as in the reference)
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < n; v++)
    for (int k = 0; k < 3; k++) {
      mn[k] = jsm::jmin(mn[k], rest(v, k));
      mx[k] = jsm::jmax(mx[k], rest(v, k));
    }
  double sc[3];
  for (int k = 0; k < 3; k++) sc[k] = jsm::jmax(mx[k] - mn[k], 1e-12);
  std::vector<uint32_t> code(n), order(n);
  for (uint32_t v = 0; v < n; v++) {
    order[v] = v;
    code[v] = sw_morton((rest(v, 0) - mn[0]) / sc[0], (rest(v, 1) - mn[1]) / sc[1], (rest(v, 2) - mn[2]) / sc[2]);
  }
  std::sort(order.begin(), order.end(),
            [&](uint32_t a, uint32_t b) { return code[a] != code[b] ? code[a] < code[b] : a < b; });
  const uint32_t NONE = 0xffffffffu;
  std::vector<uint32_t> dom(n, NONE);
  std::vector<std::vector<uint32_t>> mem;
  std::vector<uint32_t> queue;
  for (uint32_t seed : order) {
    if (dom[seed] != NONE) continue;
    const uint32_t id = (uint32_t)mem.size();
    std::vector<uint32_t> cur{seed};
    dom[seed] = id;
    queue.assign(1, seed);
    size_t head = 0;
    while ((int)cur.size() < target && head < queue.size()) {
      const uint32_t u = queue[head++];
      for (uint32_t w : adj[u]) {
        if ((int)cur.size() >= target) break;
        if (dom[w] != NONE) continue;
        dom[w] = id;
        cur.push_back(w);
        queue.push_back(w);
      }
    }
    mem.push_back(std::move(cur));
  }
  // merge sub-minimum domains into the smallest adjacent domain
  for (;;) {
    long small = -1;
    for (size_t k = 0; k < mem.size(); k++)
      if (!mem[k].empty() && (int)mem[k].size() < min_size && (int)n >= min_size)
        if (small < 0 || mem[k].size() < mem[small].size()) small = (long)k;
    if (small < 0) break;
    std::vector<uint32_t> nb;
    for (uint32_t t = 0; t < m; t++) {
      const uint32_t ds[3] = {dom[idx[t * 3]], dom[idx[t * 3 + 1]], dom[idx[t * 3 + 2]]};
      if (ds[0] != (uint32_t)small && ds[1] != (uint32_t)small && ds[2] != (uint32_t)small) continue;
      for (uint32_t b : ds)
        if (b != (uint32_t)small) nb.push_back(b);
    }
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    long best = -1;
    for (uint32_t c : nb) {
      if (mem[c].empty()) continue;
      if (best < 0 || mem[c].size() < mem[best].size()) best = (long)c;
    }
    if (best < 0) break;
    for (uint32_t v : mem[small]) {
      dom[v] = (uint32_t)best;
      mem[best].push_back(v);
    }
    mem[small].clear();
  }
  std::vector<uint32_t> remap(mem.size(), NONE);
  out.clear();
  for (size_t k = 0; k < mem.size(); k++) {
    if (mem[k].empty()) continue;
    remap[k] = (uint32_t)out.size();
    std::vector<uint32_t> s = mem[k];
    std::sort(s.begin(), s.end());
    out.push_back(std::move(s));
  }
  domain_of.assign(n, 0);
  local_of.assign(n, 0);
  for (uint32_t v = 0; v < n; v++) domain_of[v] = remap[dom[v]];
  for (auto& s : out)
    for (size_t l = 0; l < s.size(); l++) local_of[s[l]] = (uint32_t)l;
}

// ---------------------------------------------------------------- device

constexpr int kSwBlock = 128;

struct SwArgs {
  uint32_t D, dmax, DW;
  const uint32_t* members;
  const uint32_t* size;
  const uint32_t* domain_of;
  const uint32_t* local_of;
  double* A;
  double* L;
  float* L32;
  float* L32T;
  float* T32;
  float* A32;
  uint32_t nbt;
  double* W;
  uint8_t* flag;
};

// The FP32 factor of a domain of at most 12 blocks of 16 rows as 16x16 tiles
// (a, b), a >= b, ordered by block columns: tile (a, b) starts at
// tile_index(nbt, a, b) * 256 floats of the domain's nbt (nbt + 1) / 2 tiles.
// An off-diagonal tile holds L_ab by columns (entry (i, j) at j * 16 + i);
// a diagonal tile holds the inverse of the lower triangle L_aa by rows
// (entry (i, j) at i * 16 + j).  Rows beyond the domain are zero.
__host__ __device__ __forceinline__ uint32_t tile_count(uint32_t nbt) { return nbt * (nbt + 1) / 2; }
__host__ __device__ __forceinline__ uint32_t tile_index(uint32_t nbt, uint32_t a, uint32_t b) {
  return b * nbt - b * (b - 1) / 2 + (a - b);
}

__device__ __forceinline__ double* dom_mat(const SwArgs& s, uint32_t d) { return s.A + (size_t)d * s.DW * s.DW; }
__host__ __device__ __forceinline__ size_t tri_size(uint32_t dw) { return (size_t)dw * (dw + 1) / 2; }
// column j of a packed lower triangle stored by columns (rows j..dw-1)
__device__ __forceinline__ size_t col_off(uint32_t dw, uint32_t j) { return (size_t)j * dw - (size_t)j * (j - 1) / 2; }
// row i of a packed lower triangle stored by rows (columns 0..i)
__device__ __forceinline__ size_t row_off(uint32_t i) { return (size_t)i * (i + 1) / 2; }

// Local Hessian of domain d: membrane 3x3 blocks of intra-domain corner
// pairs and the Jacobi diagonal (schwarz.ts:170-203); the symmetrisation
// (:204-210) happens as the factor kernel loads it.
constexpr int kSwAsmBlock = 256;

// The local Hessian of domain d into Ad (FP64, or FP32 for the mixed
// solve's factor, which accumulates in FP32 and, with kUpper, only the
// entries on or right of the diagonal: the FP32 factor reads A_ki, i >= k).
template <typename AT, bool kUpper>
__global__ void __launch_bounds__(kSwAsmBlock) k_sw_assemble(SwArgs s, DMesh me, DMat mat,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ diag, AT* Abase) {
  const uint32_t d = blockIdx.x;
  const uint32_t dv = s.size[d], dw = dv * 3;
  AT* A = Abase + (size_t)d * s.DW * s.DW;
  // clear the matrix with consecutive threads on consecutive entries
  for (uint32_t e = threadIdx.x; e < dw * dw; e += blockDim.x) A[e] = 0;
  __syncthreads();
  // one thread per row (vertex l = r / 3, component ra = r % 3): it evaluates
  // its row of every corner Hessian of the vertex and adds it in corner
  // order, then the diagonal
  for (uint32_t r = threadIdx.x; r < dw; r += blockDim.x) {
    const uint32_t l = r / 3;
    const int ra = (int)(r % 3);
    const uint32_t v = s.members[(size_t)d * s.dmax + l];
    AT* Ar = A + (size_t)r * dw;
    for (uint32_t q = me.vt_off[v]; q < me.vt_off[v + 1]; q++) {
      const uint4 cr = me.vt_corner[q];
      const uint32_t t = cr.w >> 2;
      const int ci = cr.w & 3;
      const uint32_t cvs[3] = {cr.x, cr.y, cr.z};
      double X[3][3];
      for (int k = 0; k < 3; k++) ld3(x, cvs[k], X[k]);
      double h9[9];
      tri_hessian_rows_range(X[0], X[1], X[2], me.inv_dm + (size_t)t * 4, me.area[t], mat, ci, ra, ra + 1, h9);
      for (int cj = 0; cj < 3; cj++) {
        const uint32_t w = cvs[cj];
        if (s.domain_of[w] != d) continue;
        const uint32_t lw = s.local_of[w];
        for (int b = 0; b < 3; b++)
          if (!kUpper || lw * 3 + b >= r) Ar[lw * 3 + b] += (AT)h9[cj * 3 + b];
      }
    }
    Ar[r] += (AT)diag[(size_t)v * 3 + ra];
  }
}

// The FP32 factor of the mixed solve: packed by columns (forward
// substitution) and by rows (back substitution), diagonal as reciprocal.
__device__ __forceinline__ void store32(const SwArgs& s, uint32_t d, uint32_t dw, uint32_t i, uint32_t k, double v) {
  const float f = i == k ? (float)(1.0 / v) : (float)v;
  s.L32[(size_t)d * tri_size(s.DW) + col_off(dw, k) + (i - k)] = f;
  s.L32T[(size_t)d * tri_size(s.DW) + row_off(i) + k] = f;
}

constexpr int kSwFactorBlock = 256;

// Symmetrise (schwarz.ts:204-210), then the reference's left-looking
// Cholesky (:212-235) or its Jacobi fallback (:236-249), one work-group per
// domain.  The triangle being factored lives in LDS (kLds) or in the W
// workspace, packed by columns, which is also the layout the substitutions
// read: at step k thread i reads its row's L_it (consecutive threads,
// consecutive addresses) and the broadcast L_kt.  Thread t owns rows t and
// t + 256.  Row k's pivot sum  A_kk - sum_s L_ks^2  is kept by its owner as
// a running value, updated right after each L_ks is computed, i.e. in the
// reference's order s = 0, 1, ..., so that the pivot is ready when column k
// starts.
template <bool kLds>
__global__ void __launch_bounds__(kSwFactorBlock) k_sw_factor(SwArgs s, const double* __restrict__ diag,
                                                              int want32) {
  extern __shared__ double sw_lds[];
  __shared__ int ok;
  __shared__ double tiny;
  const uint32_t d = blockIdx.x;
  const uint32_t dv = s.size[d], dw = dv * 3;
  const double* A = dom_mat(s, d);
  double* T = kLds ? sw_lds : s.W + (size_t)d * tri_size(s.DW);
  constexpr uint32_t kRows = 2;  // rows per thread: dw <= 2 * kSwFactorBlock (checked by the host)
  // load the lower triangle, symmetrised: s = 0.5 (A_cr + A_rc)
  for (uint32_t c = 0; c < dw; c++)
    for (uint32_t r = c + threadIdx.x; r < dw; r += blockDim.x)
      T[col_off(dw, c) + (r - c)] =
          r == c ? A[(size_t)r * dw + r] : 0.5 * (A[(size_t)c * dw + r] + A[(size_t)r * dw + c]);
  __syncthreads();
  if (threadIdx.x == 0) {
    double scale = 0;
    for (uint32_t i = 0; i < dw; i++) scale = jsm::jmax(scale, jsm::fabs_(T[col_off(dw, i)]));
    tiny = 1e-20 * (1 + scale);
    ok = 1;
  }
  double piv[kRows];
#pragma unroll
  for (uint32_t q = 0; q < kRows; q++) {
    const uint32_t r = threadIdx.x + q * blockDim.x;
    piv[q] = r < dw ? T[col_off(dw, r)] : 0.0;
  }
  __syncthreads();
  for (uint32_t k = 0; k < dw; k++) {
    // the owner of row k finishes the pivot
    if (k % blockDim.x == threadIdx.x) {
      const double sum = piv[k / blockDim.x];
      if (!(sum > tiny)) ok = 0;
      else T[col_off(dw, k)] = jsm::sqrt_(sum);
    }
    __syncthreads();
    if (!ok) break;
    const double lkk = T[col_off(dw, k)];
#pragma unroll
    for (uint32_t q = 0; q < kRows; q++) {
      const uint32_t i = threadIdx.x + q * blockDim.x;
      if (i <= k || i >= dw) continue;
      double s2 = T[col_off(dw, k) + (i - k)];
      // L_it at off_t + (i - t), L_kt at off_t + (k - t), off_t = col_off(t);
      // loads issued eight ahead of the (ordered) subtractions
      size_t off = 0;
      uint32_t t = 0;
      for (; t + 8 <= k; t += 8) {
        double a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          a[u] = T[off + (i - (t + u))];
          b[u] = T[off + (k - (t + u))];
          off += dw - (t + u);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) s2 -= a[u] * b[u];
      }
      for (; t < k; t++) {
        s2 -= T[off + (i - t)] * T[off + (k - t)];
        off += dw - t;
      }
      const double lik = s2 / lkk;
      T[col_off(dw, k) + (i - k)] = lik;
      piv[q] -= lik * lik;
    }
    __syncthreads();
  }
  double* Lg = s.L + (size_t)d * tri_size(s.DW);
  float* L32 = s.L32 + (size_t)d * tri_size(s.DW);
  if (ok) {
    for (size_t e = threadIdx.x; e < tri_size(dw); e += blockDim.x) Lg[e] = T[e];
    if (want32)
      for (uint32_t k = 0; k < dw; k++)
        for (uint32_t i = k + threadIdx.x; i < dw; i += blockDim.x) store32(s, d, dw, i, k, T[col_off(dw, k) + (i - k)]);
  } else {
    // symmetric Jacobi-diagonal fallback: the inverse diagonal in the
    // diagonal slots
    for (uint32_t i = threadIdx.x; i < dw; i += blockDim.x) {
      const uint32_t v = s.members[(size_t)d * s.dmax + i / 3];
      const double dd = diag[(size_t)v * 3 + i % 3];
      const double inv = dd > 1e-12 ? 1 / dd : 0;
      Lg[col_off(dw, i)] = inv;
      if (want32) {
        L32[col_off(dw, i)] = (float)inv;
        s.L32T[(size_t)d * tri_size(s.DW) + row_off(i) + i] = (float)inv;
      }
    }
  }
  if (threadIdx.x == 0) s.flag[d] = ok ? 1 : 0;
}

__device__ __forceinline__ double readlane_f64(double x, uint32_t l) {
  const long long u = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float readlane_f32(float x, uint32_t l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}
// The same factorisation right-looking, with the triangle in registers
// (domains up to 16 NB rows): 256 threads as a 16 x 16 grid, thread
// (tr, tc) owning elements (16 a + tr, 16 b + tc), a >= b.  Step t: the 16
// threads of column t (one wave's lanes) take the pivot by a lane read and
// divide the column, which goes to LDS; after one barrier every thread
// subtracts L_it L_kt from its elements (i, k) with k > t.
//
// kF32 = false: the FP64 factor of the reference.  Each element receives
// exactly the reference's operations in the reference's order: its
// subtractions for s = 0, 1, ..., k - 1 (mul, then sub), then the division by
// L_kk (on the diagonal, the square root), so it is bit-identical to the
// left-looking loop of schwarz.ts:212-235.
// kF32 = true: the FP32 factor of the mixed solve (reciprocal pivots) from
// the FP32 assembly (A_ki for i >= k), written as the tiles of k_sw_apply32t.
//
// Elements above the diagonal and rows or columns past the domain are
// updated too, with whatever the column buffer holds: they are never read
// and never written out, and sparing them the per-element predicate halves
// the update.
template <int NB, bool kF32>
__global__ void __launch_bounds__(256) k_sw_factor_rl(SwArgs s, const double* __restrict__ diag, int want32) {
  using T = typename std::conditional<kF32, float, double>::type;
  __shared__ T colbuf[2][NB * 16];
  __shared__ double red[256];
  __shared__ int ok;
  __shared__ double tiny;
  const uint32_t d = blockIdx.x;
  const uint32_t dv = s.size[d], dw = dv * 3;
  const uint32_t tr = threadIdx.x % 16, tc = threadIdx.x / 16;
  const double* A = dom_mat(s, d);
  const float* A32 = kF32 ? s.A32 + (size_t)d * s.DW * s.DW : nullptr;  // the FP32 assembly
  T e[NB][NB];
  double dmax = 0;
#pragma unroll
  for (int a = 0; a < NB; a++)
#pragma unroll
    for (int b = 0; b <= a; b++) {
      const uint32_t i = 16 * a + tr, k = 16 * b + tc;
      double v = 0;
      if (i < dw && k <= i) {
        // symmetrised (schwarz.ts:204-210): 0.5 (A_ki + A_ik) in FP64; the
        // FP32 factor reads A_ki alone (the two halves differ by FP64
        // rounding, far below FP32's), one coalesced read instead of two
        if constexpr (kF32) v = A32[(size_t)k * dw + i];
        else v = i == k ? A[(size_t)i * dw + i] : 0.5 * (A[(size_t)k * dw + i] + A[(size_t)i * dw + k]);
        if (i == k) dmax = jsm::jmax(dmax, jsm::fabs_(v));
      }
      e[a][b] = (T)v;
    }
  red[threadIdx.x] = dmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    // Math.max over |A_ii| (schwarz.ts:213-214): order-free, NaN propagates
    double scale = 0;
    for (int q = 0; q < 256; q++) scale = jsm::jmax(scale, red[q]);
    tiny = 1e-20 * (1 + scale);
    ok = 1;
  }
  __syncthreads();
  const double tiny_v = tiny;
  bool alive = true;
#pragma unroll
  for (int bt = 0; bt < NB; bt++) {
    for (uint32_t ct = 0; ct < 16 && alive; ct++) {
      const uint32_t t = 16 * bt + ct;
      if (t >= dw) break;
      T* cb = colbuf[t & 1];
      if (tc == ct) {
        // lanes of column t; the pivot's owner is lane (17 ct) % 64
        T sum;
        if constexpr (kF32) sum = readlane_f32(e[bt][bt], (17 * ct) % kWave);
        else sum = readlane_f64(e[bt][bt], (17 * ct) % kWave);
        if (!((double)sum > tiny_v)) {
          if (tr == ct) ok = 0;
        } else {
          if constexpr (kF32) {
            const float rp = 1.0f / sqrtf(sum);
            if (tr == ct) e[bt][bt] = rp;
#pragma unroll
            for (int a = bt; a < NB; a++) {
              const uint32_t i = 16 * a + tr;
              if (i > t && i < dw) {
                e[a][bt] = e[a][bt] * rp;
                cb[i] = e[a][bt];
              }
            }
          } else {
            const double lt = jsm::sqrt_(sum);
            if (tr == ct) e[bt][bt] = lt;
#pragma unroll
            for (int a = bt; a < NB; a++) {
              const uint32_t i = 16 * a + tr;
              if (i > t && i < dw) {
                e[a][bt] = e[a][bt] / lt;
                cb[i] = e[a][bt];
              }
            }
          }
        }
      }
      __syncthreads();
      if (!ok) {
        alive = false;
        break;
      }
      T lr[NB];
#pragma unroll
      for (int a = bt; a < NB; a++) lr[a] = cb[16 * a + tr];
#pragma unroll
      for (int b = bt; b < NB; b++) {
        if (16 * b + tc <= t) continue;
        const T lk = cb[16 * b + tc];
#pragma unroll
        for (int a = b; a < NB; a++) e[a][b] -= lr[a] * lk;
      }
    }
  }
  __syncthreads();
  if constexpr (kF32) {
    // the tiles of k_sw_apply32t: tile (a, b) is this work-group's e[a][b],
    // one entry per thread, stored by columns (thread tc * 16 + tr writes
    // entry tc * 16 + tr: coalesced); a diagonal tile keeps the reciprocal
    // pivots on its diagonal until k_sw_tile_inv inverts it.  The Jacobi
    // fallback becomes the diagonal tiles diag(1 / sqrt(d)) (the two sweeps
    // then scale by 1 / d) with zero off-diagonal tiles.
    const uint32_t nbt = s.nbt;
    float* Tb = s.T32 + (size_t)d * tile_count(nbt) * 256;
    const bool good = ok != 0;
#pragma unroll
    for (int a = 0; a < NB; a++)
#pragma unroll
      for (int b = 0; b <= a; b++) {
        if ((uint32_t)a >= nbt) continue;
        const uint32_t i = 16 * a + tr, k = 16 * b + tc;
        float v = 0.f;
        if (good) {
          if (i < dw && k <= i) v = e[a][b];
        } else if (a == b && tr == tc && i < dw) {
          const uint32_t mv = s.members[(size_t)d * s.dmax + i / 3];
          const double dd = diag[(size_t)mv * 3 + i % 3];
          v = dd > 1e-12 ? (float)sqrt(1.0 / dd) : 0.f;
        }
        Tb[(size_t)tile_index(nbt, (uint32_t)a, (uint32_t)b) * 256 + tc * 16 + tr] = v;
      }
    if (threadIdx.x == 0) s.flag[d] = good ? 1 : 0;
    return;
  }
  const size_t base = (size_t)d * tri_size(s.DW);
  if (ok) {
#pragma unroll
    for (int a = 0; a < NB; a++)
#pragma unroll
      for (int b = 0; b <= a; b++) {
        const uint32_t i = 16 * a + tr, k = 16 * b + tc;
        if (i < dw && k <= i) {
          s.L[base + col_off(dw, k) + (i - k)] = e[a][b];
          if (want32) store32(s, d, dw, i, k, e[a][b]);
        }
      }
  } else {
    // symmetric Jacobi-diagonal fallback (schwarz.ts:236-249)
    for (uint32_t i = threadIdx.x; i < dw; i += blockDim.x) {
      const uint32_t v = s.members[(size_t)d * s.dmax + i / 3];
      const double dd = diag[(size_t)v * 3 + i % 3];
      const double inv = dd > 1e-12 ? 1 / dd : 0;
      s.L[base + col_off(dw, i)] = inv;
      if (want32) {
        s.L32[base + col_off(dw, i)] = (float)inv;
        s.L32T[base + row_off(i) + i] = (float)inv;
      }
    }
  }
  if (threadIdx.x == 0) s.flag[d] = ok ? 1 : 0;
}

// Each diagonal tile of the FP32 factor (reciprocal pivots on the diagonal,
// L's strictly lower entries below, by columns) becomes the inverse of its
// lower triangle, by rows: column k of the inverse by forward substitution
// in FP64 (x_k = 1 / L_kk, x_i = -(sum_{k<=j<i} L_ij x_j) / L_ii), one
// thread per (block, column).  Zero pivots (rows beyond the domain) give
// zero rows and columns.
__global__ void __launch_bounds__(192) k_sw_tile_inv(SwArgs s) {
  __shared__ float t[12][256];
  const uint32_t d = blockIdx.x, nbt = s.nbt;
  float* Tb = s.T32 + (size_t)d * tile_count(nbt) * 256;
  for (uint32_t e = threadIdx.x; e < nbt * 256; e += blockDim.x)
    t[e / 256][e % 256] = Tb[(size_t)tile_index(nbt, e / 256, e / 256) * 256 + e % 256];
  __syncthreads();
  const uint32_t a = threadIdx.x / 16, k = threadIdx.x % 16;
  if (a >= nbt) return;
  const float* L = t[a];
  double x[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    double sum = 0;
#pragma unroll
    for (int j = 0; j < i; j++)
      if ((uint32_t)j >= k) sum += (double)L[j * 16 + i] * x[j];
    const double p = L[i * 16 + i];
    x[i] = (uint32_t)i < k ? 0.0 : (uint32_t)i == k ? p : -sum * p;
  }
  float* D = Tb + (size_t)tile_index(nbt, a, a) * 256;
#pragma unroll
  for (int i = 0; i < 16; i++) D[i * 16 + k] = (float)x[i];
}

// The substitutions keep row i of a domain in lane i % 64, register slot
// i / 64 (NC slots cover the widest domain), and read every column of the
// factor contiguously, in blocks of K columns: the loads of the next block
// are issued before the current one is used, which hides the memory latency
// behind K columns of work.
template <class T>
__device__ __forceinline__ T bcast(T x, uint32_t l) {
  if constexpr (sizeof(T) == 8) return readlane_f64(x, l);
  else return readlane_f32(x, l);
}

// this lane's slot holding row j (meaningful in lane j % 64 only).  The
// slot is picked with integer masks: a compare-and-select chain over the
// slots would be folded back into a variable index into the array, which
// the compiler can only serve from scratch memory.
template <class T>
using BitsOf = typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned>::type;

template <int NC, class T>
__device__ __forceinline__ T slot_of(const T (&v)[NC], uint32_t j) {
  using U = BitsOf<T>;
  const uint32_t c0 = j / kWave;
  U x = 0;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const U m = (U)0 - (U)(c0 == (uint32_t)c);
    x |= __builtin_bit_cast(U, v[c]) & m;
  }
  return __builtin_bit_cast(T, x);
}
template <int NC, class T>
__device__ __forceinline__ void set_slot(T (&v)[NC], uint32_t j, T x) {
  using U = BitsOf<T>;
  const uint32_t c0 = j / kWave;
  const bool mine = (uint32_t)threadIdx.x == j % kWave;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const U m = (U)0 - (U)(mine && c0 == (uint32_t)c);
    v[c] = __builtin_bit_cast(T, (__builtin_bit_cast(U, v[c]) & ~m) | (__builtin_bit_cast(U, x) & m));
  }
}

// Column j's entries for this lane's rows i >= j (0 elsewhere, and for j
// outside the domain).
template <int NC, class T>
__device__ __forceinline__ void load_col(const T* __restrict__ M, uint32_t dw, int j, T (&out)[NC]) {
  const bool in = j >= 0 && (uint32_t)j < dw;
  const T* col = M + (in ? col_off(dw, (uint32_t)j) : 0);
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int i = c * kWave + (int)threadIdx.x;
    out[c] = (in && i >= j && (uint32_t)i < dw) ? col[i - j] : (T)0;
  }
}

template <int NC, int K, class T>
__device__ __forceinline__ void load_block(const T* __restrict__ M, uint32_t dw, int j0, int dir,
                                           T (&b)[K][NC]) {
#pragma unroll
  for (int k = 0; k < K; k++) load_col<NC, T>(M, dw, j0 + dir * k, b[k]);
}

// Forward substitution: y_j /= L_jj, then y_i -= L_ij y_j for i > j, column
// by column; for each row the same operations in the same order as the
// reference's row loop (schwarz.ts:262-268).
template <int NC, int K, class T>
__device__ __forceinline__ void forward_sub(const T* __restrict__ M, uint32_t dw, T (&y)[NC]) {
  T cur[K][NC], nxt[K][NC];
  load_block<NC, K, T>(M, dw, 0, 1, cur);
  for (uint32_t jb = 0; jb < dw; jb += K) {
    load_block<NC, K, T>(M, dw, (int)(jb + K), 1, nxt);
#pragma unroll
    for (int k = 0; k < K; k++) {
      const uint32_t j = jb + k;
      if (j < dw) {
        // the owner lane divides; the quotient is broadcast
        const T yj = bcast(slot_of<NC, T>(y, j) / slot_of<NC, T>(cur[k], j), j % kWave);
        set_slot<NC, T>(y, j, yj);
#pragma unroll
        for (int c = 0; c < NC; c++)
          if (c * kWave + threadIdx.x > j) y[c] -= cur[k][c] * yj;
      }
    }
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int c = 0; c < NC; c++) cur[k][c] = nxt[k][c];
  }
}

// z = A_d^-1 r on every domain, FP64, in the reference's operation order
// (solveLocal + applySchwarz, schwarz.ts:255-309).  One wave per domain.
// Back substitution row by row in decreasing i (:269-274): the wave forms
// the products L_ji x_j, lane 0 subtracts them in increasing j from LDS.
template <int NC>
__global__ void __launch_bounds__(kWave) k_sw_apply64(SwArgs s, const double* __restrict__ r, double* z,
                                                      const int32_t* done) {
  constexpr int K = NC <= 3 ? 8 : 4;
  if (done && *done) return;
  const uint32_t d = blockIdx.x;
  const uint32_t dv = s.size[d], dw = dv * 3;
  const double* M = s.L + (size_t)d * tri_size(s.DW);
  __shared__ __align__(16) double prod[NC * kWave];
  const int lane = threadIdx.x;
  const uint32_t* mem = s.members + (size_t)d * s.dmax;
  double y[NC], sol[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const uint32_t i = c * kWave + lane;
    y[c] = i < dw ? r[(size_t)mem[i / 3] * 3 + i % 3] : 0.0;
    sol[c] = 0;
  }
  if (s.flag[d]) {
    forward_sub<NC, K, double>(M, dw, y);
    double cur[K][NC], nxt[K][NC];
    load_block<NC, K, double>(M, dw, (int)dw - 1, -1, cur);
    for (uint32_t ib = 0; ib < dw; ib += K) {
      load_block<NC, K, double>(M, dw, (int)dw - 1 - (int)(ib + K), -1, nxt);
#pragma unroll
      for (int k = 0; k < K; k++) {
        if (ib + k >= dw) break;
        const uint32_t ii = dw - 1 - (ib + k);
        // products of rows below ii (others unused); lane 0 subtracts them
        // in increasing row order, reading them from LDS eight at a time
        // with the next eight already requested
#pragma unroll
        for (int c = 0; c < NC; c++) prod[c * kWave + lane] = cur[k][c] * sol[c];
        const double yi = bcast(slot_of<NC, double>(y, ii), ii % kWave);
        const double lii = bcast(slot_of<NC, double>(cur[k], ii), ii % kWave);
        __builtin_amdgcn_wave_barrier();
        double acc = yi;
        if (lane == 0) {
          uint32_t j = ii + 1;
          if ((j & 1) && j < dw) acc -= prod[j++];
          if (j + 8 <= dw) {
            const double2* q = reinterpret_cast<const double2*>(prod + j);
            double2 a0 = q[0], a1 = q[1], a2 = q[2], a3 = q[3];
            for (; j + 16 <= dw; j += 8) {
              q = reinterpret_cast<const double2*>(prod + j + 8);
              const double2 b0 = q[0], b1 = q[1], b2 = q[2], b3 = q[3];
              __builtin_amdgcn_sched_barrier(0);
              acc -= a0.x; acc -= a0.y; acc -= a1.x; acc -= a1.y;
              acc -= a2.x; acc -= a2.y; acc -= a3.x; acc -= a3.y;
              a0 = b0; a1 = b1; a2 = b2; a3 = b3;
            }
            acc -= a0.x; acc -= a0.y; acc -= a1.x; acc -= a1.y;
            acc -= a2.x; acc -= a2.y; acc -= a3.x; acc -= a3.y;
            j += 8;
          }
          for (; j < dw; j++) acc -= prod[j];
        }
        acc = bcast(acc, 0);
        set_slot<NC, double>(sol, ii, acc / lii);
        __builtin_amdgcn_wave_barrier();
      }
#pragma unroll
      for (int k = 0; k < K; k++)
#pragma unroll
        for (int c = 0; c < NC; c++) cur[k][c] = nxt[k][c];
    }
  } else {
    // the fallback's dense product with the (diagonal) inverse, row by row
    for (uint32_t j = 0; j < dw; j++) {
      const double rj = r[(size_t)mem[j / 3] * 3 + j % 3];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const uint32_t i = c * kWave + lane;
        const double inv = i < dw ? M[col_off(dw, i)] : 0.0;
        sol[c] += (j == i ? inv : 0.0) * rj;
      }
    }
  }
  // out.fill(0) then out += sol for the (disjoint) members
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const uint32_t i = c * kWave + lane;
    if (i < dw) z[(size_t)mem[i / 3] * 3 + i % 3] = 0.0 + sol[c];
  }
}

// Row ii's entries for this lane's rows i <= ii of a triangle packed by rows.
template <int NC, class T>
__device__ __forceinline__ void load_row(const T* __restrict__ M, uint32_t dw, int ii, T (&out)[NC]) {
  const bool in = ii >= 0 && (uint32_t)ii < dw;
  const T* row = M + (in ? row_off((uint32_t)ii) : 0);
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int i = c * kWave + (int)threadIdx.x;
    out[c] = (in && i <= ii) ? row[i] : (T)0;
  }
}

// The same operator with the FP32 factors (mixed-precision inner solve), in
// any order: both substitutions column-oriented (forward over L by columns,
// back over L^T, i.e. L by rows), the diagonal stored as reciprocals, so a
// step is one multiply, one broadcast and a multiply-subtract per row.
template <int NC>
__global__ void __launch_bounds__(kWave) k_sw_apply32(SwArgs s, const float* __restrict__ r, float* z,
                                                      const int32_t* done) {
  constexpr int K = NC <= 3 ? 16 : 8;
  if (done && *done) return;
  const uint32_t d = blockIdx.x;
  const uint32_t dv = s.size[d], dw = dv * 3;
  const float* M = s.L32 + (size_t)d * tri_size(s.DW);
  const float* MT = s.L32T + (size_t)d * tri_size(s.DW);
  const int lane = threadIdx.x;
  const uint32_t* mem = s.members + (size_t)d * s.dmax;
  float y[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const uint32_t i = c * kWave + lane;
    y[c] = i < dw ? r[(size_t)mem[i / 3] * 3 + i % 3] : 0.0f;
  }
  if (s.flag[d]) {
    {
      float cur[K][NC], nxt[K][NC];
      load_block<NC, K, float>(M, dw, 0, 1, cur);
      for (uint32_t jb = 0; jb < dw; jb += K) {
        load_block<NC, K, float>(M, dw, (int)(jb + K), 1, nxt);
#pragma unroll
        for (int k = 0; k < K; k++) {
          const uint32_t j = jb + k;
          if (j < dw) {
            const float yj = bcast(slot_of<NC, float>(y, j) * slot_of<NC, float>(cur[k], j), j % kWave);
            set_slot<NC, float>(y, j, yj);
#pragma unroll
            for (int c = 0; c < NC; c++)
              if (c * kWave + (uint32_t)lane > j) y[c] -= cur[k][c] * yj;
          }
        }
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
          for (int c = 0; c < NC; c++) cur[k][c] = nxt[k][c];
      }
    }
    float cur[K][NC], nxt[K][NC];
#pragma unroll
    for (int k = 0; k < K; k++) load_row<NC, float>(MT, dw, (int)dw - 1 - k, cur[k]);
    for (uint32_t ib = 0; ib < dw; ib += K) {
#pragma unroll
      for (int k = 0; k < K; k++) load_row<NC, float>(MT, dw, (int)dw - 1 - (int)(ib + K) - k, nxt[k]);
#pragma unroll
      for (int k = 0; k < K; k++) {
        if (ib + k < dw) {
          const uint32_t ii = dw - 1 - (ib + k);
          const float xi = bcast(slot_of<NC, float>(y, ii) * slot_of<NC, float>(cur[k], ii), ii % kWave);
          set_slot<NC, float>(y, ii, xi);
#pragma unroll
          for (int c = 0; c < NC; c++)
            if (c * kWave + (uint32_t)lane < ii) y[c] -= cur[k][c] * xi;
        }
      }
#pragma unroll
      for (int k = 0; k < K; k++)
#pragma unroll
        for (int c = 0; c < NC; c++) cur[k][c] = nxt[k][c];
    }
  } else {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const uint32_t i = c * kWave + lane;
      if (i < dw) y[c] = M[col_off(dw, i)] * y[c];
    }
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const uint32_t i = c * kWave + lane;
    if (i < dw) z[(size_t)mem[i / 3] * 3 + i % 3] = y[c];
  }
}

// The mixed solve's Schwarz apply over the tiles (one wave per domain):
// both substitutions by 16-row blocks, each block a small product with the
// inverted diagonal tile and every off-diagonal tile one 16x16
// matrix-vector product, so the only dependency chain is the block order.
// Lane l holds, of every tile it reads, the float4 at 4 l: column c = l / 4
// and rows 4 q .. 4 q + 3 (q = l % 4) of an off-diagonal tile, row c and
// columns 4 q .. 4 q + 3 of a diagonal tile.  A block vector is kept either
// by columns (every lane holds entry c) or by rows (every lane holds entries
// 4 q .. 4 q + 3); a sum over the lanes that share c takes two butterfly
// steps (xor 1, 2), over those that share q four (xor 4 .. 32), and the
// off-diagonal contributions of a block are summed per lane first and
// reduced once, when the block is reached.
__device__ __forceinline__ float xsum(float v, int m) { return v + __shfl_xor(v, m, kWave); }
__device__ __forceinline__ float sum_rows(float v) { return xsum(xsum(v, 1), 2); }
__device__ __forceinline__ float sum_cols(float v) { return xsum(xsum(xsum(xsum(v, 4), 8), 16), 32); }

template <int NBT>
__global__ void __launch_bounds__(kWave) k_sw_apply32t(SwArgs s, const float* __restrict__ r, float* z,
                                                       const int32_t* done) {
  if (done && *done) return;
  const uint32_t d = blockIdx.x;
  const uint32_t dw = s.size[d] * 3;
  const uint32_t l = threadIdx.x, q = l & 3, c = l >> 2;
  const float4* __restrict__ T = reinterpret_cast<const float4*>(s.T32) + (size_t)d * tile_count(NBT) * 64 + l;
  const uint32_t* mem = s.members + (size_t)d * s.dmax;
  // forward: y = L^-1 r, block column by block column
  float acc[NBT][4], y[NBT];
#pragma unroll
  for (int a = 0; a < NBT; a++) acc[a][0] = acc[a][1] = acc[a][2] = acc[a][3] = 0.f;
#pragma unroll
  for (int J = 0; J < NBT; J++) {
    float v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t i = 16 * J + 4 * q + k;
      const float ri = i < dw ? r[(size_t)mem[i / 3] * 3 + i % 3] : 0.f;
      v[k] = ri - sum_cols(acc[J][k]);  // by rows
    }
    const float4 D = T[(size_t)tile_index(NBT, J, J) * 64];
    const float yj = sum_rows(D.x * v[0] + D.y * v[1] + D.z * v[2] + D.w * v[3]);  // y_J[c]
    y[J] = yj;
#pragma unroll
    for (int I = J + 1; I < NBT; I++) {
      const float4 t = T[(size_t)tile_index(NBT, I, J) * 64];
      acc[I][0] += t.x * yj;
      acc[I][1] += t.y * yj;
      acc[I][2] += t.z * yj;
      acc[I][3] += t.w * yj;
    }
  }
  // back: x = L^-T y, block row by block row from the last
  float bcc[NBT];
#pragma unroll
  for (int a = 0; a < NBT; a++) bcc[a] = 0.f;
#pragma unroll
  for (int J = NBT - 1; J >= 0; J--) {
    const float v = y[J] - sum_rows(bcc[J]);  // by columns
    const float4 D = T[(size_t)tile_index(NBT, J, J) * 64];
    const float x0 = sum_cols(D.x * v), x1 = sum_cols(D.y * v), x2 = sum_cols(D.z * v), x3 = sum_cols(D.w * v);
    if (c == 0) {
      const float xs[4] = {x0, x1, x2, x3};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint32_t i = 16 * J + 4 * q + k;
        if (i < dw) z[(size_t)mem[i / 3] * 3 + i % 3] = xs[k];
      }
    }
#pragma unroll
    for (int I = 0; I < J; I++) {
      const float4 t = T[(size_t)tile_index(NBT, J, I) * 64];
      bcc[I] += t.x * x0 + t.y * x1 + t.z * x2 + t.w * x3;
    }
  }
}

// ---------------------------------------------------------------- facade

Schwarz::Schwarz(const cs_scene_desc& d, int target, int min_size, int max_size, hipStream_t st,
                 int coarse_mode, int coarse_iters, double coarse_omega)
    : im_(new SchwarzImpl) {
  (void)max_size;  // growth stops at the target size (schwarz.ts:138)
  SchwarzImpl& S = *im_;
  S.st = st;
  S.n = d.n_verts;
  if (target < 1 || min_size < 0) throw std::invalid_argument("Schwarz domain sizes must be positive");
  std::vector<std::vector<uint32_t>> mem;
  build_domains(d, target, min_size, mem, S.h_domain_of, S.h_local_of);
  S.D = (uint32_t)mem.size();
  for (auto& m : mem) S.dmax = std::max<uint32_t>(S.dmax, (uint32_t)m.size());
  S.DW = S.dmax * 3;
  // the factor kernel keeps up to two rows per thread
  if (S.DW > 2 * kSwFactorBlock) throw std::invalid_argument("Schwarz domains above 170 vertices are not supported");
  std::vector<uint32_t> flat((size_t)std::max<uint32_t>(S.D, 1) * std::max<uint32_t>(S.dmax, 1), 0xffffffffu),
      sz(std::max<uint32_t>(S.D, 1), 0);
  for (uint32_t k = 0; k < S.D; k++) {
    sz[k] = (uint32_t)mem[k].size();
    for (size_t l = 0; l < mem[k].size(); l++) flat[(size_t)k * S.dmax + l] = mem[k][l];
  }
  S.members.upload(flat.data(), flat.size(), st);
  S.size.upload(sz.data(), sz.size(), st);
  S.domain_of.upload(S.h_domain_of.data(), S.n, st);
  S.local_of.upload(S.h_local_of.data(), S.n, st);
  S.L.alloc((size_t)std::max<uint32_t>(S.D, 1) * tri_size(S.DW));
  S.flag.alloc(std::max<uint32_t>(S.D, 1));
  // the triangle being factored stays in LDS when it fits the device's
  // per-work-group maximum; otherwise in a global workspace
  int dev = 0, optin = 0, base = 0;
  CS_HIP(hipGetDevice(&dev));
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess) optin = 0;
  CS_HIP(hipDeviceGetAttribute(&base, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const size_t need = tri_size(S.DW) * sizeof(double);
  const size_t stat = 256;  // the kernel's own shared scalars, with room
  S.lds_max = (size_t)base > stat ? (size_t)base - stat : 0;
  if (need > S.lds_max && (size_t)optin > (size_t)base &&
      hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sw_factor<true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)((size_t)optin - stat)) == hipSuccess)
    S.lds_max = (size_t)optin - stat;
  if (need > S.lds_max) S.W.alloc((size_t)std::max<uint32_t>(S.D, 1) * tri_size(S.DW));
  if (coarse_mode != CS_COARSE_NONE && S.D)
    S.coarse.reset(new Coarse(d, mem, S.h_domain_of, coarse_mode, coarse_iters, coarse_omega, st));
}

Schwarz::~Schwarz() { delete im_; }

static SwArgs sw_args(SchwarzImpl& S) {
  SwArgs a;
  a.D = S.D; a.dmax = S.dmax; a.DW = S.DW;
  a.members = S.members.ptr; a.size = S.size.ptr; a.domain_of = S.domain_of.ptr; a.local_of = S.local_of.ptr;
  a.A = S.A.ptr; a.L = S.L.ptr; a.L32 = S.L32.ptr; a.L32T = S.L32T.ptr; a.W = S.W.ptr; a.flag = S.flag.ptr;
  a.T32 = S.T32.ptr; a.nbt = S.nbt; a.A32 = S.A32.ptr;
  return a;
}

void Schwarz::factor(const DMesh& me, const DMat& mat, const double* x, const double* diag, bool fp32) {
  SchwarzImpl& S = *im_;
  if (!S.D) return;
  if (S.coarse) S.coarse->assemble(me, mat, x, diag);
  const uint32_t nb = (S.DW + 15) / 16;
  if (fp32 && nb <= 12 && !S.T32.ptr) {
    S.nbt = nb;
    S.T32.alloc((size_t)S.D * tile_count(nb) * 256);
    S.A32.alloc((size_t)S.D * S.DW * S.DW);
  } else if (fp32 && nb > 12 && !S.L32.ptr) {
    S.L32.alloc((size_t)S.D * tri_size(S.DW));
    S.L32T.alloc((size_t)S.D * tri_size(S.DW));
  }
  // the FP64 domain matrices only for the FP64 factor (allocated on first use)
  if (!(fp32 && nb <= 12) && !S.A.ptr) S.A.alloc((size_t)S.D * S.DW * S.DW);
  const SwArgs a = sw_args(S);
  if (fp32 && nb <= 12)
    hipLaunchKernelGGL((k_sw_assemble<float, true>), dim3(S.D), dim3(kSwAsmBlock), 0, S.st, a, me, mat, x, diag, S.A32.ptr);
  else
    hipLaunchKernelGGL((k_sw_assemble<double, false>), dim3(S.D), dim3(kSwAsmBlock), 0, S.st, a, me, mat, x, diag, S.A.ptr);
  if (fp32 && nb <= 12) {
#define SW_RL32(C, NB)                                                                                       \
  case C:                                                                                                    \
    hipLaunchKernelGGL((k_sw_factor_rl<NB, true>), dim3(S.D), dim3(256), 0, S.st, a, diag, 0);                        \
    break;
    switch (nb) {
      SW_RL32(1, 1) SW_RL32(2, 2) SW_RL32(3, 3) SW_RL32(4, 4) SW_RL32(5, 5) SW_RL32(6, 6) SW_RL32(7, 7)
      SW_RL32(8, 9) SW_RL32(9, 9) SW_RL32(10, 10) SW_RL32(11, 11) default: SW_RL32(12, 12)
    }
#undef SW_RL32
    hipLaunchKernelGGL(k_sw_tile_inv, dim3(S.D), dim3(192), 0, S.st, a);
    CS_LAUNCH_CHECK();
    return;
  }
  const int w32 = fp32 ? 1 : 0;
  const size_t need = tri_size(S.DW) * sizeof(double);
#define SW_RL(C, NB)                                                                                         \
  case C:                                                                                                    \
    hipLaunchKernelGGL((k_sw_factor_rl<NB, false>), dim3(S.D), dim3(256), 0, S.st, a, diag, w32);                     \
    break;
  // (the 8-block instance would keep its tile in scratch memory; 9 blocks
  // cover the same domains from registers)
  switch (nb) {
    SW_RL(1, 1) SW_RL(2, 2) SW_RL(3, 3) SW_RL(4, 4) SW_RL(5, 5) SW_RL(6, 6) SW_RL(7, 7) SW_RL(8, 9) SW_RL(9, 9)
    SW_RL(10, 10) SW_RL(11, 11) SW_RL(12, 12)
    default:
      // wider domains: the triangle in LDS (or the workspace), left-looking
      if (need <= S.lds_max)
        hipLaunchKernelGGL(k_sw_factor<true>, dim3(S.D), dim3(kSwFactorBlock), need, S.st, a, diag, w32);
      else
        hipLaunchKernelGGL(k_sw_factor<false>, dim3(S.D), dim3(kSwFactorBlock), 0, S.st, a, diag, w32);
  }
#undef SW_RL
  CS_LAUNCH_CHECK();
}

void Schwarz::flags(uint8_t* out) const {
  if (im_->D) CS_HIP(hipMemcpy(out, im_->flag.ptr, im_->D, hipMemcpyDeviceToHost));
}

void Schwarz::apply64(const double* r, double* z, const int32_t* done) {
  SchwarzImpl& S = *im_;
  if (!S.D) return;
  const SwArgs a = sw_args(S);
  switch ((S.DW + kWave - 1) / kWave) {
    case 1: hipLaunchKernelGGL(k_sw_apply64<1>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    case 2: hipLaunchKernelGGL(k_sw_apply64<2>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    case 3: hipLaunchKernelGGL(k_sw_apply64<3>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    case 4: hipLaunchKernelGGL(k_sw_apply64<4>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    case 5: hipLaunchKernelGGL(k_sw_apply64<5>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    case 6: hipLaunchKernelGGL(k_sw_apply64<6>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    case 7: hipLaunchKernelGGL(k_sw_apply64<7>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    default: hipLaunchKernelGGL(k_sw_apply64<8>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
  }
  CS_LAUNCH_CHECK();
  if (S.coarse) S.coarse->apply64(r, z, done);
}

void Schwarz::apply32(const float* r, float* z, const int32_t* done) {
  SchwarzImpl& S = *im_;
  if (!S.D) return;
  const SwArgs a = sw_args(S);
  if (S.T32.ptr) {
    const dim3 g(S.D), b(kWave);
    switch (S.nbt) {
      case 1: hipLaunchKernelGGL(k_sw_apply32t<1>, g, b, 0, S.st, a, r, z, done); break;
      case 2: hipLaunchKernelGGL(k_sw_apply32t<2>, g, b, 0, S.st, a, r, z, done); break;
      case 3: hipLaunchKernelGGL(k_sw_apply32t<3>, g, b, 0, S.st, a, r, z, done); break;
      case 4: hipLaunchKernelGGL(k_sw_apply32t<4>, g, b, 0, S.st, a, r, z, done); break;
      case 5: hipLaunchKernelGGL(k_sw_apply32t<5>, g, b, 0, S.st, a, r, z, done); break;
      case 6: hipLaunchKernelGGL(k_sw_apply32t<6>, g, b, 0, S.st, a, r, z, done); break;
      case 7: hipLaunchKernelGGL(k_sw_apply32t<7>, g, b, 0, S.st, a, r, z, done); break;
      case 8: hipLaunchKernelGGL(k_sw_apply32t<8>, g, b, 0, S.st, a, r, z, done); break;
      case 9: hipLaunchKernelGGL(k_sw_apply32t<9>, g, b, 0, S.st, a, r, z, done); break;
      case 10: hipLaunchKernelGGL(k_sw_apply32t<10>, g, b, 0, S.st, a, r, z, done); break;
      case 11: hipLaunchKernelGGL(k_sw_apply32t<11>, g, b, 0, S.st, a, r, z, done); break;
      default: hipLaunchKernelGGL(k_sw_apply32t<12>, g, b, 0, S.st, a, r, z, done); break;
    }
    CS_LAUNCH_CHECK();
    if (S.coarse) S.coarse->apply32(r, z, done);
    return;
  }
  switch ((S.DW + kWave - 1) / kWave) {
    case 1: hipLaunchKernelGGL(k_sw_apply32<1>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    case 2: hipLaunchKernelGGL(k_sw_apply32<2>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    case 3: hipLaunchKernelGGL(k_sw_apply32<3>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    case 4: hipLaunchKernelGGL(k_sw_apply32<4>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    case 5: hipLaunchKernelGGL(k_sw_apply32<5>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    case 6: hipLaunchKernelGGL(k_sw_apply32<6>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    case 7: hipLaunchKernelGGL(k_sw_apply32<7>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
    default: hipLaunchKernelGGL(k_sw_apply32<8>, dim3(S.D), dim3(kWave), 0, S.st, a, r, z, done); break;
  }
  CS_LAUNCH_CHECK();
  if (S.coarse) S.coarse->apply32(r, z, done);
}

int schwarz_partition_host(const cs_scene_desc& d, int target, int min_size, uint32_t* domain_of, uint32_t* local_of) {
  std::vector<std::vector<uint32_t>> mem;
  std::vector<uint32_t> dof, lof;
  build_domains(d, target, min_size, mem, dof, lof);
  std::copy(dof.begin(), dof.end(), domain_of);
  std::copy(lof.begin(), lof.end(), local_of);
  return (int)mem.size();
}

uint32_t Schwarz::domain_count() const { return im_->D; }
const Coarse* Schwarz::coarse() const { return im_->coarse.get(); }
uint32_t Schwarz::max_domain() const { return im_->dmax; }

void Schwarz::partition(uint32_t* domain_of, uint32_t* local_of) const {
  std::copy(im_->h_domain_of.begin(), im_->h_domain_of.end(), domain_of);
  std::copy(im_->h_local_of.begin(), im_->h_local_of.end(), local_of);
}

}  // namespace cloth
