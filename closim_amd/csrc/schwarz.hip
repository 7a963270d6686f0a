// schwarz.hip — one-level additive Schwarz (schwarz.ts:44-309) on the device.
//
// Domains: built once on the host exactly as buildSchwarzDomains does
// (Morton-ordered seeds over the rest positions, BFS growth along mesh
// connectivity to the target size, merge of sub-minimum domains into the
// smallest adjacent one, ids compacted in creation order, members sorted).
// Factors (per Newton iteration): one work-group per domain assembles the
// dense local Hessian row by row (each thread owns the rows of one member
// vertex, so every entry receives its triangle contributions in ascending
// triangle order, as in the reference), embeds the Jacobi diagonal,
// symmetrises, and runs the reference's left-looking Cholesky (row sums in
// the reference's order, rows below the pivot in parallel).  Apply: one wave
// per domain.  The FP64 apply keeps the reference's operation order (forward
// substitution column by column, which subtracts in the same order as the
// row loop; back substitution row by row); the FP32 apply used inside the
// mixed-precision solve substitutes column by column both ways.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "elements.h"
#include "schwarz.h"

namespace cloth {

struct SchwarzImpl {
  hipStream_t st;
  uint32_t n = 0, D = 0, dmax = 0, DW = 0;
  std::vector<uint32_t> h_domain_of, h_local_of;
  DevBuf<uint32_t> members, size, domain_of, local_of;
  DevBuf<double> A;      // D * DW * DW, row stride dw of each domain
  DevBuf<float> A32;     // FP32 copy of the factors
  DevBuf<uint8_t> flag;  // 1 = Cholesky, 0 = Jacobi fallback
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

// Rows A = ci*3 + ra (ra = 0..2), all columns B, of triangleHessian9
// (membrane-blocks.ts:21-98), same loop nest and accumulation order.
__device__ void tri_hessian_rows(const double* x0, const double* x1, const double* x2, const float* inv,
                                 double area, const DMat& mat, int ci, double* h27) {
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
    const int A = ci * 3 + ra, av = A / 3, ai = A % 3;
    for (int B = 0; B < 9; B++) {
      const int bv = B / 3, bi = B % 3;
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
      h27[ra * 9 + B] = s * sum;
    }
  }
}

constexpr int kSwBlock = 128;

struct SwArgs {
  uint32_t D, dmax, DW;
  const uint32_t* members;
  const uint32_t* size;
  const uint32_t* domain_of;
  const uint32_t* local_of;
  double* A;
  float* A32;
  uint8_t* flag;
};

__device__ __forceinline__ double* dom_mat(const SwArgs& s, uint32_t d) { return s.A + (size_t)d * s.DW * s.DW; }

// Local Hessian of domain d: membrane 3x3 blocks of intra-domain corner
// pairs, Jacobi diagonal, symmetrised (schwarz.ts:170-210).
__global__ void __launch_bounds__(kSwBlock) k_sw_assemble(SwArgs s, DMesh me, DMat mat, const double* __restrict__ x,
                                                          const double* __restrict__ diag) {
  const uint32_t d = blockIdx.x;
  const uint32_t dv = s.size[d], dw = dv * 3;
  double* A = dom_mat(s, d);
  for (uint32_t l = threadIdx.x; l < dv; l += blockDim.x) {
    for (uint32_t r = l * 3; r < l * 3 + 3; r++)
      for (uint32_t c = 0; c < dw; c++) A[(size_t)r * dw + c] = 0;
    const uint32_t v = s.members[(size_t)d * s.dmax + l];
    for (uint32_t q = me.vt_off[v]; q < me.vt_off[v + 1]; q++) {
      const uint4 cr = me.vt_corner[q];
      const uint32_t t = cr.w >> 2;
      const int ci = cr.w & 3;
      const uint32_t cvs[3] = {cr.x, cr.y, cr.z};
      double X[3][3];
      for (int k = 0; k < 3; k++) ld3(x, cvs[k], X[k]);
      double h27[27];
      tri_hessian_rows(X[0], X[1], X[2], me.inv_dm + (size_t)t * 4, me.area[t], mat, ci, h27);
      for (int cj = 0; cj < 3; cj++) {
        const uint32_t w = cvs[cj];
        if (s.domain_of[w] != d) continue;
        const uint32_t lw = s.local_of[w];
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) A[(size_t)(l * 3 + a) * dw + (lw * 3 + b)] += h27[a * 9 + cj * 3 + b];
      }
    }
    for (int a = 0; a < 3; a++) A[(size_t)(l * 3 + a) * dw + (l * 3 + a)] += diag[(size_t)v * 3 + a];
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < dw; i += blockDim.x)
    for (uint32_t j = i + 1; j < dw; j++) {
      const double v = 0.5 * (A[(size_t)i * dw + j] + A[(size_t)j * dw + i]);
      A[(size_t)i * dw + j] = v;
      A[(size_t)j * dw + i] = v;
    }
}

// Dense Cholesky in place (schwarz.ts:212-249), or the Jacobi fallback.
__global__ void __launch_bounds__(kSwBlock) k_sw_cholesky(SwArgs s, const double* __restrict__ diag) {
  const uint32_t d = blockIdx.x;
  const uint32_t dv = s.size[d], dw = dv * 3;
  double* L = dom_mat(s, d);
  __shared__ int ok;
  __shared__ double tiny;
  if (threadIdx.x == 0) {
    double scale = 0;
    for (uint32_t i = 0; i < dw; i++) scale = jsm::jmax(scale, jsm::fabs_(L[(size_t)i * dw + i]));
    tiny = 1e-20 * (1 + scale);
    ok = 1;
  }
  __syncthreads();
  for (uint32_t k = 0; k < dw; k++) {
    if (threadIdx.x == 0) {
      double sum = L[(size_t)k * dw + k];
      for (uint32_t q = 0; q < k; q++) sum -= L[(size_t)k * dw + q] * L[(size_t)k * dw + q];
      if (!(sum > tiny)) ok = 0;
      else L[(size_t)k * dw + k] = jsm::sqrt_(sum);
    }
    __syncthreads();
    if (!ok) break;
    const double lkk = L[(size_t)k * dw + k];
    for (uint32_t i = k + 1 + threadIdx.x; i < dw; i += blockDim.x) {
      double s2 = L[(size_t)i * dw + k];
      for (uint32_t q = 0; q < k; q++) s2 -= L[(size_t)i * dw + q] * L[(size_t)k * dw + q];
      L[(size_t)i * dw + k] = s2 / lkk;
    }
    __syncthreads();
  }
  if (!ok) {
    // symmetric Jacobi-diagonal fallback (schwarz.ts:236-249)
    for (uint32_t e = threadIdx.x; e < dw * dw; e += blockDim.x) L[e] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < dw; i += blockDim.x) {
      const uint32_t v = s.members[(size_t)d * s.dmax + i / 3];
      const double dd = diag[(size_t)v * 3 + i % 3];
      L[(size_t)i * dw + i] = dd > 1e-12 ? 1 / dd : 0;
    }
  }
  if (threadIdx.x == 0) s.flag[d] = ok ? 1 : 0;
}

__global__ void __launch_bounds__(kSwBlock) k_sw_to32(SwArgs s) {
  const uint32_t d = blockIdx.x;
  const uint32_t dv = s.size[d], dw = dv * 3;
  const double* L = dom_mat(s, d);
  float* L32 = s.A32 + (size_t)d * s.DW * s.DW;
  for (uint32_t e = threadIdx.x; e < dw * dw; e += blockDim.x) L32[e] = (float)L[e];
}

// z = A_d^-1 r on every domain, FP64 (solveLocal + applySchwarz,
// schwarz.ts:255-309).  One wave per domain; rows i of the domain live in
// lane i % 64.
__global__ void __launch_bounds__(kWave) k_sw_apply64(SwArgs s, const double* __restrict__ r, double* z,
                                                      const int32_t* done) {
  if (done && *done) return;
  const uint32_t d = blockIdx.x;
  const uint32_t dv = s.size[d], dw = dv * 3;
  const double* M = dom_mat(s, d);
  extern __shared__ double sh[];
  double* y = sh;            // dw
  double* sol = sh + s.DW;   // dw
  const int lane = threadIdx.x;
  const uint32_t* mem = s.members + (size_t)d * s.dmax;
  if (s.flag[d]) {
    // forward: y_i = (rhs_i - sum_{j<i} M_ij y_j) / M_ii, the subtractions in
    // increasing j (column by column, rows in parallel)
    for (uint32_t i = lane; i < dw; i += kWave) y[i] = r[(size_t)mem[i / 3] * 3 + i % 3];
    __builtin_amdgcn_wave_barrier();
    for (uint32_t j = 0; j < dw; j++) {
      if ((uint32_t)lane == j % kWave) y[j] = y[j] / M[(size_t)j * dw + j];
      __builtin_amdgcn_wave_barrier();
      const double yj = y[j];
      for (uint32_t i = j + 1 + (((uint32_t)lane - (j + 1)) & (kWave - 1)); i < dw; i += kWave)
        y[i] -= M[(size_t)i * dw + j] * yj;
      __builtin_amdgcn_wave_barrier();
    }
    // back: x_i = (y_i - sum_{j>i} M_ji x_j) / M_ii, j increasing, one lane
    if (lane == 0) {
      for (uint32_t ii = dw; ii-- > 0;) {
        double acc = y[ii];
        for (uint32_t j = ii + 1; j < dw; j++) acc -= M[(size_t)j * dw + ii] * sol[j];
        sol[ii] = acc / M[(size_t)ii * dw + ii];
      }
    }
    __builtin_amdgcn_wave_barrier();
  } else {
    for (uint32_t i = lane; i < dw; i += kWave) {
      double acc = 0;
      for (uint32_t j = 0; j < dw; j++) acc += M[(size_t)i * dw + j] * r[(size_t)mem[j / 3] * 3 + j % 3];
      sol[i] = acc;
    }
    __builtin_amdgcn_wave_barrier();
  }
  // out.fill(0) then out += sol for the (disjoint) members
  for (uint32_t i = lane; i < dw; i += kWave) z[(size_t)mem[i / 3] * 3 + i % 3] = 0.0 + sol[i];
}

// The same apply with the FP32 factors, both substitutions column by column.
__global__ void __launch_bounds__(kWave) k_sw_apply32(SwArgs s, const float* __restrict__ r, float* z,
                                                      const int32_t* done) {
  if (done && *done) return;
  const uint32_t d = blockIdx.x;
  const uint32_t dv = s.size[d], dw = dv * 3;
  const float* M = s.A32 + (size_t)d * s.DW * s.DW;
  extern __shared__ float shf[];
  float* y = shf;
  const int lane = threadIdx.x;
  const uint32_t* mem = s.members + (size_t)d * s.dmax;
  if (s.flag[d]) {
    for (uint32_t i = lane; i < dw; i += kWave) y[i] = r[(size_t)mem[i / 3] * 3 + i % 3];
    __builtin_amdgcn_wave_barrier();
    for (uint32_t j = 0; j < dw; j++) {
      if ((uint32_t)lane == j % kWave) y[j] = y[j] / M[(size_t)j * dw + j];
      __builtin_amdgcn_wave_barrier();
      const float yj = y[j];
      for (uint32_t i = j + 1 + (((uint32_t)lane - (j + 1)) & (kWave - 1)); i < dw; i += kWave)
        y[i] -= M[(size_t)i * dw + j] * yj;
      __builtin_amdgcn_wave_barrier();
    }
    for (uint32_t j = dw; j-- > 0;) {
      if ((uint32_t)lane == j % kWave) y[j] = y[j] / M[(size_t)j * dw + j];
      __builtin_amdgcn_wave_barrier();
      const float xj = y[j];
      for (uint32_t i = lane; i < j; i += kWave) y[i] -= M[(size_t)j * dw + i] * xj;
      __builtin_amdgcn_wave_barrier();
    }
    for (uint32_t i = lane; i < dw; i += kWave) z[(size_t)mem[i / 3] * 3 + i % 3] = y[i];
  } else {
    for (uint32_t i = lane; i < dw; i += kWave) {
      float acc = 0;
      for (uint32_t j = 0; j < dw; j++) acc += M[(size_t)i * dw + j] * r[(size_t)mem[j / 3] * 3 + j % 3];
      z[(size_t)mem[i / 3] * 3 + i % 3] = acc;
    }
  }
}

// ---------------------------------------------------------------- facade

Schwarz::Schwarz(const cs_scene_desc& d, int target, int min_size, int max_size, hipStream_t st)
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
  if (S.DW * 2 * sizeof(double) > 60 * 1024) throw std::invalid_argument("Schwarz domain too large for the apply kernel");
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
  S.A.alloc((size_t)std::max<uint32_t>(S.D, 1) * S.DW * S.DW);
  S.flag.alloc(std::max<uint32_t>(S.D, 1));
}

Schwarz::~Schwarz() { delete im_; }

static SwArgs sw_args(SchwarzImpl& S) {
  SwArgs a;
  a.D = S.D; a.dmax = S.dmax; a.DW = S.DW;
  a.members = S.members.ptr; a.size = S.size.ptr; a.domain_of = S.domain_of.ptr; a.local_of = S.local_of.ptr;
  a.A = S.A.ptr; a.A32 = S.A32.ptr; a.flag = S.flag.ptr;
  return a;
}

void Schwarz::factor(const DMesh& me, const DMat& mat, const double* x, const double* diag, bool want32) {
  SchwarzImpl& S = *im_;
  if (!S.D) return;
  if (want32) S.A32.alloc((size_t)S.D * S.DW * S.DW);
  const SwArgs a = sw_args(S);
  hipLaunchKernelGGL(k_sw_assemble, dim3(S.D), dim3(kSwBlock), 0, S.st, a, me, mat, x, diag);
  hipLaunchKernelGGL(k_sw_cholesky, dim3(S.D), dim3(kSwBlock), 0, S.st, a, diag);
  if (want32) hipLaunchKernelGGL(k_sw_to32, dim3(S.D), dim3(kSwBlock), 0, S.st, a);
  CS_LAUNCH_CHECK();
}

void Schwarz::apply64(const double* r, double* z, const int32_t* done) {
  SchwarzImpl& S = *im_;
  if (!S.D) return;
  hipLaunchKernelGGL(k_sw_apply64, dim3(S.D), dim3(kWave), 2 * S.DW * sizeof(double), S.st, sw_args(S), r, z, done);
  CS_LAUNCH_CHECK();
}

void Schwarz::apply32(const float* r, float* z, const int32_t* done) {
  SchwarzImpl& S = *im_;
  if (!S.D) return;
  hipLaunchKernelGGL(k_sw_apply32, dim3(S.D), dim3(kWave), S.DW * sizeof(float), S.st, sw_args(S), r, z, done);
  CS_LAUNCH_CHECK();
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
uint32_t Schwarz::max_domain() const { return im_->dmax; }

void Schwarz::partition(uint32_t* domain_of, uint32_t* local_of) const {
  std::copy(im_->h_domain_of.begin(), im_->h_domain_of.end(), domain_of);
  std::copy(im_->h_local_of.begin(), im_->h_local_of.end(), local_of);
}

}  // namespace cloth
