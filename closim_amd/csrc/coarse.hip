// coarse.hip — coarse level of the two-level additive Schwarz (mas.ts,
// coarse-csr.ts) on the device.
//
// Host, once: the block pattern (buildCoarsePattern, coarse-csr.ts:30-50),
// each block's mirror (d2, d1), and for every block the list of the
// (triangle, ci, cj) corner pairs that land in it, in the reference's loop
// order (triangle, then ci, then cj).
// Assembly, per Newton iteration (FP64): the membrane Hessian rows of every
// (triangle, corner) once (the Schwarz local matrices' arithmetic), then one
// thread per block entry sums its corner pairs in that order, scaled by
// 1/sqrt(|d1||d2|), and the Jacobi diagonal spread on the diagonal blocks
// after them: the raw blocks of assembleCoarseRaw (coarse-csr.ts:70-110) in
// its operation order.  Dc = the diagonal of the diagonal blocks (the sums of
// buildCoarseDiag, mas.ts:68-101, in the same order); the symmetrised blocks
// (:117-138) and the diagonal-block inverses (invertSym33, :170-205, with the
// Dc fallback) follow.
// Apply: restriction by domain (member order), the coarse operator, and the
// prolongation add, all FP64; the coarse PCG (coarsePcg, :243-303) runs in one
// work-group of 1024 threads with fixed-tree reductions.
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "coarse.h"
#include "elements.h"
#include "hessian.h"

namespace cloth {

constexpr int kCoPcgBlock = 1024;

struct CoArgs {
  uint32_t n, m, D, nnz;
  const uint32_t* mem_off;  // D + 1
  const uint32_t* mem_v;    // members, member order
  const uint32_t* dom_of;   // vertex -> domain
  const double* sd;         // 1 / sqrt(|d|)
  const double* sinv;       // 1 / |d|
  const uint32_t* off;      // block rows
  const uint32_t* col;      // block columns
  const uint32_t* mir;      // position of block (d2, d1)
  const uint32_t* selfp;    // position of block (d, d)
  const uint32_t* brow;     // block row of each block
  const uint32_t* co_off;   // nnz + 1
  const uint32_t* co_rec;   // t << 4 | ci << 2 | cj
  const double* sblk;       // 1 / sqrt(|d1| |d2|)
  double* hrows;            // 3m x 27
  double* raw;
  double* sym;
  double* dc;
  double* binv;  // 9 D
  double* rc;    // 3 D
  double* c;     // 3 D (the coarse solution)
  double* r;     // coarse PCG vectors, 3 D each
  double* z;
  double* p;
  double* Ap;
};

struct CoarseImpl {
  hipStream_t st;
  int mode = 0, iters = 8;
  double omega = 0.5;
  uint32_t n = 0, m = 0, D = 0, nnz = 0;
  std::vector<uint32_t> h_off, h_col;
  DevBuf<uint32_t> mem_off, mem_v, dom_of, off, col, mir, selfp, brow, co_off, co_rec;
  DevBuf<double> sd, sinv, sblk, hrows, raw, sym, dc, binv, rc, c, r, z, p, Ap;
  CoArgs args() const {
    CoArgs a;
    a.n = n; a.m = m; a.D = D; a.nnz = nnz;
    a.mem_off = mem_off.ptr; a.mem_v = mem_v.ptr; a.dom_of = dom_of.ptr; a.sd = sd.ptr; a.sinv = sinv.ptr;
    a.off = off.ptr; a.col = col.ptr; a.mir = mir.ptr; a.selfp = selfp.ptr; a.brow = brow.ptr;
    a.co_off = co_off.ptr; a.co_rec = co_rec.ptr; a.sblk = sblk.ptr;
    a.hrows = hrows.ptr; a.raw = raw.ptr; a.sym = sym.ptr; a.dc = dc.ptr; a.binv = binv.ptr;
    a.rc = rc.ptr; a.c = c.ptr; a.r = r.ptr; a.z = z.ptr; a.p = p.ptr; a.Ap = Ap.ptr;
    return a;
  }
};

// ---------------------------------------------------------------- assembly

// Rows of corner ci of triangle t (t * 3 + ci = thread), 27 per thread.
__global__ void __launch_bounds__(kBlock) k_co_rows(CoArgs a, DMesh me, DMat mat, const double* __restrict__ x) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.m * 3) return;
  const uint32_t t = i / 3;
  const int ci = (int)(i % 3);
  double X[3][3];
  for (int k = 0; k < 3; k++) ld3(x, me.idx[(size_t)t * 3 + k], X[k]);
  tri_hessian_rows(X[0], X[1], X[2], me.inv_dm + (size_t)t * 4, me.area[t], mat, ci, a.hrows + (size_t)i * 27);
}

// Entry e = a3 * 3 + b of block q (thread q * 9 + e): its corner pairs in the
// reference's order, then (diagonal entries of diagonal blocks) the Jacobi
// diagonal spread over the domain's members in member order.
__global__ void __launch_bounds__(kBlock) k_co_blocks(CoArgs a, const double* __restrict__ diag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nnz * 9) return;
  const uint32_t q = i / 9, e = i % 9, ra = e / 3, rb = e % 3;
  const double s = a.sblk[q];
  double sum = 0;
  for (uint32_t k = a.co_off[q]; k < a.co_off[q + 1]; k++) {
    const uint32_t rec = a.co_rec[k];
    const uint32_t t = rec >> 4, ci = (rec >> 2) & 3, cj = rec & 3;
    sum += s * a.hrows[((size_t)t * 3 + ci) * 27 + ra * 9 + cj * 3 + rb];
  }
  const uint32_t d = a.brow[q];
  if (ra == rb && a.col[q] == d) {
    const double sv = a.sinv[d];
    for (uint32_t k = a.mem_off[d]; k < a.mem_off[d + 1]; k++) sum += sv * diag[(size_t)a.mem_v[k] * 3 + ra];
  }
  a.raw[i] = sum;
}

__global__ void __launch_bounds__(kBlock) k_co_sym(CoArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nnz * 9) return;
  const uint32_t q = i / 9, e = i % 9, ra = e / 3, rb = e % 3;
  a.sym[i] = 0.5 * (a.raw[i] + a.raw[(size_t)a.mir[q] * 9 + rb * 3 + ra]);
}

// Dc and the diagonal-block inverse of every domain (invertSym33 on the
// block as coarsePcg reads it, Dc as the fallback pivots).
__global__ void __launch_bounds__(kBlock) k_co_diag(CoArgs a) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= a.D) return;
  const double* B = a.sym + (size_t)a.selfp[d] * 9;
  const double* R = a.raw + (size_t)a.selfp[d] * 9;
  const double f0 = R[0], f1 = R[4], f2 = R[8];
  a.dc[d * 3] = f0; a.dc[d * 3 + 1] = f1; a.dc[d * 3 + 2] = f2;
  const double B00 = B[0], B01 = 0.5 * (B[1] + B[3]), B02 = 0.5 * (B[2] + B[6]);
  const double B11 = B[4], B12 = 0.5 * (B[5] + B[7]), B22 = B[8];
  const double tr = fabs(B00) + fabs(B11) + fabs(B22);
  const double tiny = 1e-20 * (1 + tr);
  bool ok = true;
  if (!(B00 > tiny)) ok = false;
  if (ok && !(B00 * B11 - B01 * B01 > tiny * (1 + fabs(B00) + fabs(B11)))) ok = false;
  const double det = B00 * (B11 * B22 - B12 * B12) - B01 * (B01 * B22 - B12 * B02) + B02 * (B01 * B12 - B11 * B02);
  if (ok && !(det > tiny * (1 + tr) * (1 + tr))) ok = false;
  double* o = a.binv + (size_t)d * 9;
  if (ok) {
    const double s = 1 / det;
    o[0] = (B11 * B22 - B12 * B12) * s;
    o[1] = (B02 * B12 - B01 * B22) * s;
    o[2] = (B01 * B12 - B11 * B02) * s;
    o[3] = o[1];
    o[4] = (B00 * B22 - B02 * B02) * s;
    o[5] = (B01 * B02 - B00 * B12) * s;
    o[6] = o[2];
    o[7] = o[5];
    o[8] = (B00 * B11 - B01 * B01) * s;
  } else {
    for (int k = 0; k < 9; k++) o[k] = 0;
    o[0] = f0 > 1e-12 ? 1 / f0 : 0;
    o[4] = f1 > 1e-12 ? 1 / f1 : 0;
    o[8] = f2 > 1e-12 ? 1 / f2 : 0;
  }
}

// ---------------------------------------------------------------- apply

__device__ __forceinline__ bool co_done(const int32_t* done) {
  return done && __hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// rc = R r: thread d * 3 + axis, members in order (restrictCoarse, mas.ts:23-40).
template <class T>
__global__ void __launch_bounds__(kBlock) k_co_restrict(CoArgs a, const T* __restrict__ r, const int32_t* done) {
  if (co_done(done)) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.D * 3) return;
  const uint32_t d = i / 3, ax = i % 3;
  const double s = a.sd[d];
  double acc = 0;
  for (uint32_t k = a.mem_off[d]; k < a.mem_off[d + 1]; k++) acc += (double)r[(size_t)a.mem_v[k] * 3 + ax] * s;
  a.rc[i] = acc;
}

// The single-sweep corrections: c = omega rc / Dc (mode 1) or the diagonal
// block inverse times rc (mode 2), one thread per coarse dof.
__global__ void __launch_bounds__(kBlock) k_co_local(CoArgs a, int mode, double omega, const int32_t* done) {
  if (co_done(done)) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.D * 3) return;
  const uint32_t d = i / 3, ax = i % 3;
  if (mode == CS_COARSE_JACOBI) {
    const double dd = a.dc[i];
    a.c[i] = dd > 1e-12 ? (omega * a.rc[i]) / dd : 0;
  } else {
    const double* f = a.binv + (size_t)d * 9 + ax * 3;
    a.c[i] = f[0] * a.rc[d * 3] + f[1] * a.rc[d * 3 + 1] + f[2] * a.rc[d * 3 + 2];
  }
}

// z += P c: thread per vertex (prolongateAdd, mas.ts:43-58).
template <class T>
__global__ void __launch_bounds__(kBlock) k_co_prolong(CoArgs a, T* z, const int32_t* done) {
  if (co_done(done)) return;
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.n) return;
  const uint32_t d = a.dom_of[v];
  const double s = a.sd[d];
  for (int ax = 0; ax < 3; ax++) {
    const size_t j = (size_t)v * 3 + ax;
    if constexpr (sizeof(T) == sizeof(double)) z[j] += a.c[d * 3 + ax] * s;
    else z[j] = (T)((double)z[j] + a.c[d * 3 + ax] * s);
  }
}

// Sum over the work-group (fixed tree), every thread gets the total.
__device__ __forceinline__ double co_block_sum(double v, double* lds) {
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_down(v, o, kWave);
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double t = 0;
  for (int k = 0; k < kCoPcgBlock / kWave; k++) t += lds[k];
  return t;
}

__device__ __forceinline__ void co_bjac(const CoArgs& a, const double* r, double* z) {
  for (uint32_t i = threadIdx.x; i < a.D * 3; i += kCoPcgBlock) {
    const uint32_t d = i / 3, ax = i % 3;
    const double* f = a.binv + (size_t)d * 9 + ax * 3;
    z[i] = f[0] * r[d * 3] + f[1] * r[d * 3 + 1] + f[2] * r[d * 3 + 2];
  }
}

// coarsePcg (coarse-csr.ts:243-303) on the symmetrised blocks: x = 0,
// block-Jacobi preconditioner, relative tolerance `tol`, pAp breakdown latch;
// the solution goes to c.  One work-group.
__global__ void __launch_bounds__(kCoPcgBlock) k_co_pcg(CoArgs a, int iters, double tol, const int32_t* done) {
  if (co_done(done)) return;
  __shared__ double lds[kCoPcgBlock / kWave];
  const uint32_t n = a.D * 3;
  double rz = 0, bb = 0;
  for (uint32_t i = threadIdx.x; i < n; i += kCoPcgBlock) {
    a.c[i] = 0;
    a.r[i] = a.rc[i];
    bb += a.rc[i] * a.rc[i];
  }
  __syncthreads();
  co_bjac(a, a.r, a.z);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += kCoPcgBlock) {
    a.p[i] = a.z[i];
    rz += a.r[i] * a.z[i];
  }
  rz = co_block_sum(rz, lds);
  bb = co_block_sum(bb, lds);
  const double absTol = fmax(tol * fmax(sqrt(bb), 1e-12), 1e-12);
  if (sqrt(bb) < absTol) return;  // r = b at the start
  for (int k = 0; k < iters; k++) {
    __syncthreads();  // p complete
    double pap = 0;
    for (uint32_t d1 = threadIdx.x; d1 < a.D; d1 += kCoPcgBlock) {
      double y[3] = {0, 0, 0};
      for (uint32_t q = a.off[d1]; q < a.off[d1 + 1]; q++) {
        const double* B = a.sym + (size_t)q * 9;
        const double* xp = a.p + (size_t)a.col[q] * 3;
        for (int ax = 0; ax < 3; ax++) {
          double s = 0;
          for (int b = 0; b < 3; b++) s += B[ax * 3 + b] * xp[b];
          y[ax] += s;
        }
      }
      for (int ax = 0; ax < 3; ax++) {
        a.Ap[d1 * 3 + ax] = y[ax];
        pap += a.p[d1 * 3 + ax] * y[ax];
      }
    }
    pap = co_block_sum(pap, lds);
    if (!(pap > 1e-30)) return;
    const double alpha = rz / pap;
    double rr = 0;
    for (uint32_t i = threadIdx.x; i < n; i += kCoPcgBlock) {
      a.c[i] += alpha * a.p[i];
      const double ri = a.r[i] - alpha * a.Ap[i];
      a.r[i] = ri;
      rr += ri * ri;
    }
    rr = co_block_sum(rr, lds);
    if (sqrt(rr) < absTol) return;
    __syncthreads();  // r complete
    co_bjac(a, a.r, a.z);
    double rzn = 0;
    for (uint32_t i = threadIdx.x; i < n; i += kCoPcgBlock) rzn += a.r[i] * a.z[i];
    rzn = co_block_sum(rzn, lds);
    const double beta = rzn / rz;
    for (uint32_t i = threadIdx.x; i < n; i += kCoPcgBlock) a.p[i] = a.z[i] + beta * a.p[i];
    rz = rzn;
  }
}

// ---------------------------------------------------------------- facade

Coarse::Coarse(const cs_scene_desc& d, const std::vector<std::vector<uint32_t>>& members,
               const std::vector<uint32_t>& domain_of, int mode, int iters, double omega, hipStream_t st)
    : im_(new CoarseImpl) {
  CoarseImpl& C = *im_;
  if (mode < CS_COARSE_JACOBI || mode > CS_COARSE_PCG) {
    delete im_;
    throw std::invalid_argument("unknown coarse level");
  }
  C.st = st;
  C.mode = mode;
  C.iters = iters > 0 ? iters : 8;
  C.omega = omega > 0 ? omega : 0.5;
  C.n = d.n_verts;
  C.m = d.n_tris;
  C.D = (uint32_t)members.size();
  const uint32_t D = C.D;
  // members (CSR) and the transfer scales
  std::vector<uint32_t> moff(D + 1, 0), mv;
  std::vector<double> sd(D), sinv(D);
  for (uint32_t k = 0; k < D; k++) {
    moff[k + 1] = moff[k] + (uint32_t)members[k].size();
    mv.insert(mv.end(), members[k].begin(), members[k].end());
    sd[k] = 1 / std::sqrt((double)members[k].size());
    sinv[k] = 1.0 / (double)members[k].size();
  }
  // pattern: domains sharing a triangle, columns sorted per row
  std::vector<std::vector<uint32_t>> rows(D);
  for (uint32_t t = 0; t < C.m; t++) {
    uint32_t ds[3];
    for (int k = 0; k < 3; k++) ds[k] = domain_of[d.indices[(size_t)t * 3 + k]];
    for (uint32_t x : ds)
      for (uint32_t y : ds) rows[x].push_back(y);
  }
  C.h_off.assign(D + 1, 0);
  C.h_col.clear();
  for (uint32_t k = 0; k < D; k++) {
    auto& r = rows[k];
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    C.h_off[k + 1] = C.h_off[k] + (uint32_t)r.size();
    C.h_col.insert(C.h_col.end(), r.begin(), r.end());
  }
  C.nnz = (uint32_t)C.h_col.size();
  auto pos = [&](uint32_t d1, uint32_t d2) -> uint32_t {
    const auto b = C.h_col.begin() + C.h_off[d1], e = C.h_col.begin() + C.h_off[d1 + 1];
    const auto it = std::lower_bound(b, e, d2);
    if (it == e || *it != d2) throw std::logic_error("coarse pattern misses a block");
    return (uint32_t)(it - C.h_col.begin());
  };
  std::vector<uint32_t> mir(C.nnz), selfp(D), brow(C.nnz);
  std::vector<double> sblk(C.nnz);
  for (uint32_t d1 = 0; d1 < D; d1++) {
    selfp[d1] = pos(d1, d1);
    for (uint32_t q = C.h_off[d1]; q < C.h_off[d1 + 1]; q++) {
      const uint32_t d2 = C.h_col[q];
      mir[q] = pos(d2, d1);
      brow[q] = d1;
      sblk[q] = 1 / std::sqrt((double)members[d1].size() * (double)members[d2].size());
    }
  }
  // corner pairs per block in (triangle, ci, cj) order (a counting sort keeps it)
  std::vector<uint32_t> qof((size_t)C.m * 9), coff(C.nnz + 1, 0);
  for (uint32_t t = 0; t < C.m; t++)
    for (int ci = 0; ci < 3; ci++)
      for (int cj = 0; cj < 3; cj++) {
        const uint32_t q = pos(domain_of[d.indices[(size_t)t * 3 + ci]], domain_of[d.indices[(size_t)t * 3 + cj]]);
        qof[(size_t)t * 9 + ci * 3 + cj] = q;
        coff[q + 1]++;
      }
  for (uint32_t q = 0; q < C.nnz; q++) coff[q + 1] += coff[q];
  std::vector<uint32_t> cur(coff.begin(), coff.end() - 1), crec((size_t)C.m * 9);
  for (uint32_t t = 0; t < C.m; t++)
    for (uint32_t k = 0; k < 9; k++) crec[cur[qof[(size_t)t * 9 + k]]++] = (t << 4) | ((k / 3) << 2) | (k % 3);
  if (C.m >= (1u << 28)) throw std::invalid_argument("too many triangles for the coarse level");
  C.mem_off.upload(moff.data(), moff.size(), st);
  C.mem_v.upload(mv.data(), std::max<size_t>(mv.size(), 1), st);
  C.dom_of.upload(domain_of.data(), C.n, st);
  C.sd.upload(sd.data(), D, st);
  C.sinv.upload(sinv.data(), D, st);
  C.off.upload(C.h_off.data(), C.h_off.size(), st);
  C.col.upload(C.h_col.data(), C.nnz, st);
  C.mir.upload(mir.data(), C.nnz, st);
  C.selfp.upload(selfp.data(), D, st);
  C.brow.upload(brow.data(), C.nnz, st);
  C.sblk.upload(sblk.data(), C.nnz, st);
  C.co_off.upload(coff.data(), coff.size(), st);
  C.co_rec.upload(crec.data(), std::max<size_t>(crec.size(), 1), st);
  C.hrows.alloc((size_t)std::max<uint32_t>(C.m, 1) * 81);
  C.raw.alloc((size_t)C.nnz * 9);
  C.sym.alloc((size_t)C.nnz * 9);
  C.dc.alloc((size_t)D * 3);
  C.binv.alloc((size_t)D * 9);
  for (auto* b : {&C.rc, &C.c, &C.r, &C.z, &C.p, &C.Ap}) b->alloc((size_t)D * 3);
  CS_HIP(hipStreamSynchronize(st));  // the host vectors above go out of scope
}

Coarse::~Coarse() { delete im_; }

void Coarse::assemble(const DMesh& me, const DMat& mat, const double* x, const double* diag) {
  CoarseImpl& C = *im_;
  if (!C.D) return;
  const CoArgs a = C.args();
  if (C.m) hipLaunchKernelGGL(k_co_rows, dim3(grid_for((size_t)C.m * 3)), dim3(kBlock), 0, C.st, a, me, mat, x);
  hipLaunchKernelGGL(k_co_blocks, dim3(grid_for((size_t)C.nnz * 9)), dim3(kBlock), 0, C.st, a, diag);
  hipLaunchKernelGGL(k_co_sym, dim3(grid_for((size_t)C.nnz * 9)), dim3(kBlock), 0, C.st, a);
  hipLaunchKernelGGL(k_co_diag, dim3(grid_for(C.D)), dim3(kBlock), 0, C.st, a);
  CS_LAUNCH_CHECK();
}

template <class T>
static void co_apply(CoarseImpl& C, const T* r, T* z, const int32_t* done) {
  if (!C.D) return;
  const CoArgs a = C.args();
  const dim3 b(kBlock), gc(grid_for((size_t)C.D * 3));
  hipLaunchKernelGGL(k_co_restrict<T>, gc, b, 0, C.st, a, r, done);
  if (C.mode == CS_COARSE_PCG)
    hipLaunchKernelGGL(k_co_pcg, dim3(1), dim3(kCoPcgBlock), 0, C.st, a, C.iters, 1e-3, done);
  else
    hipLaunchKernelGGL(k_co_local, gc, b, 0, C.st, a, C.mode, C.omega, done);
  hipLaunchKernelGGL(k_co_prolong<T>, dim3(grid_for(C.n)), b, 0, C.st, a, z, done);
  CS_LAUNCH_CHECK();
}

void Coarse::apply64(const double* r, double* z, const int32_t* done) { co_apply<double>(*im_, r, z, done); }
void Coarse::apply32(const float* r, float* z, const int32_t* done) { co_apply<float>(*im_, r, z, done); }

int Coarse::mode() const { return im_->mode; }
uint32_t Coarse::domains() const { return im_->D; }
uint64_t Coarse::nnz() const { return im_->nnz; }

void Coarse::pattern(uint32_t* off, uint32_t* col) const {
  if (off) std::copy(im_->h_off.begin(), im_->h_off.end(), off);
  if (col) std::copy(im_->h_col.begin(), im_->h_col.end(), col);
}

void Coarse::values(double* raw, double* sym, double* dc) const {
  const CoarseImpl& C = *im_;
  CS_HIP(hipStreamSynchronize(C.st));
  if (raw && C.nnz) CS_HIP(hipMemcpy(raw, C.raw.ptr, (size_t)C.nnz * 9 * sizeof(double), hipMemcpyDeviceToHost));
  if (sym && C.nnz) CS_HIP(hipMemcpy(sym, C.sym.ptr, (size_t)C.nnz * 9 * sizeof(double), hipMemcpyDeviceToHost));
  if (dc && C.D) CS_HIP(hipMemcpy(dc, C.dc.ptr, (size_t)C.D * 3 * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace cloth
