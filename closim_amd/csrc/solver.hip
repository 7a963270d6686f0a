// solver.hip — HipSolver: the implicit cloth step on one MI355X behind the
// C ABI of include/cloth_hip.h (drop-in for CpuSolver, cpu-solver.ts:7-49).
//
// Control flow follows implicitStep (newton.ts:27-232) on the host; all data
// stays resident in HBM and only scalars cross PCIe (gradient norm, PCG
// completion, g.dx, max |dx|, trial energies, contact counts/validity).
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <mutex>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cloth_hip.h"
#include "contact.h"
#include "fem.h"
#include "mixed.h"
#include "pcg.h"
#include "lbvh.h"
#include "schwarz.h"

namespace cloth {

thread_local std::string g_last_error;

constexpr int kArmijoBatch = 8;  // largest batch of line-search trials evaluated before one host read

// Scalars of the adjoint's block-Jacobi PCG, kept on the device so that its
// iterations run without a host read between them.
struct AdjPcg {
  double rz[2];  // r.z of the current iterate and of the next, alternating with the iteration's parity
  double pap, rr, bnorm2, rel, tol;
  int32_t it;    // iterations taken (each with p.Ap > 0)
  int32_t done;  // 0 running, 1 p.Ap <= 0 (not positive definite along p), 2 converged or rhs = 0
};

struct NewtonScalars {
  PcgScalars pcg;
  double grad_sumsq;
  double e2[2];
  double merit;
  double gtdx;
  double maxdx;
  double max_strain;
  double adj[4];  // adjoint dot products: r.z, p.Ap, r.r, a.a
  AdjPcg apcg;    // the adjoint PCG's scalars (full_pcg)
  // device Newton control (CS_FLAG_DEVICE_NEWTON)
  double tscale;                   // trust-region scale applied to dx
  int32_t fallback;                // 1 when the descent safeguard took -g / diag
  double etrial[kArmijoBatch][2];  // inertia and internal energy of each trial of a batch
  double mtrial[kArmijoBatch];     // merit (barrier) energy of each trial
};

// Device Newton control: the descent safeguard (newton.ts:154-160) and the
// trust region (:167-179) decided on the device from the reductions the
// preceding kernels wrote, so dx and g.dx are final without a host read.
__global__ void k_ctl_descent(NewtonScalars* s, const int32_t* pcg_done, int gate_breakdown) {
  const double g = s->gtdx;
  s->fallback = (!isfinite(g) || g >= 0 || (gate_breakdown && *pcg_done == 1)) ? 1 : 0;
}

__global__ void k_fallback_pred(const double* __restrict__ g, const double* __restrict__ diag, double* dx, size_t n,
                                const NewtonScalars* s) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n && s->fallback) dx[i] = -g[i] / diag[i];
}

__global__ void k_ctl_trust(NewtonScalars* s, double trust) {
  const double m = s->maxdx;
  if (m > trust) {
    const double sc = trust / m;
    s->tscale = sc;
    s->gtdx *= sc;
  } else {
    s->tscale = 1.0;
  }
}

__global__ void k_scale_pred(double* v, size_t n, const NewtonScalars* s, double trust) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n && s->maxdx > trust) v[i] *= s->tscale;
}

struct NonFinite {};

__global__ void k_predictor(const double* x0, const double* v0, double h, double gx, double gy, double gz,
                            double* yhat, double* x, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // yHat = x + h v + h^2 g (newton.ts:42-47)
  const double g3[3] = {gx, gy, gz};
  for (int k = 0; k < 3; k++) {
    const size_t j = (size_t)i * 3 + k;
    const double y = x0[j] + h * v0[j] + h * h * g3[k];
    yhat[j] = y;
    x[j] = y;
  }
}

__global__ void k_enforce_pins(const uint32_t* ids, const double* pos, uint32_t np, double* x) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const size_t v = ids[k];
  x[v * 3] = pos[k * 3];
  x[v * 3 + 1] = pos[k * 3 + 1];
  x[v * 3 + 2] = pos[k * 3 + 2];
}

__global__ void k_init_diag(const float* mass, double inv_h2, double beta, double* diag, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = mass[i] * inv_h2 + beta;  // newton.ts:103-109
  diag[(size_t)i * 3] = d;
  diag[(size_t)i * 3 + 1] = d;
  diag[(size_t)i * 3 + 2] = d;
}

__global__ void k_velocity(const double* x, const double* x0, double h, double damp, double* v, size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n3) v[i] = ((x[i] - x0[i]) / h) * damp;  // newton.ts:222-226
}

// ---- adjoint kernels (step_adjoint) ----------------------------------------

// out = x + (eps / |p|_inf) p, or x where p = 0 (|p|_inf read on the device)
__global__ void k_adj_shift(const double* x, const double* p, const double* pmax, double eps, double* out,
                            size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const double m = *pmax;
  const double s = m > 0 ? eps / m : 0.0;
  if (i < n3) out[i] = x[i] + s * p[i];
}

// out += (gp - gm) |p|_inf / (2 eps), nothing where p = 0
__global__ void k_adj_fd_add(const double* gp, const double* gm, const double* pmax, double two_eps, double* out,
                             size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const double m = *pmax;
  if (!(m > 0)) return;
  const double scale = m / two_eps;
  if (i < n3) out[i] += (gp[i] - gm[i]) * scale;
}

// PCG start: nothing to do for a zero right-hand side
__global__ void k_adj_pcg_init(AdjPcg* a, double tol) {
  a->tol = tol;
  a->it = 0;
  a->done = a->bnorm2 > 0 ? 0 : 2;
  a->rel = a->bnorm2 > 0 ? 1.0 : 0.0;
}

// sol += alpha p, r -= alpha Ap, alpha = r.z / p.Ap; skipped once the solve
// has ended or where p.Ap <= 0
__global__ void k_adj_pcg_step(const AdjPcg* a, int cur, double* sol, double* r, const double* p, const double* ap,
                               size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (a->done || !(a->pap > 0)) return;
  const double alpha = a->rz[cur] / a->pap;
  if (i < n3) {
    sol[i] += alpha * p[i];
    r[i] += -alpha * ap[i];
  }
}

// after r.r: count the iteration, its relative residual, and the stops
__global__ void k_adj_pcg_decide(AdjPcg* a) {
  if (a->done) return;
  if (!(a->pap > 0)) {
    a->done = 1;
    return;
  }
  a->it++;
  a->rel = sqrt(a->rr / a->bnorm2);
  if (a->rel <= a->tol) a->done = 2;
}

// p = z + beta p, beta = (r.z)_next / (r.z)_current
__global__ void k_adj_pcg_dir(const AdjPcg* a, int cur, double* p, const double* z, size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (a->done) return;
  const double beta = a->rz[cur ^ 1] / a->rz[cur];
  if (i < n3) p[i] = z[i] + beta * p[i];
}

// z = B^-1 r per vertex with the block-Jacobi inverses (pins zeroed)
__global__ void k_adj_pre_block(const double* r, const double* binv, const uint8_t* pinmask, double* z, uint32_t n) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const double* f = binv + (size_t)v * 9;
  const double r0 = r[(size_t)v * 3], r1 = r[(size_t)v * 3 + 1], r2 = r[(size_t)v * 3 + 2];
  const bool pinned = pinmask[v] != 0;
  z[(size_t)v * 3] = pinned ? 0.0 : f[0] * r0 + f[1] * r1 + f[2] * r2;
  z[(size_t)v * 3 + 1] = pinned ? 0.0 : f[3] * r0 + f[4] * r1 + f[5] * r2;
  z[(size_t)v * 3 + 2] = pinned ? 0.0 : f[6] * r0 + f[7] * r1 + f[8] * r2;
}

__global__ void k_adj_axpy(double* y, double a, const double* x, size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n3) y[i] += a * x[i];
}

__global__ void k_adj_xpby(double* p, const double* z, double beta, size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n3) p[i] = z[i] + beta * p[i];
}

// right-hand side: dL/dx_{n+1} plus the velocity's share, free DOFs only
__global__ void k_adj_rhs(const double* gx, const double* gv, double c, const uint8_t* pinmask, double* a,
                          size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n3) a[i] = pinmask[i / 3] ? 0.0 : gx[i] + c * gv[i];
}

// -g on the free DOFs (the Newton right-hand side of the root polish)
__global__ void k_neg_free(const double* g, const uint8_t* pinmask, double* out, size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n3) out[i] = pinmask[i / 3] ? 0.0 : -g[i];
}

// dL/dx_n = lambda M / h^2 - (damp / h) dL/dv_{n+1};  dL/dv_n = lambda M / h
__global__ void k_adj_out(const double* lam, const double* gv, const float* mass, double h, double damp,
                          double* dxp, double* dvp, size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n3) return;
  const double m = mass[i / 3];
  dxp[i] = lam[i] * m / (h * h) - (damp / h) * gv[i];
  dvp[i] = lam[i] * m / h;
}

// Vertex-owner tiles of the FP32 HVP (DTiles): vertices in Morton order of
// the rest positions are taken greedily into a tile until its owned vertices,
// its triangles (all incident to an owned vertex) or its local vertices would
// exceed kTileOwn / kTileTris / kTileLoc.  Within a tile the triangles are in
// ascending order, the local vertices are the owned ones followed by the
// others in order of appearance, and each owned vertex lists its corners in
// ascending triangle order.  Returns false (no tiles) when one vertex alone
// exceeds a tile.
struct TileHost {
  std::vector<TileHdr> hdr;
  std::vector<uint32_t> loc;
  std::vector<float4> inv;
  std::vector<uint2> lidx;
  std::vector<float> area;
  std::vector<uint16_t> off, ent;
};

static bool build_tiles(const cs_scene_desc& d, const std::vector<uint32_t>& voff, const std::vector<uint32_t>& vrec,
                        TileHost& T) {
  const uint32_t n = d.n_verts, m = d.n_tris;
  T = TileHost{};
  if (!n) return false;
  // Morton order of the rest positions (10 bits per axis over their box)
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  auto rp = [&](uint32_t v, int k) -> double {
    return d.rest_positions ? (double)d.rest_positions[(size_t)v * 3 + k] : d.positions[(size_t)v * 3 + k];
  };
  for (uint32_t v = 0; v < n; v++)
    for (int k = 0; k < 3; k++) {
      const double c = rp(v, k);
      if (c == c) { lo[k] = std::min(lo[k], c); hi[k] = std::max(hi[k], c); }
    }
  std::vector<std::pair<uint32_t, uint32_t>> key(n);
  for (uint32_t v = 0; v < n; v++) {
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
      const double span = hi[k] - lo[k];
      double f = span > 0 ? (rp(v, k) - lo[k]) / span : 0.0;
      f = f == f ? std::min(std::max(f, 0.0), 1.0) : 0.0;
      q[k] = std::min<uint32_t>(1023u, (uint32_t)(f * 1024.0));
    }
    uint32_t code = 0;
    for (int b = 9; b >= 0; b--) code = (code << 3) | (((q[0] >> b) & 1) << 2) | (((q[1] >> b) & 1) << 1) | ((q[2] >> b) & 1);
    key[v] = {code, v};
  }
  std::sort(key.begin(), key.end());
  std::vector<int64_t> tstamp(m, -1), vstamp(n, -1);
  std::vector<uint32_t> own, tris;
  uint32_t nloc = 0;
  int64_t tile = 0;
  auto close = [&]() {
    if (own.empty()) return;
    std::sort(tris.begin(), tris.end());
    TileHdr h{};
    h.loc = (uint32_t)T.loc.size();
    h.n_own = (uint32_t)own.size();
    h.tri = (uint32_t)T.inv.size();
    h.n_tri = (uint32_t)tris.size();
    h.ent = (uint32_t)T.ent.size();
    h.off = (uint32_t)T.off.size();
    std::unordered_map<uint32_t, uint32_t> lid;
    for (uint32_t v : own) {
      lid.emplace(v, (uint32_t)lid.size());
      T.loc.push_back(v);
    }
    for (uint32_t t : tris)
      for (int c = 0; c < 3; c++) {
        const uint32_t u = d.indices[(size_t)t * 3 + c];
        if (lid.emplace(u, (uint32_t)lid.size()).second) T.loc.push_back(u);
      }
    h.n_loc = (uint32_t)lid.size();
    for (uint32_t t : tris) {
      const uint32_t a = lid[d.indices[(size_t)t * 3]], b = lid[d.indices[(size_t)t * 3 + 1]],
                     c = lid[d.indices[(size_t)t * 3 + 2]];
      T.lidx.push_back(make_uint2(a | (b << 16), c));
      T.inv.push_back(make_float4(d.inv_dm[(size_t)t * 4], d.inv_dm[(size_t)t * 4 + 1], d.inv_dm[(size_t)t * 4 + 2],
                                  d.inv_dm[(size_t)t * 4 + 3]));
      T.area.push_back(d.areas[t]);
    }
    uint32_t e = 0;
    for (uint32_t v : own) {
      T.off.push_back((uint16_t)e);
      for (uint32_t q = voff[v]; q < voff[v + 1]; q++) {
        const uint32_t t = vrec[q] >> 2, c = vrec[q] & 3;
        const uint32_t slot = (uint32_t)(std::lower_bound(tris.begin(), tris.end(), t) - tris.begin());
        T.ent.push_back((uint16_t)(slot * 4 + c));
        e++;
      }
    }
    T.off.push_back((uint16_t)e);
    T.hdr.push_back(h);
    own.clear();
    tris.clear();
    nloc = 0;
    tile++;
  };
  for (const auto& kv : key) {
    const uint32_t v = kv.second;
    for (int attempt = 0; attempt < 2; attempt++) {
      // what v would add to the current tile
      uint32_t add_t = 0, add_l = vstamp[v] == tile ? 0 : 1;
      std::vector<uint32_t> nl;
      for (uint32_t q = voff[v]; q < voff[v + 1]; q++) {
        const uint32_t t = vrec[q] >> 2;
        if (tstamp[t] == tile) continue;
        add_t++;
        for (int c = 0; c < 3; c++) {
          const uint32_t u = d.indices[(size_t)t * 3 + c];
          if (u != v && vstamp[u] != tile && std::find(nl.begin(), nl.end(), u) == nl.end()) nl.push_back(u);
        }
      }
      add_l += (uint32_t)nl.size();
      if (own.size() + 1 <= kTileOwn && tris.size() + add_t <= kTileTris && nloc + add_l <= kTileLoc) {
        own.push_back(v);
        if (vstamp[v] != tile) { vstamp[v] = tile; nloc++; }
        for (uint32_t q = voff[v]; q < voff[v + 1]; q++) {
          const uint32_t t = vrec[q] >> 2;
          if (tstamp[t] == tile) continue;
          tstamp[t] = tile;
          tris.push_back(t);
        }
        for (uint32_t u : nl) { vstamp[u] = tile; nloc++; }
        break;
      }
      if (attempt == 1 || own.empty()) {
        T = TileHost{};
        return false;  // v alone does not fit a tile
      }
      close();
    }
  }
  close();
  return true;
}

// Packed batch (cs_batch_step_packed): every instance steps in its own host
// thread; at each Newton iteration its PCG system is handed over here, and
// when every instance still stepping has handed one over, the last to arrive
// solves them all together (MixedSolve::solve_packed) and releases the others.
// An instance that finishes its step leaves, so the rest never wait for it.
struct PackedBatch {
  std::mutex mu;
  std::condition_variable cv;
  int active = 0;    // instances still inside their step
  uint64_t gen = 0;  // completed packed solves
  std::vector<PackedJob> pending;
  std::vector<int*> outs;
  std::vector<std::exception_ptr*> errs;
  hipStream_t st = nullptr;
  PackedScratch scratch;

  int solve(MixedSolve* ms, const MixedInputs& in) {
    std::unique_lock<std::mutex> lk(mu);
    int out = -1;
    std::exception_ptr err;
    pending.push_back(PackedJob{ms, in, 0});
    outs.push_back(&out);
    errs.push_back(&err);
    const uint64_t mine = gen;
    if ((int)pending.size() == active) run(lk);
    else cv.wait(lk, [&] { return gen != mine; });
    if (err) std::rethrow_exception(err);
    return out;
  }
  void leave() {
    std::unique_lock<std::mutex> lk(mu);
    active--;
    if (!pending.empty() && (int)pending.size() == active) run(lk);
  }
  // with lk held; every instance still stepping is waiting in solve()
  void run(std::unique_lock<std::mutex>& lk) {
    std::vector<PackedJob> jobs;
    std::vector<int*> o;
    std::vector<std::exception_ptr*> e;
    jobs.swap(pending);
    o.swap(outs);
    e.swap(errs);
    lk.unlock();
    std::exception_ptr err;
    try {
      MixedSolve::solve_packed(jobs, st, scratch);
    } catch (...) {
      err = std::current_exception();
    }
    lk.lock();
    for (size_t i = 0; i < jobs.size(); i++) {
      *o[i] = jobs[i].iters;
      if (err) *e[i] = err;
    }
    gen++;
    cv.notify_all();
  }
};

struct Solver {
  int device = 0;
  hipStream_t st = nullptr;
  cs_options opt;
  uint32_t n = 0, m = 0, nh = 0;
  cs_material mat;
  double gravity[3];
  bool has_contact = false;
  PackedBatch* packer = nullptr;  // set for the duration of a packed batch step
  cs_step_stats last;
  bool has_last = false;

  // mesh
  DevBuf<uint32_t> idx, vt_off, vt_rec, hv, vh_off, vh_rec;
  DevBuf<uint4> vt_corner;
  // vertex-owner tiles of the FP32 HVP
  DevBuf<TileHdr> tl_hdr;
  DevBuf<uint32_t> tl_loc;
  DevBuf<float4> tl_inv;
  DevBuf<uint2> tl_lidx;
  DevBuf<float> tl_area;
  DevBuf<uint16_t> tl_off, tl_ent;
  uint32_t tl_count = 0;
  uint64_t tl_tris = 0, tl_locs = 0, tl_ents = 0;
  DevBuf<float> inv_dm, area, mass;
  DevBuf<double> hrest, hlen, hasum;
  // state
  DevBuf<double> pos, vel, x0, yhat, x, xt, xe, g, b, dx, diag, ceg;
  // pcg
  DevBuf<double> r, z, p, Ap, binv;
  DevBuf<jsm::dd> partials;
  DevBuf<double> mpartials;
  DevBuf<NewtonScalars> scal;
  NewtonScalars* hscal = nullptr;  // pinned host mirror
  // pins (Map insertion order)
  std::vector<uint32_t> pin_ids;
  std::vector<double> pin_pos;
  DevBuf<uint32_t> d_pin_ids;
  DevBuf<double> d_pin_pos;
  DevBuf<uint8_t> pinmask;
  bool pins_dirty = true;
  // subsystems
  std::unique_ptr<ContactSystem> contact;
  std::unique_ptr<MixedSolve> mixed;
  std::unique_ptr<Schwarz> schwarz;
  // timing
  hipEvent_t ev_pool[2 * 4096];
  int ev_used = 0;
  bool ev_init = false;
  double kt_ms[CS_KERNEL_COUNT] = {0};
  uint64_t kt_count[CS_KERNEL_COUNT] = {0};
  std::vector<std::pair<int, int>> ev_pending[CS_KERNEL_COUNT];

  ~Solver() {
    if (st) (void)hipStreamSynchronize(st);
    if (hscal) (void)hipHostFree(hscal);
    if (ev_init)
      for (auto& e : ev_pool) (void)hipEventDestroy(e);
    contact.reset();
    mixed.reset();
    if (st) (void)hipStreamDestroy(st);
  }

  DMesh dmesh() const {
    DMesh d;
    d.n = n; d.m = m; d.nh = nh;
    d.idx = idx.ptr; d.inv_dm = inv_dm.ptr; d.area = area.ptr; d.mass = mass.ptr;
    d.vt_off = vt_off.ptr; d.vt_rec = vt_rec.ptr; d.vt_corner = vt_corner.ptr;
    d.hv = hv.ptr; d.hrest = hrest.ptr; d.hlen = hlen.ptr; d.hasum = hasum.ptr;
    d.vh_off = vh_off.ptr; d.vh_rec = vh_rec.ptr;
    d.pinmask = pinmask.ptr;
    d.tl = DTiles{tl_hdr.ptr, tl_count, tl_loc.ptr, tl_inv.ptr, tl_lidx.ptr, tl_area.ptr, tl_off.ptr, tl_ent.ptr};
    return d;
  }
  DMat dmat() const {
    return DMat{mat.thickness, mat.stretch_warp, mat.stretch_weft, mat.stretch_coupling, mat.shear,
                0.5 * (mat.bend_warp + mat.bend_weft)};
  }

  template <typename T>
  T read(const T* dptr) {
    T v;
    CS_HIP(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
    return v;
  }
  void sync_scalars() {
    CS_HIP(hipMemcpyAsync(hscal, scal.ptr, sizeof(NewtonScalars), hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
  }
  NewtonScalars* ds() { return scal.ptr; }

  // ---- timing --------------------------------------------------------------
  int ev_mark() {
    if (!(opt.flags & CS_FLAG_TIMING)) return -1;
    if (!ev_init) {
      for (auto& e : ev_pool) CS_HIP(hipEventCreate(&e));
      ev_init = true;
    }
    if (ev_used >= (int)(sizeof(ev_pool) / sizeof(ev_pool[0]))) flush_timing();
    const int k = ev_used++;
    CS_HIP(hipEventRecord(ev_pool[k], st));
    return k;
  }
  void ev_span(int kernel, int a, int b) {
    if (a >= 0 && b >= 0) ev_pending[kernel].push_back({a, b});
  }
  void flush_timing() {
    if (!ev_init) return;
    CS_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < CS_KERNEL_COUNT; k++) {
      for (auto& pr : ev_pending[k]) {
        float ms = 0;
        CS_HIP(hipEventElapsedTime(&ms, ev_pool[pr.first], ev_pool[pr.second]));
        kt_ms[k] += ms;
        kt_count[k]++;
      }
      ev_pending[k].clear();
    }
    ev_used = 0;
  }

  // ---- setup ---------------------------------------------------------------
  void build_csr(const uint32_t* elems, uint32_t count, int corners, uint32_t nv, std::vector<uint32_t>& off,
                 std::vector<uint32_t>& rec) {
    // incidence.ts:20-53: records appended in ascending element order.
    off.assign(nv + 1, 0);
    for (uint32_t e = 0; e < count; e++)
      for (int c = 0; c < corners; c++) off[elems[(size_t)e * corners + c] + 1]++;
    for (uint32_t v = 0; v < nv; v++) off[v + 1] += off[v];
    rec.assign((size_t)count * corners, 0);
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (uint32_t e = 0; e < count; e++)
      for (int c = 0; c < corners; c++) rec[cur[elems[(size_t)e * corners + c]]++] = (e << 2) | (uint32_t)c;
  }

  void init(const cs_scene_desc* d, const cs_options* o) {
    opt = *o;
    device = opt.device;
    CS_HIP(hipSetDevice(device));
    CS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    n = d->n_verts; m = d->n_tris; nh = d->n_hinges;
    mat = d->material;
    for (int i = 0; i < 3; i++) gravity[i] = d->gravity[i];
    for (uint32_t k = 0; k < (size_t)m * 3; k++)
      if (d->indices[k] >= n) throw std::invalid_argument("triangle index out of range");
    for (uint32_t k = 0; k < (size_t)nh * 4; k++)
      if (d->hinge_verts[k] >= n) throw std::invalid_argument("hinge vertex out of range");
    idx.upload(d->indices, (size_t)m * 3, st);
    inv_dm.upload(d->inv_dm, (size_t)m * 4, st);
    area.upload(d->areas, m, st);
    mass.upload(d->masses, n, st);
    std::vector<uint32_t> off, rec;
    build_csr(d->indices, m, 3, n, off, rec);
    vt_off.upload(off.data(), off.size(), st);
    vt_rec.upload(rec.data(), rec.size(), st);
    {
      std::vector<uint4> cor(rec.size());
      for (size_t q = 0; q < rec.size(); q++) {
        const size_t t = rec[q] >> 2;
        cor[q] = make_uint4(d->indices[t * 3], d->indices[t * 3 + 1], d->indices[t * 3 + 2], rec[q]);
      }
      vt_corner.upload(cor.data(), cor.size(), st);
      TileHost T;
      if (o->precision == CS_PREC_MIXED && build_tiles(*d, off, rec, T)) {
        tl_hdr.upload(T.hdr.data(), T.hdr.size(), st);
        tl_loc.upload(T.loc.data(), T.loc.size(), st);
        tl_inv.upload(T.inv.data(), std::max<size_t>(T.inv.size(), 1), st);
        tl_lidx.upload(T.lidx.data(), std::max<size_t>(T.lidx.size(), 1), st);
        tl_area.upload(T.area.data(), std::max<size_t>(T.area.size(), 1), st);
        tl_off.upload(T.off.data(), T.off.size(), st);
        tl_ent.upload(T.ent.data(), std::max<size_t>(T.ent.size(), 1), st);
        tl_count = (uint32_t)T.hdr.size();
        tl_tris = T.inv.size();
        tl_locs = T.loc.size();
        tl_ents = T.ent.size();
        CS_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
      }
    }
    if (nh) {
      hv.upload(d->hinge_verts, (size_t)nh * 4, st);
      hrest.upload(d->hinge_rest_angle, nh, st);
      hlen.upload(d->hinge_edge_len, nh, st);
      hasum.upload(d->hinge_area_sum, nh, st);
    } else {
      hv.alloc(4); hrest.alloc(1); hlen.alloc(1); hasum.alloc(1);
    }
    build_csr(nh ? d->hinge_verts : nullptr, nh, 4, n, off, rec);
    vh_off.upload(off.data(), off.size(), st);
    vh_rec.upload(rec.data(), rec.size(), st);
    const size_t n3 = (size_t)n * 3;
    pos.upload(d->positions, n3, st);
    vel.upload(d->velocities, n3, st);
    for (auto* bf : {&x0, &yhat, &x, &xt, &xe, &g, &b, &dx, &diag, &ceg, &r, &z, &p, &Ap}) bf->alloc(n3);
    partials.alloc(8 * kMaxRedBlocks);
    mpartials.alloc(4 * kMaxRedBlocks);
    scal.alloc(1);
    scal.zero(st);
    CS_HIP(hipHostMalloc(&hscal, sizeof(NewtonScalars)));
    std::memset(hscal, 0, sizeof(NewtonScalars));
    pin_ids.assign(d->pin_ids, d->pin_ids + d->n_pins);
    pin_pos.assign(d->pin_pos, d->pin_pos + (size_t)d->n_pins * 3);
    for (uint32_t id : pin_ids)
      if (id >= n) throw std::invalid_argument("pin id out of range");
    pinmask.alloc(n);
    if (opt.precond == CS_PRECOND_BLOCK_JACOBI) binv.alloc((size_t)n * 9);
    has_contact = d->contact_enabled != 0;
    if (has_contact) contact.reset(new ContactSystem(*d, st));
    if (opt.precision == CS_PREC_MIXED) mixed.reset(new MixedSolve(*d, opt, st));
    if (opt.precond == CS_PRECOND_SCHWARZ)
      schwarz.reset(new Schwarz(*d, opt.schwarz_target, opt.schwarz_min, opt.schwarz_max, st, opt.coarse,
                                opt.coarse_iters, opt.coarse_omega));
    CS_HIP(hipStreamSynchronize(st));
  }

  void upload_pins() {
    if (!pins_dirty) return;
    std::vector<uint8_t> mask(n, 0);
    for (uint32_t id : pin_ids) mask[id] = 1;
    CS_HIP(hipMemcpyAsync(pinmask.ptr, mask.data(), n, hipMemcpyHostToDevice, st));
    if (!pin_ids.empty()) {
      d_pin_ids.upload(pin_ids.data(), pin_ids.size(), st);
      d_pin_pos.upload(pin_pos.data(), pin_pos.size(), st);
    }
    CS_HIP(hipStreamSynchronize(st));
    pins_dirty = false;
  }

  void enforce_pins(double* xv) {
    if (pin_ids.empty()) return;
    hipLaunchKernelGGL(k_enforce_pins, dim3(grid_for(pin_ids.size())), dim3(kBlock), 0, st, d_pin_ids.ptr,
                       d_pin_pos.ptr, (uint32_t)pin_ids.size(), xv);
    CS_LAUNCH_CHECK();
  }

  // totalEnergy (newton.ts:76-88).
  double total_energy(const double* xx, double h) {
    launch_energy(st, dmesh(), dmat(), xx, yhat.ptr, h, partials.ptr, ds()->e2);
    double merit = 0;
    if (contact) merit = contact->merit_energy(xx);
    sync_scalars();
    double e = hscal->e2[0];
    e += hscal->e2[1];
    if (contact) e += merit;
    return e;
  }

  // ---- adjoint of one step ---------------------------------------------------
  //
  // The step returns the root x1 of g(x; x0, v0, theta) = M (x - yhat) / h^2 +
  // grad E_int(x; theta) (+ contact) on the free DOFs, yhat = x0 + h v0 +
  // h^2 gravity, and v1 = damp (x1 - x0) / h.  For a loss L(x1, v1) the
  // implicit function theorem gives, with H = dg/dx at x1 and
  // a = dL/dx1 + (damp / h) dL/dv1 on the free DOFs,
  //   H lambda = a,   dL/dtheta_j = -lambda . dg/dtheta_j,
  //   dL/dx0 = lambda M / h^2 - (damp / h) dL/dv1 - lambda . dg_fric/dx0,   dL/dv0 = lambda M / h,
  // the last term from the lagged friction, whose slip is measured from x0.
  // H p is the analytic inertia + membrane product plus the central
  // difference of the bending and contact gradients along p (full_hvp), so H
  // includes the FD bending gradient the step uses, which the reference's
  // membrane-only HVP leaves out; block-Jacobi PCG (full_pcg) solves to `tol`
  // relative residual.  The membrane energy is linear in C00, C11, C01, G and the
  // hinge energy in kb = (bendWarp + bendWeft) / 2, so dg/dtheta_j is the
  // internal gradient of the material with theta_j = 1 and the others 0.
  // With contact, the active set is frozen at x1 (lagged friction as the
  // solver holds it); the velocity filter is not differentiated.
  DevBuf<double> adj_binv;
  DevBuf<double> adj_x1, adj_lam, adj_r, adj_z, adj_p, adj_ap, adj_gp, adj_gm, adj_xs, adj_gx, adj_gv, adj_a;

  void adj_gradient(const double* xx, double* out) {
    const double* cg = nullptr;
    if (contact) {
      contact->energy_grad(xx, ceg.ptr);
      cg = ceg.ptr;
    }
    launch_gradient(st, dmesh(), dmat(), xx, yhat.ptr, 1.0 / (hadj * hadj), cg, out, partials.ptr,
                    &ds()->grad_sumsq);
  }
  double hadj = 0;

  void adj_alloc() {
    const size_t n3 = (size_t)n * 3;
    for (DevBuf<double>* b : {&adj_x1, &adj_lam, &adj_r, &adj_z, &adj_p, &adj_ap, &adj_gp, &adj_gm, &adj_xs,
                              &adj_gx, &adj_gv, &adj_a})
      b->alloc(n3);
  }

  // H p at `base`: the inertia, membrane and bending terms analytically (the
  // membrane Hessian in closed form, the hinge Hessian-vector product in dual
  // numbers, bend.h), plus the central difference (1e-6 m along p / |p|_inf)
  // of the contact gradient; H is the derivative of the whole Newton
  // gradient (the step's FD bending gradient differs from the analytic one by
  // its truncation and round-off only).
  DevBuf<double> bend_term;
  void full_hvp(const double* base, const double* p, double* out) {
    const size_t n3 = (size_t)n * 3;
    launch_hvp64_analytic(st, dmesh(), dmat(), 1.0 / (hadj * hadj), base, p, out);
    const double kb = 0.5 * (mat.bend_warp + mat.bend_weft);
    if (kb > 0 && nh) {
      bend_term.alloc((size_t)nh * 12);
      launch_bend_hvp(st, dmesh(), kb, base, p, bend_term.ptr, out, true);
    }
    if (!contact) return;
    // |p|_inf stays on the device (k_adj_shift / k_adj_fd_add read it)
    launch_max_abs(st, p, n3, mpartials.ptr, &ds()->maxdx);
    const double eps = 1e-6;  // metres along p / |p|_inf
    const double* pmax = &ds()->maxdx;
    hipLaunchKernelGGL(k_adj_shift, dim3(grid_for(n3)), dim3(kBlock), 0, st, base, p, pmax, eps, adj_xs.ptr, n3);
    contact_gradient(adj_xs.ptr, adj_gp.ptr);
    hipLaunchKernelGGL(k_adj_shift, dim3(grid_for(n3)), dim3(kBlock), 0, st, base, p, pmax, -eps, adj_xs.ptr, n3);
    contact_gradient(adj_xs.ptr, adj_gm.ptr);
    hipLaunchKernelGGL(k_adj_fd_add, dim3(grid_for(n3)), dim3(kBlock), 0, st, adj_gp.ptr, adj_gm.ptr, pmax, 2 * eps,
                       out, n3);
    CS_LAUNCH_CHECK();
  }

  // the gradient term full_hvp differentiates numerically: contact alone
  // (the internal terms with every modulus zero), pins zeroed
  void contact_gradient(const double* xx, double* out) {
    contact->grad_only(xx, ceg.ptr);
    const DMat none{mat.thickness, 0, 0, 0, 0, 0};
    launch_gradient(st, dmesh(), none, xx, yhat.ptr, 0.0, ceg.ptr, out, partials.ptr, &ds()->grad_sumsq);
  }

  double adj_dot(const double* a, const double* b, double* slot) {
    launch_dot(st, a, b, (size_t)n * 3, partials.ptr, slot);
    return read(slot);
  }

  // Block-Jacobi PCG on H(base) sol = rhs over the free DOFs,
  // from sol = 0, to `tol` relative residual; returns the relative residual.
  // `rhs` must be zero on pinned DOFs.  Uses adj_r/z/p/ap/xs/gp/gm.
  double full_pcg(const double* base, const double* rhs, double* sol, double tol, int max_iters, int* iters) {
    const size_t n3 = (size_t)n * 3;
    AdjPcg* A = &ds()->apcg;
    CS_HIP(hipMemsetAsync(sol, 0, n3 * sizeof(double), st));
    CS_HIP(hipMemcpyAsync(adj_r.ptr, rhs, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    launch_dot(st, rhs, rhs, n3, partials.ptr, &A->bnorm2);
    // block-Jacobi: the 3x3 membrane vertex blocks at `base` plus the step's
    // diagonal, inverted (Jacobi where a block is not positive definite)
    adj_binv.alloc((size_t)n * 9);
    launch_block_factors(st, dmesh(), dmat(), base, diag.ptr, adj_binv.ptr, nullptr);
    auto precondition = [&]() {
      hipLaunchKernelGGL(k_adj_pre_block, dim3(grid_for(n)), dim3(kBlock), 0, st, adj_r.ptr, adj_binv.ptr,
                         pinmask.ptr, adj_z.ptr, n);
    };
    precondition();
    CS_HIP(hipMemcpyAsync(adj_p.ptr, adj_z.ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    launch_dot(st, adj_r.ptr, adj_z.ptr, n3, partials.ptr, &A->rz[0]);
    hipLaunchKernelGGL(k_adj_pcg_init, dim3(1), dim3(1), 0, st, A, tol);
    CS_LAUNCH_CHECK();
    // Iterations are enqueued in chunks and the host reads the stop flag once
    // per chunk; an iteration enqueued after the stop changes nothing (its
    // update kernels return at once).
    const int kChunk = 16;
    int cur = 0, queued = 0;
    while (queued < max_iters && read(&A->done) == 0) {
      const int m = std::min(kChunk, max_iters - queued);
      for (int k = 0; k < m; k++, queued++, cur ^= 1) {
        full_hvp(base, adj_p.ptr, adj_ap.ptr);
        launch_dot(st, adj_p.ptr, adj_ap.ptr, n3, partials.ptr, &A->pap);
        hipLaunchKernelGGL(k_adj_pcg_step, dim3(grid_for(n3)), dim3(kBlock), 0, st, A, cur, sol, adj_r.ptr,
                           adj_p.ptr, adj_ap.ptr, n3);
        launch_dot(st, adj_r.ptr, adj_r.ptr, n3, partials.ptr, &A->rr);
        hipLaunchKernelGGL(k_adj_pcg_decide, dim3(1), dim3(1), 0, st, A);
        precondition();
        launch_dot(st, adj_r.ptr, adj_z.ptr, n3, partials.ptr, &A->rz[cur ^ 1]);
        hipLaunchKernelGGL(k_adj_pcg_dir, dim3(grid_for(n3)), dim3(kBlock), 0, st, A, cur, adj_p.ptr, adj_z.ptr,
                           n3);
        CS_LAUNCH_CHECK();
      }
    }
    const AdjPcg a = read(A);
    if (iters) *iters = a.it;
    return a.rel;
  }

  void step_adjoint(double h, const double* hx0, const double* hv0, const double* hx1, const double* hgx,
                    const double* hgv, double tol, int max_iters, double* hdx0, double* hdv0, double* hdtheta,
                    int* iters_out, double* resid_out) {
    upload_pins();
    const size_t n3 = (size_t)n * 3;
    adj_alloc();
    hadj = h;
    CS_HIP(hipMemcpyAsync(x0.ptr, hx0, n3 * sizeof(double), hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(adj_gv.ptr, hv0, n3 * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_predictor, dim3(grid_for(n)), dim3(kBlock), 0, st, x0.ptr, adj_gv.ptr, h, gravity[0],
                       gravity[1], gravity[2], yhat.ptr, xt.ptr, n);
    CS_HIP(hipMemcpyAsync(adj_x1.ptr, hx1, n3 * sizeof(double), hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(adj_gx.ptr, hgx, n3 * sizeof(double), hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(adj_gv.ptr, hgv, n3 * sizeof(double), hipMemcpyHostToDevice, st));
    if (contact) {
      contact->begin_step(x0.ptr);
      contact->update_active_set(adj_x1.ptr);
    }
    const double damp = 1 - jsm::jmin(jsm::jmax(mat.damping, 0), 0.1);
    hipLaunchKernelGGL(k_adj_rhs, dim3(grid_for(n3)), dim3(kBlock), 0, st, adj_gx.ptr, adj_gv.ptr, damp / h,
                       pinmask.ptr, adj_a.ptr, n3);
    // Jacobi preconditioner: the step's diagonal M / h^2 + beta (+ contact)
    const double invH2 = 1 / (h * h);
    const double beta =
        jsm::jmax(jsm::jmax(mat.stretch_warp, mat.stretch_weft), mat.shear) * mat.thickness * 0.1 + 1e-6;
    hipLaunchKernelGGL(k_init_diag, dim3(grid_for(n)), dim3(kBlock), 0, st, mass.ptr, invH2, beta, diag.ptr, n);
    if (contact) contact->add_diag_estimate(diag.ptr, adj_x1.ptr);
    CS_LAUNCH_CHECK();
    int it = 0;
    const double rel = full_pcg(adj_x1.ptr, adj_a.ptr, adj_lam.ptr, tol, max_iters, &it);
    if (iters_out) *iters_out = it;
    if (resid_out) *resid_out = rel;
    // parameter sensitivities: dg/dtheta_j = internal gradient with theta_j = 1
    const DMat base = dmat();
    const DMat unit[5] = {{base.thick, 1, 0, 0, 0, 0}, {base.thick, 0, 1, 0, 0, 0}, {base.thick, 0, 0, 1, 0, 0},
                          {base.thick, 0, 0, 0, 1, 0}, {base.thick, 0, 0, 0, 0, 1}};
    for (int j = 0; j < 5; j++) {
      launch_gradient(st, dmesh(), unit[j], adj_x1.ptr, yhat.ptr, 0.0, nullptr, adj_gp.ptr, partials.ptr,
                      &ds()->grad_sumsq);
      hdtheta[j] = -adj_dot(adj_lam.ptr, adj_gp.ptr, &ds()->adj[0]);
    }
    hipLaunchKernelGGL(k_adj_out, dim3(grid_for(n3)), dim3(kBlock), 0, st, adj_lam.ptr, adj_gv.ptr, mass.ptr, h,
                       damp, adj_gp.ptr, adj_gm.ptr, n3);
    CS_LAUNCH_CHECK();
    // the lagged friction's slip is measured from x_prev: its share of dL/dx_prev
    if (contact) contact->friction_xstep_adjoint(adj_x1.ptr, adj_lam.ptr, adj_gp.ptr);
    CS_HIP(hipMemcpyAsync(hdx0, adj_gp.ptr, n3 * sizeof(double), hipMemcpyDeviceToHost, st));
    CS_HIP(hipMemcpyAsync(hdv0, adj_gm.ptr, n3 * sizeof(double), hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
  }

  // ---- root polish (CS_FLAG_POLISH_ROOT; not in the reference) ----------------
  //
  // Near the root the Armijo test compares energies whose difference is below
  // their round-off and rejects every trial, which ends Newton at a gradient
  // norm set by the energy's magnitude rather than by grad_tol.  With the flag,
  // an iteration whose every Armijo trial failed takes instead a full Newton
  // step of the complete Hessian (full_pcg: analytic membrane, central
  // differences of bending and contact, which the step's own HVP leaves out),
  // kept when it passes the contact check and at least halves the Newton
  // gradient norm; Newton then goes on until grad_tol, newton_iters or a step
  // that does not (the root's round-off floor).  Differentiable rollouts need this: the implicit-function
  // gradient is the derivative of the exact root.
  bool polish_step(double h, double gradNorm, double& energy) {
    const size_t n3 = (size_t)n * 3;
    adj_alloc();
    hadj = h;
    // Newton direction of the complete Hessian: H(x) d = -g(x)
    hipLaunchKernelGGL(k_neg_free, dim3(grid_for(n3)), dim3(kBlock), 0, st, g.ptr, pinmask.ptr, adj_a.ptr, n3);
    CS_LAUNCH_CHECK();
    (void)full_pcg(x.ptr, adj_a.ptr, adj_lam.ptr, 1e-8, std::max(opt.pcg_max_iters, 1000), nullptr);
    launch_trial(st, x.ptr, adj_lam.ptr, 1.0, xt.ptr, n3);
    enforce_pins(xt.ptr);
    if (contact) {
      contact->set_segment(x.ptr, xt.ptr);
      if (!contact->check_trial(xt.ptr)) return false;
    }
    adj_gradient(xt.ptr, adj_gp.ptr);
    sync_scalars();
    // a step that does not halve |grad| has reached the root's round-off
    const double gNew = std::sqrt(hscal->grad_sumsq);
    if (!(gNew < 0.5 * gradNorm)) return false;
    std::swap(x.ptr, xt.ptr);
    energy = total_energy(x.ptr, h);
    if (contact) contact->commit(x.ptr);
    return true;
  }

  // ---- the step --------------------------------------------------------------
  cs_step_stats step(double h) {
    upload_pins();
    const size_t n3 = (size_t)n * 3;
    const int newtonIters = opt.newton_iters, pcgMaxIters = opt.pcg_max_iters, lsIters = opt.line_search_iters;
    const double gradTol = opt.grad_tol;
    const bool polish = (opt.flags & CS_FLAG_POLISH_ROOT) != 0;
    const bool devctl = (opt.flags & CS_FLAG_DEVICE_NEWTON) != 0;
    // the WebGPU Newton's descent gate (gpu-solver.ts:1168-1188,
    // newton-control.wgsl:58-73): a PCG stopped on non-positive curvature
    // also selects the Jacobi-scaled steepest descent
    const bool gate_breakdown = (opt.flags & CS_FLAG_BREAKDOWN_DESCENT) != 0;
    CS_HIP(hipMemcpyAsync(x0.ptr, pos.ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_predictor, dim3(grid_for(n)), dim3(kBlock), 0, st, x0.ptr, vel.ptr, h, gravity[0],
                       gravity[1], gravity[2], yhat.ptr, x.ptr, n);
    CS_LAUNCH_CHECK();
    enforce_pins(x.ptr);
    int lsTrials = 0;
    if (contact) {
      contact->begin_step(x0.ptr);
      contact->set_segment(x0.ptr, x.ptr);  // predictor probes lie on x0 -> yhat
      if (!contact->check_trial(x.ptr)) {
        // predictor bisection (newton.ts:53-74)
        double lo = 0, hi = 1;
        for (int bb = 0; bb < 8; bb++) {
          const double mid = 0.5 * (lo + hi);
          launch_lerp(mid, xt.ptr);
          enforce_pins(xt.ptr);
          if (contact->check_trial(xt.ptr)) lo = mid;
          else hi = mid;
          if (hi - lo < 1e-4) break;
        }
        launch_lerp(lo, xt.ptr);
        std::swap(x.ptr, xt.ptr);
        enforce_pins(x.ptr);
      }
      contact->update_active_set(x.ptr);
    }
    double energy = total_energy(x.ptr, h);
    double gradNorm = INFINITY, maxStrain = 0;
    int pcgItersTotal = 0, used = 0;
    bool converged = false;
    const double invH2 = 1 / (h * h);
    const double beta = jsm::jmax(jsm::jmax(mat.stretch_warp, mat.stretch_weft), mat.shear) * mat.thickness * 0.1 + 1e-6;
    hipLaunchKernelGGL(k_init_diag, dim3(grid_for(n)), dim3(kBlock), 0, st, mass.ptr, invH2, beta, diag.ptr, n);
    CS_LAUNCH_CHECK();
    for (int ni = 0; ni < newtonIters; ni++) {
      used++;
      const double* cg = nullptr;
      if (contact) {
        // the first iteration's x is the one the active set was just built at
        // (newton.ts:99 and :124 build it twice with the same inputs)
        if (ni > 0) contact->update_active_set(x.ptr);
        contact->energy_grad(x.ptr, ceg.ptr);
        contact->add_diag_estimate(diag.ptr, x.ptr);
        cg = ceg.ptr;
      }
      const int e0 = ev_mark();
      launch_gradient(st, dmesh(), dmat(), x.ptr, yhat.ptr, invH2, cg, g.ptr, partials.ptr, &ds()->grad_sumsq);
      ev_span(CS_KERNEL_GRADIENT, e0, ev_mark());
      launch_max_strain(st, dmesh(), x.ptr, mpartials.ptr, &ds()->max_strain);
      sync_scalars();
      maxStrain = hscal->max_strain;
      gradNorm = std::sqrt(hscal->grad_sumsq);
      if (!std::isfinite(gradNorm)) throw NonFinite{};
      if (gradNorm < gradTol) {
        converged = true;
        break;
      }
      launch_neg(st, g.ptr, b.ptr, n3);
      int iters = 0;
      if (mixed) {
        iters = mixed_solve(invH2, pcgMaxIters);
      } else {
        iters = pcg_solve(invH2, pcgMaxIters);
      }
      pcgItersTotal += iters;
      const double TRUST = 0.002;
      double gtdx = 0;
      if (devctl) {
        // descent safeguard and trust region decided on the device; g.dx is
        // read with the first line-search batch
        launch_dot(st, g.ptr, dx.ptr, n3, partials.ptr, &ds()->gtdx);
        hipLaunchKernelGGL(k_ctl_descent, dim3(1), dim3(1), 0, st, ds(), pcg_done_flag(), gate_breakdown ? 1 : 0);
        hipLaunchKernelGGL(k_fallback_pred, dim3(grid_for(n3)), dim3(kBlock), 0, st, g.ptr, diag.ptr, dx.ptr, n3,
                           ds());
        launch_dot(st, g.ptr, dx.ptr, n3, partials.ptr, &ds()->gtdx);
        launch_max_abs(st, dx.ptr, n3, mpartials.ptr, &ds()->maxdx);
        hipLaunchKernelGGL(k_ctl_trust, dim3(1), dim3(1), 0, st, ds(), TRUST);
        hipLaunchKernelGGL(k_scale_pred, dim3(grid_for(n3)), dim3(kBlock), 0, st, dx.ptr, n3, ds(), TRUST);
        CS_LAUNCH_CHECK();
      } else {
        // descent safeguard (newton.ts:154-160)
        launch_dot(st, g.ptr, dx.ptr, n3, partials.ptr, &ds()->gtdx);
        gtdx = read(&ds()->gtdx);
        const bool breakdown = gate_breakdown && read(pcg_done_flag()) == 1;
        if (!std::isfinite(gtdx) || gtdx >= 0 || breakdown) {
          launch_fallback_dir(st, g.ptr, diag.ptr, dx.ptr, n3);
          launch_dot(st, g.ptr, dx.ptr, n3, partials.ptr, &ds()->gtdx);
          gtdx = read(&ds()->gtdx);
        }
        // trust region (newton.ts:167-179)
        launch_max_abs(st, dx.ptr, n3, mpartials.ptr, &ds()->maxdx);
        const double maxDx = read(&ds()->maxdx);
        if (maxDx > TRUST) {
          const double s = TRUST / maxDx;
          launch_scale(st, dx.ptr, s, n3);
          gtdx *= s;
        }
      }
      // every trial x + alpha dx (and the accepted iterate) lies on x -> x + dx
      if (contact) {
        launch_trial(st, x.ptr, dx.ptr, 1.0, xe.ptr, n3);
        enforce_pins(xe.ptr);
        contact->set_segment(x.ptr, xe.ptr);
      }
      // Armijo (newton.ts:184-219)
      double alpha = 1;
      bool accepted = false;
      if (devctl) accepted = armijo_batched(h, lsIters, energy, alpha, gtdx, lsTrials);
      for (int li = 0; li < lsIters && !devctl; li++) {
        lsTrials++;
        launch_trial(st, x.ptr, dx.ptr, alpha, xt.ptr, n3);
        enforce_pins(xt.ptr);
        if (contact && !contact->check_trial(xt.ptr)) {
          contact->add_rejection();
          alpha *= 0.5;
          continue;
        }
        const double eNew = total_energy(xt.ptr, h);
        if (!std::isfinite(eNew)) {
          if (contact) contact->add_rejection();
          alpha *= 0.5;
          continue;
        }
        if (eNew <= energy + 1e-4 * alpha * gtdx) {
          std::swap(x.ptr, xt.ptr);
          energy = eNew;
          accepted = true;
          if (contact) contact->commit(x.ptr);
          break;
        }
        if (contact) contact->add_rejection();
        alpha *= 0.5;
      }
      if (!accepted && polish && polish_step(h, gradNorm, energy)) continue;
      if (!accepted) break;
      if (alpha < 1e-6) break;
    }
    const double damp = 1 - jsm::jmin(jsm::jmax(mat.damping, 0), 0.1);
    hipLaunchKernelGGL(k_velocity, dim3(grid_for(n3)), dim3(kBlock), 0, st, x.ptr, x0.ptr, h, damp, vel.ptr, n3);
    CS_LAUNCH_CHECK();
    if (contact) contact->kill_approach_velocity(x.ptr, vel.ptr);
    CS_HIP(hipMemcpyAsync(pos.ptr, x.ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    enforce_pins(pos.ptr);
    CS_HIP(hipStreamSynchronize(st));
    flush_timing();
    cs_step_stats s;
    std::memset(&s, 0, sizeof(s));
    s.newton_iters = newtonIters;
    s.pcg_iters = pcgItersTotal;
    s.grad_norm = gradNorm;
    s.energy = energy;
    s.max_strain = maxStrain;
    s.converged = converged ? 1 : 0;
    s.has_contact = contact ? 1 : 0;
    if (contact) s.contact = contact->diagnostics();
    else {
      s.contact.min_distance = INFINITY;
      s.contact.min_toi = INFINITY;
    }
    s.newton_used = used;
    s.line_search_trials = lsTrials;
    return s;
  }

  // xt = x0 + t (x - x0) (predictor bisection probe, newton.ts:62-63)
  void launch_lerp(double t, double* out);

  // The Armijo loop of newton.ts:184-219 with its trials evaluated in
  // batches of 1, 2, 4, 8: each batch launches every trial's positions, contact
  // validation and energies, then reads them with one synchronisation; the
  // trials are then judged in order with the reference's rules, up to the
  // first accepted one, so the result, the trial count and the contact
  // diagnostics are those of the sequential loop (trials past the accepted
  // one were computed but are never judged).  g.dx comes from the device
  // control with the first batch.  Returns whether a trial was accepted;
  // `alpha` and `energy` as the sequential loop leaves them.
  DevBuf<double> tbuf[kArmijoBatch];
  bool armijo_batched(double h, int lsIters, double& energy, double& alpha, double& gtdx, int& lsTrials) {
    const size_t n3 = (size_t)n * 3;
    // the full step is usually accepted, and a trial computed past the
    // accepted one costs a whole contact check: the first batch is one trial,
    // and later batches grow only when speculation is asked for
    const bool speculate = (opt.flags & CS_FLAG_ARMIJO_SPECULATE) != 0;
    int li = 0, batch = 1;
    bool first = true;
    while (li < lsIters) {
      const int K = std::min(std::min(batch, lsIters - li), kArmijoBatch);
      double a = alpha;
      for (int j = 0; j < K; j++) {
        tbuf[j].alloc(n3);
        launch_trial(st, x.ptr, dx.ptr, a, tbuf[j].ptr, n3);
        enforce_pins(tbuf[j].ptr);
        if (contact) contact->trial_enqueue(tbuf[j].ptr, j);
        launch_energy(st, dmesh(), dmat(), tbuf[j].ptr, yhat.ptr, h, partials.ptr, ds()->etrial[j]);
        if (contact) contact->merit_enqueue(tbuf[j].ptr, &ds()->mtrial[j]);
        a *= 0.5;
      }
      CS_HIP(hipMemcpyAsync(hscal, scal.ptr, sizeof(NewtonScalars), hipMemcpyDeviceToHost, st));
      if (contact) contact->trial_fetch(K);  // synchronises the stream, the scalars included
      else CS_HIP(hipStreamSynchronize(st));
      if (first) {
        gtdx = hscal->gtdx;
        first = false;
      }
      for (int j = 0; j < K; j++) {
        lsTrials++;
        if (contact && !contact->trial_result(j)) {
          contact->add_rejection();
          alpha *= 0.5;
          continue;
        }
        double eNew = hscal->etrial[j][0];
        eNew += hscal->etrial[j][1];
        if (contact) eNew += hscal->mtrial[j];
        if (!std::isfinite(eNew)) {
          if (contact) contact->add_rejection();
          alpha *= 0.5;
          continue;
        }
        if (eNew <= energy + 1e-4 * alpha * gtdx) {
          std::swap(x.ptr, tbuf[j].ptr);
          energy = eNew;
          if (contact) contact->commit(x.ptr);
          return true;
        }
        if (contact) contact->add_rejection();
        alpha *= 0.5;
      }
      li += K;
      if (speculate) batch *= 2;
    }
    return false;
  }

  PcgArgs pcg_args(double invH2) {
    PcgArgs a;
    a.me = dmesh();
    a.mat = dmat();
    a.inv_h2 = invH2;
    a.x = x.ptr;
    a.b = b.ptr;
    a.diag = diag.ptr;
    a.binv = binv.ptr;
    a.sol = dx.ptr;
    a.r = r.ptr; a.z = z.ptr; a.p = p.ptr; a.Ap = Ap.ptr;
    a.hvp_mode = opt.hvp_mode == CS_HVP_ANALYTIC ? HVP_ANALYTIC : HVP_FD;
    a.psd = (opt.flags & CS_FLAG_PSD_NEWTON) && a.hvp_mode == HVP_ANALYTIC ? 1 : 0;
    a.precond = opt.precond == CS_PRECOND_BLOCK_JACOBI ? PRE_BLOCK
                : opt.precond == CS_PRECOND_SCHWARZ       ? PRE_SCHWARZ
                                                          : PRE_JACOBI;
    a.sw = schwarz.get();
    a.tol = 1e-3;  // newton.ts:148
    a.s = &ds()->pcg;
    a.partials = partials.ptr;
    if (contact) a.contact = contact->hvp_hook();
    return a;
  }

  // the latch of the last inner solve (0 cap, 1 non-positive curvature, 2 converged)
  const int32_t* pcg_done_flag() { return mixed ? mixed->done_flag() : &ds()->pcg.done; }

  int mixed_solve(double invH2, int maxIters) {
    if (packer) {
      MixedInputs in = mixed_inputs(invH2, maxIters);
      CS_HIP(hipStreamSynchronize(st));  // b, the diagonal and the block inverses are ready
      return packer->solve(mixed.get(), in);
    }
    MixedInputs in = mixed_inputs(invH2, maxIters);
    int mark = -1;
    return mixed->solve(in, [&](int, bool begin) {
      if (begin) mark = ev_mark();
      else ev_span(CS_KERNEL_HVP, mark, ev_mark());
    });
  }

  MixedInputs mixed_inputs(double invH2, int maxIters) {
    MixedInputs in;
    in.me = dmesh();
    in.mat = dmat();
    in.inv_h2 = invH2;
    in.x = x.ptr;
    in.b = b.ptr;
    in.diag = diag.ptr;
    in.binv = nullptr;
    if (opt.precond == CS_PRECOND_BLOCK_JACOBI) {
      const int e0 = ev_mark();
      launch_block_factors(st, in.me, in.mat, x.ptr, diag.ptr, binv.ptr, nullptr);
      ev_span(CS_KERNEL_PRECOND, e0, ev_mark());
      in.binv = binv.ptr;
    } else if (schwarz) {
      const int e0 = ev_mark();
      schwarz->factor(in.me, in.mat, x.ptr, diag.ptr, true);
      ev_span(CS_KERNEL_PRECOND, e0, ev_mark());
      in.sw = schwarz.get();
    }
    in.dx = dx.ptr;
    if (contact) in.contact = contact->hvp_hook();
    in.max_iters = maxIters;
    in.tol = 1e-3;
    in.graph = (opt.flags & (CS_FLAG_TIMING | CS_FLAG_EAGER)) == 0;
    in.psd = (opt.flags & CS_FLAG_PSD_NEWTON) != 0;
    return in;
  }

  PcgArgs last_pcg{};
  bool has_last_pcg = false;

  // Algorithmic (compulsory) HBM bytes of one HVP y = (M/h^2 + H_mem +
  // H_bar) p: FP32 40n + 32m + 64C (x, p, y 12 B each and the mass per
  // vertex; three indices, Dm^-1 and the area per triangle; a 64 B record per
  // active contact), FP64 80n + 52m + 96C.  Index maps and per-element
  // scratch are not counted.
  double hvp_bytes() const {
    const double C = contact ? (double)contact->active_count() : 0.0;
    if (opt.precision == CS_PREC_MIXED) return 40.0 * n + 32.0 * m + 64.0 * C;
    return 80.0 * n + 52.0 * m + 96.0 * C;
  }

  // Bytes the HVP kernel of this build reads and writes per launch, its
  // index maps and prepared records included (per vertex p, Ap, mass, pin
  // flag [, x and the CSR offset]; per triangle Dm^-1 and three 16 B corner
  // records [+ the 48 B F/S record of the chunked kernel, or x gathers and
  // the area]; with contacts the incidence offsets and one running sum per
  // contact vertex (FP32) or the per-vertex contact sums (FP64)).
  double hvp_read_bytes() const {
    const bool mixed_p = opt.precision == CS_PREC_MIXED;
    const double fs = mixed_p ? 4.0 : 8.0;
    double b = (double)n * (9 * fs + 4 + 1 + 4) + (double)m * (16 + 4 + 3 * 16);
    // tiles: per tile triangle Dm^-1, area and local corners; per local vertex
    // its id, x and p; per owned corner a 2 B entry; per vertex mass, pin,
    // entry offset and Ap; 32 B per tile header
    if (mixed_p && tl_count)
      b = (double)tl_tris * 28 + (double)tl_locs * (4 + 6 * fs) + (double)tl_ents * 2 + (double)n * (4 + 1 + 2 + 3 * fs) +
          32.0 * tl_count;
    if (contact && contact->active_count())
      b += mixed_p ? 4.0 * n + 3 * fs * (double)contact->incidence_vertices() : 3 * fs * n;
    return b;
  }

  double time_hvp(int reps) {
    if (opt.precision == CS_PREC_MIXED) return mixed->time_hvp(reps);
    if (!has_last_pcg) return 0.0;
    return pcg_time_hvp(st, last_pcg, reps);
  }

  int pcg_solve(double invH2, int maxIters) {
    PcgArgs a = pcg_args(invH2);
    last_pcg = a;
    has_last_pcg = true;
    if (a.precond == PRE_BLOCK) {
      const int e0 = ev_mark();
      launch_block_factors(st, a.me, a.mat, x.ptr, diag.ptr, binv.ptr, nullptr);
      ev_span(CS_KERNEL_PRECOND, e0, ev_mark());
    } else if (a.precond == PRE_SCHWARZ) {
      const int e0 = ev_mark();
      schwarz->factor(a.me, a.mat, x.ptr, diag.ptr, false);
      ev_span(CS_KERNEL_PRECOND, e0, ev_mark());
    }
    pcg_init(st, a);
    const int chunk = 8;
    for (int k = 0; k < maxIters; k++) {
      const int e0 = ev_mark();
      pcg_iteration_hvp(st, a, k);
      ev_span(CS_KERNEL_HVP, e0, ev_mark());
      pcg_iteration_rest(st, a, k);
      if ((k + 1) % chunk == 0 || k + 1 == maxIters) {
        const PcgScalars ps = read(&ds()->pcg);
        if (ps.done) break;
      }
    }
    const PcgScalars ps = read(&ds()->pcg);
    if (k_first_done_zero(ps)) return maxIters;
    return ps.iters;
  }
  static bool k_first_done_zero(const PcgScalars& ps) { return ps.done == 0; }
};

__global__ void k_lerp(const double* x0, const double* x1, double t, double* out, size_t n3) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n3) out[i] = x0[i] + t * (x1[i] - x0[i]);
}

void Solver::launch_lerp(double t, double* out) {
  const size_t n3 = (size_t)n * 3;
  hipLaunchKernelGGL(k_lerp, dim3(grid_for(n3)), dim3(kBlock), 0, st, x0.ptr, x.ptr, t, out, n3);
  CS_LAUNCH_CHECK();
}

}  // namespace cloth

// ============================================================== C ABI
using cloth::g_last_error;

struct cs_solver {
  cloth::Solver s;
};

namespace {
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const cloth::NonFinite&) {
    return fail(CS_E_NONFINITE, "non-finite gradient \xe2\x80\x94 blow up; reduce dt/stiffness");
  } catch (const cloth::OverflowError& e) {
    return fail(CS_E_OVERFLOW, e.what());
  } catch (const cloth::UnsupportedError& e) {
    return fail(CS_E_UNSUPPORTED, e.what());
  } catch (const cloth::HipError& e) {
    return fail(CS_E_HIP, e.what());
  } catch (const std::invalid_argument& e) {
    return fail(CS_E_INVALID, e.what());
  } catch (const std::bad_alloc&) {
    return fail(CS_E_NOMEM, "out of host memory");
  } catch (const std::exception& e) {
    return fail(CS_E_HIP, e.what());
  }
}
}  // namespace

extern "C" {

void cs_options_default(cs_options* o) {
  std::memset(o, 0, sizeof(*o));
  o->newton_iters = 10;
  o->pcg_max_iters = 60;
  o->line_search_iters = 10;
  o->grad_tol = 1e-5;
  o->precision = CS_PREC_FP64;
  o->hvp_mode = CS_HVP_FD;
  o->precond = CS_PRECOND_JACOBI;
  o->schwarz_target = 16;
  o->schwarz_min = 8;
  o->schwarz_max = 32;
  o->device = 0;
  o->flags = 0;
  o->coarse = CS_COARSE_NONE;
  o->coarse_iters = 8;
  o->coarse_omega = 0.5;
}

int cs_abi_version(void) { return CS_ABI_VERSION; }

const char* cs_last_error(void) { return g_last_error.c_str(); }

int cs_device_info(int device, char* name, size_t cap) {
  return guarded([&]() -> int {
    int count = 0;
    CS_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(CS_E_HIP, "no such HIP device");
    hipDeviceProp_t prop;
    CS_HIP(hipGetDeviceProperties(&prop, device));
    if (name && cap) {
      std::snprintf(name, cap, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    }
    return CS_OK;
  });
}

int cs_create(const cs_scene_desc* scene, const cs_options* opt, cs_solver** out) {
  if (!scene || !out) return fail(CS_E_INVALID, "null argument");
  cs_options def;
  cs_options_default(&def);
  if (!opt) opt = &def;
  if (scene->n_verts == 0 || scene->n_tris == 0 || !scene->positions || !scene->velocities || !scene->indices ||
      !scene->inv_dm || !scene->areas || !scene->masses)
    return fail(CS_E_INVALID, "scene is missing mesh arrays");
  if (scene->n_hinges && (!scene->hinge_verts || !scene->hinge_rest_angle || !scene->hinge_edge_len ||
                          !scene->hinge_area_sum))
    return fail(CS_E_INVALID, "scene is missing hinge arrays");
  if (scene->n_pins && (!scene->pin_ids || !scene->pin_pos)) return fail(CS_E_INVALID, "scene is missing pin arrays");
  if (opt->newton_iters < 0 || opt->pcg_max_iters < 0 || opt->line_search_iters < 0)
    return fail(CS_E_INVALID, "negative iteration count");
  if (opt->precision != CS_PREC_FP64 && opt->precision != CS_PREC_MIXED)
    return fail(CS_E_INVALID, "unknown precision");
  if (opt->precond < CS_PRECOND_JACOBI || opt->precond > CS_PRECOND_SCHWARZ)
    return fail(CS_E_INVALID, "unknown preconditioner");
  if (opt->precond == CS_PRECOND_SCHWARZ && (opt->schwarz_target < 1 || opt->schwarz_min < 0))
    return fail(CS_E_INVALID, "Schwarz domain sizes must be positive");
  if (opt->coarse < CS_COARSE_NONE || opt->coarse > CS_COARSE_PCG) return fail(CS_E_INVALID, "unknown coarse level");
  if (opt->coarse != CS_COARSE_NONE && opt->precond != CS_PRECOND_SCHWARZ)
    return fail(CS_E_INVALID, "a coarse level needs the Schwarz preconditioner");
  if (opt->coarse_iters < 0 || !(opt->coarse_omega >= 0)) return fail(CS_E_INVALID, "invalid coarse parameters");
  *out = nullptr;
  auto* h = new cs_solver;
  const int rc = guarded([&]() -> int {
    h->s.init(scene, opt);
    return CS_OK;
  });
  if (rc != CS_OK) {
    delete h;
    return rc;
  }
  *out = h;
  return CS_OK;
}

void cs_destroy(cs_solver* s) { delete s; }

int cs_step(cs_solver* s, double dt) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(CS_E_INVALID, "dt must be positive and finite");
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(s->s.device));
    s->s.last = s->s.step(dt);
    s->s.has_last = true;
    return CS_OK;
  });
}

int cs_tuned_launch(const cs_solver* s, char* buf, size_t cap) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  const std::string t = s->s.contact ? s->s.contact->tuned() : std::string();
  if (buf && cap) std::snprintf(buf, cap, "%s", t.c_str());
  return CS_OK;
}

int cs_hvp_bytes(const cs_solver* s, double* algorithmic, double* read) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (algorithmic) *algorithmic = s->s.hvp_bytes();
  if (read) *read = s->s.hvp_read_bytes();
  return CS_OK;
}

int cs_batch_step(cs_solver* const* solvers, uint32_t count, double dt, int threads, int32_t* status) {
  if (!solvers && count) return fail(CS_E_INVALID, "null solver array");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(CS_E_INVALID, "dt must be positive and finite");
  for (uint32_t i = 0; i < count; i++)
    if (!solvers[i]) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  {
    std::vector<cs_solver*> seen(solvers, solvers + count);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
      return fail(CS_E_INVALID, "a solver appears twice in the batch");
  }
  // largest instances first, so the tail of the queue is short work
  std::vector<uint32_t> order(count);
  for (uint32_t i = 0; i < count; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return solvers[a]->s.n > solvers[b]->s.n; });
  std::vector<int32_t> st(count, CS_OK);
  std::vector<std::string> msg(count);
  std::atomic<uint32_t> next{0};
  auto worker = [&]() {
    for (uint32_t k = next.fetch_add(1); k < count; k = next.fetch_add(1)) {
      const uint32_t i = order[k];
      st[i] = cs_step(solvers[i], dt);
      if (st[i] != CS_OK) msg[i] = g_last_error;
    }
  };
  const uint32_t nt = (uint32_t)std::max(1, std::min<int>(threads < 1 ? 1 : threads, (int)std::max<uint32_t>(count, 1)));
  std::vector<std::thread> pool;
  pool.reserve(nt - 1);
  for (uint32_t t = 1; t < nt; t++) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  int first = CS_OK;
  for (uint32_t i = 0; i < count; i++) {
    if (status) status[i] = st[i];
    if (st[i] != CS_OK && first == CS_OK) {
      first = st[i];
      g_last_error = "instance " + std::to_string(i) + ": " + msg[i];
    }
  }
  return first;
}

int cs_batch_step_packed(cs_solver* const* solvers, uint32_t count, double dt, int32_t* status) {
  if (!solvers && count) return fail(CS_E_INVALID, "null solver array");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(CS_E_INVALID, "dt must be positive and finite");
  for (uint32_t i = 0; i < count; i++) {
    if (!solvers[i]) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
    const cloth::Solver& S = solvers[i]->s;
    if (!S.mixed || S.schwarz) return fail(CS_E_INVALID, "a packed batch takes mixed-precision solvers with the "
                                                         "Jacobi or block-Jacobi preconditioner");
    if (S.device != solvers[0]->s.device) return fail(CS_E_INVALID, "a packed batch runs on one device");
  }
  {
    std::vector<cs_solver*> seen(solvers, solvers + count);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
      return fail(CS_E_INVALID, "a solver appears twice in the batch");
  }
  if (!count) return CS_OK;
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(solvers[0]->s.device));
    cloth::PackedBatch pb;
    pb.active = (int)count;
    CS_HIP(hipStreamCreateWithFlags(&pb.st, hipStreamNonBlocking));
    for (uint32_t i = 0; i < count; i++) solvers[i]->s.packer = &pb;
    std::vector<int32_t> st(count, CS_OK);
    std::vector<std::string> msg(count);
    auto worker = [&](uint32_t i) {
      st[i] = cs_step(solvers[i], dt);
      if (st[i] != CS_OK) msg[i] = g_last_error;
      pb.leave();
    };
    // one thread per instance: an instance waits for the others at every
    // packed solve, so none may wait for a free thread
    std::vector<std::thread> pool;
    pool.reserve(count);
    for (uint32_t i = 0; i < count; i++) pool.emplace_back(worker, i);
    for (auto& th : pool) th.join();
    for (uint32_t i = 0; i < count; i++) solvers[i]->s.packer = nullptr;
    (void)hipStreamDestroy(pb.st);
    int first = CS_OK;
    for (uint32_t i = 0; i < count; i++) {
      if (status) status[i] = st[i];
      if (st[i] != CS_OK && first == CS_OK) {
        first = st[i];
        g_last_error = "instance " + std::to_string(i) + ": " + msg[i];
      }
    }
    return first;
  });
}

int cs_step_adjoint(cs_solver* s, double dt, const double* x_prev, const double* v_prev, const double* x_next,
                    const double* dL_dx_next, const double* dL_dv_next, double tol, int max_iters,
                    double* dL_dx_prev, double* dL_dv_prev, double* dL_dtheta, int* iters, double* rel_residual) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(CS_E_INVALID, "dt must be positive and finite");
  if (!x_prev || !v_prev || !x_next || !dL_dx_next || !dL_dv_next || !dL_dx_prev || !dL_dv_prev || !dL_dtheta)
    return fail(CS_E_INVALID, "null argument");
  if (!(tol > 0) || max_iters < 1) return fail(CS_E_INVALID, "tol must be positive and max_iters >= 1");
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(s->s.device));
    s->s.step_adjoint(dt, x_prev, v_prev, x_next, dL_dx_next, dL_dv_next, tol, max_iters, dL_dx_prev, dL_dv_prev,
                      dL_dtheta, iters, rel_residual);
    return CS_OK;
  });
}

int cs_set_material(cs_solver* s, const cs_material* mat) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!mat) return fail(CS_E_INVALID, "null material");
  s->s.mat = *mat;
  return CS_OK;
}

int cs_set_static_mesh(cs_solver* s, const float* pos, uint32_t n_verts, const uint32_t* idx, uint32_t n_tris) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!s->s.contact) return fail(CS_E_INVALID, "scene has no contact system");
  if (n_tris && (!pos || !idx)) return fail(CS_E_INVALID, "null argument");
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(s->s.device));
    s->s.contact->set_static_mesh(pos, n_verts, idx, n_tris);
    return CS_OK;
  });
}

int cs_pin_vertex(cs_solver* s, uint32_t vid, const double* pos) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  cloth::Solver& S = s->s;
  if (vid >= S.n) return fail(CS_E_INVALID, "vertex id out of range");
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(S.device));
    double p[3];
    if (pos) {
      p[0] = pos[0]; p[1] = pos[1]; p[2] = pos[2];
    } else {
      CS_HIP(hipMemcpyAsync(p, S.pos.ptr + (size_t)vid * 3, 3 * sizeof(double), hipMemcpyDeviceToHost, S.st));
      CS_HIP(hipStreamSynchronize(S.st));
    }
    auto it = std::find(S.pin_ids.begin(), S.pin_ids.end(), vid);
    if (it != S.pin_ids.end()) {
      const size_t k = it - S.pin_ids.begin();
      std::copy(p, p + 3, S.pin_pos.begin() + k * 3);
    } else {
      S.pin_ids.push_back(vid);
      S.pin_pos.insert(S.pin_pos.end(), p, p + 3);
    }
    S.pins_dirty = true;
    S.upload_pins();
    S.enforce_pins(S.pos.ptr);
    CS_HIP(hipStreamSynchronize(S.st));
    return CS_OK;
  });
}

int cs_unpin_vertex(cs_solver* s, uint32_t vid) {
  if (!s) return CS_OK;  // unpinVertex uses optional chaining (cpu-solver.ts:36)
  cloth::Solver& S = s->s;
  auto it = std::find(S.pin_ids.begin(), S.pin_ids.end(), vid);
  if (it != S.pin_ids.end()) {
    const size_t k = it - S.pin_ids.begin();
    S.pin_ids.erase(it);
    S.pin_pos.erase(S.pin_pos.begin() + k * 3, S.pin_pos.begin() + k * 3 + 3);
    S.pins_dirty = true;
  }
  return CS_OK;
}

int cs_get_positions(cs_solver* s, double* out, size_t n3) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!out || n3 != (size_t)s->s.n * 3) return fail(CS_E_INVALID, "positions buffer must hold 3n doubles");
  return guarded([&]() -> int {
    CS_HIP(hipMemcpyAsync(out, s->s.pos.ptr, n3 * sizeof(double), hipMemcpyDeviceToHost, s->s.st));
    CS_HIP(hipStreamSynchronize(s->s.st));
    return CS_OK;
  });
}

int cs_get_velocities(cs_solver* s, double* out, size_t n3) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!out || n3 != (size_t)s->s.n * 3) return fail(CS_E_INVALID, "velocities buffer must hold 3n doubles");
  return guarded([&]() -> int {
    CS_HIP(hipMemcpyAsync(out, s->s.vel.ptr, n3 * sizeof(double), hipMemcpyDeviceToHost, s->s.st));
    CS_HIP(hipStreamSynchronize(s->s.st));
    return CS_OK;
  });
}

int cs_set_state(cs_solver* s, const double* x, const double* v, size_t n3) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (n3 != (size_t)s->s.n * 3) return fail(CS_E_INVALID, "state must hold 3n doubles");
  return guarded([&]() -> int {
    if (x) CS_HIP(hipMemcpyAsync(s->s.pos.ptr, x, n3 * sizeof(double), hipMemcpyHostToDevice, s->s.st));
    if (v) CS_HIP(hipMemcpyAsync(s->s.vel.ptr, v, n3 * sizeof(double), hipMemcpyHostToDevice, s->s.st));
    CS_HIP(hipStreamSynchronize(s->s.st));
    return CS_OK;
  });
}

int cs_last_stats(const cs_solver* s, cs_step_stats* out) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!out) return fail(CS_E_INVALID, "null stats");
  if (!s->s.has_last) {
    std::memset(out, 0, sizeof(*out));
    return CS_OK;
  }
  *out = s->s.last;
  return CS_OK;
}

int cs_kernel_timing(const cs_solver* s, int kernel, double* total_ms, uint64_t* count, double* bytes) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (kernel < 0 || kernel >= CS_KERNEL_COUNT) return fail(CS_E_INVALID, "unknown kernel family");
  if (total_ms) *total_ms = s->s.kt_ms[kernel];
  if (count) *count = s->s.kt_count[kernel];
  if (bytes) *bytes = kernel == CS_KERNEL_HVP ? s->s.hvp_bytes() : 0.0;
  return CS_OK;
}

int cs_time_hvp(cs_solver* s, int reps, double* mean_ms, double* bytes) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (reps < 1 || reps > 100000) return fail(CS_E_INVALID, "reps must be in [1, 100000]");
  return guarded([&]() -> int {
    cloth::Solver& S = s->s;
    CS_HIP(hipSetDevice(S.device));
    const double ms = S.time_hvp(reps);
    if (ms <= 0) return fail(CS_E_INVALID, "no HVP operands yet: run cs_step first");
    if (mean_ms) *mean_ms = ms;
    if (bytes) *bytes = S.hvp_bytes();
    return CS_OK;
  });
}

int cs_reset_timing(cs_solver* s) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  for (int k = 0; k < CS_KERNEL_COUNT; k++) {
    s->s.kt_ms[k] = 0;
    s->s.kt_count[k] = 0;
  }
  return CS_OK;
}

int cs_schwarz_partition(const cs_scene_desc* scene, int target, int min_size, int max_size, uint32_t* domain_of,
                         uint32_t* local_of, uint32_t* domains) {
  (void)max_size;
  if (!scene || !scene->indices || (!scene->positions && !scene->rest_positions) || !domain_of || !local_of)
    return fail(CS_E_INVALID, "null argument");
  if (target < 1 || min_size < 0) return fail(CS_E_INVALID, "Schwarz domain sizes must be positive");
  for (size_t k = 0; k < (size_t)scene->n_tris * 3; k++)
    if (scene->indices[k] >= scene->n_verts) return fail(CS_E_INVALID, "triangle index out of range");
  return guarded([&]() -> int {
    const int D = cloth::schwarz_partition_host(*scene, target, min_size, domain_of, local_of);
    if (domains) *domains = (uint32_t)D;
    return CS_OK;
  });
}

int cs_hvp_tiles(const cs_scene_desc* scene, uint64_t* sizes, uint32_t* hdr, uint32_t* loc, uint32_t* lidx,
                 uint16_t* off, uint16_t* ent) {
  if (!scene || !scene->indices || !scene->inv_dm || !scene->areas || !sizes) return fail(CS_E_INVALID, "null argument");
  if (!scene->positions && !scene->rest_positions) return fail(CS_E_INVALID, "null argument");
  for (size_t k = 0; k < (size_t)scene->n_tris * 3; k++)
    if (scene->indices[k] >= scene->n_verts) return fail(CS_E_INVALID, "triangle index out of range");
  return guarded([&]() -> int {
    const uint32_t n = scene->n_verts, m = scene->n_tris;
    std::vector<uint32_t> voff(n + 1, 0), vrec((size_t)m * 3);
    for (uint32_t e = 0; e < m; e++)
      for (int c = 0; c < 3; c++) voff[scene->indices[(size_t)e * 3 + c] + 1]++;
    for (uint32_t v = 0; v < n; v++) voff[v + 1] += voff[v];
    std::vector<uint32_t> cur(voff.begin(), voff.end() - 1);
    for (uint32_t e = 0; e < m; e++)
      for (int c = 0; c < 3; c++) vrec[cur[scene->indices[(size_t)e * 3 + c]]++] = (e << 2) | (uint32_t)c;
    cloth::TileHost T;
    const bool ok = cloth::build_tiles(*scene, voff, vrec, T);
    const uint64_t want[5] = {T.hdr.size(), T.loc.size(), T.lidx.size(), T.off.size(), T.ent.size()};
    const bool fill = hdr && loc && lidx && off && ent;
    if (fill)
      for (int k = 0; k < 5; k++)
        if (sizes[k] < want[k]) return fail(CS_E_OVERFLOW, "tile buffers are too small (sizes holds the sizes needed)");
    for (int k = 0; k < 5; k++) sizes[k] = ok ? want[k] : 0;
    if (!ok || !fill) return CS_OK;
    for (size_t t = 0; t < T.hdr.size(); t++) std::memcpy(hdr + t * 8, &T.hdr[t], sizeof(cloth::TileHdr));
    std::copy(T.loc.begin(), T.loc.end(), loc);
    for (size_t t = 0; t < T.lidx.size(); t++) {
      lidx[t * 2] = T.lidx[t].x;
      lidx[t * 2 + 1] = T.lidx[t].y;
    }
    std::copy(T.off.begin(), T.off.end(), off);
    std::copy(T.ent.begin(), T.ent.end(), ent);
    return CS_OK;
  });
}

int cs_hvp_layout(cs_solver* s, uint32_t* tiles) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (tiles) *tiles = s->s.tl_count;
  return CS_OK;
}

int cs_probe_hvp(cs_solver* s, const double* x, const double* p, double inv_h2, double* out) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!s->s.mixed) return fail(CS_E_INVALID, "solver was not created in mixed precision");
  if (!x || !p || !out) return fail(CS_E_INVALID, "null argument");
  return guarded([&]() -> int {
    cloth::Solver& S = s->s;
    CS_HIP(hipSetDevice(S.device));
    S.upload_pins();
    const size_t n3 = (size_t)S.n * 3;
    CS_HIP(hipMemcpyAsync(S.x.ptr, x, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    CS_HIP(hipMemcpyAsync(S.p.ptr, p, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    S.mixed->probe_hvp(S.dmesh(), S.dmat(), inv_h2, S.x.ptr, S.p.ptr, S.Ap.ptr);
    CS_HIP(hipMemcpy(out, S.Ap.ptr, n3 * sizeof(double), hipMemcpyDeviceToHost));
    return CS_OK;
  });
}

int cs_probe_hvp_contact(cs_solver* s, const double* x0, const double* x, const double* p, double inv_h2,
                         double* out, uint32_t* active) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!s->s.mixed) return fail(CS_E_INVALID, "solver was not created in mixed precision");
  if (!s->s.contact) return fail(CS_E_INVALID, "scene has no contact system");
  if (!x0 || !x || !p || !out) return fail(CS_E_INVALID, "null argument");
  return guarded([&]() -> int {
    cloth::Solver& S = s->s;
    CS_HIP(hipSetDevice(S.device));
    S.upload_pins();
    const size_t n3 = (size_t)S.n * 3;
    CS_HIP(hipMemcpyAsync(S.x0.ptr, x0, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    CS_HIP(hipMemcpyAsync(S.x.ptr, x, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    CS_HIP(hipMemcpyAsync(S.p.ptr, p, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    S.contact->begin_step(S.x0.ptr);
    S.contact->update_active_set(S.x.ptr);
    S.contact->energy_grad(S.x.ptr, S.ceg.ptr);  // the frozen coefficients
    const cloth::DContactHvp h = S.contact->hvp_hook();
    S.mixed->probe_hvp(S.dmesh(), S.dmat(), inv_h2, S.x.ptr, S.p.ptr, S.Ap.ptr, &h);
    CS_HIP(hipMemcpy(out, S.Ap.ptr, n3 * sizeof(double), hipMemcpyDeviceToHost));
    if (active) *active = S.contact->active_count();
    return CS_OK;
  });
}

int cs_coarse_pattern(cs_solver* s, uint32_t* off, uint32_t* col, uint64_t cap, uint64_t* nnz, uint32_t* domains) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!nnz || !domains) return fail(CS_E_INVALID, "null argument");
  const cloth::Coarse* c = s->s.schwarz ? s->s.schwarz->coarse() : nullptr;
  if (!c) return fail(CS_E_INVALID, "solver was not created with a coarse level");
  *nnz = c->nnz();
  *domains = c->domains();
  if (col && cap < c->nnz()) return fail(CS_E_OVERFLOW, "coarse blocks exceed the buffer capacity");
  c->pattern(off, col);
  return CS_OK;
}

int cs_probe_two_level(cs_solver* s, const double* x, const double* diag, const double* r, double* z64, double* z32,
                       double* raw, double* sym, double* dc) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!x || !diag || !r) return fail(CS_E_INVALID, "null argument");
  const cloth::Coarse* c = s->s.schwarz ? s->s.schwarz->coarse() : nullptr;
  if (!c) return fail(CS_E_INVALID, "solver was not created with a coarse level");
  return guarded([&]() -> int {
    cloth::Solver& S = s->s;
    CS_HIP(hipSetDevice(S.device));
    const size_t n3 = (size_t)S.n * 3;
    CS_HIP(hipMemcpyAsync(S.x.ptr, x, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    CS_HIP(hipMemcpyAsync(S.diag.ptr, diag, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    CS_HIP(hipMemcpyAsync(S.r.ptr, r, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    S.schwarz->factor(S.dmesh(), S.dmat(), S.x.ptr, S.diag.ptr, false);
    S.schwarz->apply64(S.r.ptr, S.z.ptr, nullptr);
    CS_HIP(hipStreamSynchronize(S.st));
    if (z64) CS_HIP(hipMemcpy(z64, S.z.ptr, n3 * sizeof(double), hipMemcpyDeviceToHost));
    c->values(raw, sym, dc);
    if (z32) {
      std::vector<float> h32(r, r + n3), o(n3);
      cloth::DevBuf<float> r32, o32;
      r32.upload(h32.data(), n3, S.st);
      o32.alloc(n3);
      S.schwarz->factor(S.dmesh(), S.dmat(), S.x.ptr, S.diag.ptr, true);
      S.schwarz->apply32(r32.ptr, o32.ptr, nullptr);
      CS_HIP(hipStreamSynchronize(S.st));
      CS_HIP(hipMemcpy(o.data(), o32.ptr, n3 * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n3; i++) z32[i] = o[i];
    }
    return CS_OK;
  });
}

int cs_probe_schwarz(cs_solver* s, const double* x, const double* diag, const double* r, double* z64, double* z32,
                     uint8_t* flags64, uint8_t* flags32, uint32_t* domains) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!s->s.schwarz) return fail(CS_E_INVALID, "solver was not created with the Schwarz preconditioner");
  if (!x || !diag || !r) return fail(CS_E_INVALID, "null argument");
  return guarded([&]() -> int {
    cloth::Solver& S = s->s;
    CS_HIP(hipSetDevice(S.device));
    const size_t n3 = (size_t)S.n * 3;
    CS_HIP(hipMemcpyAsync(S.x.ptr, x, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    CS_HIP(hipMemcpyAsync(S.diag.ptr, diag, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    CS_HIP(hipMemcpyAsync(S.r.ptr, r, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    if (domains) *domains = S.schwarz->domain_count();
    S.schwarz->factor(S.dmesh(), S.dmat(), S.x.ptr, S.diag.ptr, false);
    S.schwarz->apply64(S.r.ptr, S.z.ptr, nullptr);
    CS_HIP(hipStreamSynchronize(S.st));
    if (flags64) S.schwarz->flags(flags64);
    if (z64) CS_HIP(hipMemcpy(z64, S.z.ptr, n3 * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<float> h32(r, r + n3);
    cloth::DevBuf<float> r32, o32;
    r32.upload(h32.data(), n3, S.st);
    o32.alloc(n3);
    S.schwarz->factor(S.dmesh(), S.dmat(), S.x.ptr, S.diag.ptr, true);
    S.schwarz->apply32(r32.ptr, o32.ptr, nullptr);
    CS_HIP(hipStreamSynchronize(S.st));
    if (flags32) S.schwarz->flags(flags32);
    if (z32) {
      CS_HIP(hipMemcpy(h32.data(), o32.ptr, n3 * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n3; i++) z32[i] = h32[i];
    }
    return CS_OK;
  });
}

int cs_probe_active_set(cs_solver* s, const double* x0, const double* x, int32_t* ids6, double* vals7,
                        uint64_t cap, uint64_t* count, uint64_t* n_candidates) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!s->s.contact) return fail(CS_E_INVALID, "scene has no contact system");
  if (!x0 || !x) return fail(CS_E_INVALID, "null positions");
  return guarded([&]() -> int {
    cloth::Solver& S = s->s;
    CS_HIP(hipSetDevice(S.device));
    const size_t n3 = (size_t)S.n * 3;
    CS_HIP(hipMemcpyAsync(S.x0.ptr, x0, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    CS_HIP(hipMemcpyAsync(S.x.ptr, x, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    S.contact->begin_step(S.x0.ptr);
    S.contact->update_active_set(S.x.ptr);
    const uint64_t c = S.contact->copy_active(ids6, vals7, cap);
    if (count) *count = c;
    if (n_candidates) *n_candidates = S.contact->candidate_count();
    return CS_OK;
  });
}

int cs_probe_candidates(cs_solver* s, const double* x0, const double* x, int32_t pad_is_dmin, uint32_t* self_pairs,
                        uint64_t self_cap, uint64_t* n_self, uint32_t* static_pairs, uint64_t static_cap,
                        uint64_t* n_static) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!s->s.contact) return fail(CS_E_INVALID, "scene has no contact system");
  if (!x0 || !x || !n_self || !n_static) return fail(CS_E_INVALID, "null argument");
  return guarded([&]() -> int {
    cloth::Solver& S = s->s;
    CS_HIP(hipSetDevice(S.device));
    const size_t n3 = (size_t)S.n * 3;
    CS_HIP(hipMemcpyAsync(S.x0.ptr, x0, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    CS_HIP(hipMemcpyAsync(S.x.ptr, x, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    S.contact->begin_step(S.x0.ptr);
    std::vector<uint32_t> a, b;
    S.contact->probe_pairs(S.x.ptr, S.contact->pad(pad_is_dmin != 0), a, b);
    *n_self = a.size() / 2;
    *n_static = b.size() / 2;
    if (self_pairs) std::copy(a.begin(), a.begin() + std::min<size_t>(a.size(), 2 * self_cap), self_pairs);
    if (static_pairs) std::copy(b.begin(), b.begin() + std::min<size_t>(b.size(), 2 * static_cap), static_pairs);
    return CS_OK;
  });
}

int cs_probe_contact_adjoint(cs_solver* s, const double* x_step, const double* x, const double* lam, double* grad,
                             double* vjp) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!s->s.contact) return fail(CS_E_INVALID, "scene has no contact system");
  if (!x_step || !x || !grad || (lam && !vjp)) return fail(CS_E_INVALID, "null argument");
  return guarded([&]() -> int {
    cloth::Solver& S = s->s;
    CS_HIP(hipSetDevice(S.device));
    const size_t n3 = (size_t)S.n * 3;
    S.adj_alloc();
    CS_HIP(hipMemcpyAsync(S.x0.ptr, x_step, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    CS_HIP(hipMemcpyAsync(S.x.ptr, x, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    S.contact->begin_step(S.x0.ptr);
    S.contact->update_active_set(S.x.ptr);
    S.contact->energy_grad(S.x.ptr, S.ceg.ptr);
    CS_HIP(hipMemcpyAsync(grad, S.ceg.ptr, n3 * sizeof(double), hipMemcpyDeviceToHost, S.st));
    if (lam) {
      CS_HIP(hipMemcpyAsync(S.adj_lam.ptr, lam, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
      CS_HIP(hipMemsetAsync(S.adj_gp.ptr, 0, n3 * sizeof(double), S.st));
      S.contact->friction_xstep_adjoint(S.x.ptr, S.adj_lam.ptr, S.adj_gp.ptr);
      CS_HIP(hipMemcpyAsync(vjp, S.adj_gp.ptr, n3 * sizeof(double), hipMemcpyDeviceToHost, S.st));
    }
    CS_HIP(hipStreamSynchronize(S.st));
    return CS_OK;
  });
}

int cs_probe_check_trial(cs_solver* s, const double* x0, const double* xt, int32_t* valid, double* min_dist,
                         double* min_toi) {
  if (!s) return fail(CS_E_NOT_INIT, "CpuSolver not initialized");
  if (!s->s.contact) return fail(CS_E_INVALID, "scene has no contact system");
  if (!x0 || !xt) return fail(CS_E_INVALID, "null positions");
  return guarded([&]() -> int {
    cloth::Solver& S = s->s;
    CS_HIP(hipSetDevice(S.device));
    const size_t n3 = (size_t)S.n * 3;
    CS_HIP(hipMemcpyAsync(S.x0.ptr, x0, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    CS_HIP(hipMemcpyAsync(S.xt.ptr, xt, n3 * sizeof(double), hipMemcpyHostToDevice, S.st));
    S.contact->begin_step(S.x0.ptr);
    const bool ok = S.contact->check_trial(S.xt.ptr);
    const cs_contact_diag d = S.contact->diagnostics();
    if (valid) *valid = ok ? 1 : 0;
    if (min_dist) *min_dist = d.min_distance;
    if (min_toi) *min_toi = d.min_toi;
    return CS_OK;
  });
}

int cs_probe_ccd(int device, const double* x0, const double* x1, uint32_t n_verts, const int32_t* kinds,
                 const int32_t* ids4, uint64_t count, double thickness, double* toi) {
  if (!x0 || !x1 || !kinds || !ids4 || !toi) return fail(CS_E_INVALID, "null argument");
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(device));
    cloth::probe_ccd(nullptr, x0, x1, n_verts, kinds, ids4, count, thickness, toi);
    return CS_OK;
  });
}

int cs_probe_closest(int device, const double* x0, uint32_t n_verts, const int32_t* kinds, const int32_t* ids4,
                     uint64_t count, double* out3) {
  if (!x0 || !kinds || !ids4 || !out3) return fail(CS_E_INVALID, "null argument");
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(device));
    cloth::probe_closest(nullptr, x0, n_verts, kinds, ids4, count, out3);
    return CS_OK;
  });
}

int cs_device_alloc(int device, size_t bytes, void** out) {
  if (!out) return fail(CS_E_INVALID, "null output pointer");
  *out = nullptr;
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(device));
    CS_HIP(hipMalloc(out, bytes ? bytes : 1));
    return CS_OK;
  });
}

int cs_device_free(int device, void* ptr) {
  if (!ptr) return CS_OK;
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(device));
    CS_HIP(hipFree(ptr));
    return CS_OK;
  });
}

int cs_device_copy(int device, void* dst, const void* src, size_t bytes, int direction) {
  if (!bytes) return CS_OK;
  if (!dst || !src) return fail(CS_E_INVALID, "null pointer");
  hipMemcpyKind kind;
  switch (direction) {
    case CS_COPY_HOST_TO_DEVICE: kind = hipMemcpyHostToDevice; break;
    case CS_COPY_DEVICE_TO_HOST: kind = hipMemcpyDeviceToHost; break;
    case CS_COPY_DEVICE_TO_DEVICE: kind = hipMemcpyDeviceToDevice; break;
    default: return fail(CS_E_INVALID, "unknown copy direction");
  }
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(device));
    CS_HIP(hipMemcpy(dst, src, bytes, kind));
    return CS_OK;
  });
}

int cs_probe_hinges(int device, const double* X12, const double* P12, const double* rest, const double* len,
                    const double* asum, double k, uint64_t count, double* grad12, double* hvp12) {
  if (count && (!X12 || !P12 || !rest || !len || !asum || !grad12 || !hvp12)) return fail(CS_E_INVALID, "null argument");
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(device));
    cloth::probe_hinges(X12, P12, rest, len, asum, k, count, grad12, hvp12);
    return CS_OK;
  });
}

int cs_ccd_queries(int device, const double* x0, const double* x1, uint32_t n_verts, const int32_t* kinds,
                   const int32_t* ids4, uint64_t count, double thickness, double* toi, double* kernel_ms) {
  if (count && (!x0 || !x1 || !kinds || !ids4 || !toi)) return fail(CS_E_INVALID, "null argument");
  if (count && (reinterpret_cast<uintptr_t>(ids4) & 15)) return fail(CS_E_INVALID, "ids4 must be 16-byte aligned");
  if (!std::isfinite(thickness)) return fail(CS_E_INVALID, "thickness must be finite");
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(device));
    cloth::ccd_queries(x0, x1, n_verts, kinds, ids4, count, thickness, toi, kernel_ms);
    return CS_OK;
  });
}

int cs_broadphase_self_pairs(int device, const double* x0, const double* x1, uint32_t n_verts,
                             const uint32_t* indices, uint32_t n_tris, double pad, uint32_t* pairs, uint64_t cap,
                             uint64_t* count, double* build_ms, double* query_ms) {
  if (!count) return fail(CS_E_INVALID, "null count");
  *count = 0;
  if (n_tris && (!x0 || !x1 || !indices)) return fail(CS_E_INVALID, "null argument");
  if (cap && !pairs) return fail(CS_E_INVALID, "null pairs buffer");
  if (n_tris >= 0x80000000u) return fail(CS_E_INVALID, "too many triangles");
  if (!std::isfinite(pad) || pad < 0) return fail(CS_E_INVALID, "pad must be finite and non-negative");
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(device));
    if (!cloth::lbvh_self_pairs(x0, x1, n_verts, indices, n_tris, pad, pairs, cap, count, build_ms, query_ms))
      return fail(CS_E_OVERFLOW, "candidate pairs exceed the buffer capacity (count holds the number needed)");
    return CS_OK;
  });
}

int cs_broadphase_self_pairs_noalloc(int device, const double* x0, const double* x1, uint32_t n_verts,
                                     const uint32_t* indices, uint32_t n_tris, double pad, uint32_t* pairs,
                                     uint64_t cap, uint64_t* count) {
  if (!count) return fail(CS_E_INVALID, "null count");
  *count = 0;
  if (n_tris && (!x0 || !x1 || !indices)) return fail(CS_E_INVALID, "null argument");
  if (cap && !pairs) return fail(CS_E_INVALID, "null pairs buffer");
  if (n_tris > cloth::kLbvhMaxTris)
    return fail(CS_E_INVALID, "n_tris exceeds the allocation-free scratch (at most " +
                                  std::to_string(cloth::kLbvhMaxTris) + " triangles)");
  if (!std::isfinite(pad) || pad < 0) return fail(CS_E_INVALID, "pad must be finite and non-negative");
  return guarded([&]() -> int {
    CS_HIP(hipSetDevice(device));
    if (!cloth::lbvh_self_pairs_noalloc(device, x0, x1, n_verts, indices, n_tris, pad, pairs, cap, count))
      return fail(CS_E_OVERFLOW, "candidate pairs exceed the buffer capacity (count holds the number needed)");
    return CS_OK;
  });
}

}  // extern "C"
