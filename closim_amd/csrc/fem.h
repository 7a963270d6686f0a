// fem.h — element kernels of the cloth step (device side) and their launchers.
//
// Layout in HBM (one cloth instance):
//   x, p, ...      interleaved xyz doubles (3n), the reference's layout
//   idx            u32 [3m] triangle corners
//   inv_dm         f32 [4m] [a,b,c,d] (the reference stores Float32)
//   area           f32 [m];  mass f32 [n]
//   vt_off/vt_rec  CSR vertex -> (triangle << 2 | corner), ascending triangle
//   vt_corner      the same CSR with each triangle's three corner ids inlined
//                  (uint4), so an HVP gather needs no second indirection
//   hv, h*         hinge corners u32 [4h] + rest/len/areaSum doubles
//   vh_off/vh_rec  CSR vertex -> (hinge << 2 | corner), ascending hinge
// Every per-vertex sum walks its CSR list in element order, which reproduces
// the reference's element-order scatter (fem.ts:38-41, bending.ts:29-39)
// bit for bit without scratch arrays or atomics.
#pragma once

#include "common.h"

namespace cloth {

// Vertex-owner tiles of the FP32 HVP (built once on the host): each tile owns
// a compact set of vertices (Morton order of the rest positions) and lists
// every triangle incident to them, so a work-group evaluates each of its
// triangles once into LDS and sums the owned vertices' corners there, and
// every vertex's product is finished by exactly one work-group.
constexpr uint32_t kTileOwn = 256, kTileTris = 640, kTileLoc = 384;
struct TileHdr {
  uint32_t loc, n_own, n_loc, tri, n_tri, ent, off, pad;
};
struct DTiles {
  const TileHdr* hdr;
  uint32_t count;
  const uint32_t* loc;    // local vertices per tile, owned first
  const float4* inv;      // Dm^-1 per tile triangle
  const uint2* lidx;      // local corners (a | b << 16, c)
  const float* area;      // area per tile triangle
  const uint16_t* off;    // per owned vertex: its first entry (n_own + 1 per tile)
  const uint16_t* ent;    // slot * 4 + corner, ascending triangle order
};

struct DMesh {
  uint32_t n, m, nh;
  const uint32_t* idx;
  const float* inv_dm;
  const float* area;
  const float* mass;
  const uint32_t* vt_off;
  const uint32_t* vt_rec;
  const uint4* vt_corner;   // per CSR entry: the triangle's corners i0, i1, i2 and its record
  const uint32_t* hv;
  const double* hrest;
  const double* hlen;
  const double* hasum;
  const uint32_t* vh_off;
  const uint32_t* vh_rec;
  const uint8_t* pinmask;  // 1 = vertex is pinned (makePinFilter, scene.ts:46-55)
  DTiles tl;           // count 0 when the mesh was not tiled
};

struct DMat {
  double thick, c00, c11, c01, g, kb;  // kb = 0.5 * (bendWarp + bendWeft)
};

// Float32 mesh mirror for the mixed-precision inner solve.
struct DMesh32 {
  uint32_t n, m;
  const uint32_t* idx;
  const float* inv_dm;
  const float* area;
  const float* mass;
  const uint32_t* vt_off;
  const uint32_t* vt_rec;
  const uint4* vt_corner;
  const uint8_t* pinmask;
  DTiles tl;
};

struct Scalars;  // device scalar block (solver.hip)

// g = M (x - yHat) / h^2 + grad_membrane + grad_bending (+ ceg), pins zeroed;
// result[slot] = correctly rounded sum of g^2.
void launch_gradient(hipStream_t st, const DMesh& me, const DMat& mat, const double* x, const double* yhat,
                     double inv_h2, const double* ceg, double* g, jsm::dd* partials, double* out_sumsq);

// out = M p / h^2 + H_membrane p (+ contact), pins zeroed, for PCG iteration;
// the FD operator reads h from *h_fd; partial sums of p.out are reduced into
// the PCG scalars.  See pcg.hip for the driver.
enum HvpMode { HVP_FD = 0, HVP_ANALYTIC = 1 };

// Energies at x: inertia sum and internal (membrane + bending) sum, each
// correctly rounded; written to out2[0], out2[1].
void launch_energy(hipStream_t st, const DMesh& me, const DMat& mat, const double* x, const double* yhat,
                   double h, jsm::dd* partials, double* out2);

// Analytic hinge bending (bend.h; not in the reference, which differences the
// energy): out (+)= the bending gradient, or its Hessian-vector product along
// p, at x for hinge stiffness k; `term` holds 12 doubles per hinge.  Pinned
// rows are zero (or left unchanged when accumulating).
void launch_bend_grad(hipStream_t st, const DMesh& me, double k, const double* x, double* term, double* out,
                      bool accumulate);
void launch_bend_hvp(hipStream_t st, const DMesh& me, double k, const double* x, const double* p, double* term,
                     double* out, bool accumulate);
// Independent hinges (12 positions each) for component tests, host arrays.
void probe_hinges(const double* X12, const double* P12, const double* rest, const double* len, const double* asum,
                  double k, uint64_t count, double* grad12, double* hvp12);

// maxStrain over triangles (JS Math.max semantics).
void launch_max_strain(hipStream_t st, const DMesh& me, const double* x, double* partials, double* out);

}  // namespace cloth
