// coarse.h — the coarse level of the two-level additive Schwarz
// preconditioner (mas.ts, coarse-csr.ts): one constant vector per Schwarz
// domain and axis, normalised restriction R[(d*3+a)][v*3+a] = 1/sqrt(|d|)
// (v in d), P = R^T, and Ac = R A P assembled from the membrane Hessian's
// corner-pair blocks plus the Jacobi diagonal spread, stored as 3x3 block-CSR
// over the domains that share a triangle.  The preconditioner becomes
// z = Schwarz1(r) + P C(R r) with C one of
//   CS_COARSE_JACOBI  omega Dc^-1, Dc = diag(Ac)          (applyMasTwoLevel)
//   CS_COARSE_BLOCK   the inverse of Ac's diagonal block   (applyCoarseBlockJacobi)
//   CS_COARSE_PCG     K block-Jacobi PCG iterations on Ac  (coarsePcg)
// Barrier / contact terms stay in the exact fine operator only.
#pragma once

#include <vector>

#include "../../include/cloth_hip.h"
#include "fem.h"

namespace cloth {

struct CoarseImpl;

class Coarse {
 public:
  // members[d] in member order, domain_of[v]; mode CS_COARSE_*, `iters` the
  // coarse PCG iterations, `omega` the damped-Jacobi weight.
  Coarse(const cs_scene_desc& d, const std::vector<std::vector<uint32_t>>& members,
         const std::vector<uint32_t>& domain_of, int mode, int iters, double omega, hipStream_t st);
  ~Coarse();
  // Ac, Dc and the diagonal-block inverses at x (FP64), on the stream.
  void assemble(const DMesh& me, const DMat& mat, const double* x, const double* diag);
  // z += P C(R r), skipped when *done (the PCG latch) is set.
  void apply64(const double* r, double* z, const int32_t* done);
  void apply32(const float* r, float* z, const int32_t* done);
  int mode() const;
  uint32_t domains() const;
  uint64_t nnz() const;
  // host copies: block rows (D + 1) and columns (nnz); the last assembly's
  // raw and symmetrised blocks (9 nnz each) and Dc (3D); null skips
  void pattern(uint32_t* off, uint32_t* col) const;
  void values(double* raw, double* sym, double* dc) const;

 private:
  CoarseImpl* im_;
};

}  // namespace cloth
