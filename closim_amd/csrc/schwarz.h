// schwarz.h — one-level additive Schwarz preconditioner (schwarz.ts:44-309)
// on the device: non-overlapping vertex domains built once on the host,
// per-domain dense Hessians assembled from the membrane 9x9 blocks and
// Cholesky-factored once per Newton iteration, applied as
// z = sum_d R_d^T A_d^-1 R_d r.
#pragma once

#include "../../include/cloth_hip.h"
#include "coarse.h"
#include "fem.h"

namespace cloth {

struct SchwarzImpl;

// buildSchwarzDomains on the host only; returns the domain count.
int schwarz_partition_host(const cs_scene_desc& d, int target, int min_size, uint32_t* domain_of, uint32_t* local_of);

class Schwarz {
 public:
  // Domains from the rest positions (mesh.restPositions) and the triangles;
  // coarse_mode CS_COARSE_* adds the coarse level.
  Schwarz(const cs_scene_desc& d, int target, int min_size, int max_size, hipStream_t st,
          int coarse_mode = CS_COARSE_NONE, int coarse_iters = 0, double coarse_omega = 0);
  ~Schwarz();
  // buildSchwarzFactors (schwarz.ts:154-252) at x with the Jacobi diagonal:
  // the FP64 factors in the reference's operation order, or (fp32) the FP32
  // factors of the mixed-precision solve only.  With a coarse level, also
  // Ac at x (FP64).
  void factor(const DMesh& me, const DMat& mat, const double* x, const double* diag, bool fp32);
  // applySchwarz (schwarz.ts:290-309), FP64, in the reference's operation
  // order; with a coarse level, then z += P C(R r).
  void apply64(const double* r, double* z, const int32_t* done);
  // The same operator with the FP32 factors (mixed-precision inner solve).
  void apply32(const float* r, float* z, const int32_t* done);
  const Coarse* coarse() const;
  uint32_t domain_count() const;
  // 1 = Cholesky, 0 = Jacobi fallback, per domain, of the last factor call
  void flags(uint8_t* out) const;
  uint32_t max_domain() const;
  // Host copies of the partition (parity probes).
  void partition(uint32_t* domain_of, uint32_t* local_of) const;

 private:
  SchwarzImpl* im_;
};

}  // namespace cloth
