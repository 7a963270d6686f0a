// mixed.h — mixed-precision inner solve: FP32 analytic HVP, preconditioner
// apply and PCG vectors with FP32 accumulation, inside the FP64 Newton loop
// (the Newton gradient, energies, CCD and the Armijo test stay FP64).
#pragma once

#include <vector>

#include "../../include/cloth_hip.h"
#include "pcg.h"

namespace cloth {

// Device tables of the packed multi-instance solve (kept across solves).
struct PackedScratch {
  void* args = nullptr;  // MxArgs per instance
  void* tabs = nullptr;  // block tables
  int* caps = nullptr;   // iteration cap per instance
  int2* flags = nullptr; // (done, iters) per instance
  size_t args_bytes = 0, tabs_bytes = 0, n_inst = 0;
  ~PackedScratch();
};

struct MixedInputs {
  DMesh me;
  DMat mat;
  double inv_h2;
  const double* x;     // FP64 Newton iterate
  const double* b;     // FP64 right-hand side -g
  const double* diag;  // FP64 Jacobi diagonal
  const double* binv;  // FP64 block-Jacobi inverses [9n] or null
  double* dx;          // FP64 Newton step (output)
  DContactHvp contact; // FP64 frozen contacts (act + coeff)
  int max_iters;
  double tol;
  Schwarz* sw = nullptr;  // additive Schwarz with FP32 factors (else binv / diag)
  bool graph = false;     // replay the iterations as a captured HIP graph (no per-kernel timing)
  bool psd = false;       // stress projected to PSD in the HVP (CS_FLAG_PSD_NEWTON)
};

struct MixedImpl;
struct PackedScratch;
class MixedSolve;

// One instance's system in a packed solve (solve_packed): its solver, its
// inputs (Jacobi or block-Jacobi preconditioner; no Schwarz) and, on return,
// its PCG iteration count.
struct PackedJob {
  MixedSolve* solver;
  MixedInputs in;
  int iters = 0;
};

class MixedSolve {
 public:
  MixedSolve(const cs_scene_desc& d, const cs_options& o, hipStream_t st);
  ~MixedSolve();
  // Returns the PCG iteration count; `hvp_timer(k, begin)` brackets the HVP
  // launch of iteration k for kernel timing.
  template <typename F>
  int solve(const MixedInputs& in, F&& hvp_timer);
  int solve_plain(const MixedInputs& in, void (*timer)(void*, int, bool), void* ctx);
  // Mean duration (ms) of `reps` back-to-back HVP launches on the operands of
  // the last solve, bracketed by HIP events on the solver stream.
  double time_hvp(int reps);
  // One launch of the solve's HVP kernel (no contact) at (x, p) into out
  // (FP64 device arrays); for component parity tests.
  // One HVP launch at (x, p) into out (FP64); with `contact`, the frozen
  // barrier term of its active set in FP32 as the solve applies it.
  void probe_hvp(const DMesh& me, const DMat& mat, double inv_h2, const double* x, const double* p, double* out,
                 const DContactHvp* contact = nullptr);
  // Solves several instances' systems together on `st`: every PCG stage is
  // one launch over all instances' blocks (a block table maps each block to
  // its instance, local block and local grid), each instance with its own
  // arithmetic, partial slots and reduction trees, so every result is the
  // one solve_plain gives; `scratch` holds the tables between calls.
  static void solve_packed(std::vector<PackedJob>& jobs, hipStream_t st, PackedScratch& scratch);
  // The device latch of the last solve: 0 iteration cap, 1 stopped on
  // non-positive curvature (p.Ap <= 1e-30), 2 converged.
  const int32_t* done_flag() const;
  void set_material(const cs_material&) {}

 private:
  MixedImpl* im_;
};

template <typename F>
int MixedSolve::solve(const MixedInputs& in, F&& hvp_timer) {
  auto thunk = [](void* ctx, int k, bool begin) { (*static_cast<F*>(ctx))(k, begin); };
  return solve_plain(in, thunk, &hvp_timer);
}

}  // namespace cloth
