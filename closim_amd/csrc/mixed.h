// mixed.h — mixed-precision inner solve: FP32 analytic HVP, preconditioner
// apply and PCG vectors with FP32 accumulation, inside the FP64 Newton loop
// (the Newton gradient, energies, CCD and the Armijo test stay FP64).
#pragma once

#include "../../include/cloth_hip.h"
#include "pcg.h"

namespace cloth {

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
