// pcg.h — device-resident preconditioned CG (pcg.ts:11-95) for the cloth step.
#pragma once

#include "fem.h"

namespace cloth {

// Frozen active contact (contact-assembly.ts:22-30) as stored on the device,
// in the reference's active-list order.
struct Frozen {
  int32_t kind, p, a, b, c, d;  // 0 VT, 1 EE, 2 floor; negative ids = static vertices
  double w0, w1, w2, w3, nx, ny, nz;
};

// Contact Hessian-vector hook: per active contact c the HVP adds
// kappa max(b'', 0) (J_c . p) J_c^T, gathered per vertex through a CSR in
// active-list order.
struct DContactHvp {
  uint32_t C = 0;
  uint32_t E = 0;                    // incidence entries (sorted by vertex, then list order)
  uint32_t nv = 0;                   // vertices with incidence entries
  const Frozen* act = nullptr;
  const uint8_t* inr = nullptr;      // 1 when the contact is inside dHat
  const double* coeff = nullptr;     // kappa * max(b''(d), 0), 0 when d >= dHat
  const uint32_t* v_off = nullptr;   // n+1
  const uint32_t* v_ent = nullptr;   // [E] contact << 3 | slot
  const uint32_t* ent_vtx = nullptr; // [E] vertex of each entry
  const uint32_t* cverts = nullptr;  // [nv] the vertices with entries
  double* term = nullptr;            // [3E] per-entry terms of the current p
  double* vsum = nullptr;            // [3n] contact HVP per vertex (0 off cverts)
};

// Device scalars of one PCG solve (written by block 0 of the kernel that
// takes the decision).
struct PcgScalars {
  double rz[2];  // r.z of iteration k in rz[k & 1]
  double pAp, residual, abs_tol, bnorm;
  int32_t done, iters;
};

enum PrecondKind { PRE_JACOBI = 0, PRE_BLOCK = 1, PRE_SCHWARZ = 2 };

class Schwarz;

struct PcgArgs {
  DMesh me;
  DMat mat;
  double inv_h2;
  const double* x;     // Newton iterate (HVP linearisation point)
  const double* b;     // right-hand side (-g)
  const double* diag;  // Jacobi diagonal
  const double* binv;  // block-Jacobi inverses [9n] (PRE_BLOCK)
  double* sol;         // PCG x (the Newton step)
  double* r;
  double* z;
  double* p;
  double* Ap;
  int hvp_mode;        // HVP_FD / HVP_ANALYTIC
  int psd;             // HVP_ANALYTIC with the stress projected to PSD (stress_psd)
  int precond;         // PrecondKind
  double tol;
  PcgScalars* s;
  jsm::dd* partials;   // >= 4 * kMaxRedBlocks + kMaxHvpBlocks (the p.Ap slots last)
  DContactHvp contact;
  Schwarz* sw;         // PRE_SCHWARZ: applied between the kernels (host side only)
};

void pcg_init(hipStream_t st, const PcgArgs& a);
// One PCG iteration k (0-based): HVP, then update + preconditioner +
// residual test + direction update.
void pcg_iteration(hipStream_t st, const PcgArgs& a, int k);
void pcg_iteration_hvp(hipStream_t st, const PcgArgs& a, int k);   // k_pcg_hvp only
void pcg_iteration_rest(hipStream_t st, const PcgArgs& a, int k);  // update + direction
// Mean duration (ms) of `reps` back-to-back k_pcg_hvp launches (HIP events).
double pcg_time_hvp(hipStream_t st, const PcgArgs& a, int reps);

// Block-Jacobi factors (block-jacobi.ts:30-87) from the membrane vertex
// blocks at x and the Jacobi diagonal.
void launch_block_factors(hipStream_t st, const DMesh& me, const DMat& mat, const double* x, const double* diag,
                          double* binv, uint8_t* flag);
// out = M p / h^2 + H_membrane(x) p with the analytic membrane Hessian, pins
// zeroed (no contact): the exact part of the step adjoint's Hessian product.
void launch_hvp64_analytic(hipStream_t st, const DMesh& me, const DMat& mat, double inv_h2, const double* x,
                           const double* p, double* out);

// Vector helpers used by the Newton driver.
void launch_dot(hipStream_t st, const double* a, const double* b, size_t n, jsm::dd* partials, double* out);
void launch_max_abs(hipStream_t st, const double* a, size_t n, double* partials, double* out);
void launch_fallback_dir(hipStream_t st, const double* g, const double* diag, double* dx, size_t n);
void launch_scale(hipStream_t st, double* v, double s, size_t n);
void launch_trial(hipStream_t st, const double* x, const double* dx, double alpha, double* xt, size_t n);
void launch_neg(hipStream_t st, const double* g, double* b, size_t n);

}  // namespace cloth
