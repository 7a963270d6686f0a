// contact.h — device-resident ContactSystem (contact-assembly.ts:46-595).
#pragma once

#include <vector>

#include <stdexcept>
#include <string>

#include "../../include/cloth_hip.h"
#include "pcg.h"

namespace cloth {

struct UnsupportedError : std::runtime_error {
  explicit UnsupportedError(const std::string& s) : std::runtime_error(s) {}
};
struct OverflowError : std::runtime_error {
  explicit OverflowError(const std::string& s) : std::runtime_error(s) {}
};

struct ContactImpl;

class ContactSystem {
 public:
  ContactSystem(const cs_scene_desc& d, hipStream_t st);
  ~ContactSystem();
  void begin_step(const double* x0);                       // beginStep :75-81
  bool check_trial(const double* xt);                      // checkTrial :489-553
  // Declare the segment (xa -> xb) on which the following check_trial /
  // update_active_set positions lie until the next call or begin_step;
  // their candidate pairs are then selected from one superset traversal.
  void set_segment(const double* xa, const double* xb);
  void update_active_set(const double* x);                 // updateActiveSet :164-220
  void energy_grad(const double* x, double* ceg);          // energyGrad :262-285
  // energy_grad's gradient alone: no energy sums, no host read, the
  // diagnostics untouched (for callers that stay on the stream)
  void grad_only(const double* x, double* ceg);
  void add_diag_estimate(double* diag, const double* x);   // addDiagEstimate :436-471
  double merit_energy(const double* x);                    // meritEnergy :419-428
  void add_rejection();                                    // :83
  // checkTrial for a batched line search: enqueue trials into slots, fetch
  // the slots' counters with one synchronisation, then judge the slots in
  // order (each judged slot updates the diagnostics as check_trial does)
  void trial_enqueue(const double* xt, int slot);
  void trial_fetch(int count);
  bool trial_result(int slot);
  // the triangle pairs buildCandidates :87-121 expands at x (after begin_step)
  void probe_pairs(const double* x, double pad, std::vector<uint32_t>& self2, std::vector<uint32_t>& stat2);
  double pad(bool d_min) const;  // dMinM or dHatM
  // meritEnergy :419-428 into a device double, without synchronising
  void merit_enqueue(const double* x, double* out);
  // setStaticMesh :71-73 between steps (same triangles, new positions)
  void set_static_mesh(const float* pos, uint32_t nverts, const uint32_t* idx, uint32_t ntris);
  // out += dL/dx_step of the lagged friction for the step adjoint lambda at
  // the step's root x (x_step = the begin_step positions); no-op without
  // friction or active contacts.
  void friction_xstep_adjoint(const double* x, const double* lam, double* out);
  void commit(const double* x);                            // commit :475-485
  void kill_approach_velocity(const double* x, double* v); // newton.ts:238-303
  cs_contact_diag diagnostics() const;
  DContactHvp hvp_hook();
  uint64_t active_count() const;
  // Copy the active list (kind,p,a,b,c,d + w0..w3,nx,ny,nz) to host arrays.
  uint64_t copy_active(int32_t* ids6, double* vals7, uint64_t cap);
  uint64_t candidate_count() const;
  // grid caps the launch autotuner chose ("tuning" until it has measured)
  std::string tuned() const;
  uint64_t incidence_count() const;
  uint64_t incidence_vertices() const;
  ContactImpl* impl() { return im_; }

 private:
  ContactImpl* im_;
};

// Batched CCD over device arrays (ids4 16-byte aligned), synchronous on the
// null stream; *kernel_ms (optional) is the kernel's HIP-event time.
void ccd_queries(const double* x0, const double* x1, uint32_t n, const int32_t* kinds, const int32_t* ids4,
                 uint64_t count, double thickness, double* toi, double* kernel_ms);
// Batched device primitives for parity probes.
void probe_ccd(hipStream_t st, const double* x0, const double* x1, uint32_t n, const int32_t* kinds,
               const int32_t* ids4, uint64_t count, double thickness, double* toi);
void probe_closest(hipStream_t st, const double* x0, uint32_t n, const int32_t* kinds, const int32_t* ids4,
                   uint64_t count, double* out3);

}  // namespace cloth
