// mixed.h — mixed-precision inner solve: FP32 analytic HVP, preconditioner
// and PCG vectors inside the FP64 Newton loop.
#pragma once

#include "../../include/cloth_hip.h"
#include "pcg.h"

namespace cloth {

struct Solver;
struct MixedImpl;

class MixedSolve {
 public:
  MixedSolve(const cs_scene_desc& d, const cs_options& o, hipStream_t st);
  ~MixedSolve();
  int solve(Solver& s, double inv_h2);  // writes the FP64 Newton step into s.dx
  void set_material(const cs_material& m);

 private:
  MixedImpl* im_;
};

}  // namespace cloth
