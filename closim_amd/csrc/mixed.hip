// mixed.hip — placeholder until the FP32 inner solve lands.
#include "mixed.h"

namespace cloth {
struct MixedImpl {};
MixedSolve::MixedSolve(const cs_scene_desc&, const cs_options&, hipStream_t) : im_(nullptr) {
  throw std::invalid_argument("mixed precision is not available in this build");
}
MixedSolve::~MixedSolve() {}
int MixedSolve::solve(Solver&, double) { return 0; }
void MixedSolve::set_material(const cs_material&) {}
}  // namespace cloth
