// contact.hip — placeholder until the device contact system lands.
#include "contact.h"

namespace cloth {
struct ContactImpl {};
ContactSystem::ContactSystem(const cs_scene_desc&, hipStream_t) : im_(nullptr) {
  throw std::invalid_argument("contact is not available in this build");
}
ContactSystem::~ContactSystem() {}
void ContactSystem::begin_step(const double*) {}
bool ContactSystem::check_trial(const double*) { return true; }
void ContactSystem::update_active_set(const double*) {}
void ContactSystem::energy_grad(const double*, double*) {}
void ContactSystem::add_diag_estimate(double*, const double*) {}
double ContactSystem::merit_energy(const double*) { return 0; }
void ContactSystem::add_rejection() {}
void ContactSystem::commit(const double*) {}
void ContactSystem::kill_approach_velocity(const double*, double*) {}
cs_contact_diag ContactSystem::diagnostics() const { return cs_contact_diag{}; }
DContactHvp ContactSystem::hvp_hook() { return DContactHvp{}; }
void ContactSystem::launch_hvp_coeffs(const double*, const int32_t*) {}
uint64_t ContactSystem::active_count() const { return 0; }
}  // namespace cloth
