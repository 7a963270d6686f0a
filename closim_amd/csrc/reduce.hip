// reduce.hip — finish kernel for grid reductions whose value the host reads.
#include "common.h"

namespace cloth {

__global__ void __launch_bounds__(kBlock) k_finish(FinishJobs jobs) {
  const FinishJob& J = jobs.j[blockIdx.x];
  double r;
  switch (J.kind) {
    case FIN_DD:
      r = jsm::dd_value(merge_partials_dd<kBlock>(static_cast<const jsm::dd*>(J.partials), J.count));
      break;
    case FIN_JSMAX:
      r = merge_partials_ext<kBlock, JsMax>(static_cast<const double*>(J.partials), J.count);
      if (J.clamp0) r = JsMax::op(0.0, r);
      break;
    case FIN_GTMAX:
      r = merge_partials_ext<kBlock, GtMax>(static_cast<const double*>(J.partials), J.count);
      if (J.clamp0) r = GtMax::op(0.0, r);
      break;
    default:
      r = merge_partials_ext<kBlock, LtMin>(static_cast<const double*>(J.partials), J.count);
      if (J.clamp0) r = LtMin::op(0.0, r);
      break;
  }
  if (threadIdx.x == 0) *J.out = r;
}

void launch_finish(hipStream_t st, const FinishJobs& jobs, int njobs) {
  if (njobs < 1 || njobs > 4) throw std::invalid_argument("launch_finish: 1..4 jobs");
  hipLaunchKernelGGL(k_finish, dim3(njobs), dim3(kBlock), 0, st, jobs);
  CS_LAUNCH_CHECK();
}

}  // namespace cloth
