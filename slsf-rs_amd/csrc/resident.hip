// resident.hip -- placeholder until the resident kernel lands: no group is
// planned, so FP32 batches run on the HBM kernels.
#include "resident.h"
#include "../../include/slsf.h"

namespace slsf {
bool resident_supports(int) { return false; }
int resident_plan(const std::vector<RepDev>&, ResidentPlan& plan) {
  plan.covers_all = false;
  plan.groups.clear();
  return 0;
}
int resident_run(ResidentPlan&, const GlobalState&, const ResidentRunArgs&, hipStream_t,
                 ResidentRunResult*) {
  return SLSF_E_STATE;
}
}  // namespace slsf
