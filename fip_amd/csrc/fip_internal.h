// Library-internal helpers shared by the single-domain (fip_lib.cpp) and the ensemble
// (fip_ens.cpp) host code. Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/fip.h"
#include "fip_ctl.h"

namespace fip_internal {

int fail(int code, const std::string& msg);
// The domain, the solver flag and the stream of a simulation.
const fip_grid& grid(const fip_sim* s);
int staggered(const fip_sim* s);
int device(const fip_sim* s);
// Re-reads the iteration from the control block (after the ensemble advanced the domain)
// and marks the simulation as started.
int refresh(fip_sim* s);
// Apply stages of a BC list: stage[i] for BC i (-1 when the BC covers no position) and the
// last position applied, lim[i] (the reference's thread count caps a BC's range). Returns
// the number of stages.
int bc_stages(const fip_bc* bcs, int n, int W, int H, std::vector<int>& stage,
              std::vector<int>& lim);

}  // namespace fip_internal
