// Library-internal helpers shared by the single-domain (fip_lib.cpp) and the ensemble
// (fip_ens.cpp) host code. Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "../../include/fip.h"
#include "../../include/fip_ensemble.h"
#include "fip_ctl.h"

namespace fip_internal {

int fail(int code, const std::string& msg);
// false when the environment sets FIP_GRAPHS=0: the device-resident loops then launch every
// kernel from the host instead of replaying captured hipGraphs (results are the same; for
// profilers that cannot trace graph launches).
bool graphs_enabled();
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

// Work launched on the simulation's stream after every step of the device-resident loop
// (captured into its graphs): hook(stream, par) for the step of parity par. An empty
// hook removes it.
using StepHook = std::function<int(hipStream_t, int)>;
void set_post_hook(fip_sim* s, StepHook hook);
hipStream_t stream(const fip_sim* s);
int cur_side(const fip_sim* s);  // the current ("1") side as the host knows it
fip_grid& grid_mut(fip_sim* s);

// ensembles
void ens_set_post_hook(fip_ens* e, StepHook hook);
const fip_ens_scn* ens_scn(const fip_ens* e);
int ens_count(const fip_ens* e);
fip_sim* ens_member(const fip_ens* e, int s);
hipStream_t ens_stream(const fip_ens* e);

// Called by fip_destroy / fip_ens_destroy: closures drop their attachments to the object
// being destroyed (without touching it).
void closure_forget(const fip_sim* s, const fip_ens* e);

}  // namespace fip_internal
