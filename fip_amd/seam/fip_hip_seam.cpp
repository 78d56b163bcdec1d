// Binds the reference's GPU seam to libfiphip: the file a maintainer adds to the
// reference's sources (with include/fip.h and fip_hip_seam.h) in place of src/kernels.cu
// and src/bc_kernel_gen.cpp, linking -lfiphip. Plain C++ over the C ABI of include/fip.h;
// it defines what those two files and initDeviceBuffers (src/simdata.cpp:54-96) defined:
//   launchComputeKernel, launchFloodPlainKernel          src/kernels.h:3-4
//   BoundaryConditionKernel::generate, ::launch          src/bc_kernel_gen.h:12-16
//   initDeviceBuffers                                     src/simdata.cpp:54, called at :288
// The host loop of src/main.cpp stays: controlTimestep() reads SimData::dt_cfl_mins, which
// here is one float holding the last step's CFL minimum (already reduced on the device),
// with SimData::num_compute_blocks = 1. Errors end the program as the reference's do.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bc_kernel_gen.h"
#include "fip.h"
#include "fip_hip_seam.h"
#include "kernels.h"
#include "simdata.h"

namespace {
fip_sim* g_sim = nullptr;
float g_cfl_min = 1.f;             // the CFL table of src/simdata.cpp:75,91-94: starts at 1
std::vector<float> g_cell_cache;   // sample(): the field last downloaded
int g_cell_field = -1, g_cell_iteration = -1;

void check(int rc, const char* what) {
  if (rc != FIP_OK) {
    std::fprintf(stderr, "%s: %s\n", what, fip_last_error());
    std::exit(EXIT_FAILURE);
  }
}
}  // namespace

// src/simdata.cpp:54-96, same signature: the device state is created from the host arrays
// loadSimData read.
void initDeviceBuffers(const float* h_ah, const float* h_qx, const float* h_qy, const float* h_rei,
                       const float* h_sohle) {
  using namespace SimData;
  fip_desc d{};
  d.W = (int)W;
  d.H = (int)H;
  d.dx = dx;
  d.dt = dt;
  d.no_data = invalid_terrain;
  d.kst = kSt_fixed;
  d.bed = h_sohle;
  d.h = h_ah;
  d.qx = h_qx;
  d.qy = h_qy;
  d.kst_var = h_rei;
  d.variable_dt = is_variable_dt ? 1 : 0;
  d.staggered = is_staggered ? 1 : 0;
  d.mode = FIP_MODE_MASKED;
  d.variant = FIP_EXACT;
  d.tile_rows = 16;
  d.device = 0;
  check(fip_create(&d, &g_sim), "fip_create");
  dt_cfl_mins = &g_cfl_min;
  num_compute_blocks = 1;
}

// src/kernels.cu:441-455: one step with the host's dt; duration += dt on the device, which
// the host reads back, and the step's CFL minimum for the next controlTimestep().
void launchComputeKernel() {
  check(fip_set_dt(g_sim, SimData::dt), "fip_set_dt");
  check(fip_compute_step(g_sim), "fip_compute_step");
  check(fip_get_time(g_sim, &SimData::duration, nullptr, nullptr), "fip_get_time");
  check(fip_get_cfl(g_sim, &g_cfl_min), "fip_get_cfl");
}

// src/kernels.cu:457-463
void launchFloodPlainKernel() { check(fip_floodplain(g_sim), "fip_floodplain"); }

namespace BoundaryConditionKernel {
// src/bc_kernel_gen.cpp:154-192: the boundary kernel of the BC list, compiled with hipRTC.
// Parameters come with every launch from the host, as in the reference.
void* generate() {
  std::vector<fip_bc> v;
  for (BoundaryCondition* bc : SimData::boundary_conditions) {
    fip_bc b{};
    b.type = (int)bc->getType();  // RandType z, q, close, open = FIP_BC_Z .. FIP_BC_OPEN
    b.side = (int)bc->getSide();  // RandSide left, right, top, bottom = FIP_SIDE_*
    b.from = bc->getFrom();
    b.to = bc->getTo();
    v.push_back(b);
  }
  check(fip_bc_generate(g_sim, v.data(), (int)v.size()), "fip_bc_generate");
  return g_sim;
}

// src/bc_kernel_gen.cpp:194-204: args hold (q, z, from, to) per BC (src/main.cpp:88-95)
void launch(void*, const std::vector<JitParam>& args) {
  std::vector<float> p;
  for (size_t i = 0; i + 3 < args.size(); i += 4) {
    p.push_back(args[i].f);      // qval
    p.push_back(args[i + 1].f);  // zval
  }
  check(fip_bc_launch(g_sim, p.data()), "fip_bc_launch");
}
}  // namespace BoundaryConditionKernel

float fipHipCell(int field, int x, int y) {
  // one download per field and step serves every sample point of that step
  if (field != g_cell_field || SimData::iteration != g_cell_iteration) {
    g_cell_cache.resize(SimData::W * SimData::H);
    check(fip_download(g_sim, field, FIP_SIDE_CUR, g_cell_cache.data()), "fip_download");
    g_cell_field = field;
    g_cell_iteration = SimData::iteration;
  }
  return g_cell_cache[(size_t)y * SimData::W + x];
}

void fipHipDownload(int field, int side, float* dst) {
  check(fip_download(g_sim, field, side, dst), "fip_download");
}

void fipHipShutdown() {
  fip_destroy(g_sim);
  g_sim = nullptr;
}
