// fip: drop-in for the reference's `./fip parameters.yml` (src/main.cpp:204-263, headless
// mode), running the time step on an MI355X through libfiphip's C ABI.
//
// Output is the reference's: the sampling CSV on stdout (src/main.cpp:45-83), the
// save_state dumps {name}_{ah,qx,qy,flood_plains}_{(int)duration}_{UTC %Y-%m-%d-%H-%M}
// holding size_t W, size_t H and W*H float32 of the pre-step buffers (src/main.cpp:154-202),
// and "Finished N iterations after S s (X s/it)" on stderr (src/main.cpp:253).
//
// Options the reference does not have are environment variables:
//   FIP_MODE=masked|dense (default masked), FIP_VARIANT=exact|fast (default exact),
//   FIP_TILE_ROWS=16|32|64 (default 16), FIP_DEVICE=<ordinal> (default 0),
//   FIP_WAVES_PER_BLOCK, FIP_BLOCKS_PER_CU (the stencil's launch shape; in masked mode
//   default 2 and 4: persistent work-groups, the fastest shape of a whole 12 h run of the
//   5667 x 6500 workload). Results depend on the mode, variant and tile rows only.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "config.h"
#include "fip.h"

namespace {

[[noreturn]] void die(const std::string& what) {
  std::cerr << what << std::endl;
  std::exit(EXIT_FAILURE);
}

void check(int rc, const char* call) {
  if (rc != FIP_OK) die(std::string("HIP error in ") + call + ": " + fip_last_error());
}

int env_int(const char* name, int fallback) {
  const char* v = std::getenv(name);
  return v && *v ? std::atoi(v) : fallback;
}

std::string env_str(const char* name, const char* fallback) {
  const char* v = std::getenv(name);
  return v && *v ? std::string(v) : std::string(fallback);
}

// save2DArrayToFile (src/main.cpp:154-167)
void save2D(const std::string& fileName, const float* arr, size_t width, size_t height) {
  std::ofstream file(fileName, std::ios::binary);
  if (!file) die("Failed to open file " + fileName);
  file.write(reinterpret_cast<const char*>(&width), sizeof(size_t));
  file.write(reinterpret_cast<const char*>(&height), sizeof(size_t));
  file.write(reinterpret_cast<const char*>(arr), (std::streamsize)(width * height * sizeof(float)));
}

std::string utc_stamp() {
  std::time_t now = std::time(nullptr);
  std::tm tm{};
  gmtime_r(&now, &tm);
  char buf[64];
  std::strftime(buf, sizeof buf, "%Y-%m-%d-%H-%M", &tm);
  return buf;
}

// dumpDeviceBuffers (src/main.cpp:180-202): the "2" buffers hold the state before the
// last step; flood_plains is the running maximum.
void dump(fip_sim* sim, const fipcli::Config& c) {
  double duration;
  check(fip_get_time(sim, &duration, nullptr, nullptr), "fip_get_time");
  std::vector<float> buf(c.W * c.H);
  const std::string stamp = utc_stamp();
  const std::string t = std::to_string((int)duration);
  const struct {
    int field, side;
    const char* tag;
  } parts[] = {{FIP_FIELD_H, FIP_SIDE_OTHER, "_ah_"},
               {FIP_FIELD_QX, FIP_SIDE_OTHER, "_qx_"},
               {FIP_FIELD_QY, FIP_SIDE_OTHER, "_qy_"},
               {FIP_FIELD_HMAX, FIP_SIDE_CUR, "_flood_plains_"}};
  for (const auto& p : parts) {
    check(fip_download(sim, p.field, p.side, buf.data()), "fip_download");
    save2D(c.name + p.tag + t + "_" + stamp, buf.data(), c.W, c.H);
  }
}

}  // namespace

int main(int argc, char** argv) {
  std::string simDataFile = "inde.yml";
  if (argc < 2)
    std::cerr << "Warning: no config file given, using default " << simDataFile << std::endl;
  else
    simDataFile = argv[1];

  fipcli::Config c;
  try {
    c = fipcli::load(simDataFile);
  } catch (std::exception& ex) {
    std::cerr << "Config error: " << ex.what() << std::endl;
    std::exit(EXIT_FAILURE);
  }

  fip_desc d{};
  d.W = (int)c.W;
  d.H = (int)c.H;
  d.dx = c.dx;
  d.dt = c.dt;
  d.no_data = c.no_data;
  d.kst = c.kSt;
  d.bed = c.bed.data();
  d.h = c.h.data();
  d.qx = c.qx.data();
  d.qy = c.qy.data();
  d.kst_var = c.has_kst_var ? c.kst_var.data() : nullptr;
  d.variable_dt = c.variable_dt ? 1 : 0;
  d.staggered = c.solver == 0 ? 1 : 0;
  d.mode = env_str("FIP_MODE", "masked") == "dense" ? FIP_MODE_DENSE : FIP_MODE_MASKED;
  d.variant = env_str("FIP_VARIANT", "exact") == "fast" ? FIP_FAST : FIP_EXACT;
  d.tile_rows = env_int("FIP_TILE_ROWS", 16);
  d.device = env_int("FIP_DEVICE", 0);
  fip_sim* sim = nullptr;
  check(fip_create(&d, &sim), "fip_create");
  if (d.mode == FIP_MODE_MASKED)
    check(fip_set_launch(sim, env_int("FIP_WAVES_PER_BLOCK", 2), env_int("FIP_BLOCKS_PER_CU", 4)), "fip_set_launch");
  else
    check(fip_set_launch(sim, env_int("FIP_WAVES_PER_BLOCK", 0), -1), "fip_set_launch");

  // BoundaryConditionKernel::generate (src/main.cpp:224)
  std::vector<fip_bc> bcs;
  std::vector<std::vector<float>> tables;
  for (const auto& b : c.bcs) {
    fip_bc x{};
    x.type = b.type;
    x.side = b.side;
    x.from = b.from;
    x.to = b.to;
    x.z = b.z;
    x.q = b.q;
    auto put = [&](const fipcli::TimeSeries& ts, const float** t, const float** v, int* n) {
      std::vector<float> tt, vv;
      for (const auto& kv : ts.series) {
        tt.push_back(kv.first);
        vv.push_back(kv.second);
      }
      tables.push_back(std::move(tt));
      *t = tables.back().data();
      tables.push_back(std::move(vv));
      *v = tables.back().data();
      *n = (int)ts.series.size();
    };
    if (b.series && b.has_z) put(b.zs, &x.zt, &x.zv, &x.zn);
    if (b.series && b.has_q) put(b.qs, &x.qt, &x.qv, &x.qn);
    bcs.push_back(x);
  }
  check(fip_bc_generate(sim, bcs.data(), (int)bcs.size()), "fip_bc_generate");

  const int npts = (int)c.points.size();
  std::vector<int> sx, sy, st;
  for (const auto& p : c.points) {
    sx.push_back(p.x);
    sy.push_back(p.y);
    st.push_back(p.type);
  }
  if (npts > 0) {  // sampleHeader (src/main.cpp:45-51)
    std::cout << "t;";
    for (const auto& p : c.points) std::cout << p.label << ";";
    std::cout << std::endl;
  }
  check(fip_set_run(sim, c.target_duration, c.sampling_interval, sx.data(), sy.data(), st.data(),
                    npts, c.save_state.data(), (int)c.save_state.size()),
        "fip_set_run");

  std::vector<double> ts(256);
  std::vector<float> vals(256 * (npts > 0 ? npts : 1));
  auto drain = [&]() {
    int n = 0;
    do {
      check(fip_read_samples(sim, ts.data(), vals.data(), 256, &n), "fip_read_samples");
      for (int r = 0; r < n; r++) {  // sample() (src/main.cpp:53-83)
        std::cout << ts[r] << ";";
        for (int p = 0; p < npts; p++) std::cout << vals[(size_t)r * npts + p] << ";";
        std::cout << std::endl;
      }
    } while (n == 256);
  };

  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    int ev = FIP_EV_NONE;
    long long done = 0;
    check(fip_simulate(sim, 1LL << 40, &ev, &done), "fip_simulate");
    drain();
    if (ev == FIP_EV_DUMP) {
      dump(sim, c);
      continue;
    }
    if (ev == FIP_EV_FINISHED) break;
  }  // FIP_EV_SAMPLES / FIP_EV_NONE: the log was drained, continue
  check(fip_sync(sim), "fip_sync");
  const float seconds =
      std::chrono::duration<float>(std::chrono::steady_clock::now() - t0).count();
  int iteration = 0;
  check(fip_get_time(sim, nullptr, nullptr, &iteration), "fip_get_time");
  std::cerr << std::dec << "Finished " << iteration << " iterations after " << seconds << " s ("
            << std::fixed << seconds / iteration << " s/it)" << std::endl;
  fip_destroy(sim);
  return 0;
}
