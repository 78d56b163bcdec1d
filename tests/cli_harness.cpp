// Test-only C entry points over the fip CLI's host code (fip_amd/cli), compiled by
// tests/test_cli_host.py so the parsing can be compared with the reference's own code.
#include <cstring>
#include <sstream>
#include <stdexcept>

#include "config.h"

static void put(char* err, int cap, const char* s) {
  std::strncpy(err, s, cap - 1);
  err[cap - 1] = 0;
}

extern "C" {
int cli_series_interp(const char* path, int tcol, int dcol, const float* t, int n, float* out,
                      char* err, int cap) {
  try {
    fipcli::TimeSeries ts = fipcli::TimeSeries::fromCsv(path, tcol, dcol);
    for (int i = 0; i < n; i++) out[i] = ts.interp(t[i]);
    return 0;
  } catch (std::exception& e) {
    put(err, cap, e.what());
    return -1;
  }
}

int cli_math_expr(const char* expr, double* out, char* err, int cap) {
  try {
    *out = fipcli::solveSimpleMathExpression(expr);
    return 0;
  } catch (std::exception& e) {
    put(err, cap, e.what());
    return -1;
  }
}

// Loads a parameters.yml and writes a JSON summary into buf.
int cli_load(const char* path, char* buf, int cap) {
  try {
    fipcli::Config c = fipcli::load(path);
    std::ostringstream o;
    o.precision(9);
    o << "{\"name\":\"" << c.name << "\",\"W\":" << c.W << ",\"H\":" << c.H << ",\"dx\":" << c.dx
      << ",\"dt\":" << c.dt << ",\"duration\":" << c.target_duration << ",\"kst\":" << c.kSt
      << ",\"no_data\":" << c.no_data << ",\"variable_dt\":" << (c.variable_dt ? 1 : 0)
      << ",\"solver\":" << c.solver << ",\"kst_var\":" << (c.has_kst_var ? 1 : 0);
    double bs = 0, hs = 0;
    for (float v : c.bed) bs += v;
    for (float v : c.h) hs += v;
    o << ",\"bed_sum\":" << bs << ",\"h_sum\":" << hs << ",\"bcs\":[";
    for (size_t i = 0; i < c.bcs.size(); i++) {
      const auto& b = c.bcs[i];
      o << (i ? "," : "") << "[" << b.type << "," << b.side << "," << b.from << "," << b.to << ","
        << (b.series ? 1 : 0) << "," << b.z << "," << b.q << "," << b.zs.series.size() << ","
        << b.qs.series.size() << "]";
    }
    o << "],\"points\":[";
    for (size_t i = 0; i < c.points.size(); i++)
      o << (i ? "," : "") << "[" << c.points[i].x << "," << c.points[i].y << ",\""
        << c.points[i].label << "\"," << c.points[i].type << "]";
    o << "],\"interval\":" << c.sampling_interval << ",\"save\":[";
    for (size_t i = 0; i < c.save_state.size(); i++) o << (i ? "," : "") << c.save_state[i];
    o << "]}";
    put(buf, cap, o.str().c_str());
    return 0;
  } catch (std::exception& e) {
    put(buf, cap, e.what());
    return -1;
  }
}
}
