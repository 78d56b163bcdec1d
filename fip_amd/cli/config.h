// parameters.yml handling of the fip CLI: the configuration surface of the reference
// (src/simdata.cpp:98-310, src/timeseries.cpp, src/utils.h) rebuilt for this program.
#pragma once
#include <map>
#include <string>
#include <vector>

namespace fipcli {

// TimeSeries (src/timeseries.h:7-18): a std::map<float,float> read from a ';' CSV.
struct TimeSeries {
  std::map<float, float> series;
  float interp(float time) const;
  // TimeSeries::fromCsv (src/timeseries.cpp:28-85)
  static TimeSeries fromCsv(const std::string& file, int timeColumn, int dataColumn,
                            float timeScale = 1.f, bool skipHeader = true);
};

struct BCConfig {
  int type = 0;  // z q close open
  int side = 0;  // left right top bottom
  int from = 0, to = 0;
  bool series = false;
  float z = 0.f, q = 0.f;
  bool has_z = false, has_q = false;
  TimeSeries zs, qs;
};

struct SamplePoint {
  int x, y;
  std::string label;
  int type;  // z h qx qy
};

struct Config {
  std::string name;
  double target_duration = 0;
  size_t W = 0, H = 0;
  float dt = 0, dx = 0;
  bool variable_dt = true;
  int solver = 0;  // rmg cn
  float kSt = 30.f;
  float no_data = -9999.f;
  std::vector<float> bed, h, qx, qy, kst_var;  // W*H; h/qx/qy zero when not given
  bool has_kst_var = false;
  std::vector<BCConfig> bcs;
  int sampling_interval = 0;
  std::vector<SamplePoint> points;
  std::vector<double> save_state;
};

// SimData::loadSimData (src/simdata.cpp:238-310) up to, not including, the device upload.
// Validation problems are printed to stderr as the reference prints them; any failure
// throws std::exception whose what() the caller prints after "Config error: ".
Config load(const std::string& path);

// utils::solveSimpleMathExpression (src/utils.h:29-43)
double solveSimpleMathExpression(const std::string& expression);

}  // namespace fipcli
