#include "config.h"

#include <algorithm>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "yaml_lite.h"

namespace fipcli {

namespace {

const char* const kTypes[] = {"z", "q", "close", "open"};
const char* const kSides[] = {"left", "right", "top", "bottom"};
const char* const kSolvers[] = {"rmg", "cn"};
const char* const kSampleTypes[] = {"z", "h", "qx", "qy"};

// stringToEnum (src/enumhelper.h:15-23)
template <size_t N>
int to_enum(const char* const (&names)[N], const std::string& s) {
  for (size_t i = 0; i < N; i++)
    if (s == names[i]) return (int)i;
  throw std::invalid_argument("Invalid enum value: " + s);
}

// utils::loadFileIntoArray (src/utils.h:45-55): raw float32 at a byte offset. A short
// file leaves the remaining cells at 0.
void load_raw(const std::string& path, std::vector<float>& dst, size_t n, long long offset) {
  std::ifstream file(path, std::ios::binary);
  if (!file) throw std::invalid_argument("Failed to open file " + path);
  dst.assign(n, 0.f);
  file.seekg(offset);
  file.read(reinterpret_cast<char*>(dst.data()), (std::streamsize)(n * sizeof(float)));
}

void load_with_offset(std::vector<float>& dst, const yl::Node& parent, const std::string& key,
                      size_t n) {
  const yl::Node& node = parent[key];
  load_raw(yl::as_string(node[0]), dst, n, yl::as_int(node[1]));
}

// validateSimDataConfig (src/simdata.cpp:152-236)
bool validate(const yl::Node& config) {
  bool ok = true;
  auto error = [&ok](const yl::Node& node, const std::string& msg) {
    ok = false;
    std::cerr << "Config error at line " << node.line + 1 << ": " << msg << std::endl;
  };
  auto inConfig = [&](const yl::Node& node, const std::string& a) {
    if (!node[a].IsDefined()) error(node, "\"" + a + "\" missing");
  };
  auto xorInConfig = [&](const yl::Node& node, const std::string& a, const std::string& b) {
    if (node[a].IsDefined() && node[b].IsDefined())
      error(node[a], "Both \"" + a + "\" and \"" + b + "\" defined");
  };
  auto goodPath = [&](const yl::Node& node, const std::string& a) {
    if (node[a].IsDefined() && (!node[a].IsSequence() || node[a].size() != 2))
      error(node[a], "file path \"" + a + "\" needs format [path, byte_offset]");
  };
  auto goodSeries = [&](const yl::Node& node, const std::string& a) {
    if (node[a].IsDefined() && (!node[a].IsSequence() || node[a].size() != 3))
      error(node[a], "timeseries \"" + a +
                         "\" needs format [csv_filename, int_time_column, int_data_column]");
  };
  for (const char* k : {"name", "terrain", "W", "H", "dt", "dx", "duration"}) inConfig(config, k);
  xorInConfig(config, "kSt_var", "kSt");
  for (const char* k : {"terrain", "z", "qx", "qy", "kSt_var"}) goodPath(config, k);
  if (!config["kSt_var"].IsDefined() && !config["kSt"].IsDefined())
    std::cerr << "Warning: both \"kSt_var\" and \"kSt\" missing, using default kSt " << 30.f
              << std::endl;
  for (const yl::Node* bc : config["boundary_conditions"].items()) {
    inConfig(*bc, "side");
    inConfig(*bc, "type");
    const int type = to_enum(kTypes, yl::as_string((*bc)["type"]));
    if (type == 0) inConfig(*bc, "z");
    if (type == 1) {
      inConfig(*bc, "z");
      inConfig(*bc, "q");
    }
    if ((*bc)["z"].IsDefined() && (*bc)["z"].IsSequence()) goodSeries(*bc, "z");
    if ((*bc)["q"].IsDefined() && (*bc)["q"].IsSequence()) goodSeries(*bc, "q");
  }
  if (config["sampling_interval"].IsDefined()) {
    if (!config["sampling"].IsDefined() || !config["sampling"].IsSequence()) {
      ok = false;
      std::cerr << "Configuration error: sampling interval specified but no sampling points given"
                << std::endl;
    }
    for (const yl::Node* sp : config["sampling"].items())
      for (const char* k : {"x", "y", "label", "type"}) inConfig(*sp, k);
  }
  if (!config["save_state"].IsDefined()) std::cerr << "Warning: no save_state times given" << std::endl;
  return ok;
}

bool is_float(const std::string& str) {
  std::istringstream iss(str);
  float f;
  iss >> std::noskipws >> f;  // leading whitespace is invalid
  return iss.eof() && !iss.fail();
}

}  // namespace

float TimeSeries::interp(float time) const {
  auto upper = series.lower_bound(time);
  if (upper == series.begin()) return upper->second;
  if (upper == series.end()) return series.rbegin()->second;
  auto lower = std::prev(upper);
  float dt = upper->first - lower->first;
  float f = (time - lower->first) / dt;
  return f * upper->second + (1.f - f) * lower->second;
}

TimeSeries TimeSeries::fromCsv(const std::string& fileName, int timeColumn, int dataColumn,
                               float timeScale, bool skipHeader) {
  std::ifstream f(fileName);
  if (!f.is_open()) throw std::invalid_argument("Could not open file " + fileName);
  TimeSeries ts;
  std::string line;
  time_t firstTimestamp = -1;
  if (skipHeader) std::getline(f, line);
  while (std::getline(f, line)) {
    if (line.empty()) continue;  // a blank line has no cells
    std::stringstream ss(line);
    std::string val;
    float time = 0.f, data = 0.f;
    int counter = 0;
    while (std::getline(ss, val, ';')) {
      if (counter == timeColumn) {
        if (is_float(val)) {
          time = std::stof(val) * timeScale;
        } else {
          std::tm t = {};
          std::istringstream ds(val);
          ds >> std::get_time(&t, "%d.%m.%Y %H:%M:%S");
          if (ds.fail())
            throw std::invalid_argument("Wrong date format in timeseries " + fileName +
                                        ". Expected %d.%m.%Y %H:%M:%S");
          auto timestamp = std::mktime(&t);
          if (firstTimestamp == -1) firstTimestamp = timestamp;
          time = static_cast<float>(timestamp - firstTimestamp);
        }
      }
      if (counter == dataColumn) data = std::stof(val);
      counter++;
    }
    ts.series.emplace(time, data);  // the first of duplicate times wins
  }
  if (ts.series.empty()) throw std::invalid_argument("Empty timeseries " + fileName);
  return ts;
}

double solveSimpleMathExpression(const std::string& expression) {
  if (expression.find_first_not_of("0123456789.* ") != std::string::npos)
    throw std::invalid_argument("Unsupported math expression " + expression);
  double result = 1.0;
  std::stringstream ss(expression);
  std::string item;
  while (std::getline(ss, item, '*')) result *= std::stof(item);
  return result;
}

Config load(const std::string& path) {
  yl::NodeP root = yl::load_file(path);
  const yl::Node& config = *root;
  if (!validate(config)) throw std::invalid_argument("one or more configuration errors found");
  Config c;
  c.name = yl::as_string(config["name"]);
  c.target_duration = solveSimpleMathExpression(yl::as_string(config["duration"]));
  c.W = yl::as_size(config["W"]);
  c.H = yl::as_size(config["H"]);
  c.dt = yl::as_float(config["dt"]);
  c.dx = yl::as_float(config["dx"]);
  c.variable_dt = yl::as_bool_or(config["variable_dt"], true);
  c.solver = config["solver"].IsDefined() ? to_enum(kSolvers, yl::as_string(config["solver"])) : 0;
  c.kSt = yl::as_float_or(config["kSt"], 30.f);
  const size_t n = c.W * c.H;
  load_with_offset(c.bed, config, "terrain", n);
  if (config["no_data"].IsDefined()) c.no_data = yl::as_float(config["no_data"]);
  c.h.assign(n, 0.f);
  c.qx.assign(n, 0.f);
  c.qy.assign(n, 0.f);
  if (config["z"].IsDefined()) {
    load_with_offset(c.h, config, "z", n);
    for (size_t i = 0; i < n; i++) c.h[i] = c.h[i] - c.bed[i];  // level -> depth
  }
  if (config["qx"].IsDefined()) load_with_offset(c.qx, config, "qx", n);
  if (config["qy"].IsDefined()) load_with_offset(c.qy, config, "qy", n);
  if (config["kSt_var"].IsDefined()) {
    load_with_offset(c.kst_var, config, "kSt_var", n);
    c.has_kst_var = true;
  }
  // createBoundaryCondition (src/simdata.cpp:98-131)
  for (const yl::Node* bp : config["boundary_conditions"].items()) {
    const yl::Node& bc = *bp;
    BCConfig b;
    b.type = to_enum(kTypes, yl::as_string(bc["type"]));
    b.side = to_enum(kSides, yl::as_string(bc["side"]));
    b.series = (bc["q"].IsDefined() && bc["q"].IsSequence()) ||
               (bc["z"].IsDefined() && bc["z"].IsSequence());
    b.from = yl::as_int_or(bc["from"], 0);
    b.to = yl::as_int_or(bc["to"], (b.side == 0 || b.side == 1) ? (int)c.H - 1 : (int)c.W - 1);
    if (b.series) {
      auto parse = [](const yl::Node& node, TimeSeries& ts) {
        if (!node.IsDefined()) return false;
        if (!node.IsSequence()) throw yl::Error("operator[] call on a scalar");
        ts = TimeSeries::fromCsv(yl::as_string(node[0]), yl::as_int(node[1]), yl::as_int(node[2]));
        return true;
      };
      b.has_q = parse(bc["q"], b.qs);
      b.has_z = parse(bc["z"], b.zs);
      // BoundaryConditionTimeseries (src/boundary_condition.cpp:57-64)
      if (b.type == 0 && !b.has_z)
        throw std::invalid_argument("BoundaryCondition error: type z but no z timeseries given");
      if (b.type == 1 && (!b.has_z || !b.has_q))
        throw std::invalid_argument("BoundaryCondition error: type q requires both z and q timeseries");
    } else {
      b.q = yl::as_float_or(bc["q"], 0.f);
      b.z = yl::as_float_or(bc["z"], 0.f);
    }
    c.bcs.push_back(std::move(b));
  }
  if (config["sampling_interval"].IsDefined()) {
    c.sampling_interval = yl::as_int(config["sampling_interval"]);
    for (const yl::Node* sp : config["sampling"].items()) {
      SamplePoint p;
      p.x = yl::as_int((*sp)["x"]);
      p.y = yl::as_int((*sp)["y"]);
      p.label = yl::as_string((*sp)["label"]);
      p.type = to_enum(kSampleTypes, yl::as_string((*sp)["type"]));
      c.points.push_back(p);
    }
  }
  if (config["save_state"].IsDefined()) {
    for (const yl::Node* s : config["save_state"].items())
      c.save_state.push_back(solveSimpleMathExpression(yl::as_string(*s)));
    std::sort(c.save_state.begin(), c.save_state.end());
  }
  return c;
}

}  // namespace fipcli
