// A YAML reader for fip's parameters.yml: block mappings and sequences, single-line or
// multi-line flow sequences/mappings, plain and quoted scalars and comments. Scalar
// conversions follow yaml-cpp 0.7.0, which the reference uses
// (deps/yaml-cpp/include/yaml-cpp/node/convert.h:140-192, deps/yaml-cpp/src/convert.cpp:41-73):
// numbers are read with a stringstream that must consume the whole scalar (.inf/.nan
// spellings accepted for floats, a leading '-' rejected for unsigned types); booleans
// accept y/n/yes/no/true/false/on/off in lower, Capitalised or UPPER case.
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace yl {

struct Node;
using NodeP = std::shared_ptr<Node>;

struct Node {
  enum Kind { Undefined, Null, Scalar, Sequence, Map };
  Kind kind = Undefined;
  std::string scalar;
  std::vector<NodeP> seq;
  std::vector<std::pair<std::string, NodeP>> map;
  int line = -1, col = -1;  // 0-based position of the node's first token

  bool IsDefined() const { return kind != Undefined; }
  bool IsSequence() const { return kind == Sequence; }
  bool IsScalar() const { return kind == Scalar; }
  size_t size() const { return kind == Sequence ? seq.size() : kind == Map ? map.size() : 0; }
  // Mapping lookup; an undefined node (not an error) when absent or not a map.
  const Node& operator[](const std::string& key) const;
  // Sequence element; an undefined node when out of range.
  const Node& operator[](size_t i) const;
  const Node& operator[](int i) const { return (*this)[(size_t)i]; }
  std::vector<const Node*> items() const;  // sequence elements (empty otherwise)
};

// Thrown for parse errors and failed conversions; what() follows yaml-cpp's
// "yaml-cpp: error at line L, column C: msg".
struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

NodeP load_file(const std::string& path);
NodeP load_string(const std::string& text);

// Conversions; as_* throw Error("... bad conversion") when the node is not a matching
// scalar, try_* return false.
bool try_int(const Node& n, long long& out);
bool try_uint(const Node& n, unsigned long long& out);
bool try_float(const Node& n, float& out);
bool try_double(const Node& n, double& out);
bool try_bool(const Node& n, bool& out);
std::string as_string(const Node& n);
int as_int(const Node& n);
size_t as_size(const Node& n);
float as_float(const Node& n);
bool as_bool(const Node& n);
int as_int_or(const Node& n, int fallback);
float as_float_or(const Node& n, float fallback);
bool as_bool_or(const Node& n, bool fallback);

}  // namespace yl
