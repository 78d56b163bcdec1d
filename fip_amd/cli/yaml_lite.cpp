#include "yaml_lite.h"

#include <cctype>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>

namespace yl {

namespace {

const Node& undefined_node() {
  static const Node u;
  return u;
}

struct Line {
  int num;     // 0-based line number
  int indent;  // leading spaces
  std::string text;
};

std::string rtrim(const std::string& s) {
  size_t e = s.find_last_not_of(" \t\r");
  return e == std::string::npos ? "" : s.substr(0, e + 1);
}

// Removes a comment: '#' at the start or after whitespace, outside quotes.
std::string strip_comment(const std::string& s) {
  char q = 0;
  for (size_t i = 0; i < s.size(); i++) {
    char c = s[i];
    if (q) {
      if (c == q) {
        if (q == '\'' && i + 1 < s.size() && s[i + 1] == '\'') {
          i++;
          continue;
        }
        q = 0;
      } else if (q == '"' && c == '\\') {
        i++;
      }
      continue;
    }
    if (c == '\'' || c == '"') {
      if (i == 0 || std::isspace((unsigned char)s[i - 1]) || s[i - 1] == '[' || s[i - 1] == '{' ||
          s[i - 1] == ',' || s[i - 1] == ':')
        q = c;
      continue;
    }
    if (c == '#' && (i == 0 || std::isspace((unsigned char)s[i - 1]))) return s.substr(0, i);
  }
  return s;
}

std::string fmt_err(int line, int col, const std::string& msg) {
  if (line < 0) return msg;
  std::ostringstream o;
  o << "yaml-cpp: error at line " << line + 1 << ", column " << col + 1 << ": " << msg;
  return o.str();
}

NodeP make(Node::Kind k, int line, int col) {
  auto n = std::make_shared<Node>();
  n->kind = k;
  n->line = line;
  n->col = col;
  return n;
}

std::string unquote(const std::string& t, int line, int col) {
  if (t.size() >= 2 && t[0] == '\'' && t.back() == '\'') {
    std::string r;
    for (size_t i = 1; i + 1 < t.size(); i++) {
      r += t[i];
      if (t[i] == '\'' && i + 2 < t.size() && t[i + 1] == '\'') i++;
    }
    return r;
  }
  if (t.size() >= 2 && t[0] == '"' && t.back() == '"') {
    std::string r;
    for (size_t i = 1; i + 1 < t.size(); i++) {
      char c = t[i];
      if (c == '\\' && i + 2 < t.size()) {
        char e = t[++i];
        switch (e) {
          case 'n': r += '\n'; break;
          case 't': r += '\t'; break;
          case '\\': r += '\\'; break;
          case '"': r += '"'; break;
          case '0': r += '\0'; break;
          default: r += e; break;
        }
      } else {
        r += c;
      }
    }
    return r;
  }
  if (!t.empty() && (t[0] == '\'' || t[0] == '"'))
    throw Error(fmt_err(line, col, "end of string not found"));
  return t;
}

bool is_null_plain(const std::string& s) {
  return s.empty() || s == "~" || s == "null" || s == "Null" || s == "NULL";
}

NodeP scalar_node(const std::string& raw, int line, int col) {
  std::string t = rtrim(raw);
  size_t b = t.find_first_not_of(" \t");
  if (b == std::string::npos) return make(Node::Null, line, col);
  col += (int)b;
  t = t.substr(b);
  const bool quoted = t[0] == '\'' || t[0] == '"';
  if (!quoted && is_null_plain(t)) return make(Node::Null, line, col);
  auto n = make(Node::Scalar, line, col);
  n->scalar = unquote(t, line, col);
  return n;
}

// Flow collections: [a, b, c] and {k: v, ...}, possibly nested.
struct Flow {
  const std::string& s;
  size_t i = 0;
  int line, col0;
  Flow(const std::string& str, int l, int c) : s(str), line(l), col0(c) {}
  void ws() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
  }
  NodeP value(const char* stops) {
    ws();
    if (i >= s.size()) return make(Node::Null, line, col0 + (int)i);
    if (s[i] == '[') return seq();
    if (s[i] == '{') return map();
    size_t st = i;
    if (s[i] == '\'' || s[i] == '"') {
      char q = s[i++];
      while (i < s.size()) {
        if (s[i] == q) {
          if (q == '\'' && i + 1 < s.size() && s[i + 1] == '\'') {
            i += 2;
            continue;
          }
          i++;
          break;
        }
        if (q == '"' && s[i] == '\\') i++;
        i++;
      }
    } else {
      while (i < s.size() && !std::strchr(stops, s[i])) {
        if (s[i] == ':' && std::strchr(stops, ':') == nullptr && i + 1 < s.size() && s[i + 1] == ' ' &&
            std::strchr(stops, '}'))
          break;
        i++;
      }
    }
    return scalar_node(s.substr(st, i - st), line, col0 + (int)st);
  }
  NodeP seq() {
    auto n = make(Node::Sequence, line, col0 + (int)i);
    i++;
    ws();
    if (i < s.size() && s[i] == ']') {
      i++;
      return n;
    }
    for (;;) {
      n->seq.push_back(value(",]"));
      ws();
      if (i < s.size() && s[i] == ',') {
        i++;
        continue;
      }
      if (i < s.size() && s[i] == ']') {
        i++;
        return n;
      }
      throw Error(fmt_err(line, col0 + (int)i, "end of sequence flow not found"));
    }
  }
  NodeP map() {
    auto n = make(Node::Map, line, col0 + (int)i);
    i++;
    ws();
    if (i < s.size() && s[i] == '}') {
      i++;
      return n;
    }
    for (;;) {
      ws();
      size_t st = i;
      while (i < s.size() && s[i] != ':' && s[i] != ',' && s[i] != '}') i++;
      std::string key = unquote(rtrim(s.substr(st, i - st)), line, col0 + (int)st);
      NodeP v;
      if (i < s.size() && s[i] == ':') {
        i++;
        v = value(",}");
      } else {
        v = make(Node::Null, line, col0 + (int)i);
      }
      n->map.emplace_back(key, v);
      ws();
      if (i < s.size() && s[i] == ',') {
        i++;
        continue;
      }
      if (i < s.size() && s[i] == '}') {
        i++;
        return n;
      }
      throw Error(fmt_err(line, col0 + (int)i, "end of map flow not found"));
    }
  }
};

struct Block {
  std::vector<Line> L;
  size_t p = 0;

  static bool seq_item(const std::string& t) { return t == "-" || (t.size() >= 2 && t[0] == '-' && t[1] == ' '); }

  // Position of the mapping ':' (followed by a space or the end), outside quotes and
  // flow brackets; npos if the line is not a mapping entry.
  static size_t map_colon(const std::string& t) {
    char q = 0;
    int depth = 0;
    for (size_t i = 0; i < t.size(); i++) {
      char c = t[i];
      if (q) {
        if (c == q) q = 0;
        continue;
      }
      if ((c == '\'' || c == '"') && i == 0) {
        q = c;
        continue;
      }
      if (c == '[' || c == '{') depth++;
      if (c == ']' || c == '}') depth--;
      if (c == ':' && depth == 0 && (i + 1 == t.size() || t[i + 1] == ' ' || t[i + 1] == '\t'))
        return i;
    }
    return std::string::npos;
  }

  NodeP inline_value(const std::string& text, int line, int col) {
    size_t b = text.find_first_not_of(" \t");
    if (b == std::string::npos) return make(Node::Null, line, col);
    if (text[b] == '[' || text[b] == '{') {
      std::string joined = text;
      auto balanced = [](const std::string& s) {
        int d = 0;
        char q = 0;
        for (char c : s) {
          if (q) {
            if (c == q) q = 0;
            continue;
          }
          if (c == '\'' || c == '"') q = c;
          if (c == '[' || c == '{') d++;
          if (c == ']' || c == '}') d--;
        }
        return d <= 0;
      };
      while (!balanced(joined) && p < L.size()) joined += " " + L[p++].text;
      Flow f(joined, line, col);
      f.i = b;
      NodeP n = f.value(",]}");
      f.ws();
      if (f.i != joined.size()) throw Error(fmt_err(line, col + (int)f.i, "illegal flow end"));
      return n;
    }
    return scalar_node(text, line, col);
  }

  NodeP node(int min_indent) {
    if (p >= L.size() || L[p].indent < min_indent) return make(Node::Null, -1, -1);
    const Line& ln = L[p];
    if (seq_item(ln.text)) return seq(ln.indent);
    if (map_colon(ln.text) != std::string::npos) return map(ln.indent);
    p++;
    return inline_value(ln.text, ln.num, ln.indent);
  }

  NodeP seq(int ind) {
    auto n = make(Node::Sequence, L[p].num, ind);
    while (p < L.size() && L[p].indent == ind && seq_item(L[p].text)) {
      Line ln = L[p];
      size_t off = 1;
      while (off < ln.text.size() && ln.text[off] == ' ') off++;
      std::string rest = ln.text.substr(off);
      const int col = ind + (int)off;
      if (rest.empty()) {
        p++;
        if (p < L.size() && L[p].indent > ind)
          n->seq.push_back(node(ind + 1));
        else
          n->seq.push_back(make(Node::Null, ln.num, col));
      } else if (seq_item(rest) || map_colon(rest) != std::string::npos) {
        // a block collection that starts on the item's line: continue it as if the item
        // text began a line of its own at column `col`
        L[p].indent = col;
        L[p].text = rest;
        n->seq.push_back(node(col));
      } else {
        p++;
        n->seq.push_back(inline_value(rest, ln.num, col));
      }
    }
    return n;
  }

  NodeP map(int ind) {
    auto n = make(Node::Map, L[p].num, ind);
    while (p < L.size() && L[p].indent == ind && !seq_item(L[p].text)) {
      const Line ln = L[p];
      size_t c = map_colon(ln.text);
      if (c == std::string::npos) throw Error(fmt_err(ln.num, ind, "illegal map value"));
      std::string key = unquote(rtrim(ln.text.substr(0, c)), ln.num, ind);
      std::string rest = ln.text.substr(c + 1);
      size_t b = rest.find_first_not_of(" \t");
      p++;
      NodeP v;
      if (b != std::string::npos) {
        v = inline_value(rest.substr(b), ln.num, ind + (int)(c + 1 + b));
      } else if (p < L.size() && L[p].indent > ind) {
        v = node(ind + 1);
      } else if (p < L.size() && L[p].indent == ind && seq_item(L[p].text)) {
        v = seq(ind);
      } else {
        v = make(Node::Null, ln.num, ind + (int)c + 1);
      }
      bool dup = false;
      for (auto& kv : n->map) dup |= kv.first == key;
      if (!dup) n->map.emplace_back(key, v);
    }
    return n;
  }
};

template <class T>
bool stream_to(const std::string& in, T& out) {
  std::stringstream stream(in);
  stream.unsetf(std::ios::dec);
  if ((stream >> std::noskipws >> out) && (stream >> std::ws).eof()) return true;
  return false;
}

bool is_inf(const std::string& s) {
  return s == ".inf" || s == ".Inf" || s == ".INF" || s == "+.inf" || s == "+.Inf" || s == "+.INF";
}
bool is_ninf(const std::string& s) { return s == "-.inf" || s == "-.Inf" || s == "-.INF"; }
bool is_nan(const std::string& s) { return s == ".nan" || s == ".NaN" || s == ".NAN"; }

template <class F>
bool float_conv(const Node& n, F& out) {
  if (n.kind != Node::Scalar) return false;
  if (stream_to(n.scalar, out)) return true;
  if (is_inf(n.scalar)) {
    out = std::numeric_limits<F>::infinity();
    return true;
  }
  if (is_ninf(n.scalar)) {
    out = -std::numeric_limits<F>::infinity();
    return true;
  }
  if (is_nan(n.scalar)) {
    out = std::numeric_limits<F>::quiet_NaN();
    return true;
  }
  return false;
}

[[noreturn]] void bad_conversion(const Node& n) { throw Error(fmt_err(n.line, n.col, "bad conversion")); }

}  // namespace

const Node& Node::operator[](const std::string& key) const {
  if (kind != Map) return undefined_node();
  for (auto& kv : map)
    if (kv.first == key) return *kv.second;
  return undefined_node();
}

const Node& Node::operator[](size_t i) const {
  if (kind != Sequence || i >= seq.size()) return undefined_node();
  return *seq[i];
}

std::vector<const Node*> Node::items() const {
  std::vector<const Node*> r;
  if (kind == Sequence)
    for (auto& e : seq) r.push_back(e.get());
  return r;
}

NodeP load_string(const std::string& text) {
  Block b;
  std::istringstream in(text);
  std::string s;
  int num = 0;
  while (std::getline(in, s)) {
    std::string t = rtrim(strip_comment(s));
    size_t ind = t.find_first_not_of(' ');
    if (ind != std::string::npos && t.substr(ind) != "---" && t.substr(ind) != "...")
      b.L.push_back({num, (int)ind, t.substr(ind)});
    num++;
  }
  if (b.L.empty()) return make(Node::Null, -1, -1);
  NodeP root = b.node(0);
  if (b.p < b.L.size())
    throw Error(fmt_err(b.L[b.p].num, b.L[b.p].indent, "end of map not found"));
  return root;
}

NodeP load_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw Error("bad file: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return load_string(ss.str());
}

bool try_int(const Node& n, long long& out) {
  if (n.kind != Node::Scalar) return false;
  int v;
  if (!stream_to(n.scalar, v)) return false;
  out = v;
  return true;
}
bool try_uint(const Node& n, unsigned long long& out) {
  if (n.kind != Node::Scalar || (!n.scalar.empty() && n.scalar[0] == '-')) return false;
  size_t v;
  if (!stream_to(n.scalar, v)) return false;
  out = v;
  return true;
}
bool try_float(const Node& n, float& out) { return float_conv(n, out); }
bool try_double(const Node& n, double& out) { return float_conv(n, out); }

bool try_bool(const Node& n, bool& out) {
  if (n.kind != Node::Scalar) return false;
  const std::string& s = n.scalar;
  if (s.empty()) return false;
  auto all = [&](size_t from, int (*pred)(int)) {
    for (size_t i = from; i < s.size(); i++)
      if (!pred((unsigned char)s[i])) return false;
    return true;
  };
  const bool flexible = all(0, islower) ||
                        (isupper((unsigned char)s[0]) && (all(1, islower) || all(1, isupper)));
  if (!flexible) return false;
  std::string l;
  for (char c : s) l += (char)std::tolower((unsigned char)c);
  static const char* T[] = {"y", "yes", "true", "on"};
  static const char* F[] = {"n", "no", "false", "off"};
  for (int i = 0; i < 4; i++) {
    if (l == T[i]) {
      out = true;
      return true;
    }
    if (l == F[i]) {
      out = false;
      return true;
    }
  }
  return false;
}

std::string as_string(const Node& n) {
  if (n.kind == Node::Null) return "null";
  if (n.kind != Node::Scalar) bad_conversion(n);
  return n.scalar;
}
int as_int(const Node& n) {
  long long v;
  if (!try_int(n, v)) bad_conversion(n);
  return (int)v;
}
size_t as_size(const Node& n) {
  unsigned long long v;
  if (!try_uint(n, v)) bad_conversion(n);
  return (size_t)v;
}
float as_float(const Node& n) {
  float v;
  if (!try_float(n, v)) bad_conversion(n);
  return v;
}
bool as_bool(const Node& n) {
  bool v;
  if (!try_bool(n, v)) bad_conversion(n);
  return v;
}
int as_int_or(const Node& n, int fallback) {
  long long v;
  return try_int(n, v) ? (int)v : fallback;
}
float as_float_or(const Node& n, float fallback) {
  float v;
  return try_float(n, v) ? v : fallback;
}
bool as_bool_or(const Node& n, bool fallback) {
  bool v;
  return try_bool(n, v) ? v : fallback;
}

}  // namespace yl
