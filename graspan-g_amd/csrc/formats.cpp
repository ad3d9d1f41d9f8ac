// Parsers for the reference's grammar and graph files.
//
// Behaviour kept from the reference:
//   * lines starting with '#' are skipped wherever a data line is expected
//     (skipComments, src/main.cpp:60-67);
//   * grammar: the first data line is the rule count (atoi), then one data
//     line per rule with 1-3 TAB-separated label names; label ids are given
//     1,2,... in order of first appearance, left to right
//     (src/main.cpp:69-97);
//   * graph: header "numEdges<TAB>numVars", then numEdges data lines
//     "src<TAB>dst<TAB>labelname" (src/main.cpp:106-139, 177-221).
// Deliberate differences (the reference has undefined behaviour there):
//   * a blank line, a missing line, a non-numeric field, more than 3 names on
//     a rule line, an unknown label or a vertex id >= numVars is an error;
//   * edges need not be grouped by src (the closure sorts internally);
//   * a trailing '\r' (CRLF files) is stripped;
//   * more than 21 rules is accepted (the reference exits, src/kernel.cu:24-27).
#include "formats.h"

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "graspan_g.h"

namespace gg {
namespace {

// Whole-file reader with a line cursor.
class LineReader {
 public:
  bool open(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0) {
      std::fclose(f);
      return false;
    }
    buf_.resize(size_t(sz));
    size_t got = sz ? std::fread(&buf_[0], 1, size_t(sz), f) : 0;
    std::fclose(f);
    if (got != size_t(sz)) return false;
    pos_ = 0;
    return true;
  }
  // Next line (without '\n' and a trailing '\r'); false at end of file.
  bool next(const char*& b, const char*& e) {
    if (pos_ >= buf_.size()) return false;
    const char* base = buf_.data();
    const char* p = base + pos_;
    const char* end = base + buf_.size();
    const char* nl = static_cast<const char*>(std::memchr(p, '\n', size_t(end - p)));
    const char* le = nl ? nl : end;
    pos_ = size_t((nl ? nl + 1 : end) - base);
    ++line_no_;
    b = p;
    e = le;
    if (e > b && e[-1] == '\r') --e;
    return true;
  }
  // Next line that does not start with '#' (skipComments).
  bool next_data(const char*& b, const char*& e) {
    for (;;) {
      if (!next(b, e)) return false;
      if (b == e || *b != '#') return true;
    }
  }
  long line_no() const { return line_no_; }

 private:
  std::string buf_;
  size_t pos_ = 0;
  long line_no_ = 0;
};

// Splits [b,e) on '\t'.
void split_tabs(const char* b, const char* e, std::vector<std::string>& out) {
  out.clear();
  const char* s = b;
  for (const char* p = b; p <= e; ++p) {
    if (p == e || *p == '\t') {
      out.emplace_back(s, size_t(p - s));
      s = p + 1;
    }
  }
}

bool parse_u64(const char* b, const char* e, uint64_t& v) {
  if (b == e) return false;
  uint64_t x = 0;
  for (const char* p = b; p < e; ++p) {
    if (*p < '0' || *p > '9') return false;
    uint64_t d = uint64_t(*p - '0');
    if (x > (UINT64_MAX - d) / 10) return false;
    x = x * 10 + d;
  }
  v = x;
  return true;
}

std::string where(const char* path, long line) { return std::string(path) + ":" + std::to_string(line) + ": "; }

}  // namespace

int parse_grammar_file(const char* path, Grammar& g, std::string& err) {
  g = Grammar();
  LineReader rd;
  if (!rd.open(path)) {
    err = std::string("cannot read grammar file ") + path;
    return GG_E_IO;
  }
  const char *b, *e;
  if (!rd.next_data(b, e)) {
    err = where(path, rd.line_no()) + "missing rule count";
    return GG_E_FORMAT;
  }
  uint64_t nr;
  if (!parse_u64(b, e, nr) || nr > 1000000) {
    err = where(path, rd.line_no()) + "bad rule count '" + std::string(b, e) + "'";
    return GG_E_FORMAT;
  }
  std::vector<std::string> items;
  for (uint64_t r = 0; r < nr; ++r) {
    if (!rd.next_data(b, e)) {
      err = where(path, rd.line_no()) + "expected " + std::to_string(nr) + " rules, found " + std::to_string(r);
      return GG_E_FORMAT;
    }
    split_tabs(b, e, items);
    if (items.size() > 3) {
      err = where(path, rd.line_no()) + "more than 3 labels in a rule";
      return GG_E_FORMAT;
    }
    std::array<int32_t, 3> rule{0, 0, 0};
    for (size_t j = 0; j < items.size(); ++j) {
      if (items[j].empty()) {
        err = where(path, rd.line_no()) + "empty label name";
        return GG_E_FORMAT;
      }
      auto it = g.ids.find(items[j]);
      int id;
      if (it == g.ids.end()) {
        if (int(g.names.size()) >= kMaxLabels) {
          err = where(path, rd.line_no()) + "more than 255 labels";
          return GG_E_FORMAT;
        }
        g.names.push_back(items[j]);
        id = int(g.names.size());
        g.ids.emplace(items[j], id);
      } else {
        id = it->second;
      }
      rule[j] = id;
    }
    g.rules.push_back(rule);
  }
  return GG_OK;
}

int grammar_from_rules(int num_rules, const int32_t* rules, int num_labels, Grammar& g, std::string& err) {
  g = Grammar();
  if (num_rules < 0 || num_labels < 1 || num_labels > kMaxLabels || (num_rules > 0 && rules == nullptr)) {
    err = "bad grammar arguments";
    return GG_E_ARG;
  }
  for (int l = 1; l <= num_labels; ++l) {
    g.names.push_back(std::to_string(l));
    g.ids.emplace(std::to_string(l), l);
  }
  for (int r = 0; r < num_rules; ++r) {
    int32_t a = rules[3 * r], b = rules[3 * r + 1], c = rules[3 * r + 2];
    if (a < 1 || a > num_labels || b < 0 || b > num_labels || c < 0 || c > num_labels || (b == 0 && c != 0)) {
      err = "bad rule " + std::to_string(r);
      return GG_E_ARG;
    }
    g.rules.push_back({a, b, c});
  }
  return GG_OK;
}

namespace {
// Generators of the star label y (see plan_fixpoint_rules), or empty when y
// is not one.
std::vector<int32_t> star_generators(const int32_t* rules, int num_rules, int32_t y) {
  bool eps = false, left = true, right = true;
  std::vector<int32_t> left_gen, right_gen;
  for (int i = 0; i < num_rules; ++i) {
    const int32_t* r = rules + 3 * i;
    if (r[0] != y) continue;
    if (r[1] == 0) {
      eps = true;
      continue;
    }
    if (r[2] == 0) return {};  // a unary rule for y
    left = left && r[1] == y && r[2] != y;
    right = right && r[2] == y && r[1] != y;
    left_gen.push_back(r[2]);
    right_gen.push_back(r[1]);
  }
  std::vector<int32_t>& gen = left ? left_gen : right_gen;
  if (!eps || gen.empty() || !(left || right)) return {};
  std::sort(gen.begin(), gen.end());
  gen.erase(std::unique(gen.begin(), gen.end()), gen.end());
  return gen;
}
}  // namespace

std::vector<int32_t> plan_fixpoint_rules(const int32_t* rules, int num_rules, int max_per_kind) {
  std::vector<int32_t> out(rules, rules + 3 * std::max(num_rules, 0));
  const char* off = std::getenv("GG_STAR_REWRITE");
  if (off && off[0] == '0') return out;
  int unary = 0, binary = 0;
  for (int i = 0; i < num_rules; ++i)
    if (rules[3 * i + 1] != 0) ++(rules[3 * i + 2] ? binary : unary);
  out.clear();
  for (int i = 0; i < num_rules; ++i) {
    const int32_t x = rules[3 * i], b = rules[3 * i + 1], c = rules[3 * i + 2];
    bool done = false;
    if (b != 0 && c != 0 && x != b && x != c) {
      int own = 0;
      for (int k = 0; k < num_rules; ++k) own += rules[3 * k] == x;
      const std::vector<int32_t> gb = own == 1 ? star_generators(rules, num_rules, b) : std::vector<int32_t>{};
      const std::vector<int32_t> gc =
          own == 1 && gb.empty() ? star_generators(rules, num_rules, c) : std::vector<int32_t>{};
      const std::vector<int32_t>& gen = gb.empty() ? gc : gb;
      const bool ok = !gen.empty() && std::find(gen.begin(), gen.end(), x) == gen.end() &&
                      unary + 1 <= max_per_kind && binary - 1 + int(gen.size()) <= max_per_kind;
      if (ok) {
        const bool left = !gb.empty();
        out.insert(out.end(), {x, left ? c : b, 0});  // X ::= C  (or X ::= B)
        for (int32_t p : gen) {
          if (left)
            out.insert(out.end(), {x, p, x});  // X ::= P_k X
          else
            out.insert(out.end(), {x, x, p});  // X ::= X P_k
        }
        unary += 1;
        binary += int(gen.size()) - 1;
        done = true;
      }
    }
    if (!done) out.insert(out.end(), {x, b, c});
  }
  return out;
}

int parse_graph_file(const char* path, const Grammar& g, Graph& out, std::string& err) {
  out = Graph();
  LineReader rd;
  if (!rd.open(path)) {
    err = std::string("cannot read graph file ") + path;
    return GG_E_IO;
  }
  const char *b, *e;
  if (!rd.next_data(b, e)) {
    err = where(path, rd.line_no()) + "missing header";
    return GG_E_FORMAT;
  }
  const char* tab = static_cast<const char*>(std::memchr(b, '\t', size_t(e - b)));
  uint64_t m, n;
  if (!tab || !parse_u64(b, tab, m) || !parse_u64(tab + 1, e, n) || n > UINT32_MAX) {
    err = where(path, rd.line_no()) + "bad header '" + std::string(b, e) + "' (want numEdges<TAB>numVars)";
    return GG_E_FORMAT;
  }
  out.n = uint32_t(n);
  out.src.reserve(m);
  out.dst.reserve(m);
  out.lab.reserve(m);
  std::string name;
  for (uint64_t i = 0; i < m; ++i) {
    if (!rd.next_data(b, e)) {
      err = where(path, rd.line_no()) + "expected " + std::to_string(m) + " edges, found " + std::to_string(i);
      return GG_E_FORMAT;
    }
    const char* t1 = static_cast<const char*>(std::memchr(b, '\t', size_t(e - b)));
    const char* t2 = t1 ? static_cast<const char*>(std::memchr(t1 + 1, '\t', size_t(e - t1 - 1))) : nullptr;
    uint64_t s, d;
    if (!t1 || !t2 || !parse_u64(b, t1, s) || !parse_u64(t1 + 1, t2, d)) {
      err = where(path, rd.line_no()) + "bad edge line '" + std::string(b, e) + "'";
      return GG_E_FORMAT;
    }
    if (s >= n || d >= n) {
      err = where(path, rd.line_no()) + "vertex id out of range [0," + std::to_string(n) + ")";
      return GG_E_FORMAT;
    }
    name.assign(t2 + 1, e);
    auto it = g.ids.find(name);
    if (it == g.ids.end()) {
      err = where(path, rd.line_no()) + "label '" + name + "' is not in the grammar";
      return GG_E_FORMAT;
    }
    out.src.push_back(uint32_t(s));
    out.dst.push_back(uint32_t(d));
    out.lab.push_back(uint8_t(it->second));
  }
  return GG_OK;
}

int sniff_file_kind(const char* path) {
  LineReader rd;
  if (!rd.open(path)) return 0;
  const char *b, *e;
  if (!rd.next_data(b, e)) return 0;
  uint64_t v;
  if (parse_u64(b, e, v)) return 1;
  const char* tab = static_cast<const char*>(std::memchr(b, '\t', size_t(e - b)));
  uint64_t v2;
  if (tab && parse_u64(b, tab, v) && parse_u64(tab + 1, e, v2)) return 2;
  return 0;
}

}  // namespace gg
