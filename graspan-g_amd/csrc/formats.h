// Grammar and graph file formats of the reference (Graspan-G), parsed on the
// host. Plain C++; no HIP.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace gg {

struct Grammar {
  std::vector<std::string> names;                 // names[id - 1]
  std::unordered_map<std::string, int> ids;       // name -> id (1-based)
  std::vector<std::array<int32_t, 3>> rules;      // (lhs, rhs1, rhs2), 0 = absent
  int num_labels() const { return int(names.size()); }
};

struct Graph {
  uint32_t n = 0;
  std::vector<uint32_t> src, dst;
  std::vector<uint8_t> lab;  // label ids 1..L
  uint64_t m() const { return src.size(); }
};

// Maximum number of labels: label ids are stored as uint8.
constexpr int kMaxLabels = 255;

// Grammar file, src/main.cpp:60-97. Returns 0 or a GG_E_* status; err gets a
// message on failure.
int parse_grammar_file(const char* path, Grammar& g, std::string& err);

// Rule triples (ids 1..num_labels) -> Grammar with names "1".."L".
int grammar_from_rules(int num_rules, const int32_t* rules, int num_labels, Grammar& g, std::string& err);

// Graph file, src/main.cpp:106-139 and 177-221; labels resolved with g.
int parse_graph_file(const char* path, const Grammar& g, Graph& out, std::string& err);

// Content sniffing for the command line (README order vs code order):
// 1 = grammar (first data line is one integer), 2 = graph (first data line is
// two tab-separated integers), 0 = neither / unreadable.
int sniff_file_kind(const char* path);

}  // namespace gg
