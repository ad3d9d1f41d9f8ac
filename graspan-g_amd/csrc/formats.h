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

// The rule table the least-fixpoint engine evaluates in place of `rules`
// (flat (lhs, rhs1, rhs2) triples): same least fixpoint for every label,
// fewer join candidates. A label Y whose rules are exactly Y ::= eps and
// Y ::= Y P_k for k = 1..m (or all Y ::= P_k Y), no P_k = Y, is the
// reflexive-transitive closure P* of P = P_1 ∪ ... ∪ P_m. A binary rule
// X ::= Y C, where X has no other rule, X is not Y or C and no P_k is X, is
// replaced by X ::= C and X ::= P_k X for every k (the least X with
// X ⊇ C ∪ P·X is P*·C, Arden's lemma; the rest of the system is unchanged,
// so by Bekić's lemma every label keeps its least fixpoint); X ::= B Y, with
// Y such a star and B not, becomes X ::= B and X ::= X P_k. The join of
// X ::= Y C pairs every Y edge into a vertex with every C edge out of it;
// the star Y is dense where its generators are sparse, which is what the
// rewrite saves. A rewrite that would take the table over max_per_kind
// unary or binary rules is skipped. GG_STAR_REWRITE=0 turns it off. Applied
// twice it changes nothing more. Only the least fixpoint is preserved: the
// reference's filter != 0 schedule must run the table as given.
std::vector<int32_t> plan_fixpoint_rules(const int32_t* rules, int num_rules, int max_per_kind);

// Graph file, src/main.cpp:106-139 and 177-221; labels resolved with g.
int parse_graph_file(const char* path, const Grammar& g, Graph& out, std::string& err);

// Content sniffing for the command line (README order vs code order):
// 1 = grammar (first data line is one integer), 2 = graph (first data line is
// two tab-separated integers), 0 = neither / unreadable.
int sniff_file_kind(const char* path);

}  // namespace gg
