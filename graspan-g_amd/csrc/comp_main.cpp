// comp -- command-line driver with the reference's interface
// (README.md:21, src/main.cpp:860-921):
//
//   comp <graph_file> <grammar_file> [filter_flag] [options]
//
// The README gives graph-then-grammar; the reference's code reads
// argv[1]=grammar, argv[2]=graph (src/main.cpp:869,871). Both orders are
// accepted: the files are told apart by content and swapped with a note on
// stderr. A missing filter_flag defaults to 0 (src/run.sh:1 passes none; the
// reference then calls atoi(NULL)).
//
// stdout carries the reference's report lines (src/main.cpp:34,54,57,75,93,
// 114-115,174,914-918; src/kernel.cu:36). Options:
//   --dump FILE        closure as little-endian uint32 (src,dst,label) triples,
//                      sorted
//   --dump-text FILE   closure as "src<TAB>dst<TAB>label" lines, sorted
//   --counts           per-label edge counts on stderr
//   --stats-json FILE  run statistics as JSON
//   --budget-mb N      device memory budget
//   --device N         HIP device ordinal
//   --verbose          per-round progress on stderr
//   --partition-files DIR
//                      after the run, DIR/Partition<k> for each of the
//                      reference's vertex intervals (initializePartitions,
//                      src/main.cpp:141-175), in the record layout of its
//                      compressed partition store (storeElements_c /
//                      elements2Edges, src/main.cpp:510-600): per source vertex
//                      in the interval, per label, the destinations ascending,
//                      each as 5 bytes (u32 dst little-endian, u8 label id). At
//                      the fixpoint the delta and tmp regions are empty, so a
//                      file holds the interval's closure edges.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "formats.h"
#include "graspan_g.h"

namespace {

int die(gg_ctx* c, const char* what, int status) {
  std::fprintf(stderr, "comp: %s: %s\n", what, c ? gg_last_error(c) : "");
  if (c) gg_ctx_destroy(c);
  return status == GG_E_FORMAT || status == GG_E_IO || status == GG_E_ARG ? 2 : 1;
}

void usage() {
  std::fprintf(stderr,
               "usage: comp <graph_file> <grammar_file> [filter_flag] [--dump FILE] [--dump-text FILE]\n"
               "            [--counts] [--stats-json FILE] [--budget-mb N] [--device N] [--verbose]\n"
               "            [--partition-files DIR]\n");
}

// Writes dir/Partition<k> for the reference's intervals from the sorted
// (src, dst, label) triples t (see the option's description above).
int write_partition_files(gg_ctx* c, const char* dir, const std::vector<uint32_t>& t, uint64_t cnt) {
  int np = 0;
  int s = gg_reference_partitions(c, nullptr, nullptr, 0, &np);
  if (s != GG_OK) return s;
  std::vector<uint32_t> first(static_cast<size_t>(np)), last(static_cast<size_t>(np));
  if ((s = gg_reference_partitions(c, first.data(), last.data(), np, &np)) != GG_OK) return s;
  uint64_t i = 0;
  std::vector<uint64_t> order;
  std::vector<unsigned char> rec;
  for (int k = 0; k < np; ++k) {
    rec.clear();
    while (i < cnt && t[3 * i] < first[size_t(k)]) ++i;
    // one source vertex at a time: its edges by (label, dst); the triples
    // are sorted by (src, dst, label), so a stable sort by label is enough
    while (i < cnt && t[3 * i] <= last[size_t(k)]) {
      uint64_t j = i;
      while (j < cnt && t[3 * j] == t[3 * i]) ++j;
      order.resize(j - i);
      for (uint64_t q = i; q < j; ++q) order[q - i] = q;
      std::stable_sort(order.begin(), order.end(),
                       [&](uint64_t a, uint64_t b) { return t[3 * a + 2] < t[3 * b + 2]; });
      for (uint64_t q : order) {
        const uint32_t dst = t[3 * q + 1];
        const unsigned char b[5] = {static_cast<unsigned char>(dst), static_cast<unsigned char>(dst >> 8),
                                    static_cast<unsigned char>(dst >> 16), static_cast<unsigned char>(dst >> 24),
                                    static_cast<unsigned char>(t[3 * q + 2])};
        rec.insert(rec.end(), b, b + 5);
      }
      i = j;
    }
    const std::string path = std::string(dir) + "/Partition" + std::to_string(k);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return GG_E_IO;
    const bool ok = rec.empty() || std::fwrite(rec.data(), 1, rec.size(), f) == rec.size();
    if (std::fclose(f) != 0 || !ok) return GG_E_IO;
  }
  return GG_OK;
}

}  // namespace

int main(int argc, char** argv) {
  const clock_t start = clock();
  std::vector<const char*> pos;
  const char *dump = nullptr, *dump_text = nullptr, *stats_json = nullptr, *part_dir = nullptr;
  bool counts = false;
  gg_options opts{-1, 0, 0};
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto need = [&](const char* name) -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "comp: %s needs a value\n", name);
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--dump")
      dump = need("--dump");
    else if (a == "--dump-text")
      dump_text = need("--dump-text");
    else if (a == "--partition-files")
      part_dir = need("--partition-files");
    else if (a == "--stats-json")
      stats_json = need("--stats-json");
    else if (a == "--counts")
      counts = true;
    else if (a == "--verbose")
      opts.verbose = 1;
    else if (a == "--budget-mb")
      opts.device_budget = std::strtoull(need("--budget-mb"), nullptr, 10) << 20;
    else if (a == "--device")
      opts.device = std::atoi(need("--device"));
    else if (a == "-h" || a == "--help") {
      usage();
      return 0;
    } else if (a.size() > 2 && a[0] == '-' && a[1] == '-') {
      std::fprintf(stderr, "comp: unknown option %s\n", a.c_str());
      usage();
      return 2;
    } else
      pos.push_back(argv[i]);
  }
  if (pos.size() < 2 || pos.size() > 3) {
    usage();
    return 2;
  }
  const char* graph_path = pos[0];
  const char* grammar_path = pos[1];
  const int k0 = gg::sniff_file_kind(graph_path), k1 = gg::sniff_file_kind(grammar_path);
  if (k0 == 1 && k1 == 2) {
    std::fprintf(stderr, "comp: note: %s looks like the grammar and %s like the graph; using them that way\n",
                 graph_path, grammar_path);
    std::swap(graph_path, grammar_path);
  }
  int filter = 0;
  if (pos.size() == 3) {
    char* end = nullptr;
    long f = std::strtol(pos[2], &end, 10);
    if (!end || *end) {
      std::fprintf(stderr, "comp: filter_flag must be an integer, got '%s'\n", pos[2]);
      return 2;
    }
    filter = int(f);
  }

  gg_ctx* c = nullptr;
  if (gg_ctx_create(&opts, &c) != GG_OK) return 1;
  std::printf("===== PREPROCESSING INFO =====\n");
  int ndev = 0, sel = 0;
  char name[256] = {0};
  uint64_t tot = 0, fr = 0;
  int s = gg_device_info(c, &ndev, &sel, name, sizeof name, &tot, &fr);
  std::printf("NUM COMPUTE-CAPABLE DEVICES: %d\n", ndev);
  if (s != GG_OK) {
    std::fprintf(stderr, "There is no device supporting HIP.\n");
    return die(c, "device", s);
  }
  std::printf("SELECTED DEVICE: %d %s\n", sel, name);
  std::printf("GPU's total memory: %llu MB, free memory: %llu MB\n", (unsigned long long)(tot >> 20),
              (unsigned long long)(fr >> 20));
  if ((s = gg_grammar_load_file(c, grammar_path)) != GG_OK) return die(c, "grammar", s);
  int nr = 0, nl = 0;
  gg_grammar_info(c, &nr, &nl);
  std::printf("NUM RULES: %d\nNUM LABELS: %d\n", nr, nl);
  if ((s = gg_graph_load_file(c, graph_path)) != GG_OK) return die(c, "graph", s);
  uint32_t nv = 0;
  uint64_t ne = 0;
  gg_graph_info(c, &nv, &ne);
  std::printf("NUM VERTICES: %u\nINITIAL NUM EDGES: %llu\n", nv, (unsigned long long)ne);
  const uint64_t heap = (opts.device_budget ? opts.device_budget : fr) / 4;
  std::printf("HEAP SIZE: %llu (in 32-bit words)\n", (unsigned long long)heap);
  int nparts = 1;
  if (filter != 0) {
    // the reference's schedule: its own vertex intervals
    if ((s = gg_reference_partitions(c, nullptr, nullptr, 0, &nparts)) != GG_OK) return die(c, "partitions", s);
  }
  std::printf("INITIAL NUM PARTITIONS: %d\n", nparts);
  std::fflush(stdout);

  gg_stats st{};
  if ((s = gg_closure_run(c, filter, &st)) != GG_OK) return die(c, "closure", s);
  uint64_t total = 0;
  std::vector<uint64_t> per(size_t(nl), 0);
  if ((s = gg_closure_counts(c, per.data(), &total)) != GG_OK) return die(c, "counts", s);

  std::printf("===== SPAGPU FINISHED =====\n");
  std::printf("NUM SUPERSTEPS: %u\n", filter != 0 ? st.supersteps : st.rounds);
  std::printf("FINAL NUM PARTITIONS: %u\n", st.partitions);
  std::printf("FINAL NUM EDGES: %llu\n", (unsigned long long)total);
  std::printf("TOTAL SPAGPU TIME: %ldmin.\n", long((clock() - start) / (CLOCKS_PER_SEC * 60)));
  std::fflush(stdout);

  if (counts)
    for (int l = 1; l <= nl; ++l)
      std::fprintf(stderr, "COUNT %s %llu\n", gg_label_name(c, l), (unsigned long long)per[size_t(l - 1)]);
  if (stats_json) {
    FILE* f = std::fopen(stats_json, "w");
    if (!f) return die(c, "stats-json", GG_E_IO);
    std::fprintf(f,
                 "{\"rounds\": %u, \"retries\": %u, \"input_unique\": %llu, \"total_edges\": %llu, "
                 "\"fixpoint_ms\": %.3f, \"expand_ms\": %.3f, \"compact_ms\": %.3f, \"merge_ms\": %.3f, "
                 "\"other_ms\": %.3f, \"probe_reads\": %llu, \"candidates\": %llu, \"old_reads\": %llu, "
                 "\"new_edges\": %llu, \"device_bytes_peak\": %llu}\n",
                 st.rounds, st.retries, (unsigned long long)st.input_unique, (unsigned long long)st.total_edges,
                 st.fixpoint_ms, st.expand_ms, st.compact_ms, st.merge_ms, st.other_ms,
                 (unsigned long long)st.probe_reads, (unsigned long long)st.candidates,
                 (unsigned long long)st.old_reads, (unsigned long long)st.new_edges,
                 (unsigned long long)st.device_bytes_peak);
    std::fclose(f);
  }
  if (dump || dump_text || part_dir) {
    std::vector<uint32_t> buf(size_t(total) * 3);
    uint64_t cnt = 0;
    if ((s = gg_closure_export(c, buf.data(), total, &cnt)) != GG_OK) return die(c, "export", s);
    if (part_dir && (s = write_partition_files(c, part_dir, buf, cnt)) != GG_OK) return die(c, "partition-files", s);
    if (dump) {
      FILE* f = std::fopen(dump, "wb");
      if (!f || std::fwrite(buf.data(), 4, buf.size(), f) != buf.size()) return die(c, "dump", GG_E_IO);
      std::fclose(f);
    }
    if (dump_text) {
      FILE* f = std::fopen(dump_text, "w");
      if (!f) return die(c, "dump-text", GG_E_IO);
      for (uint64_t i = 0; i < cnt; ++i)
        std::fprintf(f, "%u\t%u\t%s\n", buf[3 * i], buf[3 * i + 1], gg_label_name(c, int(buf[3 * i + 2])));
      std::fclose(f);
    }
  }
  gg_ctx_destroy(c);
  return 0;
}
