/*
 * graspan_g.h -- C ABI of the MI355X (gfx950) CFL-reachability closure engine.
 *
 * This is the library boundary that replaces the host API of the reference
 * (Graspan-G, src/kernel.h:115-129: transferRules, allocateElementPool,
 * transferElements, spagpu_s, spagpu_b, spagpu, getDegree, mergeAndDiff) and
 * the host parsers that drive it (src/main.cpp:60-139, 177-221). The
 * reference keeps one graph per process in __constant__/__device__ globals
 * (src/kernel.cu:3-20); here every graph lives in an explicit context.
 *
 * Conventions
 *   - Every function returns GG_OK (0) or a negative GG_E* status; none of
 *     them throws or exits. gg_last_error() holds a message for the last
 *     failure on that context.
 *   - Buffers are owned by the caller; the library copies what it keeps.
 *   - Label ids are 1..L in order of first appearance in the grammar file
 *     (src/main.cpp:77-88); 0 means "absent" inside a rule triple.
 *   - A context is used by one host thread at a time. Independent contexts
 *     may run concurrently.
 *   - Functions that only parse or inspect host data (grammar, graph, label
 *     names) work without a GPU; the closure functions need one.
 */
#ifndef GRASPAN_G_H
#define GRASPAN_G_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GG_OK 0
#define GG_E_ARG -1       /* bad argument (null pointer, size out of range) */
#define GG_E_IO -2        /* file could not be opened or read */
#define GG_E_FORMAT -3    /* malformed grammar or graph file */
#define GG_E_STATE -4     /* call out of order (e.g. run before load) */
#define GG_E_DEVICE -5    /* HIP runtime error or no usable GPU */
#define GG_E_NOMEM -6     /* device or host allocation failed / over budget */
#define GG_E_UNSUPPORTED -7

typedef struct gg_ctx gg_ctx;

typedef struct gg_options {
  int device;               /* HIP device ordinal; -1 = the device with most CUs
                               (src/main.cpp:30-51) */
  uint64_t device_budget;   /* bytes of device memory the context may hold;
                               0 = no limit beyond the device's free memory */
  int verbose;              /* 1 = progress lines on stderr */
} gg_options;

/* Statistics of one closure run. Times are wall-clock milliseconds measured
 * with HIP events on the context's stream. */
typedef struct gg_stats {
  uint32_t rounds;          /* semi-naive rounds until delta was empty */
  uint32_t retries;         /* rounds re-run after growing the staging area */
  uint64_t input_unique;    /* |dedup(input) U eps loops| */
  uint64_t total_edges;     /* |closure| (the reference's FINAL NUM EDGES) */
  double fixpoint_ms;       /* first round launch .. termination */
  double expand_ms;         /* sum of the join/dedup/difference kernel */
  double compact_ms;        /* sum of the delta compaction kernels */
  double merge_ms;          /* sum of the old U delta merge kernels */
  double other_ms;          /* activation, scans, bitmaps */
  uint32_t expand_launches; /* launches of the join/dedup/difference kernel */
  /* algorithmic counters summed over rounds (same definitions as the
   * oracle): probe reads, candidates, old-row reads, new edges */
  uint64_t probe_reads;
  uint64_t candidates;
  uint64_t old_reads;
  uint64_t new_edges;
  uint64_t device_bytes_peak; /* high-water mark of the context's device arena */
  uint32_t supersteps;      /* partitioned schedule: supersteps after superstep 0
                               (the reference's NUM SUPERSTEPS); 0 otherwise */
  uint32_t partitions;      /* partitioned schedule: vertex intervals; 1 otherwise */
  uint32_t pairs_run;       /* out-of-core: partition pairs closed */
  uint32_t pairs_skipped;   /* out-of-core: pairs with no crossing edges or nothing new */
  uint64_t spill_bytes;     /* out-of-core: bytes of evicted partitions moved between HBM and host DRAM
                               (out and back in; the input's first upload is not counted) */
  uint32_t pair_reruns;     /* passes of the old x delta pair-list kernel re-run after growing its arenas */
  uint32_t uploads_ahead;   /* out-of-core: partitions uploaded during the step before their first use */
  uint32_t spills_ahead;    /* out-of-core: partitions copied out during a step to make room for the next */
  uint32_t pad0;
  uint64_t light_rows;      /* rows expanded one per lane (rounds with few candidates per row) */
  uint64_t group_rows;      /* rows expanded four per wave in a shared guessed window */
  uint64_t group_deferred;  /* of those, rows with a candidate outside the window (expanded again alone) */
} gg_stats;

/* ---- context ------------------------------------------------------------ */

/* opts may be NULL for defaults. Does not touch the GPU. */
int gg_ctx_create(const gg_options *opts, gg_ctx **out);
int gg_ctx_destroy(gg_ctx *ctx);
const char *gg_last_error(const gg_ctx *ctx);
const char *gg_version(void);

/* ---- grammar ------------------------------------------------------------ */

/* Parses a grammar file (src/main.cpp:60-97): '#' lines skipped, first data
 * line = rule count, then one line per rule of 1-3 tab-separated label names.
 * Replaces readAndTransferRules + transferRules (src/kernel.cu:22-31). More
 * than 21 rules is accepted (the reference exits); more than 3 names on a
 * line, an empty line or a short file is GG_E_FORMAT. */
int gg_grammar_load_file(gg_ctx *ctx, const char *path);

/* Sets the grammar from rule triples (lhs, rhs1, rhs2), ids in 1..num_labels,
 * 0 = absent. Label names default to "1".."L". */
int gg_grammar_set(gg_ctx *ctx, int num_rules, const int32_t *rules, int num_labels);

int gg_grammar_info(const gg_ctx *ctx, int *num_rules, int *num_labels);
/* Writes the rule triples (3 * num_rules int32). */
int gg_grammar_rules(const gg_ctx *ctx, int32_t *rules_out);
/* The rule triples the least-fixpoint engine evaluates for a rule table (host
 * only; no GPU). No counterpart in the reference, which evaluates its rule
 * table as given (src/kernel.cu:403-464). A label Y whose rules are exactly
 * Y ::= eps and Y ::= Y P_k (or all Y ::= P_k Y) is P* for P = the union of
 * its P_k; a binary rule X ::= Y C whose lhs has no other rule is evaluated
 * as X ::= C plus X ::= P_k X (X ::= B Y as X ::= B plus X ::= X P_k). Every
 * label keeps its least fixpoint (Arden's lemma), with far fewer join
 * candidates when Y is dense and P sparse. gg_closure_run (filter 0), the
 * batched closure, the serving loop and the allocation-free closure use it;
 * the partitioned and out-of-core schedules run the table as given.
 * GG_STAR_REWRITE=0 in the environment turns it off. Writes up to cap
 * triples to out (3 * cap int32); *out_rules receives their number
 * (GG_E_ARG, nothing written, when cap is smaller). */
int gg_grammar_plan_fixpoint(int num_rules, const int32_t *rules, int num_labels, int32_t *out, int cap,
                             int *out_rules);

/* Name of label id (1..L); NULL if out of range. */
const char *gg_label_name(const gg_ctx *ctx, int label_id);
/* Id of a label name, 0 if unknown. */
int gg_label_id(const gg_ctx *ctx, const char *name);

/* ---- graph -------------------------------------------------------------- */

/* Parses a graph file (src/main.cpp:106-139, 177-221): '#' lines skipped,
 * header "numEdges<TAB>numVars", then numEdges lines "src<TAB>dst<TAB>label"
 * with labels by name. Needs a grammar first. Unlike the reference, edges do
 * not have to be grouped by src; ids >= numVars, unknown labels and short or
 * blank lines are GG_E_FORMAT. Replaces getGraphInfo + readEdges +
 * createGraph + transferElements (src/main.cpp:106-333, src/kernel.cu:42-52).
 * The edges are copied to device memory here if a GPU is present, else at
 * the first gg_closure_run. */
int gg_graph_load_file(gg_ctx *ctx, const char *path);

/* Sets the graph from host COO arrays (labels are ids 1..L). */
int gg_graph_set_coo(gg_ctx *ctx, uint32_t num_vertices, uint64_t num_edges, const uint32_t *src,
                     const uint32_t *dst, const uint8_t *label);

int gg_graph_info(const gg_ctx *ctx, uint32_t *num_vertices, uint64_t *num_edges);
/* Copies the parsed input edges (host copy) into caller buffers of
 * num_edges entries each. */
int gg_graph_edges(const gg_ctx *ctx, uint32_t *src, uint32_t *dst, uint8_t *label);

/* ---- closure ------------------------------------------------------------ */

/* Runs the fixpoint on the GPU. filter_flag follows the reference's third
 * argument (src/main.cpp:890, src/kernel.cu:490-780). 0 computes the least
 * fixpoint with the device-resident semi-naive engine (its result does not
 * depend on the schedule). A non-zero value skips the unary rules and the
 * old x delta joins, which makes the result depend on the reference's
 * partitions and sweep order, so it runs gg_closure_run_partitioned. stats
 * may be NULL. The input edges stay resident on the device across runs. */
int gg_closure_run(gg_ctx *ctx, int filter_flag, gg_stats *stats);

/* Runs the reference's own schedule (initializePartitions, precomputation,
 * run_computation, spagpu_s/spagpu_b/spagpu, mergeAndDiff;
 * src/main.cpp:141-175, 764-906, src/kernel.cu:466-985) on the GPU, one
 * device operator per reference kernel launch, for any filter_flag. With
 * filter 0 it reaches the same least fixpoint as gg_closure_run, more slowly.
 * GG_E_UNSUPPORTED when the reference's small-graph split is degenerate (a
 * first vertex holding more than half of the elements, or no elements). */
int gg_closure_run_partitioned(gg_ctx *ctx, int filter_flag, gg_stats *stats);

/* Out-of-core least fixpoint (filter 0) for graphs whose closure does not fit
 * the device budget (gg_options.device_budget). The vertices are cut into
 * num_partitions equal contiguous intervals. The reference's partition-pair
 * schedule runs (superstep 0 per partition, then pairs i < j; the reference
 * keeps the partitions it is not using in Partition<k> files,
 * src/main.cpp:335-631, 764-858). Partitions are loaded into HBM when the
 * schedule needs them and stay there. When an allocation would pass the
 * budget (or hipMalloc fails), the partition that the current step does not
 * hold and whose next use in the sweep is farthest is spilled to pinned host DRAM with
 * hipMemcpyAsync on a copy stream. The finished closure is assembled in HBM
 * and must fit the budget. Pairs are chosen by their cross-partition edge counts,
 * recomputed on the device whenever a partition changes: a pair runs only
 * when edges cross between the two partitions and one of them has new edges
 * (the weights of the reference's DDM, src/ddm.cpp:27-35, which the reference
 * never calls). */
int gg_closure_run_out_of_core(gg_ctx *ctx, int num_partitions, gg_stats *stats);

/* The reference's vertex intervals for the loaded graph (host only; no GPU):
 * up to cap (first, last) pairs, *count receives their number. Follows
 * initializePartitions (src/main.cpp:141-175) over getGraphInfo's degree
 * pre-pass (src/main.cpp:106-139). */
int gg_reference_partitions(const gg_ctx *ctx, uint32_t *first, uint32_t *last, int cap, int *count);

/* Per-label edge counts of the last closure (num_labels uint64, index l-1 for
 * label l) and their sum (the reference's FINAL NUM EDGES, src/main.cpp:909-917,
 * widened to 64 bits). Either pointer may be NULL. */
int gg_closure_counts(gg_ctx *ctx, uint64_t *per_label, uint64_t *total);

/* Writes the closure as uint32 (src, dst, label) triples sorted
 * lexicographically into buf (3 * cap words). *count receives the number of
 * triples; if cap is smaller nothing is written and GG_E_ARG is returned. */
int gg_closure_export(gg_ctx *ctx, uint32_t *buf, uint64_t cap, uint64_t *count);

/* Per-vertex out-degree of the closure over all labels (num_vertices
 * uint64), the reference's degree[old] column (src/kernel.cu:818-839). */
int gg_closure_degrees(gg_ctx *ctx, uint64_t *deg);

/* Order-independent checksum of the closure: the sum modulo 2^64 of
 * gg_triple_hash(src, dst, label) over every closure edge. It lets two
 * closures too large to export be compared. No counterpart in the
 * reference, which reports only the edge count. */
int gg_closure_checksum(gg_ctx *ctx, uint64_t *sum);

/* The per-edge hash summed by gg_closure_checksum (a splitmix64 finaliser of a
 * linear mix of the three ids). */
uint64_t gg_triple_hash(uint32_t src, uint32_t dst, uint32_t label);

/* Releases the device state of the last closure (the input stays). */
int gg_closure_release(gg_ctx *ctx);

/* ---- allocation-free closure -------------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. It computes what gg_closure_run(ctx, 0, ...) computes
 * (the least fixpoint, filter 0) on the current device and the given stream
 * (NULL: the null stream), with no hipMalloc, hipHostMalloc, hipFree or event:
 * every working buffer comes from one fixed 256 MiB device region that the
 * library holds on each GPU it runs on until the process ends. The caller
 * passes the input edges in device memory (num_edges each; labels 1..L) and
 * receives the closure in its own device buffers as a CSR: row_ptr holds
 * num_labels * num_vertices + 1 uint64 offsets, row (label - 1) * num_vertices
 * + src lists that row's dst ids in col, sorted ascending. *total (and, if
 * not NULL, checksum as gg_closure_checksum and per_label counts, host
 * memory) are written; col_cap smaller than the closure returns GG_E_ARG with
 * *total set. A graph or closure that does not fit the region returns
 * GG_E_NOMEM; an edge out of range returns GG_E_ARG. err (err_cap bytes, may
 * be NULL) receives the message. Every call on a GPU shares that GPU's
 * region: calls that overlap on one GPU race, so callers must serialise
 * them per device. */
int gg_closure_run_noalloc(int num_rules, const int32_t *rules, int num_labels, uint32_t num_vertices,
                           uint64_t num_edges, const uint32_t *src, const uint32_t *dst, const uint8_t *label,
                           void *stream, uint64_t *row_ptr, uint32_t *col, uint64_t col_cap, uint64_t *total,
                           uint64_t *checksum, uint64_t *per_label, char *err, size_t err_cap);

/* ---- batched closure ---------------------------------------------------
 * No counterpart in the reference, which closes one graph per process. Many
 * independent graphs (each with its own grammar) are closed together: the
 * graphs get disjoint vertex ranges and each distinct grammar its own label
 * range, eps rules become explicit self-loops, and the disjoint union is
 * closed by one launch sequence per round (closure of a disjoint union is the
 * disjoint union of the closures). */

typedef struct gg_batch gg_batch;

/* The batch uses ctx's device and budget. */
int gg_batch_create(gg_ctx *ctx, gg_batch **out);
int gg_batch_destroy(gg_batch *batch);

/* Adds one graph: rule triples with ids 1..num_labels, edges with label ids
 * 1..num_labels, vertices 0..num_vertices-1. *graph_id receives its index. */
int gg_batch_add(gg_batch *batch, int num_rules, const int32_t *rules, int num_labels, uint32_t num_vertices,
                 uint64_t num_edges, const uint32_t *src, const uint32_t *dst, const uint8_t *label, int *graph_id);

/* Closes every added graph. Graphs are ordered by grammar and cut into
 * launch groups of at most max_vertices vertices in total; 0 plans each
 * group from device memory instead (the context's budget, else the device's
 * memory): a group may use a tuned share of it (group_mem_share in
 * tuned_gfx950.cfg beside the library, 0.4), at the peak bytes per vertex measured
 * on the previous group of the same grammar (2^21 vertices for a grammar
 * not yet seen). A single graph above the cap runs alone. A group that still
 * runs out of device memory is cut in half and run again. The 2^32 row limit
 * and the engine's rule limits always cut groups. keep_closures keeps each
 * graph's triples on the host for gg_batch_export. stats (may be NULL) sums
 * the launches. */
int gg_batch_run(gg_batch *batch, uint64_t max_vertices, int keep_closures, gg_stats *stats);

/* Number of graphs and of launch groups the last run used. */
int gg_batch_info(const gg_batch *batch, int *num_graphs, int *num_groups);

/* Results of one graph, in its own numbering: per-label counts (num_labels
 * entries), total, checksum (as gg_closure_checksum). Pointers may be NULL. */
int gg_batch_result(const gg_batch *batch, int graph_id, uint64_t *per_label, uint64_t *total, uint64_t *checksum);

/* Sorted (src, dst, label) triples of one graph (needs keep_closures). */
int gg_batch_export(const gg_batch *batch, int graph_id, uint32_t *buf, uint64_t cap, uint64_t *count);

/* ---- serving loop ---------------------------------------------------------
 * No counterpart in the reference. A request is one graph with its grammar.
 * Continuous batching at the granularity of fixpoint rounds: each worker slot
 * (one context and HIP stream) runs a session, a semi-naive fixpoint that
 * requests join at round boundaries. Before every round the slot admits
 * queued requests in arrival order while their estimated footprint fits (the
 * slot's share, budget / slots, for its session, and the budget for
 * everything in flight); an admitted request's edges enter the running
 * fixpoint in that round. After every round, each request whose graph got no
 * new edges has reached its fixpoint and its result is published at once,
 * while the others go on. A session whose next request does not fit (or that
 * holds more than half its share) admits no more and ends with its last
 * request; the slot then starts another. The slots' sessions overlap on the
 * device. All slots carve their device memory from one region of
 * budget_bytes, allocated when the server starts. A session that still runs
 * out of memory puts its unfinished requests back at the front of the queue
 * once, to run alone; a second failure is reported in the result. */

typedef struct gg_server gg_server;

typedef struct gg_server_options {
  int device;                  /* as gg_options.device */
  int slots;                   /* concurrent sessions (>= 1) */
  uint64_t budget_bytes;       /* device memory for all in-flight requests (allocated once) */
  uint64_t bytes_per_vertex;   /* footprint estimate of a request: this times
                                  its vertex count plus 64 bytes per input edge;
                                  a request's estimate stays charged until its
                                  session ends */
  uint32_t max_batch_requests; /* cap on requests admitted at one round boundary (0 = no cap) */
} gg_server_options;

typedef struct gg_server_result {
  int status;                  /* GG_OK or the error of the request's session */
  uint64_t total_edges;
  uint64_t checksum;           /* as gg_closure_checksum, in the request's numbering */
  double submit_ms;            /* times on the server's clock (ms since creation) */
  double start_ms;             /* it joined a session */
  double done_ms;              /* its result was published */
  uint32_t batch_size;         /* requests admitted at the same round boundary */
  int slot;
} gg_server_result;

int gg_server_create(const gg_server_options *opts, gg_server **out);
/* Queues one request (copied); never blocks on the GPU. Thread-safe. */
int gg_server_submit(gg_server *srv, int num_rules, const int32_t *rules, int num_labels, uint32_t num_vertices,
                     uint64_t num_edges, const uint32_t *src, const uint32_t *dst, const uint8_t *label,
                     int64_t *request_id);
/* Blocks until the request is done and copies its result; per_label (may be
 * NULL) receives num_labels counts. */
int gg_server_wait(gg_server *srv, int64_t request_id, gg_server_result *result, uint64_t *per_label);
/* Milliseconds since the server was created (the clock of the results). */
double gg_server_now_ms(const gg_server *srv);
/* Finishes queued work and stops the workers. */
int gg_server_destroy(gg_server *srv);

/* Device description used for the report lines of src/main.cpp:34,54,57.
 * name_cap bytes are written to name. */
int gg_device_info(gg_ctx *ctx, int *device_count, int *selected, char *name, size_t name_cap,
                   uint64_t *total_bytes, uint64_t *free_bytes);

#ifdef __cplusplus
}
#endif

#endif /* GRASPAN_G_H */
