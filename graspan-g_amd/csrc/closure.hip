// Device engine of the CFL-reachability closure for MI355X (gfx950, CDNA4).
//
// Replaces the reference's device code and its launch drivers
// (src/kernel.cu:85-985: compute00/compute10/compute11 with
// apply/decode/unionAll/union2, merge/diff/mergeAndDiff, computeDegree, and
// spagpu_s/spagpu_b/spagpu) with a device-resident semi-naive fixpoint over
// sorted uint32 CSR rows.
//
// Data layout in HBM (one context = one graph):
//   row id r = l * n + u for label l (0-based) and vertex u; L*n < 2^32.
//   old, delta : rowptr uint64[L*n + 1] + col uint32[], rows sorted/unique.
//   staging    : uint32[] bump-allocated output of the expand kernel.
//   seg_off/seg_cnt : per-row (offset, count) of that output.
//
// One round (rule semantics of src/kernel.cu:403-464; end-of-step update of
// src/kernel.cu:943-985):
//   cand(A,u) = U_{A::=B} delta_B(u)
//             U_{A::=B C} ( U_{v in delta_B(u)} old_C(v) U delta_C(v) )
//             U           ( U_{v in old_B(u)}   delta_C(v) )
//   delta'(A,u) = cand(A,u) - old_A(u) - delta_A(u);  old' = old U delta.
//   Kernels: k_activate (which rows can receive candidates) -> k_expand (join
//   + dedup + difference, fused, output sorted) -> one host read of the round
//   counters -> k_scan_* + k_gather (compact delta') -> k_merge (old U delta,
//   or pending U delta while delta is small next to old: the joins then read
//   old U pending).
//
// k_expand is the hot kernel. One wave64 owns one (A,u) row: it seeds an LDS
// bitmap (`seen`) over the row's candidate window with old_A(u) U delta_A(u),
// ORs every candidate into a second bitmap (`fresh`) with no-return LDS
// atomics, and writes fresh & ~seen out: the new edges, already in
// ascending order. Candidates never go to HBM. The candidate lists of a
// batch are flattened into one coalesced element stream; an element's list
// is found by an mbcnt rank over a per-tile bitmap of list starts
// (stream_tile). Windows wider than kChunkBits are processed in chunks
// (count pass, then write pass).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <atomic>
#include <thread>
#include <vector>

#include "engine.h"
#include "hash.h"

namespace gg {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kChunkBits = 16384;  // LDS window per wave: 2 KiB seen + 2 KiB new
constexpr uint32_t kChunkWords = kChunkBits / 32;
constexpr int kUnroll = 8;              // loads in flight per lane in the candidate stream
constexpr uint32_t kXcds = 8;           // MI355X: 8 XCDs, each with its own L2
constexpr int kExpandWaves = 7;         // k_expand occupancy target (waves per SIMD)
constexpr uint64_t kDeferRatio = 128;   // pending edges kept apart while under 1/128 of old
constexpr uint32_t kMaxCrossParts = 1024;  // partitions of the out-of-core schedule
constexpr uint32_t kStageBlock = 4096;  // staging words reserved per cursor atomic
constexpr uint32_t kPairBlock = 1024;   // k_pairs: pair-list words a wave reserves per arena atomic
constexpr uint32_t kPairRowBuf = 4;     // k_pairs: pairs per row and slot held in LDS before they are written
constexpr uint32_t kGuessMin = 4096;     // smallest guessed window (ids)
constexpr uint32_t kMergeBits = 16384;   // k_merge: rows spanning fewer ids are merged in an LDS bitmap
constexpr uint32_t kShortRow = 1024;     // k_merge: unchanged rows up to this length move in the block copy
constexpr uint32_t kLaneGather = 32;     // k_gather: rows of up to this many ids are packed by one lane each
constexpr uint32_t kMergeSlot = 2048;    // k_merge_rows: rows spanning fewer ids are merged 8 per wave
constexpr uint32_t kLaneMerge = 8;       // k_merge_rows: rows of up to this many elements are merged by one lane each
constexpr int kMaxRules = kMaxRulesPerKind;
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

struct DevGrammar {
  int nu = 0, nb = 0, ne = 0;
  uint8_t u_lhs[kMaxRules], u_rhs[kMaxRules];
  uint8_t b_lhs[kMaxRules], b_rhs1[kMaxRules], b_rhs2[kMaxRules];
  uint8_t e_lhs[kMaxRules];
  // old_B x delta_C goes through per-round pair lists (the keys v of old_B(u)
  // whose delta_C(v) is not empty). b_slot[i] = pair-list slot of binary rule
  // i (0xFF when C is never derived, so delta_C is empty after the first
  // round); slot_c/slot_b = C and B label of each slot.
  int nslot = 0;
  uint8_t b_slot[kMaxRules];
  uint8_t slot_c[kMaxRules], slot_b[kMaxRules];
};

// Statistics counters are added to by every block; updates of one address
// from many blocks serialise, so blocks add to one of kSpread slots (by block
// index), each on its own 128-byte line, and the host sums the slots.
constexpr int kSpread = 64;
struct alignas(128) CtrSlot {
  unsigned long long cand;     // K: candidates streamed
  unsigned long long pkeys;    // key elements scanned (binary rules)
  unsigned long long oreads;   // O: old/delta row elements read for the difference
  unsigned long long newe;     // N: new edges
  unsigned long long checksum; // closure checksum accumulator
  unsigned long long xs[4];    // built with GG_EXPAND_STATS: k_expand cycles in seeding, key lists, list streams, whole rows
  unsigned int misses;         // rows whose guessed window missed a candidate
  unsigned int pad;
  unsigned long long pairs;    // k_pairs: words reserved in this slot's arena of the pair lists
};
struct alignas(128) WorkIdx {
  unsigned int v;  // one XCD's work index into the active list, on a line of its own
};

// Round counters, one device copy per context (zeroed by a fill kernel).
// The fields from cand to xs are the sums of the slots, filled in on the host
// when the counters are read back.
struct Counters {
  unsigned long long cursor;   // staging bump pointer (new edges + unused tails of per-wave blocks)
  unsigned long long cand;
  unsigned long long pkeys;
  unsigned long long oreads;
  unsigned long long newe;
  unsigned long long checksum;
  unsigned int active;         // rows in the active list
  unsigned int active_base;    // MODE 1 expand: its rows start here in the active list
  unsigned int deferred;       // rows the light pass left to the wave kernel (k_expand_light)
  unsigned int work;           // work-stealing index
  unsigned int overflow;       // staging too small
  unsigned int misses;
  unsigned int label_flags[8]; // bit l: delta of label l is not empty
  // k_region_span: per destination region d of a partition operator, the
  // local vertices with a non-empty row span [np - span_lo_c[d], span_hi[d])
  // (both 0: no row is non-empty)
  unsigned int span_lo_c[2];
  unsigned int span_hi[2];
  unsigned long long xs[4];
  WorkIdx work_xcd[8];         // per-XCD work indices into the active list
  CtrSlot slot[kSpread];
};

__device__ inline CtrSlot& ctr_slot(Counters* c) { return c->slot[blockIdx.x % kSpread]; }

struct DCsr {
  const unsigned long long* ptr;
  const uint32_t* col;
};

// ---------------------------------------------------------------------------
// wave helpers

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Values the whole wave holds alike, moved to scalar registers (the first
// active lane's copy), which frees vector registers for occupancy.
// (The builtin works on int: each half goes back through uint32_t before it
// is widened, so no sign bit spreads into the high word.)
__device__ inline uint32_t uni_u32(uint32_t x) { return uint32_t(__builtin_amdgcn_readfirstlane(int(x))); }
__device__ inline unsigned long long uni_u64(unsigned long long x) {
  const uint32_t lo = uni_u32(uint32_t(x)), hi = uni_u32(uint32_t(x >> 32));
  return (unsigned long long)lo | ((unsigned long long)hi << 32);
}

// Wave64 scans and reductions on DPP row operations (GCN/CDNA: row_shr
// within 16-lane rows, then row_bcast:15 and row_bcast:31 across rows):
// register moves, no LDS round trip. A lane with no DPP source keeps the
// identity `id`.
template <class Op>
__device__ inline uint32_t dpp_scan_u32(uint32_t x, uint32_t id, Op op) {
  x = op(x, uint32_t(__builtin_amdgcn_update_dpp(int(id), int(x), 0x111, 0xf, 0xf, false)));  // row_shr:1
  x = op(x, uint32_t(__builtin_amdgcn_update_dpp(int(id), int(x), 0x112, 0xf, 0xf, false)));  // row_shr:2
  x = op(x, uint32_t(__builtin_amdgcn_update_dpp(int(id), int(x), 0x114, 0xf, 0xf, false)));  // row_shr:4
  x = op(x, uint32_t(__builtin_amdgcn_update_dpp(int(id), int(x), 0x118, 0xf, 0xf, false)));  // row_shr:8
  x = op(x, uint32_t(__builtin_amdgcn_update_dpp(int(id), int(x), 0x142, 0xa, 0xf, false)));  // row_bcast:15
  x = op(x, uint32_t(__builtin_amdgcn_update_dpp(int(id), int(x), 0x143, 0xc, 0xf, false)));  // row_bcast:31
  return x;
}

template <int CTRL, int ROW_MASK>
__device__ inline unsigned long long dpp_u64(unsigned long long x) {
  const uint32_t lo = uint32_t(__builtin_amdgcn_update_dpp(0, int(uint32_t(x)), CTRL, ROW_MASK, 0xf, false));
  const uint32_t hi = uint32_t(__builtin_amdgcn_update_dpp(0, int(uint32_t(x >> 32)), CTRL, ROW_MASK, 0xf, false));
  return (unsigned long long)lo | ((unsigned long long)hi << 32);
}

// Inclusive prefix sum over the wave (lane i gets x_0 + ... + x_i).
__device__ inline unsigned long long wave_incl_scan_u64(unsigned long long x) {
  x += dpp_u64<0x111, 0xf>(x);
  x += dpp_u64<0x112, 0xf>(x);
  x += dpp_u64<0x114, 0xf>(x);
  x += dpp_u64<0x118, 0xf>(x);
  x += dpp_u64<0x142, 0xa>(x);
  x += dpp_u64<0x143, 0xc>(x);
  return x;
}

__device__ inline uint32_t wave_incl_scan(uint32_t x, int) {
  return dpp_scan_u32(x, 0u, [](uint32_t p, uint32_t q) { return p + q; });
}

__device__ inline uint32_t lane63_u32(uint32_t x) { return uint32_t(__builtin_amdgcn_readlane(int(x), 63)); }
// lane k's value (k uniform), in scalar registers
__device__ inline unsigned long long lane_u64(unsigned long long x, uint32_t k) {
  return (unsigned long long)uint32_t(__builtin_amdgcn_readlane(int(uint32_t(x)), int(k))) |
         ((unsigned long long)uint32_t(__builtin_amdgcn_readlane(int(uint32_t(x >> 32)), int(k))) << 32);
}
__device__ inline unsigned long long lane63_u64(unsigned long long x) {
  return (unsigned long long)lane63_u32(uint32_t(x)) | ((unsigned long long)lane63_u32(uint32_t(x >> 32)) << 32);
}

// Reductions: the last lane of the scan holds the result; it is read into a
// scalar register, so the result is uniform.
__device__ inline unsigned long long wave_sum_u64(unsigned long long x) { return lane63_u64(wave_incl_scan_u64(x)); }
__device__ inline uint32_t wave_sum_u32(uint32_t x) { return lane63_u32(wave_incl_scan(x, 0)); }
// The same for 64-bit values that are mostly small: when every lane's value
// is below 2^26 their sum fits 32 bits, and the 32-bit scan (one DPP add per
// step, against a carry chain of three instructions) gives it.
__device__ inline unsigned long long wave_sum_u64_small(unsigned long long x) {
  if (__ballot(x >= (1ull << 26)) == 0) return wave_sum_u32(uint32_t(x));
  return wave_sum_u64(x);
}

__device__ inline uint32_t wave_min_u32(uint32_t x) {
  return lane63_u32(dpp_scan_u32(x, 0xFFFFFFFFu, [](uint32_t p, uint32_t q) { return p < q ? p : q; }));
}

__device__ inline uint32_t wave_max_u32(uint32_t x) {
  return lane63_u32(dpp_scan_u32(x, 0u, [](uint32_t p, uint32_t q) { return p > q ? p : q; }));
}

__device__ inline bool label_nonempty(const Counters* c, uint32_t l) { return (c->label_flags[l >> 5] >> (l & 31)) & 1u; }

// ---------------------------------------------------------------------------
// small utility kernels

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t count) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, uint64_t count) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// out = a + b elementwise (rowptr of old U delta = rowptr(old) + rowptr(delta)).
__global__ void k_add_ptr(unsigned long long* out, const unsigned long long* a, const unsigned long long* b, uint64_t count) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = a[i] + b[i];
}

// Exclusive scan of uint32 counts into uint64 offsets, out[count] = total.
// Three phases: per-tile sums, scan of tile sums (one block), per-tile scan.
__device__ inline unsigned long long block_excl_scan_u64(unsigned long long x, unsigned long long* s_tmp,
                                                         unsigned long long& total) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const unsigned long long incl = wave_incl_scan_u64(x);
  if (lane == kWave - 1) s_tmp[w] = incl;
  __syncthreads();
  unsigned long long off = 0, tot = 0;
  for (int i = 0; i < kWavesPerBlock; ++i) {
    if (i < w) off += s_tmp[i];
    tot += s_tmp[i];
  }
  __syncthreads();
  total = tot;
  return off + incl - x;
}

__global__ __launch_bounds__(kBlock) void k_scan_tiles(const uint32_t* in, uint64_t count, unsigned long long* tile_sums) {
  __shared__ unsigned long long s_tmp[kWavesPerBlock];
  const uint64_t base = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * kScanItems;
  unsigned long long s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < count) s += in[base + i];
  unsigned long long tot;
  block_excl_scan_u64(s, s_tmp, tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_scan_sums(unsigned long long* tile_sums, uint64_t ntiles, unsigned long long* out_total) {
  __shared__ unsigned long long s_tmp[kWavesPerBlock];
  unsigned long long carry = 0;
  for (uint64_t b = 0; b < ntiles; b += kBlock) {
    uint64_t i = b + threadIdx.x;
    unsigned long long x = i < ntiles ? tile_sums[i] : 0;
    unsigned long long tot;
    unsigned long long ex = block_excl_scan_u64(x, s_tmp, tot);
    if (i < ntiles) tile_sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *out_total = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_final(const uint32_t* in, uint64_t count, const unsigned long long* tile_off,
                                                      unsigned long long* out) {
  __shared__ unsigned long long s_tmp[kWavesPerBlock];
  const uint64_t base = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * kScanItems;
  uint32_t v[kScanItems];
  unsigned long long s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = base + i < count ? in[base + i] : 0;
    s += v[i];
  }
  unsigned long long tot;
  unsigned long long ex = block_excl_scan_u64(s, s_tmp, tot) + tile_off[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < count) out[base + i] = ex;
    ex += v[i];
  }
}

// label_flags bit l = the delta of label l is not empty.
__global__ void k_label_flags(const unsigned long long* ptr, uint32_t n, uint32_t L, Counters* ctr) {
  for (uint32_t l = threadIdx.x; l < L; l += blockDim.x)
    if (ptr[uint64_t(l + 1) * n] > ptr[uint64_t(l) * n]) atomicOr(&ctr->label_flags[l >> 5], 1u << (l & 31));
}


// ---------------------------------------------------------------------------
// initial delta: raw rows from input COO + eps loops

__global__ void k_count_raw(const uint32_t* src, const uint8_t* lab, uint64_t m, uint32_t n, const DevGrammar g,
                            uint32_t* cnt) {
  const uint64_t tid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = tid; i < m; i += stride) atomicAdd(&cnt[uint32_t(lab[i] - 1) * n + src[i]], 1u);
  const uint64_t neps = uint64_t(g.ne) * n;
  for (uint64_t i = tid; i < neps; i += stride) {
    uint32_t e = uint32_t(i / n), v = uint32_t(i % n);
    atomicAdd(&cnt[uint32_t(g.e_lhs[e]) * n + v], 1u);
  }
}

__global__ void k_scatter_raw(const uint32_t* src, const uint32_t* dst, const uint8_t* lab, uint64_t m, uint32_t n,
                              const DevGrammar g, const unsigned long long* ptr, uint32_t* fill, uint32_t* col) {
  const uint64_t tid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = tid; i < m; i += stride) {
    uint32_t r = uint32_t(lab[i] - 1) * n + src[i];
    col[ptr[r] + atomicAdd(&fill[r], 1u)] = dst[i];
  }
  const uint64_t neps = uint64_t(g.ne) * n;
  for (uint64_t i = tid; i < neps; i += stride) {
    uint32_t e = uint32_t(i / n), v = uint32_t(i % n);
    uint32_t r = uint32_t(g.e_lhs[e]) * n + v;
    col[ptr[r] + atomicAdd(&fill[r], 1u)] = v;
  }
}

// Histogram of the partitions the ids of a region point into (bounds holds
// the P+1 first vertices). Per-block LDS counters, flushed with one atomic
// each.
__global__ __launch_bounds__(kBlock) void k_cross(DCsr c, uint32_t rows, const uint32_t* bounds, uint32_t P,
                                                  unsigned long long* out) {
  __shared__ unsigned long long hist[kMaxCrossParts];
  for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const unsigned long long total = c.ptr[rows];
  for (unsigned long long i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < total;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t v = c.col[i];
    uint32_t lo = 0, hi = P;  // largest q with bounds[q] <= v
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (bounds[mid] <= v)
        lo = mid;
      else
        hi = mid;
    }
    atomicAdd(&hist[lo], 1ull);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < P; i += blockDim.x)
    if (hist[i]) atomicAdd(&out[i], hist[i]);
}

// The per-XCD work indices start again (a second expand launch over another
// row list).
__global__ void k_reset_work(Counters* ctr) {
  const int t = int(threadIdx.x);
  if (t < int(kXcds)) ctr->work_xcd[t].v = 0;
}

// Active list = every row 0..rows-1 (region operators).
// Before rows are appended to a round's active list for a second expand
// launch: that launch's rows start at the current end, and the work indices
// start again.
__global__ void k_mark_active_base(Counters* ctr) {
  const int t = int(threadIdx.x);
  if (t == 0) {
    ctr->active_base = ctr->active;
    ctr->work = 0;
  }
  if (t < int(kXcds)) ctr->work_xcd[t].v = 0;
}

// Active list of the initial dedup: rows with raw entries.
__global__ __launch_bounds__(kBlock) void k_activate_raw(const unsigned long long* ptr, uint32_t rows,
                                                         uint32_t* active, Counters* ctr) {
  // active rows gather per wave and take their place in the list with one
  // counter update per kBuf rows (updates of one address serialise)
  constexpr uint32_t kBuf = 512;
  __shared__ uint32_t s_buf[kWavesPerBlock][kBuf];
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t* buf = s_buf[threadIdx.x / kWave];
  uint32_t nbuf = 0;  // uniform
  auto flush = [&]() {
    wave_sync();
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(&ctr->active, nbuf);
    pos = uni_u32(pos);  // lane 0: every lane is active here
    for (uint32_t i = uint32_t(lane); i < nbuf; i += kWave) active[pos + i] = buf[i];
    nbuf = 0;
    wave_sync();
  };
  for (uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~uint32_t(kWave - 1); base < rows;
       base += gridDim.x * blockDim.x) {
    uint32_t r = base + lane;
    bool act = r < rows && ptr[r + 1] > ptr[r];
    unsigned long long m = __ballot(act);
    if (!m) continue;
    if (nbuf + kWave > kBuf) flush();
    if (act) buf[nbuf + __popcll(m & ((1ull << lane) - 1))] = r;
    nbuf += uint32_t(__popcll(m));
  }
  if (nbuf) flush();
}

// ---------------------------------------------------------------------------
// activation: which (A,u) rows can receive candidates this round, plus the
// key-scan part of the probe counter (same definition as the oracle).

struct ActArgs {
  DCsr old, dl;
  const unsigned long long* pend_ptr;  // may be null (no pending edges)
  const unsigned long long* b2rng;  // pair-list bounds, two per row; may be null (no pair lists this round)
  uint32_t n;
  uint32_t nlhs;
  uint8_t lhs[256];
  uint16_t u_li[kMaxRules], b_li[kMaxRules];  // per rule: index of its lhs label in lhs[] (0xFFFF: none)
  DevGrammar g;
  uint32_t* active;
  Counters* ctr;
};

// One thread per vertex u, looping over the lhs labels, so the active list
// comes out vertex-major (rows of one vertex range stay together, which keeps
// the rows they join against in one XCD's L2 when the expand kernel splits
// the list by XCD).
__global__ __launch_bounds__(kBlock) void k_activate(const ActArgs a, uint32_t sub) {
  // A block takes sub-chunks of 256 consecutive vertices, U of them at a
  // time (U <= sub, U*4*nlhs <= 1024 masks in LDS). Each wave records one
  // ballot per sub-chunk and lhs label; the block then reserves the rows of
  // all U sub-chunks with a single atomic (the counter is one address for
  // the whole device: a reservation per 256 vertices serialised on it, 2.4e5
  // of them per round at 6e7 vertices) and every wave writes its rows
  // sub-chunk by sub-chunk, label by label.
  constexpr uint32_t kMasks = 1024;
  __shared__ unsigned long long s_m[kMasks];  // [U][wave][nlhs]
  __shared__ uint32_t s_cnt[kWavesPerBlock];
  __shared__ uint32_t s_base;
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  const uint32_t nl = a.nlhs;
  const uint32_t U = max(1u, min(sub, kMasks / (uint32_t(kWavesPerBlock) * max(nl, 1u))));
  auto mask = [&](uint32_t k, uint32_t li) -> unsigned long long& {
    return s_m[(k * kWavesPerBlock + uint32_t(wv)) * nl + li];
  };
  const uint64_t span = uint64_t(blockDim.x) * U;
  for (uint64_t bbase = uint64_t(blockIdx.x) * span; bbase < a.n; bbase += uint64_t(gridDim.x) * span) {
    uint32_t wcnt = 0;
    for (uint32_t k = 0; k < U; ++k) {
      const uint64_t u64 = bbase + uint64_t(k) * blockDim.x + uint32_t(wv) * kWave + uint32_t(lane);
      const bool valid = u64 < a.n;
      const uint32_t u = valid ? uint32_t(u64) : 0u;
      for (uint32_t li = lane; li < nl; li += kWave) mask(k, li) = 0;
      wave_sync();
      // Rules go in chunks of kRuleChunk: every row-bound load of a chunk is
      // issued before any is used, and each rule's ballot is ORed into the
      // mask of its lhs label. A rule reads delta_B(u) only when delta_B is
      // not empty, and its pair list only when delta_C is not empty (each is
      // empty otherwise). 8 rules per chunk: one round trip per vertex for
      // grammars of up to 8 rules of a kind.
      constexpr int kRuleChunk = 8;
      for (int i0 = 0; i0 < a.g.nu; i0 += kRuleChunk) {
        unsigned long long lo[kRuleChunk], hi[kRuleChunk];
#pragma unroll
        for (int q = 0; q < kRuleChunk; ++q) {
          const int i = i0 + q;
          lo[q] = hi[q] = 0;
          if (i < a.g.nu && a.u_li[i] != 0xFFFFu && valid && label_nonempty(a.ctr, a.g.u_rhs[i])) {
            const uint32_t rb = uint32_t(a.g.u_rhs[i]) * a.n + u;
            lo[q] = a.dl.ptr[rb];
            hi[q] = a.dl.ptr[rb + 1];
          }
        }
#pragma unroll
        for (int q = 0; q < kRuleChunk; ++q) {
          const int i = i0 + q;
          if (i < a.g.nu && a.u_li[i] != 0xFFFFu) {
            const unsigned long long m = __ballot(hi[q] > lo[q]);
            if (lane == 0) mask(k, a.u_li[i]) |= m;
          }
        }
      }
      for (int i0 = 0; i0 < a.g.nb; i0 += kRuleChunk) {
        unsigned long long d0[kRuleChunk], d1[kRuleChunk], p0[kRuleChunk], p1[kRuleChunk];
#pragma unroll
        for (int q = 0; q < kRuleChunk; ++q) {
          const int i = i0 + q;
          d0[q] = d1[q] = p0[q] = p1[q] = 0;
          if (i < a.g.nb && a.b_li[i] != 0xFFFFu && valid) {
            const uint32_t rb = uint32_t(a.g.b_rhs1[i]) * a.n + u;
            if (label_nonempty(a.ctr, a.g.b_rhs1[i])) {
              d0[q] = a.dl.ptr[rb];
              d1[q] = a.dl.ptr[rb + 1];
            }
            const uint32_t slot = a.g.b_slot[i];
            if (slot != 0xFFu && a.b2rng && label_nonempty(a.ctr, a.g.b_rhs2[i])) {
              const uint64_t rs = uint64_t(slot) * a.n + u;
              p0[q] = a.b2rng[2 * rs];
              p1[q] = a.b2rng[2 * rs + 1];
            }
          }
        }
#pragma unroll
        for (int q = 0; q < kRuleChunk; ++q) {
          const int i = i0 + q;
          if (i < a.g.nb && a.b_li[i] != 0xFFFFu) {
            const unsigned long long m = __ballot(d1[q] > d0[q] || p1[q] > p0[q]);
            if (lane == 0) mask(k, a.b_li[i]) |= m;
          }
        }
      }
      wave_sync();
      for (uint32_t li = 0; li < nl; ++li) wcnt += uint32_t(__popcll(mask(k, li)));
    }
    if (lane == 0) s_cnt[wv] = wcnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
      for (int w = 0; w < kWavesPerBlock; ++w) tot += s_cnt[w];
      s_base = tot ? atomicAdd(&a.ctr->active, tot) : 0u;
    }
    __syncthreads();
    uint32_t pos = s_base;
    for (int w = 0; w < wv; ++w) pos += s_cnt[w];
    for (uint32_t k = 0; k < U; ++k) {
      const uint32_t u = uint32_t(bbase + uint64_t(k) * blockDim.x) + uint32_t(wv) * kWave + uint32_t(lane);
      for (uint32_t li = 0; li < nl; ++li) {
        const unsigned long long m = mask(k, li);
        if ((m >> lane) & 1ull) a.active[pos + __popcll(m & ((1ull << lane) - 1))] = uint32_t(a.lhs[li]) * a.n + u;
        pos += uint32_t(__popcll(m));
      }
    }
    __syncthreads();  // s_m, s_cnt and s_base are rewritten by the next batch
  }
}

// The round's key elements scanned (the probe counter's key part): per
// binary rule A ::= B C, every element of delta_B when delta_B is not empty,
// and every element of old_B and the pending B when delta_C is not empty
// (the rows the join reads keys from), i.e. whole-label sizes: one lane per
// rule, no per-vertex loads.
__global__ void k_probe_keys(const DevGrammar g, const unsigned long long* old_ptr, const unsigned long long* dl_ptr,
                             const unsigned long long* pend_ptr, uint32_t n, Counters* ctr) {
  const int lane = threadIdx.x & (kWave - 1);
  unsigned long long pk = 0;
  for (int i = lane; i < g.nb; i += kWave) {
    const uint64_t b0 = uint64_t(g.b_rhs1[i]) * n, b1 = b0 + n;
    const uint64_t c0 = uint64_t(g.b_rhs2[i]) * n, c1 = c0 + n;
    pk += dl_ptr[b1] - dl_ptr[b0];
    if (dl_ptr[c1] > dl_ptr[c0]) {
      pk += old_ptr[b1] - old_ptr[b0];
      if (pend_ptr) pk += pend_ptr[b1] - pend_ptr[b0];
    }
  }
  pk = wave_sum_u64(pk);
  if (lane == 0 && pk) atomicAdd(&ctr->slot[0].pkeys, pk);
}

// ---------------------------------------------------------------------------
// frontier pair lists for old_B x delta_C and the transpose of old_B

// Inclusive prefix of the lanes' list lengths and its total, for the
// flattened streams below. When no list reaches 2^26 elements the 64-lane sum
// fits 32 bits and the scan runs on 32-bit DPP steps (half the VALU work of
// the 64-bit scan); a wave with no list at all returns total 0 at once.
__device__ inline unsigned long long stream_prefix(uint32_t len, unsigned long long& total) {
  if (__ballot(len != 0) == 0) {
    total = 0;
    return 0;
  }
  if (__ballot(len >= (1u << 26)) == 0) {
    const uint32_t incl = wave_incl_scan(len, 0);
    total = lane63_u32(incl);
    return incl;
  }
  const unsigned long long incl = wave_incl_scan_u64(len);
  total = lane63_u64(incl);
  return incl;
}

// Flattened list streams. A wave's lanes each hold one list (lane k: len
// elements starting at `start` = the exclusive prefix of the lengths); all
// lists form one element stream, cut into tiles of U*kWave elements, and lane
// l takes elements t0 + q*kWave + l of a tile, so neighbouring lanes read
// neighbouring elements. An element belongs to the last non-empty list that
// starts at or before it. Per tile every list starting inside the tile sets
// its start's bit in an LDS bitmap `tmask` (one 64-bit word per q); the list
// of element t0 + q*kWave + l is then its rank among the non-empty lists:
// the bits of word q up to bit l (mbcnt) plus the lists that started before.
// The caller keeps per-rank tables (written with stream_rank) and calls
// stream_tile once per tile with every lane of the wave active;
// f(q, rank) runs on every lane, in order of q.
__device__ inline uint32_t stream_rank(uint32_t len) {
  const unsigned long long nonempty = __ballot(len != 0);
  return __builtin_amdgcn_mbcnt_hi(uint32_t(nonempty >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(nonempty), 0u));
}

// nq (uniform): the q slices of the tile that hold elements; the rest are
// skipped.
template <int U, class F>
__device__ inline void stream_tile(unsigned long long start, uint32_t len, unsigned long long t0,
                                   unsigned long long* tmask, int lane, F&& f, int nq = U) {
  const unsigned long long pos = start - t0;
  if (len && pos < uint64_t(U) * kWave)
    __hip_atomic_fetch_or(&tmask[pos / kWave], 1ull << (pos % kWave), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  // non-empty lists that start before the tile, less one: the rank of the
  // list holding element t0 - 1
  uint32_t c = uint32_t(__popcll(__ballot(len != 0 && start < t0))) - 1u;
  wave_sync();
#pragma unroll
  for (int q = 0; q < U; ++q) {
    if (q < nq) {
      const unsigned long long m = uni_u64(tmask[q]);
      const unsigned long long up = m >> 1;  // mbcnt counts bits below the lane: bits 1..l of m, plus bit 0
      f(q, __builtin_amdgcn_mbcnt_hi(uint32_t(up >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(up), c + uint32_t(m & 1))));
      c += uint32_t(__popcll(m));
    }
  }
  wave_sync();
  if (lane < U) tmask[lane] = 0;
  wave_sync();
}

// Walks the element positions of the lanes' lists [lb, lb+len) as one
// flattened stream (see stream_tile), U elements per lane per tile, and calls
// f(own, pos, ok) once per tile with U-entry arrays: ok[q] says that slice q
// of this lane holds an element, at position pos[q] of the list of lane
// own[q]. f gets a whole tile so that it can issue all U loads before it
// uses any. tab: 64-entry LDS table of the wave; tmask: U LDS words of the
// wave.
constexpr int kPosUnroll = 4;
template <int U, class F>
__device__ inline void flat_tiles(unsigned long long lb, uint32_t len, unsigned long long* tab,
                                  unsigned long long* tmask, int lane, F&& f) {
  unsigned long long total;
  const unsigned long long incl = stream_prefix(len, total);
  if (total == 0) return;
  const unsigned long long start = incl - len;
  const uint32_t rank = stream_rank(len);
  wave_sync();
  // owner lane in the top 6 bits; position of element t: low bits + t
  if (len) tab[rank] = ((lb - start) & ((1ull << 58) - 1)) | (uint64_t(lane) << 58);
  if (lane < U) tmask[lane] = 0;
  wave_sync();
  for (unsigned long long t0 = 0; t0 < total; t0 += U * kWave) {
    const unsigned long long rem = total - t0;
    const int lim = rem < U * kWave ? int(rem) : U * kWave;
    const int nq = (lim + kWave - 1) / kWave;
    int own[U];
    unsigned long long pos[U];
    bool ok[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      own[q] = 0;
      pos[q] = 0;
      ok[q] = false;
    }
    stream_tile<U>(
        start, len, t0, tmask, lane,
        [&](int q, uint32_t rk) {
          if (lane < lim - q * kWave) {
            const unsigned long long e = tab[rk];
            own[q] = int(e >> 58);
            pos[q] = (e + t0 + uint32_t(q * kWave + lane)) & ((1ull << 58) - 1);
            ok[q] = true;
          }
        },
        nq);
    f(own, pos, ok);
  }
}

// Packed form of a sorted row: one entry per non-empty 32-id word,
// low 32 bits = word index (id >> 5), high 32 bits = the word's bit mask.
// The packed entries of row r start at the row's own CSR offset (the entry
// count never exceeds the id count) and len[r] says how many there are.
struct DPacked {
  const uint32_t* len;            // per row; null when absent
  const unsigned long long* ent;
  uint32_t labels[8];             // bit l: rows of label l are packed
};

__device__ inline bool packed_label(const DPacked& p, uint32_t l) { return (p.labels[l >> 5] >> (l & 31)) & 1u; }

// Frontier bitmaps: bit v of slot j is set when delta_C(v) is not empty
// (C = slot_c[j]); words of 64 vertices, `words` per slot.
__global__ __launch_bounds__(kBlock) void k_frontier(const DevGrammar g, uint32_t n, DCsr dl, uint32_t words,
                                                     unsigned long long* bits) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (int j = 0; j < g.nslot; ++j) {
    const uint32_t cb = uint32_t(g.slot_c[j]) * n;
    for (uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~uint32_t(kWave - 1); base < n; base += stride) {
      const uint32_t v = base + lane;
      const unsigned long long m = __ballot(v < n && dl.ptr[cb + v + 1] > dl.ptr[cb + v]);
      if (lane == 0) bits[uint64_t(j) * words + base / kWave] = m;
    }
  }
}

// Slots that share their B label are paired, so that old_B is streamed once
// for both: mate[j] = the slot streamed with j (0xFF: none), 0xFE for the
// second slot of a pair (it is handled with its mate).
struct PairPlan {
  uint8_t mate[kMaxRules];
};

// Pair lists for old_B x delta_C: row (j, u) lists the v in old_B(u) whose
// delta_C(v) is not empty (B = slot_b[j], C = slot_c[j]), in no particular
// order; its bounds are rng[2r], rng[2r+1] (r = j*n+u), so rows need not be
// contiguous. One pass: a wave takes 64 rows and streams their old_B rows
// together, testing each v against the frontier bitmaps of the slot and its
// mate; per-row counts live in LDS, and the first kPairRowBuf pairs of each
// row too. The wave then reserves each slot's words in its block's arena
// (arena blockIdx % kSpread, `arena` words each; a wave takes kPairBlock
// words at a time from the arena's cursor, so the atomics stay few and spread
// over 64 lines) and writes the pairs out. Rows with more than kPairRowBuf
// pairs are streamed a second time, alone, and write their pairs directly.
// When a cursor passes `arena`, no pairs go past it and the host grows the
// arenas and runs the pass again. A slot whose delta_C is empty has no pairs:
// its rows get empty bounds.
__global__ __launch_bounds__(kBlock) void k_pairs(const DevGrammar g, const PairPlan plan, uint32_t n, DCsr old,
                                                  DCsr pend, DPacked pk, const unsigned long long* dlp,
                                                  const unsigned long long* bits, uint32_t words, Counters* ctr,
                                                  unsigned long long arena, unsigned long long* rng, uint32_t* col) {
  constexpr int U = kPosUnroll;
  __shared__ unsigned long long s_tab[kWavesPerBlock][kWave], s_tmask[kWavesPerBlock][U],
      s_at[kWavesPerBlock][2][kWave];
  __shared__ uint32_t s_cnt[kWavesPerBlock][2][kWave], s_buf[kWavesPerBlock][2][kWave * kPairRowBuf];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  const uint32_t stride = gridDim.x * blockDim.x;
  unsigned long long* cursor = &ctr_slot(ctr).pairs;
  const unsigned long long a0 = uint64_t(blockIdx.x % kSpread) * arena;
  unsigned long long blk = 0, blk_end = 0;  // the wave's reserved words, arena-relative (uniform)
  for (int j = 0; j < g.nslot; ++j) {
    if (plan.mate[j] == 0xFEu) continue;
    int js[2] = {j, plan.mate[j] == 0xFFu ? -1 : int(plan.mate[j])};
    bool live[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      live[h] = false;
      if (js[h] < 0) continue;
      const uint64_t cb = uint64_t(g.slot_c[js[h]]) * n;
      live[h] = dlp[cb + n] != dlp[cb];
      if (!live[h])
        for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += stride) {
          rng[2 * (uint64_t(js[h]) * n + u)] = 0;
          rng[2 * (uint64_t(js[h]) * n + u) + 1] = 0;
        }
    }
    if (!live[0]) {
      js[0] = js[1];
      live[0] = live[1];
      live[1] = false;
    }
    if (!live[0]) continue;
    const bool two = live[1];
    const uint32_t bb = uint32_t(g.slot_b[js[0]]) * n;
    const unsigned long long* fb0 = bits + uint64_t(js[0]) * words;
    const unsigned long long* fb1 = two ? bits + uint64_t(js[1]) * words : fb0;
    // packed old_B rows: 32 ids per entry, tested against the bitmaps at once
    const bool packed = pk.len && packed_label(pk, g.slot_b[js[0]]);
    for (uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~uint32_t(kWave - 1); base < n; base += stride) {
      const uint32_t u = base + lane;
      unsigned long long b = 0, qb = 0;
      uint32_t len0 = 0, ql0 = 0;
      if (u < n) {
        b = old.ptr[bb + u];
        len0 = packed ? pk.len[bb + u] : uint32_t(old.ptr[bb + u + 1] - b);
        if (pend.ptr) {
          qb = pend.ptr[bb + u];
          ql0 = uint32_t(pend.ptr[bb + u + 1] - qb);
        }
      }
      uint32_t len = len0, ql = ql0;
      // mode 0: both slots' pairs counted, the first kPairRowBuf of a row kept
      // in LDS; mode 1 + h: slot h's pair `at` of row k written to
      // col[s_at[h][k] + at]
      auto stream = [&](int mode) {
        if (mode == 0) s_cnt[wv][1][lane] = 0;
        s_cnt[wv][mode == 2 ? 1 : 0][lane] = 0;
        wave_sync();
        auto emit = [&](int h, int k, uint32_t id0, uint32_t hit) {
          uint32_t at = atomicAdd(&s_cnt[wv][h][k], uint32_t(__popc(hit)));
          while (hit) {
            const uint32_t id = id0 + uint32_t(__ffs(hit) - 1);
            hit &= hit - 1;
            if (mode != 0)
              col[s_at[wv][h][k] + at] = id;
            else if (at < kPairRowBuf)
              s_buf[wv][h][uint32_t(k) * kPairRowBuf + at] = id;
            ++at;
          }
        };
        const bool t0 = mode != 2, t1 = two && mode != 1;
        // a tile's U entries are loaded, then their bitmap words, then tested
        if (packed)
          flat_tiles<U>(b, len, s_tab[wv], s_tmask[wv], lane, [&](const int* own, const unsigned long long* pos,
                                                                   const bool* ok) {
            unsigned long long e[U];
            uint32_t f0[U], f1[U];
#pragma unroll
            for (int q = 0; q < U; ++q) e[q] = ok[q] ? pk.ent[pos[q]] : 0ull;
#pragma unroll
            for (int q = 0; q < U; ++q) {
              const uint32_t w = uint32_t(e[q]), sh = (w & 1u) * 32u;
              f0[q] = ok[q] && t0 ? uint32_t(fb0[w >> 1] >> sh) : 0u;
              f1[q] = ok[q] && t1 ? uint32_t(fb1[w >> 1] >> sh) : 0u;
            }
#pragma unroll
            for (int q = 0; q < U; ++q) {
              const uint32_t m = uint32_t(e[q] >> 32), h0 = m & f0[q], h1 = m & f1[q];
              if (h0) emit(0, own[q], uint32_t(e[q]) * 32u, h0);
              if (h1) emit(1, own[q], uint32_t(e[q]) * 32u, h1);
            }
          });
        // id rows: old_B(u) when it is not packed, and the pending edges of B(u)
        auto ids = [&](const uint32_t* src, unsigned long long ib, uint32_t il) {
          flat_tiles<U>(ib, il, s_tab[wv], s_tmask[wv], lane, [&](const int* own, const unsigned long long* pos,
                                                                   const bool* ok) {
            uint32_t v[U];
            unsigned long long f0[U], f1[U];
#pragma unroll
            for (int q = 0; q < U; ++q) v[q] = ok[q] ? src[pos[q]] : 0u;
#pragma unroll
            for (int q = 0; q < U; ++q) {
              f0[q] = ok[q] && t0 ? fb0[v[q] / kWave] : 0ull;
              f1[q] = ok[q] && t1 ? fb1[v[q] / kWave] : 0ull;
            }
#pragma unroll
            for (int q = 0; q < U; ++q) {
              if ((f0[q] >> (v[q] % kWave)) & 1ull) emit(0, own[q], v[q], 1u);
              if ((f1[q] >> (v[q] % kWave)) & 1ull) emit(1, own[q], v[q], 1u);
            }
          });
        };
        if (!packed) ids(old.col, b, len);
        if (pend.ptr) ids(pend.col, qb, ql);
        wave_sync();
      };
      stream(0);
      const uint32_t c0 = s_cnt[wv][0][lane], c1 = two ? s_cnt[wv][1][lane] : 0u;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (h == 1 && !two) break;
        const uint32_t c = h ? c1 : c0;
        unsigned long long tot;
        const unsigned long long ex = stream_prefix(c, tot) - c;
        unsigned long long at0 = 0;
        bool fits = true;
        if (tot) {
          if (blk + tot > blk_end) {
            const unsigned long long grab = tot > kPairBlock ? tot : kPairBlock;
            unsigned long long nb = 0;
            if (lane == 0) nb = atomicAdd(cursor, grab);
            blk = __shfl(nb, 0, kWave);
            blk_end = blk + grab;
          }
          at0 = blk + ex;
          fits = blk + tot <= arena;
          blk += tot;
        }
        if (u < n) {
          const unsigned long long lo = a0 + at0;
          rng[2 * (uint64_t(js[h]) * n + u)] = lo;
          rng[2 * (uint64_t(js[h]) * n + u) + 1] = lo + c;
        }
        if (tot && fits) {
          const bool big = c > kPairRowBuf;
          if (!big)
            for (uint32_t t = 0; t < c; ++t) col[a0 + at0 + t] = s_buf[wv][h][uint32_t(lane) * kPairRowBuf + t];
          if (__ballot(big)) {
            // only the rows that did not fit are streamed again
            s_at[wv][h][lane] = a0 + at0;
            len = big ? len0 : 0u;
            ql = big ? ql0 : 0u;
            stream(1 + h);
          }
        }
      }
      wave_sync();
    }
  }
}

// ---------------------------------------------------------------------------
// expand: fused join + dedup + difference, one wave per active row


// One operator of the reference's partitioned schedule (MODE 2 of the
// expand kernel). Rows are local to partition [first, first+np): row
// label*np + (u - first). Candidates for row (A,u) go into the seed region
// (tmp) as new ids:
//   kind 0 (compute11): for A ::= B C, v in key_B(u), dst[d]_C(v) for every
//                      dst whose partition holds v;
//   kind 1 (compute10): for A ::= B, key_B(u);
//   kind 2 (compute00): u itself, for eps rules A;
//   kind 3 (difference): key_A(u) (seed = old: the fresh ids are key - old).
struct RegionOp {
  int kind;
  uint32_t first, np;
  DCsr seed, key;
  DCsr dst[2];
  uint32_t dfirst[2], dn[2];
  int ndst;
};

struct ExpandArgs {
  DCsr old, dl, raw;
  DCsr pend;  // edges found since old was last rebuilt (ptr null: none); old ∪ pend is the closure so far
  RegionOp ro;
  DPacked pk;  // packed old rows
  DPacked dpk;  // packed delta rows (same labels; entries at the row's delta offset)
  DCsr b2;  // rows r = slot*n + u: keys v with old_B(u,v) and delta_C(v) non-empty, at col[ptr[2r], ptr[2r+1])
  uint32_t n;
  DevGrammar g;
  const uint32_t* active;
  Counters* ctr;
  uint32_t* staging;
  unsigned long long staging_cap;
  unsigned long long* seg_off;
  uint32_t* seg_cnt;
  uint32_t guess;  // ids of a window centred on the row's vertex tried first (0: none)
  uint32_t grab;   // active rows a wave takes per work-counter atomic
  // MODE 0: when set, the rows to expand are the ctr->deferred rows listed
  // here (those the light pass left), not the active list
  const uint32_t* defer;
  uint32_t group;  // k_expand_group: rows per group at most
};

// The span of destination region d of a partition operator: global vertex
// ids [lo, hi) outside which its rows are all empty (lo >= hi: all empty).
// Keys outside the span have no candidates.
__device__ inline void dst_span(const ExpandArgs& a, int d, uint32_t& lo, uint32_t& hi) {
  lo = a.ro.dfirst[d] + (a.ro.dn[d] - a.ctr->span_lo_c[d]);
  hi = a.ro.dfirst[d] + a.ctr->span_hi[d];
}

// Span of the local vertices of a region (rows label * np + local vertex)
// with a non-empty row of any label, into ctr->span_lo_c[d] / span_hi[d],
// which start at zero. Block y covers label y; each thread reduces its own
// rows, then the block, so a launch makes two counter updates per block.
__global__ __launch_bounds__(kBlock) void k_region_span(DCsr reg, uint32_t np, int d, Counters* ctr) {
  __shared__ uint32_t s_mn[kWavesPerBlock], s_mx[kWavesPerBlock];
  const unsigned long long* ptr = reg.ptr + (unsigned long long)blockIdx.y * np;
  uint32_t mn = 0xFFFFFFFFu, mx = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < np; v += gridDim.x * blockDim.x)
    if (ptr[v + 1] > ptr[v]) {
      mn = min(mn, v);
      mx = v + 1;  // v grows along the loop
    }
  mn = wave_min_u32(mn);
  mx = wave_max_u32(mx);
  const int wv = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s_mn[wv] = mn;
    s_mx[wv] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kWavesPerBlock; ++i) {
      mn = min(mn, s_mn[i]);
      mx = max(mx, s_mx[i]);
    }
    if (mx) {
      atomicMax(&ctr->span_lo_c[d], np - mn);
      atomicMax(&ctr->span_hi[d], mx);
    }
  }
}

// The same search by a whole wave (b, e and x uniform; uniform result): each
// step samples 64 evenly spaced ids and keeps the gap where they cross x, so
// a row of length len takes about log64(len) dependent loads.
__device__ inline unsigned long long wave_lower_bound(const uint32_t* col, unsigned long long b, unsigned long long e,
                                                      uint32_t x, int lane) {
  while (e - b > kWave) {
    const unsigned long long step = (e - b + kWave - 1) / kWave;
    const unsigned long long k = b + uint32_t(lane) * step;
    // the lanes whose sample is below x are a prefix of the wave
    const uint32_t c = uint32_t(__popcll(__ballot(k < e && col[k] < x)));
    if (c == 0) return b;
    // col[b + (c-1)*step] < x, and col[b + c*step] >= x or b + c*step >= e
    const unsigned long long nb = b + (c - 1) * step + 1;
    e = min(e, b + c * step);
    b = nb;
  }
  const unsigned long long k = b + uint32_t(lane);
  return b + uint32_t(__popcll(__ballot(k < e && col[k] < x)));
}

// First index k in [b, e) with col[k] >= x (e if none), col ascending.
__device__ inline unsigned long long lower_bound_u32(const uint32_t* col, unsigned long long b, unsigned long long e,
                                                     uint32_t x) {
  while (b < e) {
    const unsigned long long m = b + (e - b) / 2;
    if (col[m] < x)
      b = m + 1;
    else
      e = m;
  }
  return b;
}

// The active list of one partition operator: the rows of the partition that
// can receive candidates (see RegionOp). kind 3 and kind 1 rows need a
// non-empty key row, kind 2 rows an eps rule, kind 0 rows a key id inside a
// destination's span (ids in key rows are ascending: the first and last id
// decide most rows, a binary search the rest). Rows gather per wave in LDS
// and take their place in the list with one counter update per kBuf rows.
__global__ __launch_bounds__(kBlock) void k_region_active(ExpandArgs a, uint32_t rows) {
  constexpr uint32_t kBuf = 512;
  __shared__ uint32_t s_buf[kWavesPerBlock][kBuf];
  const RegionOp& ro = a.ro;
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t* buf = s_buf[threadIdx.x / kWave];
  uint32_t nbuf = 0;  // uniform
  uint32_t* out = const_cast<uint32_t*>(a.active);  // the list this launch writes
  uint32_t lo[2] = {0, 0}, hi[2] = {0, 0};
  for (int d = 0; d < ro.ndst; ++d) dst_span(a, d, lo[d], hi[d]);
  auto flush = [&]() {
    wave_sync();
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(&a.ctr->active, nbuf);
    pos = uni_u32(pos);  // lane 0: every lane is active here
    for (uint32_t i = uint32_t(lane); i < nbuf; i += kWave) out[pos + i] = buf[i];
    nbuf = 0;
    wave_sync();
  };
  for (uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~uint32_t(kWave - 1); base < rows;
       base += gridDim.x * blockDim.x) {
    const uint32_t r = base + lane;
    bool act = false;
    if (r < rows) {
      const uint32_t A = r / ro.np, ul = r - A * ro.np;
      if (ro.kind == 3) {
        act = ro.key.ptr[r + 1] > ro.key.ptr[r];
      } else if (ro.kind == 2) {
        for (int i = 0; i < a.g.ne; ++i) act |= a.g.e_lhs[i] == A;
      } else if (ro.kind == 1) {
        for (int i = 0; i < a.g.nu && !act; ++i) {
          if (a.g.u_lhs[i] != A) continue;
          const uint32_t rk = uint32_t(a.g.u_rhs[i]) * ro.np + ul;
          act = ro.key.ptr[rk + 1] > ro.key.ptr[rk];
        }
      } else {
        for (int i = 0; i < a.g.nb && !act; ++i) {
          if (a.g.b_lhs[i] != A) continue;
          const uint32_t rk = uint32_t(a.g.b_rhs1[i]) * ro.np + ul;
          const unsigned long long kb = ro.key.ptr[rk], ke = ro.key.ptr[rk + 1];
          if (kb == ke) continue;
          const uint32_t first = ro.key.col[kb], last = ro.key.col[ke - 1];
          for (int d = 0; d < ro.ndst && !act; ++d) {
            if (lo[d] >= hi[d] || last < lo[d] || first >= hi[d]) continue;
            const unsigned long long k = first >= lo[d] ? kb : lower_bound_u32(ro.key.col, kb + 1, ke, lo[d]);
            act = k < ke && ro.key.col[k] < hi[d];
          }
        }
      }
    }
    const unsigned long long m = __ballot(act);
    if (!m) continue;
    if (nbuf + kWave > kBuf) flush();
    if (act) buf[nbuf + __popcll(m & ((1ull << lane) - 1))] = r;
    nbuf += uint32_t(__popcll(m));
  }
  if (nbuf) flush();
}

struct WaveLds {
  uint32_t seen[kChunkWords];
  uint32_t fresh[kChunkWords];
  unsigned long long lbase[kWave];   // expand_stream: base address per non-empty list
  unsigned long long tmask[kUnroll];  // expand_stream: list starts inside the tile
  unsigned long long acc[3];          // the wave's K, O and N counts (in LDS, not in registers for the whole kernel)
#ifdef GG_EXPAND_STATS
  unsigned long long xs[4];
#endif
};
#ifdef GG_EXPAND_STATS
#define GG_CLK(v) const unsigned long long v = __builtin_amdgcn_s_memtime()
#define GG_ACC(w, i, t0) \
  if ((threadIdx.x & (kWave - 1)) == 0) (w).xs[i] += __builtin_amdgcn_s_memtime() - (t0)
#else
#define GG_CLK(v)
#define GG_ACC(w, i, t0)
#endif

// Calls elem(x) for every element of the lists held by the lanes (lane i
// holds src[lb .. lb+len)), as one flattened stream (see stream_tile) with
// kUnroll loads (kUnroll/2 for 64-bit packed words) in flight per lane before
// they are consumed. The per-rank
// table holds each non-empty list's base address. All lanes take part.
template <class T, class F>
__device__ inline void expand_stream(const T* __restrict__ src, unsigned long long lb, uint32_t len, WaveLds& w,
                                     int lane, F&& elem) {
  // loads in flight per lane: kUnroll ids, or half as many 64-bit packed
  // words, so both kinds of stream hold the same registers
  constexpr int kU = sizeof(T) == 8 ? kUnroll / 2 : kUnroll;
  constexpr uint32_t kTile = kWave * kU;
  unsigned long long total;
  const unsigned long long incl = stream_prefix(len, total);
  if (total == 0) return;
  const unsigned long long start = incl - len;
  const uint32_t rank = stream_rank(len);
  wave_sync();
  // element t of the stream lies at byte address lbase[rank of its list] + t*sizeof(T)
  if (len) w.lbase[rank] = reinterpret_cast<unsigned long long>(src) + (lb - start) * sizeof(T);
  if (lane < kU) w.tmask[lane] = 0;
  wave_sync();
  // full tiles first, with no bounds test, then the partial last tile
  const unsigned long long full = total / kTile * kTile;
  for (unsigned long long t0 = 0; t0 < full; t0 += kTile) {
    const unsigned long long tb = (t0 + uint32_t(lane)) * sizeof(T);
    T x[kU];
    stream_tile<kU>(start, len, t0, w.tmask, lane, [&](int q, uint32_t rk) {
      x[q] = *reinterpret_cast<const __attribute__((address_space(1))) T*>(w.lbase[rk] + tb +
                                                                            uint32_t(q * kWave) * sizeof(T));
    });
#pragma unroll
    for (int q = 0; q < kU; ++q) elem(x[q]);
  }
  if (full < total) {
    const unsigned long long t0 = full;
    const int lim = int(total - full);
    const int nq = (lim + kWave - 1) / kWave;
    const unsigned long long tb = (t0 + uint32_t(lane)) * sizeof(T);
    T x[kU];
    stream_tile<kU>(
        start, len, t0, w.tmask, lane,
        [&](int q, uint32_t rk) {
          x[q] = 0;
          if (lane < lim - q * kWave)  // lane against a scalar bound: no per-q lane constants
            x[q] = *reinterpret_cast<const __attribute__((address_space(1))) T*>(w.lbase[rk] + tb +
                                                                                  uint32_t(q * kWave) * sizeof(T));
        },
        nq);
#pragma unroll
    for (int q = 0; q < kU; ++q)
      if (q < nq && lane < lim - q * kWave) elem(x[q]);
  }
  wave_sync();
}

// Visits every candidate list of row (A,u): calls visit(col, lb, len) once
// per batch with one list per lane (len 0 = no list on that lane). With
// PACKED, rows of packed labels (old and delta) go to visitw(ent, pb, plen,
// len) instead: plen packed words at ent[pb ..] standing for len ids.
template <int MODE, bool PACKED, bool PEND, class V, class W>
__device__ inline void for_each_list(const ExpandArgs& a, uint32_t r, uint32_t A, uint32_t u, int lane, V&& visit,
                                     W&& visitw) {
  const uint32_t n = a.n;
  if (MODE == 1) {
    const unsigned long long b = a.raw.ptr[r], e = a.raw.ptr[r + 1];
    visit(a.raw.col, lane == 0 ? b : 0ull, lane == 0 ? uint32_t(e - b) : 0u);
    return;
  }
  if (MODE == 2) {
    const RegionOp& ro = a.ro;
    const uint32_t ul = u - ro.first;
    if (ro.kind == 3) {
      const uint32_t rk = A * ro.np + ul;
      const unsigned long long b = ro.key.ptr[rk], e = ro.key.ptr[rk + 1];
      visit(ro.key.col, lane == 0 ? b : 0ull, lane == 0 ? uint32_t(e - b) : 0u);
    } else if (ro.kind == 1) {
      for (int i = 0; i < a.g.nu; ++i) {
        if (a.g.u_lhs[i] != A) continue;
        const uint32_t rk = uint32_t(a.g.u_rhs[i]) * ro.np + ul;
        const unsigned long long b = ro.key.ptr[rk], e = ro.key.ptr[rk + 1];
        visit(ro.key.col, lane == 0 ? b : 0ull, lane == 0 ? uint32_t(e - b) : 0u);
      }
    } else if (ro.kind == 0) {
      // keys outside every destination's span have no candidates: long key
      // rows are cut to [klo, khi) first
      uint32_t slo[2] = {0, 0}, sw[2] = {0, 0};
      uint32_t klo = 0xFFFFFFFFu, khi = 0;
      for (int d = 0; d < ro.ndst; ++d) {
        uint32_t hi = 0;
        dst_span(a, d, slo[d], hi);
        if (slo[d] < hi) {
          sw[d] = hi - slo[d];
          klo = min(klo, slo[d]);
          khi = max(khi, hi);
        }
      }
      if (klo >= khi) return;
      for (int i = 0; i < a.g.nb; ++i) {
        if (a.g.b_lhs[i] != A) continue;
        const uint32_t B = a.g.b_rhs1[i], C = a.g.b_rhs2[i];
        const uint32_t rk = B * ro.np + ul;
        unsigned long long kb = ro.key.ptr[rk], ke = ro.key.ptr[rk + 1];
        if (ke - kb > 2 * kWave) {
          kb = wave_lower_bound(ro.key.col, kb, ke, klo, lane);
          ke = wave_lower_bound(ro.key.col, kb, ke, khi, lane);
        }
        for (unsigned long long k0 = kb; k0 < ke; k0 += kWave) {
          const unsigned long long k = k0 + lane;
          const uint32_t v = k < ke ? ro.key.col[k] : 0u;
          for (int d = 0; d < ro.ndst; ++d) {
            unsigned long long lb = 0;
            uint32_t len = 0;
            if (k < ke && v - slo[d] < sw[d]) {
              const uint32_t rc = C * ro.dn[d] + (v - ro.dfirst[d]);
              lb = ro.dst[d].ptr[rc];
              len = uint32_t(ro.dst[d].ptr[rc + 1] - lb);
            }
            visit(ro.dst[d].col, lb, len);
          }
        }
      }
    }
    return;
  }
  // MODE 0. The row's lists come from segments: a unary rule's Δ_B(u); per
  // binary rule, the keys Δ_B(u) (lists old_C(v) and Δ_C(v)) and the keys of
  // the frontier pair list (lists Δ_C(v)). Lane j describes segment j. The
  // bounds of every segment are loaded together, and the keys of all keyed
  // segments are flattened into batches of 64, so a row costs one round trip
  // per stage (bounds, keys, key rows, lists) whatever its number of rules.
  // a segment descriptor: kind (bits 0-1), packed old_C rows (bit 2), C (bits 8-15)
  enum { kSegU = 0, kSegD = 1, kSegP = 2 };
  uint32_t s_desc = 0, s_row = 0;
  uint32_t ns = 0;  // segments in the current group (uniform)
  auto run_group = [&]() {
    unsigned long long b = 0, e = 0;
    const uint32_t kind_l = s_desc & 3u;
    if (uint32_t(lane) < ns) {
      // pair lists keep two bounds per row (rows are not contiguous)
      const unsigned long long* ptr = kind_l == kSegP ? a.b2.ptr + s_row : a.dl.ptr;
      b = ptr[s_row];
      e = ptr[s_row + 1];
    }
    {
      const bool ul = uint32_t(lane) < ns && kind_l == kSegU;
      const bool upk = PACKED && ul && (s_desc & 4u) && a.dpk.len && e > b;
      if (PACKED) visitw(a.dpk.ent, upk ? b : 0ull, upk ? a.dpk.len[s_row] : 0u, upk ? uint32_t(e - b) : 0u);
      visit(a.dl.col, ul && !upk ? b : 0ull, ul && !upk ? uint32_t(e - b) : 0u);
    }
    const unsigned long long klen = (uint32_t(lane) < ns && kind_l != kSegU) ? e - b : 0ull;
    unsigned long long total;  // (a key list is a row: fewer than 2^32 keys)
    const unsigned long long incl = stream_prefix(uint32_t(klen), total);
    for (unsigned long long t0 = 0; t0 < total; t0 += kWave) {
      const unsigned long long t = t0 + lane;
      // segment of key t: the first j whose inclusive end is past t
      int j = 0;
      unsigned long long ex = 0;
      for (uint32_t q = 0; q + 1 < ns; ++q) {
        const unsigned long long iq =
            (unsigned long long)uint32_t(__builtin_amdgcn_readlane(int(uint32_t(incl)), int(q))) |
            ((unsigned long long)uint32_t(__builtin_amdgcn_readlane(int(uint32_t(incl >> 32)), int(q))) << 32);
        if (iq <= t) {
          j = int(q) + 1;
          ex = iq;
        }
      }
      const uint32_t desc = __shfl(s_desc, j, kWave);
      const uint32_t kind = desc & 3u, C = desc >> 8;
      const bool pk = (desc & 4u) != 0;
      const unsigned long long bj = __shfl(b, j, kWave);
      unsigned long long ob = 0, db = 0, qb = 0;
      uint32_t ol = 0, dlen = 0, pl = 0, dpl = 0, ql = 0;
      bool is_pk = false, d_pk = false;
      if (t < total) {
        const unsigned long long k = bj + (t - ex);
        const uint32_t v = kind == kSegP ? a.b2.col[k] : a.dl.col[k];
        const uint32_t rc = C * n + v;
        db = a.dl.ptr[rc];
        dlen = uint32_t(a.dl.ptr[rc + 1] - db);
        if (PACKED && pk && a.dpk.len && dlen) {
          d_pk = true;
          dpl = a.dpk.len[rc];
        }
        if (kind == kSegD) {
          ob = a.old.ptr[rc];
          ol = uint32_t(a.old.ptr[rc + 1] - ob);
          if (PACKED && pk && a.pk.len) {
            is_pk = true;
            pl = a.pk.len[rc];
          }
          if (PEND && a.pend.ptr) {
            qb = a.pend.ptr[rc];
            ql = uint32_t(a.pend.ptr[rc + 1] - qb);
          }
        }
      }
      if (PACKED) {
        visitw(a.pk.ent, is_pk ? ob : 0ull, is_pk ? pl : 0u, is_pk ? ol : 0u);
        visitw(a.dpk.ent, d_pk ? db : 0ull, d_pk ? dpl : 0u, d_pk ? dlen : 0u);
      }
      visit(a.old.col, is_pk ? 0ull : ob, is_pk ? 0u : ol);
      visit(a.dl.col, d_pk ? 0ull : db, d_pk ? 0u : dlen);
      if (PEND && a.pend.ptr) visit(a.pend.col, qb, ql);
    }
    ns = 0;
  };
  auto add = [&](uint32_t kind, uint32_t label, uint32_t row, bool pk) {
    if (uint32_t(lane) == ns) {
      s_desc = kind | (pk ? 4u : 0u) | (label << 8);
      s_row = row;
    }
    if (++ns == uint32_t(kWave)) run_group();
  };
  for (int i = 0; i < a.g.nu; ++i)
    if (a.g.u_lhs[i] == A)
      add(kSegU, 0, uint32_t(a.g.u_rhs[i]) * n + u, packed_label(a.pk, uint32_t(a.g.u_rhs[i])));
  for (int i = 0; i < a.g.nb; ++i) {
    if (a.g.b_lhs[i] != A) continue;
    const uint32_t B = a.g.b_rhs1[i], C = a.g.b_rhs2[i];
    add(kSegD, C, B * n + u, packed_label(a.pk, C));
    const uint32_t slot = a.g.b_slot[i];
    if (slot != 0xFFu && a.b2.ptr) add(kSegP, C, slot * n + u, false);
  }
  if (ns) run_group();
}

// First position in sorted ids col[b, e) whose id is >= x.
__device__ inline unsigned long long lower_pos(const uint32_t* col, unsigned long long b, unsigned long long e,
                                               unsigned long long x) {
  while (b < e) {
    const unsigned long long mid = (b + e) >> 1;
    if (col[mid] < x)
      b = mid + 1;
    else
      e = mid;
  }
  return b;
}

// Same over packed entries (word index in the low 32 bits, ascending).
__device__ inline unsigned long long lower_pos_word(const unsigned long long* ent, unsigned long long b,
                                                    unsigned long long e, uint32_t wi) {
  while (b < e) {
    const unsigned long long mid = (b + e) >> 1;
    if (uint32_t(ent[mid]) < wi)
      b = mid + 1;
    else
      e = mid;
  }
  return b;
}

// lower_pos and lower_pos_word for a caller whose keys only grow: the first
// position >= x found by galloping from b (steps 1, 2, 4, ... past b, then a
// binary search inside the last step), so that a key just past the previous
// one costs a few dependent loads, not a search over the whole rest.
__device__ inline unsigned long long gallop_pos(const uint32_t* col, unsigned long long b, unsigned long long e,
                                                unsigned long long x) {
  if (b >= e || col[b] >= x) return b;
  unsigned long long lo = b, step = 1;  // col[lo] < x
  while (lo + step < e && col[lo + step] < x) {
    lo += step;
    step <<= 1;
  }
  return lower_pos(col, lo + 1, lo + step < e ? lo + step : e, x);
}
__device__ inline unsigned long long gallop_pos_word(const unsigned long long* ent, unsigned long long b,
                                                     unsigned long long e, uint32_t wi) {
  if (b >= e || uint32_t(ent[b]) >= wi) return b;
  unsigned long long lo = b, step = 1;  // word of ent[lo] < wi
  while (lo + step < e && uint32_t(ent[lo + step]) < wi) {
    lo += step;
    step <<= 1;
  }
  return lower_pos_word(ent, lo + 1, lo + step < e ? lo + step : e, wi);
}

// Clears the window [c0, c0+span), seeds it with old_A(u) U delta_A(u) and
// streams every candidate through it. Returns the number of new edges (set
// bits of w.fresh) and stores the candidate count in kcount.
//
// ranged (rows whose ids span several windows; sorted lists only): every
// list and seed row is cut to [c0, c0+span) by binary search, so a window
// costs its own ids plus a search per list, and `next` receives the
// smallest candidate id at or past the window's end (0xFFFFFFFF if none):
// windows with no candidate are skipped.
template <int MODE, bool PEND>
__device__ __attribute__((always_inline)) inline uint32_t process_chunk(const ExpandArgs& a, uint32_t r, uint32_t A,
                                                                        uint32_t u, uint32_t c0, uint32_t span,
                                                                        bool ranged, WaveLds& w, int lane,
                                                                        unsigned long long& kcount, bool& miss,
                                                                        uint32_t& next) {
  bool miss_l = false;
  GG_CLK(t_seed);
  const uint32_t words = (span + 31) >> 5;
  const unsigned long long c1 = (unsigned long long)c0 + span;
  const uint32_t w0 = c0 >> 5;
  uint32_t nxt = 0xFFFFFFFFu;
  for (uint32_t i = lane; i < words; i += kWave) {
    w.seen[i] = 0;
    w.fresh[i] = 0;
  }
  wave_sync();
  // [b, e) of a sorted id row, cut to the window when ranged
  auto cut = [&](const uint32_t* col, unsigned long long& b, unsigned long long& e) {
    if (ranged && b < e) {
      b = lower_pos(col, b, e, c0);
      e = lower_pos(col, b, e, c1);
    }
  };
  if (MODE == 2) {
    unsigned long long sb = a.ro.seed.ptr[r], se = a.ro.seed.ptr[r + 1];
    cut(a.ro.seed.col, sb, se);
    for (unsigned long long i = sb + lane; i < se; i += kWave) {
      const uint32_t off = a.ro.seed.col[i] - c0;
      if (off < span) atomicOr(&w.seen[off >> 5], 1u << (off & 31));
    }
    wave_sync();
  }
  if (MODE == 0) {
    auto seed = [&](uint32_t x) {
      const uint32_t off = x - c0;
      if (off < span) atomicOr(&w.seen[off >> 5], 1u << (off & 31));
    };
    if (a.pk.len && packed_label(a.pk, A)) {
      // seed old_A(u) word by word (c0 is a multiple of 32)
      unsigned long long pb = a.old.ptr[r], pe = pb + a.pk.len[r];
      if (ranged && pb < pe) {
        pb = lower_pos_word(a.pk.ent, pb, pe, w0);
        pe = lower_pos_word(a.pk.ent, pb, pe, w0 + words);
      }
      for (unsigned long long i = pb + lane; i < pe; i += kWave) {
        const unsigned long long e = a.pk.ent[i];
        const uint32_t wi = uint32_t(e) - w0;
        if (wi < words) atomicOr(&w.seen[wi], uint32_t(e >> 32));
      }
    } else {
      unsigned long long ob = a.old.ptr[r], oe = a.old.ptr[r + 1];
      cut(a.old.col, ob, oe);
      for (unsigned long long i = ob + lane; i < oe; i += kWave) seed(a.old.col[i]);
    }
    unsigned long long db = a.dl.ptr[r], de = a.dl.ptr[r + 1];
    cut(a.dl.col, db, de);
    for (unsigned long long i = db + lane; i < de; i += kWave) seed(a.dl.col[i]);
    if (PEND && a.pend.ptr) {
      unsigned long long qb = a.pend.ptr[r], qe = a.pend.ptr[r + 1];
      cut(a.pend.col, qb, qe);
      for (unsigned long long i = qb + lane; i < qe; i += kWave) seed(a.pend.col[i]);
    }
    wave_sync();
  }
  // candidates go to `fresh` with plain (no-return) LDS ORs; the ids already
  // known (`seen`) are taken out once, word by word, at the end
  auto mark = [&](uint32_t x) {
    const uint32_t off = x - c0;
    if (off < span)
      __hip_atomic_fetch_or(&w.fresh[off >> 5], 1u << (off & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else
      miss_l = true;
  };
  unsigned long long kc = 0;
  auto mark_word = [&](unsigned long long e) {
    const uint32_t wi = uint32_t(e) - w0;
    if (wi < words) {
      const uint32_t bits = uint32_t(e >> 32);
      if (ranged) kc += __popc(bits);
      __hip_atomic_fetch_or(&w.fresh[wi], bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
      miss_l = true;
    }
  };
  GG_ACC(w, 0, t_seed);
  GG_CLK(t_keys);
  for_each_list<MODE, true, PEND>(
      a, r, A, u, lane,
      [&](const uint32_t* col, unsigned long long lb, uint32_t len) {
        unsigned long long b = lb, e = lb + len;
        cut(col, b, e);
        if (ranged && e < lb + len) nxt = min(nxt, col[e]);
        kc += e - b;
        GG_CLK(t_s);
        expand_stream(col, b, uint32_t(e - b), w, lane, mark);
        GG_ACC(w, 2, t_s);
      },
      [&](const unsigned long long* ent, unsigned long long pb, uint32_t pl, uint32_t len) {
        unsigned long long b = pb, e = pb + pl;
        if (ranged && b < e) {
          b = lower_pos_word(ent, b, e, w0);
          e = lower_pos_word(ent, b, e, w0 + words);
          if (e < pb + pl) nxt = min(nxt, uint32_t(ent[e]) << 5);
        } else {
          kc += len;
        }
        GG_CLK(t_s);
        expand_stream(ent, b, uint32_t(e - b), w, lane, mark_word);
        GG_ACC(w, 2, t_s);
      });
  GG_ACC(w, 1, t_keys);
  if (MODE == 2 && a.ro.kind == 2) {
    // eps: the row's own vertex
    bool has_eps = false;
    for (int i = 0; i < a.g.ne; ++i) has_eps |= a.g.e_lhs[i] == A;
    if (has_eps) {
      if (!ranged || (u >= c0 && u < c1)) {
        if (lane == 0) mark(u);
        kc += lane == 0 ? 1 : 0;
      } else if (u >= c1) {
        nxt = min(nxt, u);
      }
      wave_sync();
    }
  }
  kcount = wave_sum_u64_small(kc);
  miss = __ballot(miss_l) != 0;
  next = ranged ? wave_min_u32(nxt) : 0xFFFFFFFFu;  // (only ranged windows use it)
  wave_sync();
  uint32_t c = 0;
  for (uint32_t i = lane; i < words; i += kWave) {
    const uint32_t f = w.fresh[i] & ~w.seen[i];
    w.fresh[i] = f;
    c += __popc(f);
  }
  wave_sync();
  return wave_sum_u32(c);  // at most span
}

// Writes the set bits of the bitmap `fresh` (window base c0) in ascending
// order at staging[pos ..].
__device__ inline void emit_chunk(const ExpandArgs& a, uint32_t c0, uint32_t span, unsigned long long pos,
                                  const uint32_t* fresh, int lane) {
  const uint32_t words = (span + 31) >> 5;
  for (uint32_t w0 = 0; w0 < words; w0 += kWave) {
    const uint32_t wi = w0 + lane;
    uint32_t bits = wi < words ? fresh[wi] : 0u;
    const uint32_t pc = __popc(bits);
    const uint32_t incl = wave_incl_scan(pc, lane);
    unsigned long long p = pos + (incl - pc);
    while (bits) {
      const int b = __ffs(bits) - 1;
      bits &= bits - 1;
      a.staging[p++] = c0 + wi * 32 + uint32_t(b);
    }
    pos += lane63_u32(incl);
  }
}

// ---------------------------------------------------------------------------
// Row groups: up to kGroupRows consecutive rows of the active list per wave,
// in one guessed window [c0, c0 + span) shared by the group (centred on its
// middle row's vertex; the active list is vertex-major, so the group's rows
// are one or a few neighbouring vertices), each row with its own seen/fresh
// bitmaps (words [j*words, (j+1)*words) for row j). The seed rows and the
// lists of the whole group go through one set of flattened streams, every
// list carrying its row's bitmap offset, so what the wave kernel pays per row
// (segment bounds, key batches, stream set-ups, reductions) is paid once per
// group. A row with a candidate outside the window goes to the deferred
// list, which the wave kernel expands afterwards (its own windows); its
// counts are left to that kernel.
constexpr uint32_t kGroupRows = 8;  // at most (the tuned group_rows, and as many bitmaps as fit)
constexpr uint32_t kGroupWords = kChunkWords;  // bitmap words per wave (seen and fresh each)
constexpr int kGroupWaves = kExpandWaves;      // occupancy target (waves per SIMD)

struct GroupLds {
  uint32_t seen[kGroupWords];
  uint32_t fresh[kGroupWords];
  unsigned long long lbase[kWave];      // stream: base address per non-empty list
  uint32_t lwoff[kWave];                // stream: bitmap offset (words) per non-empty list
  unsigned long long tmask[kUnroll];    // stream: list starts inside the tile
  unsigned long long kacc[kGroupRows];  // candidates per row of the group
  unsigned long long acc[3];            // the wave's K, O and N counts
};

// expand_stream with a bitmap offset per list: lane i's list
// src[lb .. lb+len) (src may differ from lane to lane) carries woff, and
// elem(x, woff) gets the offset of the element's list.
template <class T, class F>
__device__ inline void expand_stream_w(const T* src, unsigned long long lb, uint32_t len, uint32_t woff, GroupLds& w,
                                       int lane, F&& elem) {
  constexpr int kU = sizeof(T) == 8 ? kUnroll / 2 : kUnroll;
  constexpr uint32_t kTile = kWave * kU;
  unsigned long long total;
  const unsigned long long incl = stream_prefix(len, total);
  if (total == 0) return;
  const unsigned long long start = incl - len;
  const uint32_t rank = stream_rank(len);
  wave_sync();
  if (len) {
    w.lbase[rank] = reinterpret_cast<unsigned long long>(src) + (lb - start) * sizeof(T);
    w.lwoff[rank] = woff;
  }
  if (lane < kU) w.tmask[lane] = 0;
  wave_sync();
  const unsigned long long full = total / kTile * kTile;
  for (unsigned long long t0 = 0; t0 < full; t0 += kTile) {
    const unsigned long long tb = (t0 + uint32_t(lane)) * sizeof(T);
    T x[kU];
    uint32_t wo[kU];
    stream_tile<kU>(start, len, t0, w.tmask, lane, [&](int q, uint32_t rk) {
      x[q] = *reinterpret_cast<const __attribute__((address_space(1))) T*>(w.lbase[rk] + tb +
                                                                            uint32_t(q * kWave) * sizeof(T));
      wo[q] = w.lwoff[rk];
    });
#pragma unroll
    for (int q = 0; q < kU; ++q) elem(x[q], wo[q]);
  }
  if (full < total) {
    const unsigned long long t0 = full;
    const int lim = int(total - full);
    const int nq = (lim + kWave - 1) / kWave;
    const unsigned long long tb = (t0 + uint32_t(lane)) * sizeof(T);
    T x[kU];
    uint32_t wo[kU];
    stream_tile<kU>(
        start, len, t0, w.tmask, lane,
        [&](int q, uint32_t rk) {
          x[q] = 0;
          wo[q] = 0;
          if (lane < lim - q * kWave) {
            x[q] = *reinterpret_cast<const __attribute__((address_space(1))) T*>(w.lbase[rk] + tb +
                                                                                  uint32_t(q * kWave) * sizeof(T));
            wo[q] = w.lwoff[rk];
          }
        },
        nq);
#pragma unroll
    for (int q = 0; q < kU; ++q)
      if (q < nq && lane < lim - q * kWave) elem(x[q], wo[q]);
  }
  wave_sync();
}

// The candidate lists of the group's rows (lane j < G holds row j's A and
// u): for_each_list's MODE 0 segments, every one tagged with its row, so
// visit(col, lb, len, j) and visitw(ent, pb, plen, len, j) name the row j of
// the lane's list.
template <bool PEND, class V, class W>
__device__ inline void for_each_list_group(const ExpandArgs& a, uint32_t G, uint32_t A_l, uint32_t u_l, int lane,
                                           V&& visit, W&& visitw) {
  const uint32_t n = a.n;
  // a segment descriptor: kind (bits 0-1), packed rows (bit 2), C (bits
  // 8-15), row of the group (bits 16-18)
  enum { kSegU = 0, kSegD = 1, kSegP = 2 };
  uint32_t s_desc = 0, s_row = 0;
  uint32_t ns = 0;  // segments in the current batch (uniform)
  auto run_group = [&]() {
    unsigned long long b = 0, e = 0;
    const uint32_t kind_l = s_desc & 3u, tag_l = (s_desc >> 16) & 7u;
    if (uint32_t(lane) < ns) {
      // pair lists keep two bounds per row (rows are not contiguous)
      const unsigned long long* ptr = kind_l == kSegP ? a.b2.ptr + s_row : a.dl.ptr;
      b = ptr[s_row];
      e = ptr[s_row + 1];
    }
    {
      const bool ul = uint32_t(lane) < ns && kind_l == kSegU;
      const bool upk = ul && (s_desc & 4u) && a.dpk.len && e > b;
      visitw(a.dpk.ent, upk ? b : 0ull, upk ? a.dpk.len[s_row] : 0u, upk ? uint32_t(e - b) : 0u, tag_l);
      visit(a.dl.col, ul && !upk ? b : 0ull, ul && !upk ? uint32_t(e - b) : 0u, tag_l);
    }
    const unsigned long long klen = (uint32_t(lane) < ns && kind_l != kSegU) ? e - b : 0ull;
    unsigned long long total;  // (a key list is a row: fewer than 2^32 keys)
    const unsigned long long incl = stream_prefix(uint32_t(klen), total);
    for (unsigned long long t0 = 0; t0 < total; t0 += kWave) {
      const unsigned long long t = t0 + lane;
      // segment of key t: the first j whose inclusive end is past t
      int j = 0;
      unsigned long long ex = 0;
      for (uint32_t q = 0; q + 1 < ns; ++q) {
        const unsigned long long iq =
            (unsigned long long)uint32_t(__builtin_amdgcn_readlane(int(uint32_t(incl)), int(q))) |
            ((unsigned long long)uint32_t(__builtin_amdgcn_readlane(int(uint32_t(incl >> 32)), int(q))) << 32);
        if (iq <= t) {
          j = int(q) + 1;
          ex = iq;
        }
      }
      const uint32_t desc = __shfl(s_desc, j, kWave);
      const uint32_t kind = desc & 3u, C = (desc >> 8) & 0xFFu, tag = (desc >> 16) & 7u;
      const bool pk = (desc & 4u) != 0;
      const unsigned long long bj = __shfl(b, j, kWave);
      unsigned long long ob = 0, db = 0, qb = 0;
      uint32_t ol = 0, dlen = 0, pl = 0, dpl = 0, ql = 0;
      bool is_pk = false, d_pk = false;
      if (t < total) {
        const unsigned long long k = bj + (t - ex);
        const uint32_t v = kind == kSegP ? a.b2.col[k] : a.dl.col[k];
        const uint32_t rc = C * n + v;
        db = a.dl.ptr[rc];
        dlen = uint32_t(a.dl.ptr[rc + 1] - db);
        if (pk && a.dpk.len && dlen) {
          d_pk = true;
          dpl = a.dpk.len[rc];
        }
        if (kind == kSegD) {
          ob = a.old.ptr[rc];
          ol = uint32_t(a.old.ptr[rc + 1] - ob);
          if (pk && a.pk.len) {
            is_pk = true;
            pl = a.pk.len[rc];
          }
          if (PEND && a.pend.ptr) {
            qb = a.pend.ptr[rc];
            ql = uint32_t(a.pend.ptr[rc + 1] - qb);
          }
        }
      }
      visitw(a.pk.ent, is_pk ? ob : 0ull, is_pk ? pl : 0u, is_pk ? ol : 0u, tag);
      visitw(a.dpk.ent, d_pk ? db : 0ull, d_pk ? dpl : 0u, d_pk ? dlen : 0u, tag);
      visit(a.old.col, is_pk ? 0ull : ob, is_pk ? 0u : ol, tag);
      visit(a.dl.col, d_pk ? 0ull : db, d_pk ? 0u : dlen, tag);
      if (PEND && a.pend.ptr) visit(a.pend.col, qb, ql, tag);
    }
    ns = 0;
  };
  auto add = [&](uint32_t kind, uint32_t label, uint32_t row, bool pk, uint32_t tag) {
    if (uint32_t(lane) == ns) {
      s_desc = kind | (pk ? 4u : 0u) | (label << 8) | (tag << 16);
      s_row = row;
    }
    if (++ns == uint32_t(kWave)) run_group();
  };
  for (uint32_t j = 0; j < G; ++j) {
    const uint32_t A = uint32_t(__builtin_amdgcn_readlane(int(A_l), int(j)));
    const uint32_t u = uint32_t(__builtin_amdgcn_readlane(int(u_l), int(j)));
    for (int i = 0; i < a.g.nu; ++i)
      if (a.g.u_lhs[i] == A)
        add(kSegU, 0, uint32_t(a.g.u_rhs[i]) * n + u, packed_label(a.pk, uint32_t(a.g.u_rhs[i])), j);
    for (int i = 0; i < a.g.nb; ++i) {
      if (a.g.b_lhs[i] != A) continue;
      const uint32_t B = a.g.b_rhs1[i], C = a.g.b_rhs2[i];
      add(kSegD, C, B * n + u, packed_label(a.pk, C), j);
      const uint32_t slot = a.g.b_slot[i];
      if (slot != 0xFFu && a.b2.ptr) add(kSegP, C, slot * n + u, false, j);
    }
  }
  if (ns) run_group();
}

// The active rows (MODE 0, guessed windows of a.guess ids, a power of two
// that leaves room for at least two rows' bitmaps) in groups, see above.
// Rows that do not fit their group's window are appended to `defer`
// (ctr->deferred of them, also counted as misses).
template <bool PEND>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kGroupWaves, kGroupWaves))) void
k_expand_group(const ExpandArgs a, uint32_t* defer) {
  __shared__ GroupLds s_lds[kWavesPerBlock];
  const int lane = threadIdx.x & (kWave - 1);
  GroupLds& w = s_lds[threadIdx.x / kWave];
  const uint32_t nactive = a.ctr->active;
  if (nactive == 0) return;  // (launched before the host knew the count)
  if (lane < 3) w.acc[lane] = 0;
  const uint32_t n = a.n;
  const uint32_t span = a.guess, words = span >> 5, wshift = 31u - uint32_t(__clz(int(words)));
  const uint32_t gmax = min(min(kGroupRows, max(a.group, 1u)), kGroupWords / words);
  unsigned long long blk_pos = 0, blk_end = 0;
  auto reserve = [&](uint32_t cnt) -> unsigned long long {
    if (blk_pos + cnt > blk_end) {
      const unsigned long long take = max<unsigned long long>(cnt, kStageBlock);
      unsigned long long b = 0;
      if (lane == 0) b = atomicAdd(&a.ctr->cursor, take);
      b = uni_u64(b);
      blk_pos = b;
      blk_end = b + take;
    }
    const unsigned long long p = blk_pos;
    blk_pos += cnt;
    return p;
  };
  // each XCD first drains its own eighth of the active list (see k_expand)
  uint32_t seg = blockIdx.x % kXcds, visited = 0;
  for (;;) {
    uint32_t i0 = 0, G = 0;
    while (visited < kXcds) {
      const uint32_t sb = uint32_t((unsigned long long)nactive * seg / kXcds);
      const uint32_t se = uint32_t((unsigned long long)nactive * (seg + 1) / kXcds);
      uint32_t b = 0;
      if (lane == 0) b = atomicAdd(&a.ctr->work_xcd[seg].v, gmax);
      b = sb + uni_u32(b);
      if (b < se) {
        i0 = b;
        G = min(gmax, se - b);
        break;
      }
      seg = (seg + 1) % kXcds;
      ++visited;
    }
    if (!G) break;
    uint32_t r_l = 0, A_l = 0, u_l = 0;
    if (uint32_t(lane) < G) {
      r_l = a.active[i0 + lane];
      A_l = r_l / n;
      u_l = r_l - A_l * n;
    }
    // the shared window, word-aligned, inside [0, n) (n > span)
    const uint32_t um = uint32_t(__builtin_amdgcn_readlane(int(u_l), int(G / 2)));
    uint32_t c0 = (um > span / 2 ? um - span / 2 : 0u) & ~31u;
    if (c0 + span > n) c0 = (n - span) & ~31u;
    const uint32_t w0 = c0 >> 5;
    for (uint32_t i = lane; i < G * words; i += kWave) {
      w.seen[i] = 0;
      w.fresh[i] = 0;
    }
    if (lane < int(kGroupRows)) w.kacc[lane] = 0;
    wave_sync();
    // seeds: row j's old row (lane j), delta row (lane G + j) and pending
    // row (lane 2G + j); orl = that row's length in ids (the O count)
    unsigned long long orl = 0;
    {
      const uint32_t k = uint32_t(lane);
      const bool in_old = k < G, in_dl = k >= G && k < 2 * G, in_pd = PEND && a.pend.ptr && k >= 2 * G && k < 3 * G;
      const uint32_t j = in_old ? k : (in_dl ? k - G : (in_pd ? k - 2 * G : 0u));
      const uint32_t rj = uint32_t(__shfl(int(r_l), int(j), kWave));
      const uint32_t Aj = uint32_t(__shfl(int(A_l), int(j), kWave));
      unsigned long long pb = 0, ib = 0;
      uint32_t plen = 0, ilen = 0;
      const uint32_t* src = a.dl.col;
      if (in_old) {
        const unsigned long long ob = a.old.ptr[rj], oe = a.old.ptr[rj + 1];
        orl = oe - ob;
        if (a.pk.len && packed_label(a.pk, Aj)) {
          pb = ob;
          plen = a.pk.len[rj];
        } else {
          src = a.old.col;
          ib = ob;
          ilen = uint32_t(oe - ob);
        }
      } else if (in_dl) {
        ib = a.dl.ptr[rj];
        ilen = uint32_t(a.dl.ptr[rj + 1] - ib);
        orl = ilen;
      } else if (in_pd) {
        src = a.pend.col;
        ib = a.pend.ptr[rj];
        ilen = uint32_t(a.pend.ptr[rj + 1] - ib);
        orl = ilen;
      }
      const uint32_t woff = j * words;
      expand_stream_w(a.pk.ent, pb, plen, woff, w, lane, [&](unsigned long long e, uint32_t wo) {
        const uint32_t wi = uint32_t(e) - w0;
        if (wi < words) atomicOr(&w.seen[wo + wi], uint32_t(e >> 32));
      });
      expand_stream_w(src, ib, ilen, woff, w, lane, [&](uint32_t x, uint32_t wo) {
        const uint32_t off = x - c0;
        if (off < span) atomicOr(&w.seen[wo + (off >> 5)], 1u << (off & 31));
      });
    }
    wave_sync();
    // candidates: plain LDS ORs into the rows' fresh bitmaps; a candidate
    // outside the window marks its row missed
    uint32_t miss_l = 0;
    auto mark = [&](uint32_t x, uint32_t wo) {
      const uint32_t off = x - c0;
      if (off < span)
        __hip_atomic_fetch_or(&w.fresh[wo + (off >> 5)], 1u << (off & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      else
        miss_l |= 1u << (wo >> wshift);
    };
    auto mark_word = [&](unsigned long long e, uint32_t wo) {
      const uint32_t wi = uint32_t(e) - w0;
      if (wi < words)
        __hip_atomic_fetch_or(&w.fresh[wo + wi], uint32_t(e >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      else
        miss_l |= 1u << (wo >> wshift);
    };
    for_each_list_group<PEND>(
        a, G, A_l, u_l, lane,
        [&](const uint32_t* col, unsigned long long lb, uint32_t len, uint32_t tag) {
          if (len) __hip_atomic_fetch_add(&w.kacc[tag], (unsigned long long)len, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          expand_stream_w(col, lb, len, tag * words, w, lane, mark);
        },
        [&](const unsigned long long* ent, unsigned long long pb, uint32_t pl, uint32_t len, uint32_t tag) {
          if (len) __hip_atomic_fetch_add(&w.kacc[tag], (unsigned long long)len, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          expand_stream_w(ent, pb, pl, tag * words, w, lane, mark_word);
        });
    wave_sync();
    // per row: new ids = fresh minus seen, then its output or its deferral
    for (uint32_t j = 0; j < G; ++j) {
      const uint32_t rj = uint32_t(__builtin_amdgcn_readlane(int(r_l), int(j)));
      if (__ballot((miss_l >> j) & 1u)) {
        if (lane == 0) {
          defer[atomicAdd(&a.ctr->deferred, 1u)] = rj;
          atomicAdd(&ctr_slot(a.ctr).misses, 1u);
        }
        continue;
      }
      uint32_t* fr = w.fresh + j * words;
      const uint32_t* sn = w.seen + j * words;
      uint32_t c = 0;
      for (uint32_t i = lane; i < words; i += kWave) {
        const uint32_t f = fr[i] & ~sn[i];
        fr[i] = f;
        c += __popc(f);
      }
      c = wave_sum_u32(c);
      wave_sync();
      unsigned long long pos = 0;
      if (c) {
        pos = reserve(c);
        if (pos + c <= a.staging_cap)
          emit_chunk(a, c0, span, pos, fr, lane);
        else if (lane == 0)
          atomicExch(&a.ctr->overflow, 1u);
      }
      const unsigned long long kj = w.kacc[j];
      const unsigned long long oj = lane_u64(orl, j) + lane_u64(orl, G + j) + lane_u64(orl, min(2 * G + j, 63u));
      if (lane == 0) {
        a.seg_off[rj] = pos;
        a.seg_cnt[rj] = c;
        w.acc[0] += kj;
        w.acc[2] += c;
        if (kj) w.acc[1] += oj;
      }
    }
    wave_sync();
  }
  wave_sync();
  if (lane == 0) {
    CtrSlot& cs = ctr_slot(a.ctr);
    const unsigned long long kc = w.acc[0], orr = w.acc[1], ne = w.acc[2];
    if (kc) atomicAdd(&cs.cand, kc);
    if (orr) atomicAdd(&cs.oreads, orr);
    if (ne) atomicAdd(&cs.newe, ne);
  }
}

// ---------------------------------------------------------------------------
// Light rows: one lane per row. In rounds whose rows have few candidates the
// wave-per-row kernel spends a row's time on its chain of dependent loads
// (segment bounds, keys, key rows, lists, seed rows), one row per wave at a
// time. Here each lane follows that chain for a row of its own, so a wave
// keeps 64 rows in flight. A lane gathers its row's candidates into its
// column of an LDS buffer (at most kLightCap), sorts them, and keeps each
// distinct id absent from old_A(u), delta_A(u) and the pending edges (binary
// searches that only move forward, the ids being sorted). A row with more
// candidates goes to the deferred list (ctr->deferred rows at `defer`),
// which the wave kernel expands afterwards. The lists are those of the wave
// kernel (for_each_list, MODE 0) and K, O and N count as it does.
template <bool PEND, uint32_t kLightCap>
__global__ __launch_bounds__(kBlock) void k_expand_light(const ExpandArgs a, uint32_t* defer) {
  __shared__ uint32_t s_buf[kLightCap][kBlock];  // lane t's candidates in column t
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t t = threadIdx.x;
  const uint32_t n = a.n;
  const uint32_t nactive = a.ctr->active;
  const bool pend = PEND && a.pend.ptr;
  unsigned long long k_sum = 0, o_sum = 0, n_sum = 0;
  constexpr unsigned long long kLightStage = 1024;
  unsigned long long blk_pos = 0, blk_end = 0;  // this wave's staging block
  for (uint32_t base = blockIdx.x * blockDim.x; base < nactive; base += gridDim.x * blockDim.x) {
    const uint32_t idx = base + t;
    const bool live = idx < nactive;
    const uint32_t r = live ? a.active[idx] : 0u;
    const uint32_t A = r / n, u = r - A * n;
    uint32_t cnt = 0;
    bool over = false;
    unsigned long long krow = 0;
    auto put = [&](uint32_t x) {
      if (cnt < kLightCap) s_buf[cnt][t] = x;
      ++cnt;
    };
    // ids col[b, e)
    auto take_ids = [&](const uint32_t* col, unsigned long long b, unsigned long long e) {
      krow += e - b;
      if (over || e == b) return;
      if (cnt + (e - b) > kLightCap) {
        over = true;
        return;
      }
      for (unsigned long long k = b; k < e; ++k) put(col[k]);
    };
    // a packed row: pl entries at ent[b ..] standing for `ids` ids
    auto take_words = [&](const unsigned long long* ent, unsigned long long b, uint32_t pl, unsigned long long ids) {
      krow += ids;
      if (over || ids == 0) return;
      if (cnt + ids > kLightCap) {
        over = true;
        return;
      }
      for (uint32_t q = 0; q < pl; ++q) {
        const unsigned long long e = ent[b + q];
        const uint32_t w = uint32_t(e) << 5;
        for (uint32_t bits = uint32_t(e >> 32); bits; bits &= bits - 1) put(w + uint32_t(__ffs(bits) - 1));
      }
      over = cnt > kLightCap;
    };
    if (live) {
      for (int i = 0; i < a.g.nu; ++i)
        if (a.g.u_lhs[i] == A) {
          const uint32_t row = uint32_t(a.g.u_rhs[i]) * n + u;
          take_ids(a.dl.col, a.dl.ptr[row], a.dl.ptr[row + 1]);
        }
      for (int i = 0; i < a.g.nb && !over; ++i) {
        if (a.g.b_lhs[i] != A) continue;
        const uint32_t B = a.g.b_rhs1[i], C = a.g.b_rhs2[i];
        const bool pkc = a.pk.len && packed_label(a.pk, C);
        // keys delta_B(u): lists old_C(v), delta_C(v) and the pending C(v)
        const uint32_t kr = B * n + u;
        const unsigned long long ke = a.dl.ptr[kr + 1];
        for (unsigned long long k = a.dl.ptr[kr]; k < ke && !over; ++k) {
          const uint32_t rc = C * n + a.dl.col[k];
          const unsigned long long ob = a.old.ptr[rc], oe = a.old.ptr[rc + 1];
          if (pkc)
            take_words(a.pk.ent, ob, a.pk.len[rc], oe - ob);
          else
            take_ids(a.old.col, ob, oe);
          take_ids(a.dl.col, a.dl.ptr[rc], a.dl.ptr[rc + 1]);
          if (pend) take_ids(a.pend.col, a.pend.ptr[rc], a.pend.ptr[rc + 1]);
        }
        // keys of the frontier pair list: lists delta_C(v)
        const uint32_t slot = a.g.b_slot[i];
        if (slot != 0xFFu && a.b2.ptr) {
          const unsigned long long pr = (unsigned long long)slot * n + u;
          const unsigned long long pe = a.b2.ptr[2 * pr + 1];
          for (unsigned long long k = a.b2.ptr[2 * pr]; k < pe && !over; ++k) {
            const uint32_t rc = C * n + a.b2.col[k];
            take_ids(a.dl.col, a.dl.ptr[rc], a.dl.ptr[rc + 1]);
          }
        }
      }
    }
    uint32_t c = 0;
    if (live && !over && cnt) {
      for (uint32_t i = 1; i < cnt; ++i) {
        const uint32_t x = s_buf[i][t];
        uint32_t j = i;
        for (; j > 0 && s_buf[j - 1][t] > x; --j) s_buf[j][t] = s_buf[j - 1][t];
        s_buf[j][t] = x;
      }
      const bool pka = a.pk.len && packed_label(a.pk, A);
      unsigned long long ob = a.old.ptr[r];
      const unsigned long long oe = pka ? ob + a.pk.len[r] : a.old.ptr[r + 1];
      unsigned long long db = a.dl.ptr[r];
      const unsigned long long de = a.dl.ptr[r + 1];
      unsigned long long qb = 0, qe = 0;
      if (pend) {
        qb = a.pend.ptr[r];
        qe = a.pend.ptr[r + 1];
      }
      // ids are below n < 2^32 - 1, so no id equals the first `prev`
      uint32_t prev = 0xFFFFFFFFu;
      for (uint32_t i = 0; i < cnt; ++i) {
        const uint32_t x = s_buf[i][t];
        if (x == prev) continue;
        prev = x;
        bool seen;
        if (pka) {
          ob = gallop_pos_word(a.pk.ent, ob, oe, x >> 5);
          seen = ob < oe && uint32_t(a.pk.ent[ob]) == (x >> 5) && ((a.pk.ent[ob] >> (32 + (x & 31))) & 1ull);
        } else {
          ob = gallop_pos(a.old.col, ob, oe, x);
          seen = ob < oe && a.old.col[ob] == x;
        }
        if (!seen) {
          db = gallop_pos(a.dl.col, db, de, x);
          seen = db < de && a.dl.col[db] == x;
        }
        if (!seen && pend) {
          qb = gallop_pos(a.pend.col, qb, qe, x);
          seen = qb < qe && a.pend.col[qb] == x;
        }
        if (!seen) s_buf[c++][t] = x;
      }
    }
    // the wave's new ids go to its current staging block; a wave takes
    // kLightStage words of staging at a time (one atomic on the device-wide
    // cursor per block, not per 64 rows: on large vertex counts those
    // serialised), the rest of a block it leaves is a gap that the
    // compaction skips (rows are copied by their own offsets)
    const uint32_t incl = wave_incl_scan(c, lane);
    const uint32_t tot = lane63_u32(incl);
    unsigned long long wb = 0;
    if (tot) {
      if (blk_pos + tot > blk_end) {
        const unsigned long long take = max<unsigned long long>(tot, kLightStage);
        unsigned long long b0 = 0;
        if (lane == 0) b0 = atomicAdd(&a.ctr->cursor, take);
        b0 = uni_u64(b0);
        blk_pos = b0;
        blk_end = b0 + take;
      }
      wb = blk_pos;
      blk_pos += tot;
    }
    if (tot && wb + tot > a.staging_cap) {
      if (lane == 0) atomicExch(&a.ctr->overflow, 1u);
    } else if (live && !over) {
      const unsigned long long pos = wb + (incl - c);
      for (uint32_t i = 0; i < c; ++i) a.staging[pos + i] = s_buf[i][t];
      a.seg_off[r] = pos;
      a.seg_cnt[r] = c;
      k_sum += krow;
      n_sum += c;
      if (krow) {
        o_sum += (a.old.ptr[r + 1] - a.old.ptr[r]) + (a.dl.ptr[r + 1] - a.dl.ptr[r]);
        if (pend) o_sum += a.pend.ptr[r + 1] - a.pend.ptr[r];
      }
    }
    // rows with too many candidates go to the wave kernel
    const unsigned long long dm = __ballot(live && over);
    if (dm) {
      uint32_t dpos = 0;
      if (lane == 0) dpos = atomicAdd(&a.ctr->deferred, uint32_t(__popcll(dm)));
      dpos = uni_u32(dpos);
      if (live && over) defer[dpos + uint32_t(__popcll(dm & ((1ull << lane) - 1)))] = r;
    }
  }
  k_sum = wave_sum_u64(k_sum);
  o_sum = wave_sum_u64(o_sum);
  n_sum = wave_sum_u64(n_sum);
  if (lane == 0) {
    CtrSlot& cs = ctr_slot(a.ctr);
    if (k_sum) atomicAdd(&cs.cand, k_sum);
    if (o_sum) atomicAdd(&cs.oreads, o_sum);
    if (n_sum) atomicAdd(&cs.newe, n_sum);
  }
}

// The first round's dedup of the raw input rows, one lane per row: a row of
// at most kDedupCap raw entries is sorted in the lane's LDS column and its
// distinct ids are written out (old, delta and pending rows are all empty
// before the first round). Longer rows go to `defer` (ctr->deferred of them)
// for k_expand<1>.
constexpr uint32_t kDedupCap = 16;
__global__ __launch_bounds__(kBlock) void k_dedup_light(const ExpandArgs a, uint32_t* defer) {
  __shared__ uint32_t s_buf[kDedupCap][kBlock];
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t t = threadIdx.x;
  const uint32_t nactive = a.ctr->active;
  unsigned long long n_sum = 0;
  constexpr unsigned long long kStage = 1024;
  unsigned long long blk_pos = 0, blk_end = 0;  // this wave's staging block
  for (uint32_t base = blockIdx.x * blockDim.x; base < nactive; base += gridDim.x * blockDim.x) {
    const uint32_t idx = base + t;
    const bool live = idx < nactive;
    const uint32_t r = live ? a.active[idx] : 0u;
    unsigned long long b = 0, e = 0;
    if (live) {
      b = a.raw.ptr[r];
      e = a.raw.ptr[r + 1];
    }
    const bool over = live && e - b > kDedupCap;
    uint32_t c = 0;
    if (live && !over) {
      const uint32_t m = uint32_t(e - b);
      for (uint32_t i = 0; i < m; ++i) {
        const uint32_t x = a.raw.col[b + i];
        uint32_t j = i;
        for (; j > 0 && s_buf[j - 1][t] > x; --j) s_buf[j][t] = s_buf[j - 1][t];
        s_buf[j][t] = x;
      }
      for (uint32_t i = 0; i < m; ++i) {
        const uint32_t x = s_buf[i][t];
        if (c == 0 || s_buf[c - 1][t] != x) s_buf[c++][t] = x;
      }
    }
    const uint32_t incl = wave_incl_scan(c, lane);
    const uint32_t tot = lane63_u32(incl);
    unsigned long long wb = 0;
    if (tot) {
      if (blk_pos + tot > blk_end) {
        const unsigned long long take = max<unsigned long long>(tot, kStage);
        unsigned long long b0 = 0;
        if (lane == 0) b0 = atomicAdd(&a.ctr->cursor, take);
        b0 = uni_u64(b0);
        blk_pos = b0;
        blk_end = b0 + take;
      }
      wb = blk_pos;
      blk_pos += tot;
    }
    if (tot && wb + tot > a.staging_cap) {
      if (lane == 0) atomicExch(&a.ctr->overflow, 1u);
    } else if (live && !over && c) {
      const unsigned long long pos = wb + (incl - c);
      for (uint32_t i = 0; i < c; ++i) a.staging[pos + i] = s_buf[i][t];
      a.seg_off[r] = pos;
      a.seg_cnt[r] = c;
      n_sum += c;
    }
    const unsigned long long dm = __ballot(over);
    if (dm) {
      uint32_t dpos = 0;
      if (lane == 0) dpos = atomicAdd(&a.ctr->deferred, uint32_t(__popcll(dm)));
      dpos = uni_u32(dpos);
      if (over) defer[dpos + uint32_t(__popcll(dm & ((1ull << lane) - 1)))] = r;
    }
  }
  n_sum = wave_sum_u64(n_sum);
  if (lane == 0 && n_sum) atomicAdd(&ctr_slot(a.ctr).newe, n_sum);
}

// PEND: the round reads pending edges beside old (a separate instance, so
// rounds without them carry none of that code)
template <int MODE, bool PEND = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kExpandWaves, kExpandWaves))) void k_expand(
    const ExpandArgs a) {
  __shared__ WaveLds s_lds[kWavesPerBlock];
  const int lane = threadIdx.x & (kWave - 1);
  WaveLds& w = s_lds[threadIdx.x / kWave];
#ifdef GG_EXPAND_STATS
  if (lane < 4) w.xs[lane] = 0;
#endif
  // the first round's dedup (MODE 1) takes the active rows from active_base:
  // graphs joining a running session append their rows after the join's
  const uint32_t abase = MODE == 1 ? a.ctr->active_base : 0u;
  const bool from_defer = (MODE == 0 || MODE == 1) && a.defer != nullptr;
  const uint32_t* rows_in = from_defer ? a.defer : a.active + abase;
  const uint32_t nactive = from_defer ? a.ctr->deferred : a.ctr->active - abase;
  if (nactive == 0) return;  // (launched before the host knew the count)
  if (lane < 3) w.acc[lane] = 0;
  // rows are taken kGrab at a time and staging space kStageBlock words at a
  // time, so the two global counters see one atomic per many rows
  uint32_t next = 0, end = 0;
  // partition operators take at least 1/8 of a wave's even share per grab:
  // their rows are many and mostly light
  const uint32_t grab =
      MODE == 2 ? max(max(a.grab, 1u), nactive / (gridDim.x * kWavesPerBlock * 8u)) : max(a.grab, 1u);
  unsigned long long blk_pos = 0, blk_end = 0;
  auto reserve = [&](uint32_t cnt) -> unsigned long long {
    if (blk_pos + cnt > blk_end) {
      const unsigned long long grab = max<unsigned long long>(cnt, kStageBlock);
      unsigned long long b = 0;
      if (lane == 0) b = atomicAdd(&a.ctr->cursor, grab);
      b = uni_u64(b);  // lane 0's value: every lane is active here
      blk_pos = b;
      blk_end = b + grab;
    }
    const unsigned long long p = blk_pos;
    blk_pos += cnt;
    return p;
  };
  // Workgroups are dispatched to the XCDs round-robin, so blockIdx % 8 names
  // this block's XCD. Each XCD first drains its own eighth of the (vertex-
  // major) active list, then helps the others.
  uint32_t seg = blockIdx.x % kXcds, visited = 0;
  for (;;) {
    if (next >= end) {
      bool got = false;
      while (visited < kXcds) {
        const uint32_t sb = uint32_t((unsigned long long)nactive * seg / kXcds);
        const uint32_t se = uint32_t((unsigned long long)nactive * (seg + 1) / kXcds);
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(&a.ctr->work_xcd[seg].v, grab);
        b = sb + uni_u32(b);
        if (b < se) {
          next = b;
          end = min(b + grab, se);
          got = true;
          break;
        }
        seg = (seg + 1) % kXcds;
        ++visited;
      }
      if (!got) break;
    }
    const uint32_t idx = next++;
    const uint32_t r = uni_u32(rows_in[idx]);
    GG_CLK(t_row);
    const uint32_t rn = MODE == 2 ? a.ro.np : a.n;
    const uint32_t A = r / rn, u = r - A * rn + (MODE == 2 ? a.ro.first : 0u);
    unsigned long long krow = 0;
    uint32_t cnt = 0;
    unsigned long long pos = 0;
    // attempt 0: the a.guess ids centred on u (rows of graphs with id
    // locality fit; a candidate outside it sends the row to attempt 1);
    // attempt 1: the exact candidate window, in chunks when it is wider.
    bool empty = false;
    for (int attempt = (MODE == 0 && a.guess && a.n > a.guess) ? 0 : 1; attempt < 2; ++attempt) {
      uint32_t lo = 0, hi = a.n - 1;
      if (attempt == 0) {
        const uint32_t half = a.guess / 2;
        lo = (u > half ? u - half : 0u) & ~31u;
        if (lo + a.guess > a.n) lo = (a.n - a.guess) & ~31u;
        hi = min(a.n - 1, lo + a.guess - 1);
      } else if (a.n > kChunkBits) {
        uint32_t mn = 0xFFFFFFFFu, mx = 0;
        unsigned long long kk = 0;
        // (packed lists through their entries: a packed row's ids are not
        // kept during the fixpoint)
        for_each_list<MODE, MODE == 0, PEND>(a, r, A, u, lane, [&](const uint32_t* col, unsigned long long lb, uint32_t len) {
          if (MODE == 1) {
            // raw rows are unsorted: reduce over every element
            const unsigned long long b = (unsigned long long)uint32_t(__builtin_amdgcn_readlane(int(uint32_t(lb)), 0)) |
                                         ((unsigned long long)uint32_t(__builtin_amdgcn_readlane(int(uint32_t(lb >> 32)), 0)) << 32);
            const uint32_t l = uint32_t(__builtin_amdgcn_readlane(int(len), 0));
            for (uint32_t t = lane; t < l; t += kWave) {
              const uint32_t x = col[b + t];
              mn = min(mn, x);
              mx = max(mx, x);
            }
            kk += len;
          } else if (len) {
            mn = min(mn, col[lb]);
            mx = max(mx, col[lb + len - 1]);
            kk += len;
          }
        }, [&](const unsigned long long* ent, unsigned long long pb, uint32_t pl, uint32_t len) {
          if (pl) {
            mn = min(mn, uint32_t(ent[pb]) << 5);
            mx = max(mx, min((uint32_t(ent[pb + pl - 1]) << 5) | 31u, a.n - 1));
            kk += len;
          }
        });
        if (MODE == 2 && a.ro.kind == 2) {
          bool has_eps = false;
          for (int i = 0; i < a.g.ne; ++i) has_eps |= a.g.e_lhs[i] == A;
          if (has_eps && lane == 0) {
            mn = min(mn, u);
            mx = max(mx, u);
            kk += 1;
          }
        }
        kk = wave_sum_u64_small(kk);
        if (kk == 0) {
          empty = true;
          break;
        }
        lo = wave_min_u32(mn) & ~31u;  // word-aligned, for packed rows
        hi = wave_max_u32(mx);
      }
      const uint32_t nchunks = (hi - lo) / kChunkBits + 1;
      // windows wider than one chunk: a counting pass, one reservation, then
      // a writing pass; with sorted lists (every mode but the raw rows) each
      // pass visits only the windows that hold candidates
      const bool ranged = MODE != 1 && nchunks > 1;
      bool missed = false;
      uint32_t total = 0;
      unsigned long long p = 0;
      for (int pass = nchunks > 1 ? 0 : 1; pass < 2; ++pass) {
        if (pass == 1 && nchunks > 1) {
          cnt = total;
          if (total) p = pos = reserve(total);
          if (total && pos + total > a.staging_cap) {
            if (lane == 0) atomicExch(&a.ctr->overflow, 1u);
            break;
          }
          if (!total) break;
        }
        uint32_t c0 = lo;
        for (uint32_t ch = 0;; ++ch) {
          const uint32_t span = min(kChunkBits, hi - c0 + 1);
          unsigned long long kch = 0;
          bool miss = false;
          uint32_t nxt = 0xFFFFFFFFu;
          const uint32_t c = process_chunk<MODE, PEND>(a, r, A, u, c0, span, ranged, w, lane, kch, miss, nxt);
          if (attempt == 0 && miss) {
            missed = true;
            break;
          }
          // the next window: past the last one, or (ranged) at the next
          // candidate; c0 stays a multiple of 32 and grows every step
          bool last;
          uint32_t c0n = 0;
          if (ranged) {
            last = nxt == 0xFFFFFFFFu || nxt > hi;
            c0n = nxt & ~31u;
          } else {
            last = ch + 1 >= nchunks;
            c0n = c0 + kChunkBits;
          }
          if (pass == 0) {
            total += c;
            if (ranged)
              krow += kch;
            else if (ch == 0)
              krow = kch;
            if (last) break;
            c0 = c0n;
            continue;
          }
          if (nchunks == 1) {
            krow = kch;
            cnt = c;
            if (c) {
              pos = reserve(c);
              if (pos + c <= a.staging_cap)
                emit_chunk(a, c0, span, pos, w.fresh, lane);
              else if (lane == 0)
                atomicExch(&a.ctr->overflow, 1u);
            }
          } else {
            emit_chunk(a, c0, span, p, w.fresh, lane);
            p += c;
          }
          if (last) break;
          c0 = c0n;
        }
        if (missed) break;
      }
      if (!missed) break;
      if (lane == 0) atomicAdd(&ctr_slot(a.ctr).misses, 1u);
      krow = 0;
      cnt = 0;
      pos = 0;
    }
    GG_ACC(w, 3, t_row);
    if (empty) continue;
    if (lane == 0) {
      a.seg_off[r] = pos;
      a.seg_cnt[r] = cnt;
      __hip_atomic_fetch_add(&w.acc[0], krow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(&w.acc[2], (unsigned long long)cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (MODE == 0 && krow) {
        unsigned long long orr = (a.old.ptr[r + 1] - a.old.ptr[r]) + (a.dl.ptr[r + 1] - a.dl.ptr[r]);
        if (PEND && a.pend.ptr) orr += a.pend.ptr[r + 1] - a.pend.ptr[r];
        __hip_atomic_fetch_add(&w.acc[1], orr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
  }
  wave_sync();
  if (lane == 0) {
    CtrSlot& cs = ctr_slot(a.ctr);
    const unsigned long long kc = w.acc[0], orr = w.acc[1], ne = w.acc[2];
    if (kc) atomicAdd(&cs.cand, kc);
    if (orr) atomicAdd(&cs.oreads, orr);
    if (ne) atomicAdd(&cs.newe, ne);
#ifdef GG_EXPAND_STATS
    for (int i = 0; i < 4; ++i) atomicAdd(&cs.xs[i], w.xs[i]);
#endif
  }
}

__device__ inline uint32_t pack_union(const uint32_t* a, unsigned long long ia, unsigned long long ea,
                                      const uint32_t* b, unsigned long long ib, unsigned long long eb,
                                      unsigned long long* ent, uint32_t* win, int lane);
template <class T>
__device__ inline void flat_copy(const T* __restrict__ src, unsigned long long sb, T* __restrict__ dst,
                                 unsigned long long db, uint32_t len, unsigned long long* tmask,
                                 unsigned long long* sbase, unsigned long long* dbase, int lane);

// Copies each active row's staged output into the delta' CSR.
__global__ __launch_bounds__(kBlock) void k_gather(const uint32_t* active, const Counters* ctr, const uint32_t* seg_cnt,
                                                   const unsigned long long* seg_off, const uint32_t* staging,
                                                   const unsigned long long* out_ptr, uint32_t* out_col, uint32_t n,
                                                   DPacked plab, unsigned long long* out_pk, uint32_t* out_pklen) {
  // rows of packed labels are also packed, at the row's own offset (out_pk
  // null: no packing)
  __shared__ uint32_t s_win[kWavesPerBlock][kWave], s_word[kWavesPerBlock][kWave];
  __shared__ unsigned long long s_tmask[kWavesPerBlock][4], s_sb[kWavesPerBlock][kWave], s_db[kWavesPerBlock][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wi = threadIdx.x / kWave;
  uint32_t* win = s_win[wi];
  uint32_t* wword = s_word[wi];
  const unsigned long long upto = (2ull << lane) - 1;  // lanes 0..lane (lane 63: all, by unsigned wrap)
  const uint32_t nactive = ctr->active;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  // rows go to the waves in batches (up to 64; fewer when the list is short);
  // a batch's row descriptors are loaded at once, one row per lane, and its
  // ids move in one flattened copy
  const uint32_t per = max(1u, min(uint32_t(kWave), nactive / nwaves));
  for (uint32_t i0 = wave * per; i0 < nactive; i0 += nwaves * per) {
    const uint32_t nb = min(per, nactive - i0);
    uint32_t r_l = 0, c_l = 0;
    unsigned long long s_l = 0, d_l = 0;
    if (uint32_t(lane) < nb) {
      r_l = active[i0 + lane];
      c_l = seg_cnt[r_l];
      if (c_l) {
        s_l = seg_off[r_l];
        d_l = out_ptr[r_l];
      }
    }
    flat_copy(staging, s_l, out_col, d_l, c_l, s_tmask[wi], s_sb[wi], s_db[wi], lane);
    if (!out_pk) continue;
    const bool pkr = c_l != 0 && packed_label(plab, r_l / n);
    // rows of up to kLaneGather ids are packed by their own lane (the wave's
    // rows at once: the loop runs as long as the longest of them), the
    // others by the whole wave one after another below
    const bool mine = pkr && c_l <= kLaneGather;
    if (__ballot(mine)) {
      if (mine) {
        const uint32_t* ids = staging + s_l;
        unsigned long long* ent = out_pk + d_l;
        uint32_t wd = ids[0] >> 5, mask = 0, ne = 0;
        for (uint32_t i = 0; i < c_l; i += 4) {
          uint32_t x[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) x[q] = i + q < c_l ? ids[i + q] : 0u;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (i + q >= c_l) break;
            if ((x[q] >> 5) != wd) {
              ent[ne++] = ((unsigned long long)mask << 32) | wd;
              wd = x[q] >> 5;
              mask = 0;
            }
            mask |= 1u << (x[q] & 31);
          }
        }
        ent[ne++] = ((unsigned long long)mask << 32) | wd;
        out_pklen[r_l] = ne;
      }
    }
    unsigned long long todo = __ballot(pkr && !mine);
    while (todo) {
      const int k = __ffsll(todo) - 1;
      todo &= todo - 1;
      const uint32_t r = uint32_t(__builtin_amdgcn_readlane(int(r_l), k));
      const uint32_t c = uint32_t(__builtin_amdgcn_readlane(int(c_l), k));
      const unsigned long long s = lane_u64(s_l, uint32_t(k)), d = lane_u64(d_l, uint32_t(k));
      if (c <= uint32_t(kWave)) {
        // up to 64 sorted ids, one per lane: a lane whose word differs from
        // the previous lane's starts an entry; the entry index is the number
        // of starts up to the lane, the bits OR into that entry in LDS
        const bool in = uint32_t(lane) < c;
        const uint32_t x = in ? staging[s + uint32_t(lane)] : 0u;
        const uint32_t wd = x >> 5;
        const uint32_t prev = uint32_t(__shfl_up(int(wd), 1, kWave));
        const bool head = in && (lane == 0 || prev != wd);
        const unsigned long long hm = __ballot(head);
        const uint32_t ei = uint32_t(__popcll(hm & upto)) - 1u;
        win[lane] = 0;
        wave_sync();
        if (in) atomicOr(&win[ei], 1u << (x & 31));
        if (head) wword[ei] = wd;
        wave_sync();
        const uint32_t ne = uint32_t(__popcll(hm));
        if (uint32_t(lane) < ne) out_pk[d + uint32_t(lane)] = ((unsigned long long)win[lane] << 32) | wword[lane];
        if (lane == 0) out_pklen[r] = ne;
        wave_sync();
        continue;
      }
      const uint32_t pc = pack_union(staging, s, s + c, staging, s, s, out_pk + d, win, lane);
      if (lane == 0) out_pklen[r] = pc;
    }
  }
}

// Number of elements of sorted row [b, e) smaller than x.
__device__ inline uint32_t lower_rank(const uint32_t* col, unsigned long long b, unsigned long long e, uint32_t x) {
  unsigned long long lo = b, hi = e;
  while (lo < hi) {
    const unsigned long long mid = (lo + hi) >> 1;
    if (col[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return uint32_t(lo - b);
}

// Packs the union of the sorted id lists a[ia .. ea) and b[ib .. eb) into
// (word index, mask) entries at ent[0 ..]; returns the entry count. Ids are
// gathered 64 at a time into a 64-word LDS window (2048 ids) that is flushed
// in word order.
__device__ inline uint32_t pack_union(const uint32_t* a, unsigned long long ia, unsigned long long ea,
                                      const uint32_t* b, unsigned long long ib, unsigned long long eb,
                                      unsigned long long* ent, uint32_t* win, int lane) {
  const unsigned long long below = (1ull << lane) - 1;
  uint32_t written = 0;
  auto absorb = [&](const uint32_t* col, unsigned long long& i, unsigned long long e, uint32_t c0,
                    unsigned long long c1) {
    for (;;) {
      const unsigned long long t = i + lane;
      const uint32_t x = t < e ? col[t] : 0u;
      const bool in = t < e && x < c1;
      if (in) atomicOr(&win[(x - c0) >> 5], 1u << (x & 31));
      const uint32_t used = uint32_t(__popcll(__ballot(in)));
      i += used;
      if (used < kWave) break;
    }
  };
  while (ia < ea || ib < eb) {
    const uint32_t xa = ia < ea ? a[ia] : 0xFFFFFFFFu;
    const uint32_t xb = ib < eb ? b[ib] : 0xFFFFFFFFu;
    const uint32_t c0 = min(xa, xb) & ~31u;
    const unsigned long long c1 = (unsigned long long)c0 + 32u * kWave;
    win[lane] = 0;
    wave_sync();
    absorb(a, ia, ea, c0, c1);
    absorb(b, ib, eb, c0, c1);
    wave_sync();
    const uint32_t word = win[lane];
    const unsigned long long nz = __ballot(word != 0);
    if (word) ent[written + __popcll(nz & below)] = ((unsigned long long)word << 32) | ((c0 >> 5) + uint32_t(lane));
    written += uint32_t(__popcll(nz));
    wave_sync();
  }
  return written;
}

// Copies the lanes' segments src[sb .. sb+len) to dst[db ..] (one segment
// per lane, all flattened into one coalesced element stream, see
// stream_tile). sbase/dbase: 64-entry LDS tables of the wave; tmask: 4
// LDS words of the wave.
template <class T>
__device__ inline void flat_copy(const T* __restrict__ src, unsigned long long sb, T* __restrict__ dst,
                                 unsigned long long db, uint32_t len, unsigned long long* tmask,
                                 unsigned long long* sbase, unsigned long long* dbase, int lane) {
  constexpr int kU = 4;
  unsigned long long total;
  const unsigned long long incl = stream_prefix(len, total);
  if (total == 0) return;
  const unsigned long long start = incl - len;
  const uint32_t rank = stream_rank(len);
  wave_sync();
  if (len) {  // element t of segment k: src[sbase[k] + t] -> dst[dbase[k] + t]
    sbase[rank] = sb - start;
    dbase[rank] = db - start;
  }
  if (lane < kU) tmask[lane] = 0;
  wave_sync();
  for (unsigned long long t0 = 0; t0 < total; t0 += kWave * kU) {
    const unsigned long long rem = total - t0;
    const int lim = rem < kWave * kU ? int(rem) : kWave * kU;
    const unsigned long long t = t0 + uint32_t(lane);
    T x[kU];
    unsigned long long d[kU];
    const int nq = (lim + kWave - 1) / kWave;
    stream_tile<kU>(
        start, len, t0, tmask, lane,
        [&](int q, uint32_t rk) {
          d[q] = ~0ull;
          if (lane < lim - q * kWave) {
            x[q] = src[sbase[rk] + t + uint32_t(q * kWave)];
            d[q] = dbase[rk] + t + uint32_t(q * kWave);
          }
        },
        nq);
#pragma unroll
    for (int q = 0; q < kU; ++q)
      if (q < nq && d[q] != ~0ull) dst[d[q]] = x[q];
  }
  wave_sync();
}

// Copies n elements with the whole wave as 16-byte vectors (only the
// element's own alignment is assumed: global dwordx4 accesses need 4 bytes),
// 4 vectors in flight per lane, so a wave keeps 16 KB moving; the last
// n mod (16/sizeof(T)) elements go one by one.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef u32x4 u32x4_a4 __attribute__((aligned(4)));
template <class T>
__device__ inline void wave_copy(const T* __restrict__ src, T* __restrict__ dst, uint32_t n, int lane) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte elements");
  constexpr uint32_t E = 16 / sizeof(T);
  const uint32_t nv = n / E;
  const u32x4_a4* __restrict__ s = reinterpret_cast<const u32x4_a4*>(src);
  u32x4_a4* __restrict__ d = reinterpret_cast<u32x4_a4*>(dst);
  uint32_t t = uint32_t(lane);
  for (; t + 3 * kWave < nv; t += 4 * kWave) {
    const u32x4 a0 = s[t], a1 = s[t + kWave], a2 = s[t + 2 * kWave], a3 = s[t + 3 * kWave];
    d[t] = a0;
    d[t + kWave] = a1;
    d[t + 2 * kWave] = a2;
    d[t + 3 * kWave] = a3;
  }
  for (; t < nv; t += kWave) d[t] = s[t];
  for (uint32_t i = nv * E + uint32_t(lane); i < n; i += kWave) dst[i] = src[i];
}

// old' = old U delta for the labels whose delta is not empty
// (rowptr(old') = rowptr(old) + rowptr(delta)); the other labels are moved in
// bulk by k_copy_labels. A wave takes 64 consecutive rows. Short rows without
// delta only shift (by their delta rowptr): their ids and packed entries move
// in one flattened copy. Long unchanged rows, rows with delta and (before old
// is packed) the packed rows to build are appended to mlist for
// k_merge_rows, one wave per row.
// ids false: rows of packed labels (with out_pk) get packed entries only;
// their ids are written once after the fixpoint (k_unpack).
__global__ __launch_bounds__(kBlock) void k_merge(DCsr old, DCsr dl, uint32_t rows, uint32_t n, const Counters* ctr,
                                                  DPacked plab, const unsigned long long* old_pk, uint32_t* out_col,
                                                  unsigned long long* out_pk, uint32_t* pk_len, uint32_t* mlist,
                                                  uint32_t* mcnt, bool ids) {
  // ctr == nullptr: merge every row (no label is skipped)
  __shared__ unsigned long long s_tmask[kWavesPerBlock][4], s_sb[kWavesPerBlock][kWave], s_db[kWavesPerBlock][kWave];
  // listed rows gather per wave and take their place in mlist with one
  // counter update per kListBuf rows (updates of one address serialise)
  constexpr uint32_t kListBuf = 512;
  __shared__ uint32_t s_list[kWavesPerBlock][kListBuf];
  const int lane = threadIdx.x & (kWave - 1);
  const int wi = threadIdx.x / kWave;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  const uint32_t nblk = (rows + kWave - 1) / kWave;
  uint32_t nlist = 0;  // rows in s_list (uniform)
  auto flush = [&]() {
    wave_sync();
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(mcnt, nlist);
    pos = uni_u32(pos);  // lane 0: every lane is active here
    for (uint32_t i = uint32_t(lane); i < nlist; i += kWave) mlist[pos + i] = s_list[wi][i];
    nlist = 0;
    wave_sync();
  };
  for (uint32_t blk = wave; blk < nblk; blk += nwaves) {
    const uint32_t r = blk * kWave + lane;
    const uint32_t l = r / n;
    const bool live = r < rows && (!ctr || label_nonempty(ctr, l));
    unsigned long long ob = 0, oe = 0, db = 0, de = 0;
    if (live) {
      ob = old.ptr[r];
      oe = old.ptr[r + 1];
      db = dl.ptr[r];
      de = dl.ptr[r + 1];
    }
    const bool same = live && de == db && oe - ob <= kShortRow;  // short, no delta: the row only shifts
    const bool same_ids = same && (ids || !(out_pk && packed_label(plab, l)));
    // consecutive such rows are one contiguous range with one shift (no
    // delta between them): copied run by run, no per-id row search
    unsigned long long runs = __ballot(same_ids);
    while (runs) {
      const int k0 = __ffsll(runs) - 1;
      const unsigned long long from = runs >> k0;
      const int len = (~from == 0ull) ? kWave - k0 : __ffsll(~from) - 1;
      runs &= ~(len == kWave ? ~0ull : (((1ull << len) - 1) << k0));
      const unsigned long long sb = uni_u64((unsigned long long)__shfl((long long)ob, k0, kWave));
      const unsigned long long se = uni_u64((unsigned long long)__shfl((long long)oe, k0 + len - 1, kWave));
      const unsigned long long sh = uni_u64((unsigned long long)__shfl((long long)db, k0, kWave));
      if (se > sb) wave_copy(old.col + sb, out_col + sb + sh, uint32_t(se - sb), lane);
    }
    if (out_pk && old_pk) {
      const bool pkl = same && oe > ob && packed_label(plab, l);
      flat_copy(old_pk, ob, out_pk, ob + db, pkl ? pk_len[r] : 0u, s_tmask[wi], s_sb[wi], s_db[wi], lane);
    }
    const bool more = live && (!same || (out_pk && !old_pk && oe > ob && packed_label(plab, l)));
    const unsigned long long m = __ballot(more);
    if (m) {
      if (nlist + kWave > kListBuf) flush();
      if (more) s_list[wi][nlist + __popcll(m & ((1ull << lane) - 1))] = r;
      nlist += uint32_t(__popcll(m));
    }
  }
  if (nlist) flush();
}

// The rows k_merge listed, one wave per row. Old and delta are disjoint and
// sorted: when their ids fit one LDS window they are ORed into a bitmap that
// is written out in order, its non-zero words being the row's packed
// entries; wider rows place each id by its rank in the other list and are
// packed by pack_union. Packed rows record their entry count in pk_len[r].
// 7 waves per SIMD (a few registers spill; measured faster than 6 without)
__global__ __launch_bounds__(kBlock, 7) void k_merge_rows(DCsr old, DCsr dl, uint32_t n, DPacked plab,
                                                       const unsigned long long* old_pk, uint32_t* out_col,
                                                       unsigned long long* out_pk, uint32_t* pk_len,
                                                       const uint32_t* mlist, const uint32_t* mcnt, bool ids) {
  constexpr uint32_t kSlotWords = kMergeSlot / 32, kSlots = kMergeBits / kMergeSlot;
  __shared__ uint32_t s_win[kWavesPerBlock][kWave];
  __shared__ uint32_t s_bits[kWavesPerBlock][kMergeBits / 32];
  __shared__ unsigned long long s_tab[kWavesPerBlock][kWave], s_tmask[kWavesPerBlock][kPosUnroll];
  __shared__ uint32_t s_lw[kWavesPerBlock][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wi = threadIdx.x / kWave;
  uint32_t* win = s_win[wi];
  uint32_t* bits = s_bits[wi];
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  const uint32_t count = *mcnt;
  // rows go to the waves in batches (up to 64; fewer when the list is short,
  // so that every wave gets some). A batch's row bounds and first and last
  // ids are loaded at once, one row per lane, and the row loop reads them
  // with readlane instead of fetching them row after row.
  const uint32_t per = max(1u, min(uint32_t(kWave), count / nwaves));
  for (uint32_t i0 = wave * per; i0 < count; i0 += nwaves * per) {
    const uint32_t nb = min(per, count - i0);
    uint32_t rr_l = 0, lo_l = 0xFFFFFFFFu, hi_l = 0, pl_l = 0;
    unsigned long long rob_l = 0, roe_l = 0, rdb_l = 0, rde_l = 0;
    if (uint32_t(lane) < nb) {
      rr_l = mlist[i0 + lane];
      // packed entries of the old row (packed labels): the merge reads those
      const bool opk = old_pk && out_pk && packed_label(plab, rr_l / n);
      if (opk) pl_l = pk_len[rr_l];
      rob_l = old.ptr[rr_l];
      roe_l = old.ptr[rr_l + 1];
      rdb_l = dl.ptr[rr_l];
      rde_l = dl.ptr[rr_l + 1];
      if (opk && pl_l) {
        // the old row's first and last word (its ids may not be written)
        lo_l = uint32_t(old_pk[rob_l]) << 5;
        hi_l = min((uint32_t(old_pk[rob_l + pl_l - 1]) << 5) | 31u, n - 1);
      } else if (!opk && roe_l > rob_l) {
        lo_l = old.col[rob_l];
        hi_l = old.col[roe_l - 1];
      }
      if (rde_l > rdb_l) {
        lo_l = min(lo_l, dl.col[rdb_l]);
        hi_l = max(hi_l, dl.col[rde_l - 1]);
      }
    }
    // the next row's first 64 old entries (packed labels) and delta ids are
    // loaded while the current row merges; a row that fits them uses them
    // instead of loading its own
    unsigned long long e_nxt = 0;
    uint32_t d_nxt = 0;
    auto prefetch = [&](uint32_t j) {
      const uint32_t rj = uint32_t(__builtin_amdgcn_readlane(int(rr_l), int(j)));
      const bool pj = old_pk && out_pk && packed_label(plab, rj / n);
      const uint32_t oj = pj ? uint32_t(__builtin_amdgcn_readlane(int(pl_l), int(j))) : 0u;
      const unsigned long long robj = lane_u64(rob_l, j), rdbj = lane_u64(rdb_l, j);
      const uint32_t dj = uint32_t(lane_u64(rde_l, j) - rdbj);
      e_nxt = uint32_t(lane) < oj ? old_pk[robj + uint32_t(lane)] : 0ull;
      d_nxt = uint32_t(lane) < dj ? dl.col[rdbj + uint32_t(lane)] : 0u;
    };
    // packed rows with delta and at most kLaneMerge old entries and delta ids
    // (while their ids are not written) are merged by their own lane: a
    // sorted merge of the entries and the ids' words
    bool mine = false;
    if (uint32_t(lane) < nb && old_pk && out_pk && !ids && rde_l > rdb_l)
      mine = packed_label(plab, rr_l / n) && pl_l + (rde_l - rdb_l) <= kLaneMerge;
    if (__ballot(mine)) {
      if (mine) {
        const unsigned long long* oe = old_pk + rob_l;
        const uint32_t* di = dl.col + rdb_l;
        unsigned long long* out = out_pk + rob_l + rdb_l;
        const uint32_t dn = uint32_t(rde_l - rdb_l);
        uint32_t io = 0, id = 0, ne = 0;
        unsigned long long e = pl_l ? oe[0] : 0ull;
        uint32_t x = di[0];
        while (io < pl_l || id < dn) {
          // words are below 2^27: 0xFFFFFFFF marks an exhausted list
          const uint32_t we = io < pl_l ? uint32_t(e) : 0xFFFFFFFFu;
          const uint32_t wx = id < dn ? x >> 5 : 0xFFFFFFFFu;
          const uint32_t wmin = min(we, wx);
          uint32_t mask = 0;
          if (we == wmin) {
            mask = uint32_t(e >> 32);
            if (++io < pl_l) e = oe[io];
          }
          while (id < dn && (x >> 5) == wmin) {
            mask |= 1u << (x & 31);
            if (++id < dn) x = di[id];
          }
          out[ne++] = ((unsigned long long)mask << 32) | wmin;
        }
        pk_len[rr_l] = ne;
      }
    }
    // packed rows with delta (ids not written) whose old and delta ids span
    // fewer than kMergeSlot ids: merged kSlots at a time, each in its own
    // kMergeSlot-bit slot of the wave's bitmap, the set's old entries and
    // delta ids going through two flattened streams
    bool grp = false;
    if (uint32_t(lane) < nb && !mine && old_pk && out_pk && !ids && rde_l > rdb_l)
      grp = packed_label(plab, rr_l / n) && hi_l - (lo_l & ~31u) < kMergeSlot;
    for (unsigned long long gm = __ballot(grp); gm;) {
      // lane k < S takes the k-th row of the set
      int src = 0;
      uint32_t S = 0;
      for (; S < kSlots && gm; ++S) {
        const int p = __ffsll(gm) - 1;
        gm &= gm - 1;
        if (uint32_t(lane) == S) src = p;
      }
      const bool in = uint32_t(lane) < S;
      const uint32_t rr_s = uint32_t(__shfl(int(rr_l), src, kWave));
      const uint32_t pl_s = uint32_t(__shfl(int(pl_l), src, kWave));
      const uint32_t lo_s = uint32_t(__shfl(int(lo_l), src, kWave)) & ~31u;
      const unsigned long long rob_s = (unsigned long long)__shfl((long long)rob_l, src, kWave);
      const unsigned long long rdb_s = (unsigned long long)__shfl((long long)rdb_l, src, kWave);
      const unsigned long long rde_s = (unsigned long long)__shfl((long long)rde_l, src, kWave);
      for (uint32_t i = lane; i < S * kSlotWords; i += kWave) bits[i] = 0;
      // word w of slot k: bits[w + lw], lw = k * kSlotWords - first word
      s_lw[wi][lane] = uint32_t(lane) * kSlotWords - (lo_s >> 5);
      wave_sync();
      flat_tiles<kPosUnroll>(rob_s, in ? pl_s : 0u, s_tab[wi], s_tmask[wi], lane,
                             [&](const int* own, const unsigned long long* pos, const bool* ok) {
                               unsigned long long e[kPosUnroll];
#pragma unroll
                               for (int q = 0; q < kPosUnroll; ++q) e[q] = ok[q] ? old_pk[pos[q]] : 0ull;
#pragma unroll
                               for (int q = 0; q < kPosUnroll; ++q)
                                 if (ok[q])
                                   __hip_atomic_fetch_or(&bits[uint32_t(e[q]) + s_lw[wi][own[q]]], uint32_t(e[q] >> 32),
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                             });
      flat_tiles<kPosUnroll>(rdb_s, in ? uint32_t(rde_s - rdb_s) : 0u, s_tab[wi], s_tmask[wi], lane,
                             [&](const int* own, const unsigned long long* pos, const bool* ok) {
                               uint32_t x[kPosUnroll];
#pragma unroll
                               for (int q = 0; q < kPosUnroll; ++q) x[q] = ok[q] ? dl.col[pos[q]] : 0u;
#pragma unroll
                               for (int q = 0; q < kPosUnroll; ++q)
                                 if (ok[q])
                                   __hip_atomic_fetch_or(&bits[(x[q] >> 5) + s_lw[wi][own[q]]], 1u << (x[q] & 31),
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                             });
      wave_sync();
      // each row's non-empty words are its merged packed entries
      const unsigned long long below = (1ull << lane) - 1;
      for (uint32_t j = 0; j < S; ++j) {
        const uint32_t rj = uint32_t(__builtin_amdgcn_readlane(int(rr_s), int(j)));
        const uint32_t lw0 = uint32_t(__builtin_amdgcn_readlane(int(lo_s), int(j))) >> 5;
        unsigned long long* out = out_pk + lane_u64(rob_s, j) + lane_u64(rdb_s, j);
        uint32_t pe = 0;
        for (uint32_t w0 = 0; w0 < kSlotWords; w0 += kWave) {
          const uint32_t wd = bits[j * kSlotWords + w0 + uint32_t(lane)];
          const unsigned long long nz = __ballot(wd != 0);
          if (wd) out[pe + __popcll(nz & below)] = ((unsigned long long)wd << 32) | (lw0 + w0 + uint32_t(lane));
          pe += uint32_t(__popcll(nz));
        }
        if (lane == 0) pk_len[rj] = pe;
      }
      wave_sync();
    }
    // the other rows: one after another, the next one's first old entries
    // and delta ids loading while the current one merges
    unsigned long long rest = __ballot(uint32_t(lane) < nb && !mine && !grp);
    int kn = rest ? __ffsll(rest) - 1 : -1;
    if (kn >= 0) prefetch(uint32_t(kn));
    for (int kk = kn; kk >= 0; kk = kn) {
      const uint32_t k = uint32_t(kk);
      rest &= rest - 1;  // (kk is its lowest bit)
      kn = rest ? __ffsll(rest) - 1 : -1;
      const unsigned long long e_cur = e_nxt;
      const uint32_t d_cur = d_nxt;
      if (kn >= 0) prefetch(uint32_t(kn));
      const uint32_t rr = uint32_t(__builtin_amdgcn_readlane(int(rr_l), int(k)));
      const uint32_t lr = rr / n;
      const unsigned long long rob = lane_u64(rob_l, k), roe = lane_u64(roe_l, k);
      const unsigned long long rdb = lane_u64(rdb_l, k), rde = lane_u64(rde_l, k);
      uint32_t* out = out_col + rob + rdb;
      const uint32_t ol = uint32_t(roe - rob), dln = uint32_t(rde - rdb);
      const bool pkr = out_pk && packed_label(plab, lr);
      if (dln == 0 && (!pkr || old_pk)) {
        // a long unchanged row: ids and packed entries move as they are
        if (!pkr || ids) wave_copy(old.col + rob, out, ol, lane);
        if (pkr) wave_copy(old_pk + rob, out_pk + rob + rdb, uint32_t(__builtin_amdgcn_readlane(int(pl_l), int(k))), lane);
        continue;
      }
      uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(lo_l), int(k)));
      const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(hi_l), int(k)));
      lo &= ~31u;
      if (hi - lo < kMergeBits) {
        const uint32_t words = ((hi - lo) >> 5) + 1;
        for (uint32_t t = lane; t < words; t += kWave) bits[t] = 0;
        wave_sync();
        if (pkr && old_pk) {
          // the old row word by word: 32 ids per entry
          const uint32_t opl = uint32_t(__builtin_amdgcn_readlane(int(pl_l), int(k)));
          if (opl <= uint32_t(kWave)) {
            if (uint32_t(lane) < opl)
              __hip_atomic_fetch_or(&bits[uint32_t(e_cur) - (lo >> 5)], uint32_t(e_cur >> 32), __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_WORKGROUP);
          } else {
            for (uint32_t t = lane; t < opl; t += kWave) {
              const unsigned long long e = old_pk[rob + t];
              __hip_atomic_fetch_or(&bits[uint32_t(e) - (lo >> 5)], uint32_t(e >> 32), __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_WORKGROUP);
            }
          }
        } else {
          for (uint32_t t = lane; t < ol; t += kWave) {
            const uint32_t off = old.col[rob + t] - lo;
            __hip_atomic_fetch_or(&bits[off >> 5], 1u << (off & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        }
        if (dln <= uint32_t(kWave)) {
          if (uint32_t(lane) < dln) {
            const uint32_t off = d_cur - lo;
            __hip_atomic_fetch_or(&bits[off >> 5], 1u << (off & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        } else {
          for (uint32_t t = lane; t < dln; t += kWave) {
            const uint32_t off = dl.col[rdb + t] - lo;
            __hip_atomic_fetch_or(&bits[off >> 5], 1u << (off & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        }
        wave_sync();
        const unsigned long long below = (1ull << lane) - 1;
        // dense rows (4+ ids per word on average): two words per wave step,
        // one bit per lane, so each step stores its ids contiguously; sparse
        // rows: one word per lane, its bits stored one by one
        const bool wid = !pkr || ids;  // the row's ids are written (else only its packed entries)
        const bool dense = wid && uint64_t(ol) + dln >= 4ull * words;
        unsigned long long o = 0;
        uint32_t pe = 0;
        if (dense) {
          for (uint32_t j = 0; j < words; j += 2) {
            const uint32_t wj = j + uint32_t(lane >> 5);
            const uint32_t wd = wj < words ? bits[wj] : 0u;
            const bool set = (wd >> (lane & 31)) & 1u;
            const unsigned long long m = __ballot(set);
            if (set) out[o + __popcll(m & below)] = lo + wj * 32 + uint32_t(lane & 31);
            o += uint32_t(__popcll(m));
          }
        }
        for (uint32_t w0 = 0; w0 < words; w0 += kWave) {
          const uint32_t wd = w0 + lane < words ? bits[w0 + lane] : 0u;
          if (!dense && wid) {
            const uint32_t pc = __popc(wd);
            const uint32_t incl = wave_incl_scan(pc, lane);
            uint32_t x = wd;
            unsigned long long q = o + (incl - pc);
            while (x) {
              const int b = __ffs(x) - 1;
              x &= x - 1;
              out[q++] = lo + (w0 + uint32_t(lane)) * 32 + uint32_t(b);
            }
            o += lane63_u32(incl);
          }
          if (pkr) {
            const unsigned long long nz = __ballot(wd != 0);
            if (wd)
              out_pk[rob + rdb + pe + __popcll(nz & below)] =
                  ((unsigned long long)wd << 32) | ((lo >> 5) + w0 + uint32_t(lane));
            pe += uint32_t(__popcll(nz));
          }
        }
        if (pkr && lane == 0) pk_len[rr] = pe;
        wave_sync();
        continue;
      }
      if (pkr && old_pk && !ids) {
        // a wide packed row whose ids are not written: LDS windows of
        // kMergeBits ids, each starting at the smallest id not merged yet;
        // the old entries and delta ids inside a window are found by binary
        // search (both are sorted), ORed into the bitmap and its non-zero
        // words written as entries. Every window takes at least one element.
        const unsigned long long below = (1ull << lane) - 1;
        const uint32_t opl = uint32_t(__builtin_amdgcn_readlane(int(pl_l), int(k)));
        unsigned long long io = rob, id = rdb;
        const unsigned long long ie = rob + opl;
        uint32_t pe = 0;
        for (;;) {
          uint32_t c0 = 0xFFFFFFFFu;  // no real window starts there: window starts are multiples of 32
          if (io < ie) c0 = uint32_t(old_pk[io]) << 5;
          if (id < rde) c0 = min(c0, dl.col[id] & ~31u);
          c0 = uni_u32(c0);
          if (c0 == 0xFFFFFFFFu) break;
          const unsigned long long c1 = (unsigned long long)c0 + kMergeBits;
          const unsigned long long jo = uni_u64(lower_pos_word(old_pk, io, ie, uint32_t(c1 >> 5)));
          const unsigned long long jd = uni_u64(lower_pos(dl.col, id, rde, c1));
          for (uint32_t t = lane; t < kMergeBits / 32; t += kWave) bits[t] = 0;
          wave_sync();
          for (unsigned long long i = io + lane; i < jo; i += kWave) {
            const unsigned long long e = old_pk[i];
            __hip_atomic_fetch_or(&bits[uint32_t(e) - (c0 >> 5)], uint32_t(e >> 32), __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_WORKGROUP);
          }
          for (unsigned long long i = id + lane; i < jd; i += kWave) {
            const uint32_t off = dl.col[i] - c0;
            __hip_atomic_fetch_or(&bits[off >> 5], 1u << (off & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
          wave_sync();
          for (uint32_t w0 = 0; w0 < kMergeBits / 32; w0 += kWave) {
            const uint32_t wd = bits[w0 + lane];
            const unsigned long long nz = __ballot(wd != 0);
            if (wd)
              out_pk[rob + rdb + pe + __popcll(nz & below)] =
                  ((unsigned long long)wd << 32) | ((c0 >> 5) + w0 + uint32_t(lane));
            pe += uint32_t(__popcll(nz));
          }
          wave_sync();
          io = jo;
          id = jd;
        }
        if (lane == 0) pk_len[rr] = pe;
        continue;
      }
      if (dln == 0) {
        for (uint32_t t = lane; t < ol; t += kWave) out[t] = old.col[rob + t];
      } else if (ol == 0) {
        for (uint32_t t = lane; t < dln; t += kWave) out[t] = dl.col[rdb + t];
      } else {
        for (uint32_t t = lane; t < ol; t += kWave) {
          const uint32_t x = old.col[rob + t];
          out[t + lower_rank(dl.col, rdb, rde, x)] = x;
        }
        for (uint32_t t = lane; t < dln; t += kWave) {
          const uint32_t x = dl.col[rdb + t];
          out[t + lower_rank(old.col, rob, roe, x)] = x;
        }
      }
      if (pkr) {
        const uint32_t c = pack_union(old.col, rob, roe, dl.col, rdb, rde, out_pk + rob + rdb, win, lane);
        if (lane == 0) pk_len[rr] = c;
      }
    }
  }
}

// Writes the ids of every row of a packed label from its packed entries (the
// merges keep only the entries of those rows during the fixpoint). Rows go
// to the waves in batches of up to 64 (one row per lane: its offset and
// entry count load at once); the wave then expands the batch's non-empty rows one
// after another, in order, which fills exactly the row's [ptr[r], ptr[r+1]).
__global__ __launch_bounds__(kBlock) void k_unpack(const unsigned long long* ptr, uint32_t* col, DPacked plab,
                                                   uint32_t rows, uint32_t n) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  // up to 64 rows per batch, fewer when there are few rows (so that every
  // wave gets some: few long rows must not queue on a few waves)
  const uint32_t per = max(1u, min(uint32_t(kWave), rows / nwaves));
  for (uint32_t r0 = wave * per; r0 < rows; r0 += nwaves * per) {
    const uint32_t r = r0 + uint32_t(lane);
    uint32_t pl = 0;
    unsigned long long b = 0;
    if (uint32_t(lane) < per && r < rows && packed_label(plab, r / n)) {
      pl = plab.len[r];
      if (pl) b = ptr[r];
    }
    unsigned long long todo = __ballot(pl != 0);
    // a row's first 64 entries, one per lane; the next row's load while the
    // current row expands
    auto first = [&](int k) -> unsigned long long {
      const uint32_t pk = uint32_t(__builtin_amdgcn_readlane(int(pl), k));
      const unsigned long long bk = lane_u64(b, uint32_t(k));
      return uint32_t(lane) < pk ? plab.ent[bk + uint32_t(lane)] : 0ull;
    };
    unsigned long long ev_next = todo ? first(__ffsll(todo) - 1) : 0ull;
    while (todo) {
      const int k = __ffsll(todo) - 1;
      todo &= todo - 1;
      const uint32_t pk = uint32_t(__builtin_amdgcn_readlane(int(pl), k));
      const unsigned long long bk = lane_u64(b, uint32_t(k));
      const unsigned long long ev0 = ev_next;
      if (todo) ev_next = first(__ffsll(todo) - 1);
      // the row's entries 64 at a time, one per lane; then two entries per
      // wave step (moved to lanes 0-31 and 32-63 by shuffle), one bit per
      // lane, so each step stores its ids contiguously
      const unsigned long long below = (1ull << lane) - 1;
      unsigned long long o = bk;
      for (uint32_t j0 = 0; j0 < pk; j0 += kWave) {
        const uint32_t t = j0 + uint32_t(lane);
        const unsigned long long ev = j0 == 0 ? ev0 : (t < pk ? plab.ent[bk + t] : 0ull);
        const uint32_t nloc = min(uint32_t(kWave), pk - j0);
        for (uint32_t j = 0; j < nloc; j += 2) {
          const uint32_t src = j + uint32_t(lane >> 5);  // <= 64: lane 64 wraps to 0 and is masked below
          const uint32_t lo = uint32_t(__shfl(int(uint32_t(ev)), int(src), kWave));
          const uint32_t hi = uint32_t(__shfl(int(uint32_t(ev >> 32)), int(src), kWave));
          const bool set = src < nloc && ((hi >> (lane & 31)) & 1u);
          const unsigned long long m = __ballot(set);
          if (set) col[o + __popcll(m & below)] = (lo << 5) + uint32_t(lane & 31);
          o += uint32_t(__popcll(m));
        }
      }
    }
  }
}

// Labels whose delta is empty keep their rows: their id (and packed) segment
// moves in one piece, by the label's constant delta rowptr.
__global__ __launch_bounds__(kBlock) void k_copy_labels(DCsr old, DCsr dl, uint32_t n, const Counters* ctr, DPacked plab,
                                                        const unsigned long long* old_pk, uint32_t* out_col,
                                                        unsigned long long* out_pk, bool ids) {
  const uint32_t l = blockIdx.y;
  if (label_nonempty(ctr, l)) return;
  const unsigned long long b = old.ptr[uint64_t(l) * n], e = old.ptr[uint64_t(l + 1) * n];
  const unsigned long long shift = dl.ptr[uint64_t(l) * n];
  const bool pk = out_pk && old_pk && packed_label(plab, l);
  const bool wid = ids || !pk;  // else the label's ids are not kept during the fixpoint
  for (unsigned long long i = b + blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < e;
       i += uint64_t(gridDim.x) * blockDim.x) {
    if (wid) out_col[i + shift] = old.col[i];
    if (pk) out_pk[i + shift] = old_pk[i];
  }
}

// Adds to acc the hashes of the lanes' rows, one row per lane: lane k's row
// has len_p packed entries (at ent[b ..]) or len_i ids (at col[b ..]), its
// part of the hash is crow = (src - base) * kHashS ^ label * kHashL, and
// every dst is numbered from base (the same on every lane). The rows' elements
// go as one flattened stream (flat_tiles). row: 64-entry LDS table of the
// wave; bit[b] = b * kHashD.
__device__ inline void hash_rows(const unsigned long long* __restrict__ ent, const uint32_t* __restrict__ col,
                                 unsigned long long b, uint32_t len_p, uint32_t len_i, uint32_t base,
                                 unsigned long long crow, unsigned long long* tab, unsigned long long* tmask,
                                 unsigned long long* row, const unsigned long long* bit, int lane,
                                 unsigned long long& acc) {
  constexpr int U = kPosUnroll;
  row[lane] = crow;
  wave_sync();
  flat_tiles<U>(b, len_p, tab, tmask, lane, [&](const int* own, const unsigned long long* pos, const bool* ok) {
    unsigned long long en[U];
#pragma unroll
    for (int q = 0; q < U; ++q) en[q] = ok[q] ? ent[pos[q]] : 0ull;
#pragma unroll
    for (int q = 0; q < U; ++q) {
      if (!ok[q]) continue;
      const unsigned long long cr = row[own[q]];
      // (word * 32 - base) * kHashD modulo 2^64 (the word may start below
      // base); each set bit b adds b * kHashD
      const unsigned long long dk = (uint64_t(uint32_t(en[q])) * 32u - uint64_t(base)) * kHashD;
      for (uint32_t bits = uint32_t(en[q] >> 32); bits; bits &= bits - 1)
        acc += hash_mix(cr ^ (dk + bit[__ffs(bits) - 1]));
    }
  });
  flat_tiles<U>(b, len_i, tab, tmask, lane, [&](const int* own, const unsigned long long* pos, const bool* ok) {
    uint32_t x[U];
#pragma unroll
    for (int q = 0; q < U; ++q) x[q] = ok[q] ? col[pos[q]] : 0u;
#pragma unroll
    for (int q = 0; q < U; ++q)
      if (ok[q]) acc += hash_mix(row[own[q]] ^ (uint64_t(x[q] - base) * kHashD));
  });
  wave_sync();
}

// Sum of triple_hash over the closure (its ids: rows of packed labels are
// written first). A wave takes 64 consecutive rows, one per lane, and streams
// their ids together (hash_rows).
__global__ __launch_bounds__(kBlock) void k_checksum(DCsr c, uint32_t rows, uint32_t n, Counters* ctr) {
  __shared__ unsigned long long s_tab[kWavesPerBlock][kWave], s_tmask[kWavesPerBlock][kPosUnroll],
      s_row[kWavesPerBlock][kWave];
  __shared__ unsigned long long s_bit[32];
  if (threadIdx.x < 32) s_bit[threadIdx.x] = uint64_t(threadIdx.x) * kHashD;
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int wi = threadIdx.x / kWave;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  unsigned long long acc = 0;
  for (uint32_t r0 = wave * kWave; r0 < rows; r0 += nwaves * kWave) {
    const uint32_t r = r0 + uint32_t(lane);
    unsigned long long b = 0, e = 0;
    if (r < rows) {
      b = c.ptr[r];
      e = c.ptr[r + 1];
    }
    const uint32_t l = r / n, u = r - l * n;
    hash_rows(nullptr, c.col, b, 0u, uint32_t(e - b), 0u, uint64_t(u) * kHashS ^ uint64_t(l + 1) * kHashL, s_tab[wi],
              s_tmask[wi], s_row[wi], s_bit, lane, acc);
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(&ctr_slot(ctr).checksum, acc);
}

// ---------------------------------------------------------------------------
// per-graph results of a batched closure: graph g owns vertices
// [v0[g], v0[g+1]) and the combined labels [l0[g], l0[g] + lg[g]); in its own
// numbering a vertex is (id - v0[g]) and a label (combined - l0[g] + 1).

struct DSlices {
  const uint32_t* v0;    // G + 1
  const uint32_t* vg;    // graph of every vertex
  const uint32_t* l0;    // G
  const uint32_t* lg;    // G
  const uint32_t* coff;  // G: offset of graph g's per-label counts
  uint32_t G;
};

__global__ void k_slice_counts(const unsigned long long* ptr, uint32_t n, DSlices s, unsigned long long* out) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < s.G; g += gridDim.x * blockDim.x) {
    const uint32_t a = s.v0[g], b = s.v0[g + 1];
    for (uint32_t i = 0; i < s.lg[g]; ++i) {
      const uint64_t base = uint64_t(s.l0[g] + i) * n;
      out[s.coff[g] + i] = ptr[base + b] - ptr[base + a];
    }
  }
}

// plab.len set: rows of packed labels are read from their packed entries
// (their ids are not written).
// A wave takes 64 consecutive rows, one per lane (bounds, graph, the row's
// part of the hash), and streams their elements as one flattened stream
// (flat_tiles): packed entries first, then id rows. The lanes' sums belong to
// one graph; they go to its checksum when the wave moves on to another graph.
// A batch whose rows belong to several graphs (at a graph's first or last
// rows) is taken row by row.
__global__ __launch_bounds__(kBlock) void k_slice_checksum(DCsr c, DPacked plab, uint32_t rows, uint32_t n, DSlices s,
                                                           unsigned long long* sums) {
  constexpr int U = kPosUnroll;
  __shared__ unsigned long long s_tab[kWavesPerBlock][kWave], s_tmask[kWavesPerBlock][U], s_row[kWavesPerBlock][kWave];
  __shared__ unsigned long long s_bit[32];  // b * kHashD: the hash's dst part steps by it within a word
  if (threadIdx.x < 32) s_bit[threadIdx.x] = uint64_t(threadIdx.x) * kHashD;
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int wi = threadIdx.x / kWave;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
  uint32_t gcur = 0xFFFFFFFFu;
  unsigned long long acc = 0;  // this lane's part of graph gcur's sum
  auto flush = [&](uint32_t g) {
    if (g == gcur) return;
    const unsigned long long t = wave_sum_u64(acc);
    if (lane == 0 && t) atomicAdd(&sums[gcur], t);
    gcur = g;
    acc = 0;
  };
  for (uint32_t r0 = wave * kWave; r0 < rows; r0 += nwaves * kWave) {
    const uint32_t r = r0 + uint32_t(lane);
    unsigned long long b = 0, e = 0;
    if (r < rows) {
      b = c.ptr[r];
      e = c.ptr[r + 1];
    }
    const bool live = e > b;
    const uint32_t l = r / n, u = r - l * n;
    uint32_t g = 0, base = 0, len_p = 0, len_i = 0;
    unsigned long long crow = 0;
    if (live) {
      g = s.vg[u];
      base = s.v0[g];
      // the row's part of every hash: (u - base) * kHashS ^ label * kHashL
      crow = uint64_t(u - base) * kHashS ^ uint64_t(l - s.l0[g] + 1) * kHashL;
      if (plab.len && packed_label(plab, l))
        len_p = plab.len[r];
      else
        len_i = uint32_t(e - b);
    }
    const unsigned long long lm = __ballot(live);
    if (!lm) continue;
    const int f = __ffsll(lm) - 1;
    const uint32_t g0 = uint32_t(__builtin_amdgcn_readlane(int(g), f));
    if (__ballot(live && g != g0) == 0) {
      flush(g0);
      hash_rows(plab.ent, c.col, b, len_p, len_i, uint32_t(__builtin_amdgcn_readlane(int(base), f)), crow, s_tab[wi],
                s_tmask[wi], s_row[wi], s_bit, lane, acc);
      continue;
    }
    // rows of several graphs: one after another, all lanes on each row
    for (unsigned long long todo = lm; todo; todo &= todo - 1) {
      const int k = __ffsll(todo) - 1;
      const uint32_t gk = uint32_t(__builtin_amdgcn_readlane(int(g), k));
      flush(gk);
      const uint32_t bk = uint32_t(__builtin_amdgcn_readlane(int(base), k));
      const unsigned long long ck = lane_u64(crow, uint32_t(k)), rb = lane_u64(b, uint32_t(k));
      const uint32_t lp = uint32_t(__builtin_amdgcn_readlane(int(len_p), k));
      const uint32_t li = uint32_t(__builtin_amdgcn_readlane(int(len_i), k));
      for (uint32_t i = uint32_t(lane); i < lp; i += kWave) {
        const unsigned long long en = plab.ent[rb + i];
        const unsigned long long dk = (uint64_t(uint32_t(en)) * 32u - uint64_t(bk)) * kHashD;
        for (uint32_t bits = uint32_t(en >> 32); bits; bits &= bits - 1)
          acc += hash_mix(ck ^ (dk + s_bit[__ffs(bits) - 1]));
      }
      for (uint32_t i = uint32_t(lane); i < li; i += kWave) acc += hash_mix(ck ^ (uint64_t(c.col[rb + i] - bk) * kHashD));
    }
  }
  flush(0xFFFFFFFEu);
}

// ---------------------------------------------------------------------------
// host side

#define GG_HIP(expr)                                                                           \
  do {                                                                                         \
    hipError_t e__ = (expr);                                                                   \
    if (e__ != hipSuccess) {                                                                   \
      err = std::string(#expr) + ": " + hipGetErrorString(e__);                                \
      return GG_E_DEVICE;                                                                      \
    }                                                                                          \
  } while (0)

#define GG_TRY(expr)        \
  do {                      \
    int s__ = (expr);       \
    if (s__ != GG_OK) return s__; \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// Launch and schedule parameters that the autotuner (graspan_g.tune) picks
// on the device and writes to tuned_gfx950.cfg beside the library ("name
// value" lines, '#' comments). GG_TUNE_<NAME> environment variables override
// the file, read at the start of every run, so a tuner can sweep in one
// process. 0 for a *_bpc value keeps the occupancy the kernel allows.
struct Tuning {
  // k_expand takes active rows from the work counters grab_light at a time,
  // or grab_heavy at a time in a round whose previous round joined more than
  // heavy_cand candidates per active row (a few heavy rows per grab balance
  // the waves better; many light rows per grab save counter atomics)
  uint32_t grab_light = 4;
  uint32_t grab_heavy = 4;
  uint64_t heavy_cand = 4096;
  uint32_t guess_min = 4096;   // k_expand: first guessed window (ids)
  uint64_t defer_ratio = 128;  // deltas go to the pending CSR while (pending + delta) * ratio <= old
  int expand_bpc = 0;          // blocks per CU of k_expand
  int merge_bpc = 0;           // of k_merge
  int merge_rows_bpc = 0;      // of k_merge_rows
  int gather_bpc = 0;          // of k_gather
  double group_mem_share = 0.4;  // batch: share of device memory a launch group plans for
  // a round whose previous round joined at most light_cand candidates per
  // active row (or the first round) runs the lane-per-row pass first
  // (k_expand_light); 0: never
  uint64_t light_cand = 32;
  // rows per wave in k_expand_group (at most 8, and as many guessed windows'
  // bitmaps as fit the wave's LDS: 4 of 4096 ids)
  uint32_t group_rows = 4;
  uint32_t light_cap = 32;  // candidates a lane holds in the light pass: 16, 32 or 64
};

// Sets one parameter from its name; false for an unknown name.
bool set_tuning(Tuning& t, const std::string& name, const char* v) {
  if (name == "grab_light") t.grab_light = uint32_t(std::strtoul(v, nullptr, 10));
  else if (name == "grab_heavy") t.grab_heavy = uint32_t(std::strtoul(v, nullptr, 10));
  else if (name == "heavy_cand") t.heavy_cand = std::strtoull(v, nullptr, 10);
  else if (name == "guess_min") t.guess_min = uint32_t(std::strtoul(v, nullptr, 10));
  else if (name == "defer_ratio") t.defer_ratio = std::strtoull(v, nullptr, 10);
  else if (name == "expand_bpc") t.expand_bpc = std::atoi(v);
  else if (name == "merge_bpc") t.merge_bpc = std::atoi(v);
  else if (name == "merge_rows_bpc") t.merge_rows_bpc = std::atoi(v);
  else if (name == "gather_bpc") t.gather_bpc = std::atoi(v);
  else if (name == "group_mem_share") t.group_mem_share = std::atof(v);
  else if (name == "light_cand") t.light_cand = std::strtoull(v, nullptr, 10);
  else if (name == "group_rows") t.group_rows = uint32_t(std::strtoul(v, nullptr, 10));
  else if (name == "light_cap") t.light_cap = uint32_t(std::strtoul(v, nullptr, 10));
  else return false;
  return true;
}

const char* const kTuningNames[] = {"grab_light", "grab_heavy",     "heavy_cand", "guess_min",      "defer_ratio",
                                    "expand_bpc", "merge_bpc",      "merge_rows_bpc", "gather_bpc", "group_mem_share",
                                    "light_cand", "light_cap", "group_rows"};

// The tuning file beside this library (read once), then GG_TUNE_* variables.
Tuning load_tuning() {
  static Tuning from_file;
  static std::once_flag once;
  std::call_once(once, [] {
    Dl_info info{};
    if (!dladdr(reinterpret_cast<void*>(&load_tuning), &info) || !info.dli_fname) return;
    std::string path(info.dli_fname);
    const size_t slash = path.rfind('/');
    path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuned_gfx950.cfg";
    if (const char* f = std::getenv("GG_TUNE_FILE")) path = f;
    FILE* fp = std::fopen(path.c_str(), "r");
    if (!fp) return;
    char line[256];
    while (std::fgets(line, sizeof line, fp)) {
      char name[64], value[64];
      if (line[0] == '#' || std::sscanf(line, "%63s %63s", name, value) != 2) continue;
      set_tuning(from_file, name, value);
    }
    std::fclose(fp);
  });
  Tuning t = from_file;
  for (const char* n : kTuningNames) {
    std::string env = "GG_TUNE_";
    for (const char* c = n; *c; ++c) env += char(std::toupper(static_cast<unsigned char>(*c)));
    if (const char* v = std::getenv(env.c_str())) set_tuning(t, n, v);
  }
  return t;
}

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
};

// The timing events of one run, by phase (none are created when disabled).
struct RunEvents {
  bool enabled = true;
  std::vector<EventPair> expand, compact, merge, other;
  EventPair& mk(std::vector<EventPair>& v) {
    v.emplace_back();
    if (enabled) {
      (void)hipEventCreate(&v.back().a);
      (void)hipEventCreate(&v.back().b);
    }
    return v.back();
  }
  ~RunEvents() {
    for (auto* v : {&expand, &compact, &merge, &other})
      for (auto& e : *v) {
        if (e.a) (void)hipEventDestroy(e.a);
        if (e.b) (void)hipEventDestroy(e.b);
      }
  }
};

// What the round loop needs that depends only on the grammar and the vertex
// count (plan_rounds).
struct RoundPlan {
  ActArgs act{};
  PairPlan pair_plan{};
  uint64_t act_work = 0, NS = 0;
  uint32_t fr_words = 0;
  uint32_t n = 0;  // the vertex count it was made for
  uint64_t pair_arena_forced = 0;
};

}  // namespace

// A device region carved by best fit, with its free list on the host; one
// region may serve several engines (the serving loop's slots), so its list is
// guarded by a mutex. `owned`: the region was allocated for this object and is
// freed with it (the allocation-free entry point's region is a static array).
struct DeviceRegion {
  char* base = nullptr;
  size_t cap = 0;
  bool owned = false;
  std::mutex mu;
  std::map<size_t, size_t> free_list;  // offset -> bytes, neighbours merged
  ~DeviceRegion() {
    if (owned && base) (void)hipFree(base);
  }
  void put(size_t off, size_t bytes) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = free_list.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_list.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_list.erase(next);
    }
    if (it != free_list.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_list.erase(it);
      }
    }
  }
  // best fit: the smallest free block that holds the request, so that the
  // large blocks stay whole for the buffers that grow every round
  bool take(size_t bytes, size_t& off) {
    std::lock_guard<std::mutex> lk(mu);
    auto best = free_list.end();
    for (auto it = free_list.begin(); it != free_list.end(); ++it)
      if (it->second >= bytes && (best == free_list.end() || it->second < best->second)) best = it;
    if (best == free_list.end()) return false;
    off = best->first;
    const size_t left = best->second - bytes;
    free_list.erase(best);
    if (left) free_list.emplace(off + bytes, left);
    return true;
  }
};

class Engine {
 public:
  int device = 0;
  int cus = 256;
  uint64_t budget = 0;
  int verbose = 0;
  hipStream_t stream = nullptr;
  size_t used = 0, peak = 0;
  std::function<bool()> on_pressure;  // frees device memory, if it can, when an allocation does not fit
  // called on the host after each semi-naive round with the round's active
  // rows and all rows (batched lanes hand the device on in a closure's tail)
  std::function<void(uint64_t, uint64_t)> on_round;
  double alloc_ms = 0;                // host time spent allocating and freeing device memory
  uint64_t alloc_calls = 0;

  DevGrammar g{};
  int L = 0;
  std::vector<uint8_t> lhs;  // labels that are the lhs of a unary/binary rule
  bool has_grammar = false;
  // The grammar's rule table as given (the partitioned schedule runs it)
  // and as planned for the least fixpoint (plan_fixpoint_rules: same
  // closure, fewer join candidates). A run loads one of them into g, lhs
  // and pack_labels before it starts.
  struct RuleSet {
    DevGrammar g{};
    std::vector<uint8_t> lhs;
    uint32_t pack[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  };
  RuleSet rules_given, rules_planned;
  void use_rules(const RuleSet& r) {
    g = r.g;
    lhs = r.lhs;
    for (int i = 0; i < 8; ++i) pack_labels[i] = r.pack[i];
  }


  uint32_t n = 0;
  uint64_t m = 0;
  DevBuf in_src, in_dst, in_lab;
  bool has_graph = false;

  // closure state
  DevBuf old_ptr, old_col, dl_ptr, dl_col, nx_ptr, nx_col, o2_ptr, o2_col;
  // pending edges: the rounds' deltas since old was last rebuilt (unpacked
  // CSR, same row numbering); old ∪ pend is the closure so far
  DevBuf pd_ptr, pd_col, p2_ptr, p2_col;
  uint64_t pend_nnz = 0;
  // a round's delta goes to pend while (pend + delta) * defer_ratio <= old
  // (0: never); GG_DEFER_RATIO overrides the default
  uint64_t defer_ratio = kDeferRatio;
  // rows of packed labels keep their ids during the fixpoint (false: only
  // their packed entries, the ids written once after the last round);
  // GG_KEEP_PACKED_IDS=1 sets it
  bool keep_ids = false;
  // defer_unpack: run() leaves the packed labels' ids unwritten (ids_pending)
  // until a caller needs them (ensure_ids); results that can be read from the
  // packed entries (the batch's per-graph counts and checksums) never need
  // them
  bool defer_unpack = false, ids_pending = false;
  DevBuf raw_ptr, raw_col, staging, seg_off, seg_cnt, active, tiles, ctr;
  // frontier bitmaps and pair lists for old_B x delta_C
  DevBuf fr_bits, pr_ptr, pr_col;
  uint64_t pr_arena = 0;  // words per arena of the pair lists (one arena per counter slot)
  // packed old rows (labels in pack_labels)
  DevBuf pk_len, pk_ent, pk2_ent;
  // packed delta rows (current and next round), entries at the row's delta offset
  DevBuf dpk_len, dpk_ent, npk_len, npk_ent;
  bool dpk_ready = false;
  uint32_t pack_labels[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool pk_ready = false;
  // Guessed window (ids) for the next round: starts small; a round where
  // more than 1/8 of the rows missed moves it up a size (4096, 16384, off),
  // and after a round without guessing it is tried again at 16384.
  uint32_t guess_bits = kGuessMin;
  uint32_t supersteps_r = 0;  // supersteps of the last partitioned-schedule run
  // host copy of the input (the partitioned schedule builds its regions from it)
  std::vector<uint32_t> h_src, h_dst;
  std::vector<uint8_t> h_lab;
  uint64_t old_nnz = 0, dl_nnz = 0;
  bool pairs_ready = false;
  bool has_closure = false;
  Counters* h_ctr = nullptr;  // pinned

  // Fixed-region mode (the allocation-free entry point): every buffer is
  // carved from [arena, arena + arena_cap), a device region the engine does
  // not own, by first fit over a free list kept on the host; the stream, the
  // input edges and the host copies of the counters belong to the caller or
  // to this object. Nothing is allocated or freed through HIP in this mode,
  // and no events are created (the statistics' times stay 0).
  char* arena = nullptr;                 // region->base when buffers come from a region
  std::shared_ptr<DeviceRegion> region;
  bool owns_stream = true;
  Counters arena_h_ctr{};
  unsigned long long arena_h_u64 = 0;
  uint32_t expand_grid_cap = 0;  // 0: as many blocks as fit the device
  uint64_t light_rows = 0;       // rows the current round's light pass expanded
  bool grouped = false;          // the current round's rows went through k_expand_group
  static constexpr uint64_t kUnknown = ~0ull;
  uint64_t last_deferred = kUnknown;  // rows the last grouped round deferred to k_expand (kUnknown: none yet)
  Tuning tune;                   // the tuned parameters of the current run
  int occ_expand = 4, occ_merge = 8, occ_merge_rows = 8, occ_gather = 7;  // blocks per CU the kernels allow

  // session state (a running fixpoint that graphs join at round boundaries)
  RoundPlan sess_rp;
  gg_stats sess_S{};
  uint64_t sess_inject = 0;  // raw edges waiting in raw_ptr/raw_col for the next round
  bool sess_timing = false;  // time a session's kernels with events (statistics only)

  ~Engine() {
    DevBuf* all[] = {&in_src, &in_dst, &in_lab, &old_ptr, &old_col, &dl_ptr, &dl_col,  &nx_ptr, &nx_col,
                     &o2_ptr, &o2_col, &raw_ptr, &raw_col, &staging, &seg_off, &seg_cnt, &active,
                     &tiles,  &ctr,    &fr_bits,
                     &pr_ptr, &pr_col, &pk_len, &pk_ent, &pk2_ent, &dpk_len, &dpk_ent, &npk_len, &npk_ent,
                     &mlist,  &mcnt,   &pd_ptr, &pd_col, &p2_ptr, &p2_col};
    if (stream) (void)hipStreamSynchronize(stream);
    if (!arena) {
      for (DevBuf* b : all)
        if (b->p) (void)hipFree(b->p);
      for (auto& kv : cache) (void)hipFree(kv.second);
    } else {
      // a shared region outlives this engine: its blocks go back to it
      for (DevBuf* b : all)
        if (b->p && static_cast<char*>(b->p) >= arena && static_cast<char*>(b->p) < arena + region->cap)
          region->put(size_t(static_cast<char*>(b->p) - arena), b->bytes);
    }
    if (h_ctr && h_ctr != &arena_h_ctr) (void)hipHostFree(h_ctr);
    if (h_u64 && h_u64 != &arena_h_u64) (void)hipHostFree(h_u64);
    if (stream && owns_stream) (void)hipStreamDestroy(stream);
  }



  // Device memory. Buffers that are released or outgrown go to a per-engine
  // block cache and are handed out again for requests up to half their size:
  // the round loop's ping-pong buffers grow every round and then trade the
  // same blocks, instead of returning multi-GB blocks to the driver and
  // allocating fresh ones each round (large fresh allocations were seen to
  // cost seconds per closure once memory had been freed before). The cache is
  // trimmed (hipFree) when an allocation would not fit the budget or the
  // device, and when a context releases its closure.
  std::multimap<size_t, void*> cache;  // size -> block
  size_t cached = 0;

  void trim_cache() {
    if (cache.empty()) return;
    (void)hipStreamSynchronize(stream);
    for (auto& kv : cache) (void)hipFree(kv.second);
    cache.clear();
    cached = 0;
  }

  void put_block(DevBuf& b) {
    if (!b.p) return;
    if (arena) {
      // A shared region hands a block to whichever engine asks next, on its
      // own stream: this engine's stream must be done with the block first
      // (kernels launched before the release may still read or write it).
      // The fixed scratch region of the allocation-free entry point has one
      // engine per call on the caller's stream and needs no wait.
      if (region->owned) (void)hipStreamSynchronize(stream);
      region->put(size_t(static_cast<char*>(b.p) - arena), b.bytes);
    } else
      cache.emplace(b.bytes, b.p);
    if (!arena) cached += b.bytes;
    used -= b.bytes;
    b.p = nullptr;
    b.bytes = 0;
  }

  int arena_alloc(DevBuf& b, size_t bytes, std::string& err) {
    put_block(b);
    bytes = (std::max<size_t>(bytes, 256) + 255) & ~size_t(255);
    size_t off = 0;
    // engines that share a region may each be held to a budget of it
    if (budget && used + bytes > budget) {
      err = "device budget exceeded: need " + std::to_string(used + bytes) + " bytes of the region, budget " +
            std::to_string(budget);
      return GG_E_NOMEM;
    }
    if (region->take(bytes, off)) {
      b.p = arena + off;
      b.bytes = bytes;
      used += bytes;
      peak = std::max(peak, used);
      return GG_OK;
    }
    if (region->owned)
      err = "the device region (" + std::to_string(region->cap) + " bytes) cannot hold a further " +
            std::to_string(bytes) + " bytes (" + std::to_string(used) + " in use by this engine)";
    else
      err = "the fixed scratch region (" + std::to_string(region->cap) + " bytes) cannot hold a further " +
            std::to_string(bytes) + " bytes (" + std::to_string(used) + " in use): the graph or its closure is "
            "too large for the allocation-free entry point";
    return GG_E_NOMEM;
  }

  int alloc(DevBuf& b, size_t bytes, std::string& err) {
    if (b.p && b.bytes >= bytes) return GG_OK;
    if (arena) return arena_alloc(b, bytes, err);
    const auto a0 = std::chrono::steady_clock::now();
    struct Clock {
      double& ms;
      std::chrono::steady_clock::time_point t;
      ~Clock() { ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count(); }
    } clk{alloc_ms, a0};
    put_block(b);
    bytes = std::max<size_t>(bytes, 256);
    if (bytes >= (2u << 20)) bytes = (bytes + (2u << 20) - 1) & ~size_t((2u << 20) - 1);
    // a cached block of at least this size and at most twice it
    auto it = cache.lower_bound(bytes);
    if (it != cache.end() && it->first <= 2 * bytes) {
      b.p = it->second;
      b.bytes = it->first;
      cached -= it->first;
      cache.erase(it);
      used += b.bytes;
      peak = std::max(peak, used);
      return GG_OK;
    }
    ++alloc_calls;
    // under memory pressure: first the cache goes back, then the out-of-core
    // schedule may move a partition to host DRAM (on_pressure returns true
    // when it freed something)
    for (;;) {
      if (budget && used + cached + bytes > budget) {
        if (cached) {
          trim_cache();
          continue;
        }
        if (on_pressure && on_pressure()) continue;
        err = "device budget exceeded: need " + std::to_string(used + bytes) + " bytes, budget " +
              std::to_string(budget);
        return GG_E_NOMEM;
      }
      hipError_t e = hipMalloc(&b.p, bytes);
      if (e == hipSuccess) break;
      (void)hipGetLastError();
      b.p = nullptr;
      if (cached) {
        trim_cache();
        continue;
      }
      if (on_pressure && on_pressure()) continue;
      err = "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e);
      return GG_E_NOMEM;
    }
    b.bytes = bytes;
    used += bytes;
    peak = std::max(peak, used);
    return GG_OK;
  }

  // Grows geometrically (x1.5) so per-round reallocations stay rare.
  int grow(DevBuf& b, size_t bytes, std::string& err) {
    if (b.p && b.bytes >= bytes) return GG_OK;
    size_t want = std::max(bytes, b.bytes + b.bytes / 2);
    if (budget && used - b.bytes + want > budget) want = bytes;  // no slack under the budget's edge
    int s = alloc(b, want, err);
    if (s == GG_E_NOMEM && want > bytes) s = alloc(b, bytes, err);
    return s;
  }

  // The buffer's block goes to the cache (its memory stays with the engine
  // until the cache is trimmed).
  void release(DevBuf& b) { put_block(b); }

  uint32_t grid_for(uint64_t work, uint32_t per_block = kBlock) const {
    uint64_t g = (work + per_block - 1) / per_block;
    uint64_t cap = uint64_t(cus) * 16;
    return uint32_t(std::max<uint64_t>(1, std::min(g, cap)));
  }

  template <class T>
  T* P(DevBuf& b) {
    return static_cast<T*>(b.p);
  }
  Counters* C() { return static_cast<Counters*>(ctr.p); }

  void fill_u32(DevBuf& b, uint32_t v, uint64_t count) {
    k_fill_u32<<<grid_for(count), kBlock, 0, stream>>>(P<uint32_t>(b), v, count);
  }
  void fill_u64(DevBuf& b, unsigned long long v, uint64_t count) {
    k_fill_u64<<<grid_for(count), kBlock, 0, stream>>>(P<unsigned long long>(b), v, count);
  }
  void zero_ctr() {
    k_fill_u32<<<1, kBlock, 0, stream>>>(P<uint32_t>(ctr), 0u, sizeof(Counters) / 4);
  }

  // counts (rows entries) -> out_ptr (rows+1 entries, last = total)
  int scan(DevBuf& counts, uint64_t rows, DevBuf& out_ptr, std::string& err) {
    const uint64_t ntiles = (rows + kScanTile - 1) / kScanTile;
    GG_TRY(grow(tiles, (ntiles + 1) * 8, err));
    k_scan_tiles<<<uint32_t(ntiles), kBlock, 0, stream>>>(P<uint32_t>(counts), rows, P<unsigned long long>(tiles));
    k_scan_sums<<<1, kBlock, 0, stream>>>(P<unsigned long long>(tiles), ntiles, P<unsigned long long>(out_ptr) + rows);
    k_scan_final<<<uint32_t(ntiles), kBlock, 0, stream>>>(P<uint32_t>(counts), rows, P<unsigned long long>(tiles),
                                                         P<unsigned long long>(out_ptr));
    GG_HIP(hipGetLastError());
    return GG_OK;
  }

  // Reads one uint64 from device memory (synchronises the stream).
  int read_u64(const void* dptr, uint64_t& out, std::string& err) {
    unsigned long long v = 0;
    GG_HIP(hipMemcpyAsync(h_u64, dptr, 8, hipMemcpyDeviceToHost, stream));
    GG_HIP(hipStreamSynchronize(stream));
    v = *h_u64;
    out = v;
    return GG_OK;
  }
  unsigned long long* h_u64 = nullptr;  // pinned

  int read_ctr(std::string& err) {
    GG_HIP(hipMemcpyAsync(h_ctr, ctr.p, sizeof(Counters), hipMemcpyDeviceToHost, stream));
    GG_HIP(hipStreamSynchronize(stream));
    Counters& c = *h_ctr;
    c.cand = c.pkeys = c.oreads = c.newe = c.checksum = 0;
    c.misses = 0;
    for (auto& x : c.xs) x = 0;
    for (const CtrSlot& q : c.slot) {
      c.cand += q.cand;
      c.pkeys += q.pkeys;
      c.oreads += q.oreads;
      c.newe += q.newe;
      c.checksum += q.checksum;
      c.misses += q.misses;
      for (int i = 0; i < 4; ++i) c.xs[i] += q.xs[i];
    }
    return GG_OK;
  }

  DPacked plabels() const {
    DPacked p{};
    for (int i = 0; i < 8; ++i) p.labels[i] = pack_labels[i];
    return p;
  }
  // delta rows of packed labels are packed too (GG_NO_PACK turns both off)
  bool pack_delta() const {
    bool any = false;
    for (int i = 0; i < 8; ++i) any |= pack_labels[i] != 0;
    return any;
  }

  // out = old U delta row by row (k_merge, then k_merge_rows on the rows
  // with delta); ctr == nullptr merges every label
  DevBuf mlist, mcnt;
  // old' = old ∪ x (x disjoint from old, xn edges), packed rows rebuilt for
  // the packed labels. flags: the label flags of x (labels without x edges
  // move in one piece), or null to merge every row.
  int fold_into_old(DCsr xc, uint64_t xn, const Counters* flags, std::string& err) {
    const uint32_t R = rows();
    bool pack_any = false;
    for (int i = 0; i < 8; ++i) pack_any |= pack_labels[i] != 0;
    const uint64_t o2n = old_nnz + xn;
    GG_TRY(grow(o2_col, std::max<uint64_t>(o2n, 1) * 4, err));
    k_add_ptr<<<grid_for(uint64_t(R) + 1), kBlock, 0, stream>>>(P<unsigned long long>(o2_ptr), P<unsigned long long>(old_ptr),
                                                               xc.ptr, uint64_t(R) + 1);
    DPacked plab{};
    for (int i = 0; i < 8; ++i) plab.labels[i] = pack_labels[i];
    unsigned long long* opk = nullptr;
    if (pack_any) {
      GG_TRY(grow(pk2_ent, std::max<uint64_t>(o2n, 1) * 8, err));
      opk = P<unsigned long long>(pk2_ent);
    }
    const DCsr oc{P<unsigned long long>(old_ptr), P<uint32_t>(old_col)};
    GG_TRY(merge_rows(oc, xc, R, n, flags, plab, pk_ready ? P<unsigned long long>(pk_ent) : nullptr,
                      P<uint32_t>(o2_col), opk, P<uint32_t>(pk_len), err, keep_ids));
    if (old_nnz && flags)
      k_copy_labels<<<dim3(uint32_t(cus), uint32_t(L)), kBlock, 0, stream>>>(
          oc, xc, n, flags, plab, pk_ready ? P<unsigned long long>(pk_ent) : nullptr, P<uint32_t>(o2_col), opk,
          keep_ids);
    pk_ready = pack_any;
    if (pack_any) std::swap(pk_ent, pk2_ent);
    std::swap(old_ptr, o2_ptr);
    std::swap(old_col, o2_col);
    old_nnz = o2n;
    return GG_OK;
  }

  // ids false: rows of packed labels get packed entries only (see k_merge)
  int merge_rows(DCsr oc, DCsr dc, uint32_t rows, uint32_t nn, const Counters* c, DPacked plab,
                 const unsigned long long* old_pk, uint32_t* out_col, unsigned long long* out_pk, uint32_t* pklen,
                 std::string& err, bool ids = true) {
    GG_TRY(grow(mlist, uint64_t(rows) * 4, err));
    GG_TRY(alloc(mcnt, 4, err));
    fill_u32(mcnt, 0u, 1);
    k_merge<<<uint32_t(cus) * uint32_t(merge_blocks_per_cu), kBlock, 0, stream>>>(oc, dc, rows, nn, c, plab, old_pk, out_col, out_pk, pklen,
                                                      P<uint32_t>(mlist), P<uint32_t>(mcnt), ids);
    k_merge_rows<<<uint32_t(cus) * uint32_t(merge_rows_blocks_per_cu), kBlock, 0, stream>>>(oc, dc, nn, plab, old_pk, out_col, out_pk, pklen,
                                                           P<uint32_t>(mlist), P<uint32_t>(mcnt), ids);
    GG_HIP(hipGetLastError());
    return GG_OK;
  }

  DCsr pend_csr() {
    return {pend_nnz ? P<unsigned long long>(pd_ptr) : nullptr, pend_nnz ? P<uint32_t>(pd_col) : nullptr};
  }
  ExpandArgs expand_args(bool init) {
    ExpandArgs a{};
    a.old = {P<unsigned long long>(old_ptr), P<uint32_t>(old_col)};
    a.dl = {P<unsigned long long>(dl_ptr), P<uint32_t>(dl_col)};
    a.raw = {init ? P<unsigned long long>(raw_ptr) : nullptr, init ? P<uint32_t>(raw_col) : nullptr};
    a.b2 = {pairs_ready ? P<unsigned long long>(pr_ptr) : nullptr, pairs_ready ? P<uint32_t>(pr_col) : nullptr};
    a.pend = pend_csr();
    a.pk.len = (!init && pk_ready) ? P<uint32_t>(pk_len) : nullptr;
    a.pk.ent = P<unsigned long long>(pk_ent);
    for (int i = 0; i < 8; ++i) a.pk.labels[i] = pack_labels[i];
    a.dpk.len = (!init && dpk_ready) ? P<uint32_t>(dpk_len) : nullptr;
    a.dpk.ent = P<unsigned long long>(dpk_ent);
    for (int i = 0; i < 8; ++i) a.dpk.labels[i] = pack_labels[i];
    a.n = n;
    a.g = g;
    a.active = P<uint32_t>(active);
    a.ctr = C();
    a.staging = P<uint32_t>(staging);
    a.staging_cap = staging.bytes / 4;
    a.seg_off = P<unsigned long long>(seg_off);
    a.seg_cnt = P<uint32_t>(seg_cnt);
    a.guess = guess_bits;
    a.grab = grab;
    return a;
  }

  int expand_blocks_per_cu = 4;
  bool expand_group = true;  // rows in groups (k_expand_group) where the guessed windows allow
  int merge_blocks_per_cu = 8, merge_rows_blocks_per_cu = 8, gather_blocks_per_cu = 7;
  // the tuned parameters for this run: file beside the library, then GG_TUNE_*
  void apply_tuning() {
    tune = load_tuning();
    expand_blocks_per_cu = tune.expand_bpc > 0 ? std::min(tune.expand_bpc, occ_expand) : occ_expand;
    merge_blocks_per_cu = tune.merge_bpc > 0 ? std::min(tune.merge_bpc, occ_merge) : occ_merge;
    merge_rows_blocks_per_cu = tune.merge_rows_bpc > 0 ? std::min(tune.merge_rows_bpc, occ_merge_rows) : occ_merge_rows;
    gather_blocks_per_cu = tune.gather_bpc > 0 ? std::min(tune.gather_bpc, occ_gather) : occ_gather;
    if (const char* eg = std::getenv("GG_EXPAND_GROUP")) expand_group = eg[0] != '0';
    tune.grab_light = std::max(1u, tune.grab_light);
    tune.grab_heavy = std::max(1u, tune.grab_heavy);
    grab = tune.grab_light;
  }
  // rows per work-counter grab in this round (see Tuning) and the join's
  // candidates and active rows of the last round, which choose it
  uint32_t grab = 4;
  uint64_t last_cand = 0, last_active = 0;
  // the lane-per-row pass runs in the first round and after rounds with
  // few candidates per active row
  // k_expand_group: guessed windows of a power of two ids that leave room
  // for two rows' bitmaps or more (GG_EXPAND_GROUP=0 turns it off)
  bool group_rows() const {
    if (!expand_group || !guess_bits || n <= guess_bits || (guess_bits & (guess_bits - 1))) return false;
    return tune.group_rows >= 2 && kGroupWords / (guess_bits / 32) >= 2;
  }
  bool light_round() const {
    return tune.light_cand && (last_active == 0 || last_cand <= tune.light_cand * last_active);
  }
  void choose_grab() {
    grab = (last_active && last_cand / last_active > tune.heavy_cand) ? tune.grab_heavy : tune.grab_light;
  }
  uint32_t expand_grid() const {
    const uint32_t g = uint32_t(cus) * uint32_t(expand_blocks_per_cu);
    return expand_grid_cap ? std::min(g, expand_grid_cap) : g;
  }
  uint32_t gather_grid() const { return uint32_t(cus) * uint32_t(gather_blocks_per_cu); }
  uint32_t rows() const { return uint32_t(uint64_t(L) * n); }

  bool pack_any_label() const {
    bool any = false;
    for (int i = 0; i < 8; ++i) any |= pack_labels[i] != 0;
    return any;
  }
  hipError_t rec(hipEvent_t ev) { return ev ? hipEventRecord(ev, stream) : hipSuccess; }

  int run(int filter, gg_stats* st, std::string& err);
  int plan_rounds(RoundPlan& rp, std::string& err);
  int round_step(RoundPlan& rp, gg_stats& S, RunEvents& ev, uint64_t inject, std::string& err);
  int run_r(int filter, const std::vector<std::pair<uint32_t, uint32_t>>& parts, const PartOpts& po, gg_stats* st,
            std::string& err);
};

int Engine::run(int filter, gg_stats* st, std::string& err) {
  if (!has_grammar || !has_graph) {
    err = "grammar and graph must be loaded before gg_closure_run";
    return GG_E_STATE;
  }
  if (filter != 0) {
    err = "the semi-naive engine computes the least fixpoint only; filter_flag != 0 runs the partitioned schedule";
    return GG_E_ARG;
  }
  use_rules(rules_planned);
  ids_pending = false;
  pend_nnz = 0;
  apply_tuning();
  defer_ratio = tune.defer_ratio;
  if (const char* dr = std::getenv("GG_DEFER_RATIO")) defer_ratio = std::strtoull(dr, nullptr, 10);
  keep_ids = false;
  if (const char* ki = std::getenv("GG_KEEP_PACKED_IDS")) keep_ids = std::atoi(ki) != 0;
  // words per pair arena for every round's first k_pairs pass (0: sized by
  // use); a small value makes the passes overflow and rerun
  uint64_t pair_arena_forced = 0;
  if (const char* pa = std::getenv("GG_PAIR_ARENA")) pair_arena_forced = std::strtoull(pa, nullptr, 10);
  if (uint64_t(L) * n >= 0xFFFFFFFFull || uint64_t(g.nslot) * n >= 0xFFFFFFFFull) {
    err = "num_labels * num_vertices (and binary rules with a derived second label * num_vertices) must be < 2^32";
    return GG_E_UNSUPPORTED;
  }
  GG_HIP(hipSetDevice(device));
  gg_stats S{};
  const uint32_t R = rows();
  RunEvents ev;
  ev.enabled = !arena;  // fixed-region mode creates no events: its statistics carry no times

  peak = used;  // device_bytes_peak is per run
  GG_TRY(alloc(ctr, sizeof(Counters), err));
  GG_TRY(alloc(old_ptr, (uint64_t(R) + 1) * 8, err));
  GG_TRY(alloc(dl_ptr, (uint64_t(R) + 1) * 8, err));
  GG_TRY(alloc(nx_ptr, (uint64_t(R) + 1) * 8, err));
  GG_TRY(alloc(o2_ptr, (uint64_t(R) + 1) * 8, err));
  GG_TRY(alloc(seg_off, uint64_t(R) * 8, err));
  GG_TRY(alloc(seg_cnt, uint64_t(R) * 4, err));
  GG_TRY(alloc(active, uint64_t(R) * 4, err));
  if (!h_ctr) GG_HIP(hipHostMalloc((void**)&h_ctr, sizeof(Counters), hipHostMallocDefault));
  if (!h_u64) GG_HIP(hipHostMalloc((void**)&h_u64, 8, hipHostMallocDefault));

  const auto t0 = std::chrono::steady_clock::now();

  // ---- initial delta: dedup(input U eps loops) ---------------------------
  {
    EventPair& e = ev.mk(ev.other);
    GG_HIP(rec(e.a));
    GG_TRY(alloc(raw_ptr, (uint64_t(R) + 1) * 8, err));
    fill_u32(seg_cnt, 0, R);
    k_count_raw<<<grid_for(std::max<uint64_t>(m, uint64_t(g.ne) * n)), kBlock, 0, stream>>>(
        P<uint32_t>(in_src), P<uint8_t>(in_lab), m, n, g, P<uint32_t>(seg_cnt));
    GG_TRY(scan(seg_cnt, R, raw_ptr, err));
    const uint64_t raw_n = m + uint64_t(g.ne) * n;
    GG_TRY(alloc(raw_col, raw_n * 4, err));
    fill_u32(seg_cnt, 0, R);
    k_scatter_raw<<<grid_for(std::max<uint64_t>(m, uint64_t(g.ne) * n)), kBlock, 0, stream>>>(
        P<uint32_t>(in_src), P<uint32_t>(in_dst), P<uint8_t>(in_lab), m, n, g, P<unsigned long long>(raw_ptr),
        P<uint32_t>(seg_cnt), P<uint32_t>(raw_col));
    fill_u64(old_ptr, 0, uint64_t(R) + 1);
    fill_u64(dl_ptr, 0, uint64_t(R) + 1);
    // GG_STAGING_MIN_WORDS sets a smaller first staging area (tests use it to
    // exercise the grow-and-retry path).
    // (room for the raw rows plus every wave's reservation block, so the
    // first expand does not overflow)
    uint64_t stage_words = raw_n + uint64_t(expand_grid()) * kWavesPerBlock * kStageBlock * 2;
    if (const char* sw = std::getenv("GG_STAGING_MIN_WORDS")) stage_words = std::max<uint64_t>(1, std::strtoull(sw, nullptr, 10));
    release(staging);
    GG_TRY(alloc(staging, stage_words * 4, err));
    GG_HIP(rec(e.b));
    for (;;) {
      EventPair& x = ev.mk(ev.expand);
      zero_ctr();
      fill_u32(seg_cnt, 0, R);
      k_activate_raw<<<grid_for(R), kBlock, 0, stream>>>(P<unsigned long long>(raw_ptr), R, P<uint32_t>(active), C());
      GG_HIP(rec(x.a));
      // rows of a few raw entries one per lane; the others (read on the
      // device: no read-back) one per wave
      GG_TRY(grow(mlist, uint64_t(R) * 4, err));
      ExpandArgs ra = expand_args(true);
      k_dedup_light<<<grid_for(R), kBlock, 0, stream>>>(ra, P<uint32_t>(mlist));
      ra.defer = P<uint32_t>(mlist);
      k_expand<1><<<expand_grid(), kBlock, 0, stream>>>(ra);
      GG_HIP(rec(x.b));
      GG_HIP(hipGetLastError());
      GG_TRY(read_ctr(err));
      S.expand_launches += 2;
      if (!h_ctr->overflow) break;
      S.retries++;
      GG_TRY(grow(staging, size_t(h_ctr->cursor + h_ctr->cursor / 4 + 1024) * 4, err));
    }
    const uint64_t N0 = h_ctr->newe;
    EventPair& c = ev.mk(ev.compact);
    GG_HIP(rec(c.a));
    GG_TRY(scan(seg_cnt, R, dl_ptr, err));
    GG_TRY(grow(dl_col, std::max<uint64_t>(N0, 1) * 4, err));
    dpk_ready = false;
    if (pack_delta()) {
      GG_TRY(grow(dpk_len, uint64_t(R) * 4, err));
      GG_TRY(grow(dpk_ent, std::max<uint64_t>(N0, 1) * 8, err));
      dpk_ready = true;
    }
    k_gather<<<gather_grid(), kBlock, 0, stream>>>(P<uint32_t>(active), C(), P<uint32_t>(seg_cnt),
                                                   P<unsigned long long>(seg_off), P<uint32_t>(staging),
                                                   P<unsigned long long>(dl_ptr), P<uint32_t>(dl_col), n, plabels(),
                                                   dpk_ready ? P<unsigned long long>(dpk_ent) : nullptr,
                                                   P<uint32_t>(dpk_len));
    GG_HIP(rec(c.b));
    release(raw_col);
    release(raw_ptr);
    dl_nnz = N0;
    old_nnz = 0;
    S.input_unique = N0;
  }

  // ---- semi-naive rounds -------------------------------------------------
  RoundPlan rp;
  rp.pair_arena_forced = pair_arena_forced;
  pr_arena = 0;
  pk_ready = false;
  guess_bits = tune.guess_min;
  last_cand = last_active = 0;
  last_deferred = kUnknown;
  GG_TRY(plan_rounds(rp, err));
  if (pack_any_label()) {
    GG_TRY(alloc(pk_len, uint64_t(R) * 4, err));
    fill_u32(pk_len, 0, R);
  }
  while (dl_nnz > 0) GG_TRY(round_step(rp, S, ev, 0, err));
  if (pend_nnz) {  // the closure is old ∪ pend: rebuild old once
    EventPair& mg = ev.mk(ev.merge);
    GG_HIP(rec(mg.a));
    GG_TRY(fold_into_old({P<unsigned long long>(pd_ptr), P<uint32_t>(pd_col)}, pend_nnz, nullptr, err));
    pend_nnz = 0;
    GG_HIP(rec(mg.b));
  }
  ids_pending = pk_ready && !keep_ids && defer_unpack;
  if (pk_ready && !keep_ids && !defer_unpack) {  // the packed labels' ids, from their entries
    EventPair& mg = ev.mk(ev.merge);
    GG_HIP(rec(mg.a));
    DPacked plab{};
    for (int i = 0; i < 8; ++i) plab.labels[i] = pack_labels[i];
    plab.len = P<uint32_t>(pk_len);
    plab.ent = P<unsigned long long>(pk_ent);
    k_unpack<<<uint32_t(cus) * 8, kBlock, 0, stream>>>(P<unsigned long long>(old_ptr), P<uint32_t>(old_col), plab, R, n);
    GG_HIP(rec(mg.b));
    GG_HIP(hipGetLastError());
  }
  release(pd_ptr);
  release(pd_col);
  release(p2_ptr);
  release(p2_col);
  GG_HIP(hipStreamSynchronize(stream));
  const auto t1 = std::chrono::steady_clock::now();
  S.fixpoint_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  auto sum = [](std::vector<EventPair>& v) {
    double t = 0;
    for (auto& e : v) {
      float ms = 0;
      if (e.a && e.b && hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) t += ms;
    }
    return t;
  };
  S.expand_ms = sum(ev.expand);
  S.compact_ms = sum(ev.compact);
  S.merge_ms = sum(ev.merge);
  S.other_ms = sum(ev.other);
  S.total_edges = old_nnz;
  S.device_bytes_peak = peak;
  // drop per-round scratch; keep the closure (old) for queries
  release(nx_col);
  release(o2_col);
  release(staging);
  for (DevBuf* b : {&dpk_len, &dpk_ent, &npk_len, &npk_ent}) release(*b);
  dpk_ready = false;
  has_closure = true;
  if (st) *st = S;
  return GG_OK;
}

// Per-size setup of the round loop: the activation and pair-list plans and
// the per-row arrays for the current vertex count (a session that grows its
// vertex range plans again).
int Engine::plan_rounds(RoundPlan& rp, std::string& err) {
  const uint32_t R = rows();
  GG_TRY(alloc(nx_ptr, (uint64_t(R) + 1) * 8, err));
  GG_TRY(alloc(o2_ptr, (uint64_t(R) + 1) * 8, err));
  GG_TRY(alloc(seg_off, uint64_t(R) * 8, err));
  GG_TRY(alloc(seg_cnt, uint64_t(R) * 4, err));
  GG_TRY(alloc(active, uint64_t(R) * 4, err));
  ActArgs& act = rp.act;
  act = ActArgs{};
  act.n = n;
  act.nlhs = uint32_t(lhs.size());
  for (size_t i = 0; i < lhs.size(); ++i) act.lhs[i] = lhs[i];
  for (int i = 0; i < kMaxRules; ++i) act.u_li[i] = act.b_li[i] = 0xFFFF;
  for (size_t li = 0; li < lhs.size(); ++li) {
    for (int i = 0; i < g.nu; ++i)
      if (g.u_lhs[i] == lhs[li]) act.u_li[i] = uint16_t(li);
    for (int i = 0; i < g.nb; ++i)
      if (g.b_lhs[i] == lhs[li]) act.b_li[i] = uint16_t(li);
  }
  act.g = g;
  act.active = P<uint32_t>(active);
  act.ctr = C();
  rp.act_work = uint64_t(act.nlhs) * n;
  rp.NS = uint64_t(g.nslot) * n;
  const uint64_t NS = rp.NS;
  // slots with the same B are streamed together, two at a time
  PairPlan& pair_plan = rp.pair_plan;
  pair_plan = PairPlan{};
  for (int j = 0; j < kMaxRules; ++j) pair_plan.mate[j] = 0xFF;
  for (int j = 0; j < g.nslot; ++j) {
    if (pair_plan.mate[j] != 0xFF) continue;
    for (int j2 = j + 1; j2 < g.nslot; ++j2)
      if (pair_plan.mate[j2] == 0xFF && g.slot_b[j2] == g.slot_b[j]) {
        pair_plan.mate[j] = uint8_t(j2);
        pair_plan.mate[j2] = 0xFE;
        break;
      }
  }
  rp.fr_words = (n + kWave - 1) / kWave;
  const uint32_t fr_words = rp.fr_words;
  if (NS) {
    GG_TRY(alloc(pr_ptr, 2 * NS * 8, err));
    GG_TRY(alloc(fr_bits, uint64_t(g.nslot) * fr_words * 8, err));
  }
  rp.n = n;
  return GG_OK;
}

// One semi-naive round: pair lists, activation, the fused join, compaction of
// the new edges into the next delta, and old' = old ∪ delta (or the pending
// CSR). inject > 0: the raw rows in raw_ptr/raw_col (inject edges, rows of
// vertices that have no edges yet) are deduplicated in the same round, after
// the join, so that they enter the next delta as new edges (graphs joining a
// running session).
int Engine::round_step(RoundPlan& rp, gg_stats& S, RunEvents& ev, uint64_t inject, std::string& err) {
  const uint32_t R = rows();
  size_t ex0 = ev.expand.size();  // this round's expand event pairs start here
  uint64_t N = 0;
  // frontier pair lists for old_B x delta_C (old is empty in the first round)
  pairs_ready = false;
  if (rp.NS && old_nnz > 0) {
    EventPair& o = ev.mk(ev.other);
    GG_HIP(rec(o.a));
    const DCsr dlc{P<unsigned long long>(dl_ptr), P<uint32_t>(dl_col)};
    const DCsr oc{P<unsigned long long>(old_ptr), P<uint32_t>(old_col)};
    k_frontier<<<grid_for(n), kBlock, 0, stream>>>(g, n, dlc, rp.fr_words, P<unsigned long long>(fr_bits));
    DPacked opk = plabels();
    opk.len = pk_ready ? P<uint32_t>(pk_len) : nullptr;
    opk.ent = P<unsigned long long>(pk_ent);
    // one pass into kSpread arenas; a pass whose arenas overflowed is run
    // again with larger ones
    const uint32_t pgrid = grid_for(n);
    const uint64_t pwaves = uint64_t((pgrid + kSpread - 1) / kSpread) * kWavesPerBlock;
    pr_arena = rp.pair_arena_forced ? rp.pair_arena_forced : std::max<uint64_t>(pr_arena, 2 * pwaves * kPairBlock);
    for (bool first = true;; first = false) {
      GG_TRY(grow(pr_col, kSpread * pr_arena * 4, err));
      if (!rp.pair_arena_forced || !first) pr_arena = pr_col.bytes / 4 / kSpread;
      zero_ctr();
      k_pairs<<<pgrid, kBlock, 0, stream>>>(g, rp.pair_plan, n, oc, pend_csr(), opk, dlc.ptr, P<unsigned long long>(fr_bits),
                                            rp.fr_words, C(),
                                            pr_arena, P<unsigned long long>(pr_ptr), P<uint32_t>(pr_col));
      GG_HIP(hipGetLastError());
      GG_TRY(read_ctr(err));
      uint64_t most = 0;
      for (const CtrSlot& q : h_ctr->slot) most = std::max<uint64_t>(most, q.pairs);
      if (most <= pr_arena) {
        // pair lists grow over the first rounds: size the next round ahead
        if (most * 4 > pr_arena * 3) pr_arena = 2 * most;
        break;
      }
      S.pair_reruns++;
      pr_arena = 2 * most + kPairBlock;
    }
    GG_HIP(rec(o.b));
    pairs_ready = true;
  }
  choose_grab();
  // staging: room for every wave's reservation block plus a guess of the
  // output; an overflow grows it and re-runs the round
  GG_TRY(grow(staging, (uint64_t(expand_grid()) * kWavesPerBlock * kStageBlock * 2 + 4 * dl_nnz + inject) * 4, err));
  for (;;) {
    EventPair& o = ev.mk(ev.other);
    GG_HIP(rec(o.a));
    zero_ctr();
    fill_u32(seg_cnt, 0, R);
    k_label_flags<<<1, kBlock, 0, stream>>>(P<unsigned long long>(dl_ptr), n, uint32_t(L), C());
    rp.act.old = {P<unsigned long long>(old_ptr), P<uint32_t>(old_col)};
    rp.act.dl = {P<unsigned long long>(dl_ptr), P<uint32_t>(dl_col)};
    rp.act.b2rng = pairs_ready ? P<unsigned long long>(pr_ptr) : nullptr;
    rp.act.pend_ptr = pend_csr().ptr;
    if (rp.act_work) {
      // sub-chunks per counter reservation: 1 while a grid of 256-vertex
      // blocks fits the device (4096 blocks), more beyond (at most 8)
      const uint32_t sub = uint32_t(std::min<uint64_t>(8, std::max<uint64_t>(1, uint64_t(n) / (uint64_t(cus) * 16 * kBlock))));
      k_activate<<<grid_for((n + sub - 1) / sub), kBlock, 0, stream>>>(rp.act, sub);
    }
    if (g.nb)
      k_probe_keys<<<1, kWave, 0, stream>>>(g, P<unsigned long long>(old_ptr), P<unsigned long long>(dl_ptr),
                                            pend_csr().ptr, n, C());
    // The first expand launch follows the activation without a read-back:
    // its kernel reads the active row count on the device. The grid comes
    // from the last round's active rows (a round with few of them launches
    // only the blocks that can get some, kGrab rows per grab and 4 waves per
    // block, not a full grid of waves that would only contend on the work
    // counters); the waves take rows until none are left, whatever the
    // grid.
    const uint64_t est = std::max<uint64_t>(last_active, 1) * 2;
    const uint32_t want = uint32_t(std::min<uint64_t>(expand_grid(), (est + grab * kWavesPerBlock - 1) /
                                                                          (grab * kWavesPerBlock)));
    const uint32_t grid = last_active ? std::max(1u, want) : expand_grid();
    GG_HIP(rec(o.b));
    ex0 = ev.expand.size();
    EventPair* xp = &ev.mk(ev.expand);
    GG_HIP(rec(xp->a));
    light_rows = 0;
    grouped = false;
    const bool light = light_round();
    if (light) {
      // one lane per row first; the rows with more candidates than a lane
      // holds are then expanded by the wave kernel
      GG_TRY(grow(mlist, uint64_t(R) * 4, err));
      ExpandArgs la = expand_args(false);
      // lanes past the active count exit at once
      const uint32_t lg = grid_for(last_active ? std::min<uint64_t>(R, 2 * last_active) : R);
      uint32_t* dl = P<uint32_t>(mlist);
      if (tune.light_cap > 32) {
        if (pend_nnz)
          k_expand_light<true, 64><<<lg, kBlock, 0, stream>>>(la, dl);
        else
          k_expand_light<false, 64><<<lg, kBlock, 0, stream>>>(la, dl);
      } else if (tune.light_cap <= 16) {
        // half the LDS column per lane: twice the blocks per CU
        if (pend_nnz)
          k_expand_light<true, 16><<<lg, kBlock, 0, stream>>>(la, dl);
        else
          k_expand_light<false, 16><<<lg, kBlock, 0, stream>>>(la, dl);
      } else if (pend_nnz) {
        k_expand_light<true, 32><<<lg, kBlock, 0, stream>>>(la, dl);
      } else {
        k_expand_light<false, 32><<<lg, kBlock, 0, stream>>>(la, dl);
      }
      GG_HIP(rec(xp->b));
      GG_HIP(hipGetLastError());
      S.expand_launches++;
      // the rows the light pass left go to the wave kernel, with a grid for
      // their number; not after an overflow (the round runs again), whose
      // further staging reservations would only inflate the regrowth
      GG_TRY(read_ctr(err));
      xp = nullptr;
      if (h_ctr->deferred && !h_ctr->overflow) {
        la.defer = P<uint32_t>(mlist);
        const uint32_t dwant =
            uint32_t((uint64_t(h_ctr->deferred) + grab * kWavesPerBlock - 1) / (grab * kWavesPerBlock));
        const uint32_t dgrid = std::max(1u, std::min(expand_grid(), dwant));
        xp = &ev.mk(ev.expand);
        GG_HIP(rec(xp->a));
        if (pend_nnz)
          k_expand<0, true><<<dgrid, kBlock, 0, stream>>>(la);
        else
          k_expand<0><<<dgrid, kBlock, 0, stream>>>(la);
        S.expand_launches++;
      }
    } else if (group_rows()) {
      // rows in groups of up to group_rows per wave; the rows that do not
      // fit their group's window then go to the wave kernel, whose launch
      // reads their count on the device (no read-back: with none it exits)
      GG_TRY(grow(mlist, uint64_t(R) * 4, err));
      ExpandArgs xa = expand_args(false);
      xa.group = std::min({tune.group_rows, kGroupRows, kGroupWords / (guess_bits / 32)});
      const uint32_t gwant = uint32_t(std::min<uint64_t>(
          expand_grid(), (est + uint64_t(xa.group) * kWavesPerBlock - 1) / (uint64_t(xa.group) * kWavesPerBlock)));
      const uint32_t ggrid = last_active ? std::max(1u, gwant) : expand_grid();
      if (pend_nnz)
        k_expand_group<true><<<ggrid, kBlock, 0, stream>>>(xa, P<uint32_t>(mlist));
      else
        k_expand_group<false><<<ggrid, kBlock, 0, stream>>>(xa, P<uint32_t>(mlist));
      k_reset_work<<<1, kWave, 0, stream>>>(C());
      xa.defer = P<uint32_t>(mlist);
      // grid for twice the last grouped round's deferred rows (at least 8
      // blocks: a graph with id locality defers few or none, and a round
      // that defers many moves the guess off for the next one); a full grid
      // in the run's first grouped round
      const uint64_t dper = uint64_t(grab) * kWavesPerBlock;
      const uint32_t dgrid =
          last_deferred == kUnknown
              ? expand_grid()
              : uint32_t(std::max<uint64_t>(8, std::min<uint64_t>(expand_grid(), (2 * last_deferred + dper - 1) / dper)));
      if (pend_nnz)
        k_expand<0, true><<<dgrid, kBlock, 0, stream>>>(xa);
      else
        k_expand<0><<<dgrid, kBlock, 0, stream>>>(xa);
      S.expand_launches += 2;
      grouped = true;
    } else {
      if (pend_nnz)
        k_expand<0, true><<<grid, kBlock, 0, stream>>>(expand_args(false));
      else
        k_expand<0><<<grid, kBlock, 0, stream>>>(expand_args(false));
      S.expand_launches++;
    }
    if (inject) {
      // graphs joining a session: their raw rows go to the end of this
      // round's active list and are deduplicated there (their old, delta and
      // pending rows are empty, so the result is their first delta)
      if (!xp) {
        xp = &ev.mk(ev.expand);
        GG_HIP(rec(xp->a));
      }
      k_mark_active_base<<<1, kWave, 0, stream>>>(C());
      k_activate_raw<<<grid_for(R), kBlock, 0, stream>>>(P<unsigned long long>(raw_ptr), R, P<uint32_t>(active), C());
      k_expand<1><<<expand_grid(), kBlock, 0, stream>>>(expand_args(true));
      S.expand_launches++;
    }
    if (xp) GG_HIP(rec(xp->b));
    GG_HIP(hipGetLastError());
    GG_TRY(read_ctr(err));
    if (light) light_rows = h_ctr->active - h_ctr->deferred;
    if (!h_ctr->overflow) break;
    S.retries++;
    GG_TRY(grow(staging, size_t(h_ctr->cursor + h_ctr->cursor / 4 + 1024) * 4, err));
  }
  N = h_ctr->newe;
  S.light_rows += light_rows;
  if (grouped) {
    S.group_rows += inject ? h_ctr->active_base : h_ctr->active;
    S.group_deferred += h_ctr->deferred;
    last_deferred = h_ctr->deferred;
  }
  last_cand = h_ctr->cand;
  last_active = h_ctr->active;
  const uint32_t used_guess = guess_bits;
  if (guess_bits == 0)
    guess_bits = kChunkBits;
  else if (uint64_t(h_ctr->misses) * 8u > h_ctr->active)
    guess_bits = guess_bits < kChunkBits ? kChunkBits : 0u;
  S.rounds++;
  S.probe_reads += h_ctr->pkeys + h_ctr->cand;
  S.candidates += h_ctr->cand;
  S.old_reads += h_ctr->oreads;
  S.new_edges += N;
  if (verbose) {
    float xms = 0;
    for (size_t i = ex0; i < ev.expand.size(); ++i) {
      float t = 0;
      if (ev.expand[i].a && hipEventElapsedTime(&t, ev.expand[i].a, ev.expand[i].b) == hipSuccess) xms += t;
    }
    std::fprintf(stderr,
                 "[gg] round %u: active %u cand %llu new %llu guess %u misses %u expand %.2f ms"
                 " (cycles: seeding %llu lists %llu streams %llu rows %llu)\n",
                 S.rounds, h_ctr->active, (unsigned long long)h_ctr->cand, (unsigned long long)N, used_guess,
                 h_ctr->misses, xms, h_ctr->xs[0], h_ctr->xs[1], h_ctr->xs[2], h_ctr->xs[3]);
  }
  // delta' from staging
  EventPair& c = ev.mk(ev.compact);
  GG_HIP(rec(c.a));
  GG_TRY(scan(seg_cnt, R, nx_ptr, err));
  GG_TRY(grow(nx_col, std::max<uint64_t>(N, 1) * 4, err));
  const bool npk = pack_delta();
  if (npk) {
    GG_TRY(grow(npk_len, uint64_t(R) * 4, err));
    GG_TRY(grow(npk_ent, std::max<uint64_t>(N, 1) * 8, err));
  }
  if (N)
    k_gather<<<gather_grid(), kBlock, 0, stream>>>(P<uint32_t>(active), C(), P<uint32_t>(seg_cnt),
                                                   P<unsigned long long>(seg_off), P<uint32_t>(staging),
                                                   P<unsigned long long>(nx_ptr), P<uint32_t>(nx_col), n, plabels(),
                                                   npk ? P<unsigned long long>(npk_ent) : nullptr,
                                                   P<uint32_t>(npk_len));
  GG_HIP(rec(c.b));
  // old' = old ∪ delta. While the deltas are small next to old they go to
  // the pending CSR instead (pend' = pend ∪ delta), so a round that finds
  // few edges does not rewrite all of old; the joins read old ∪ pend.
  EventPair& mg = ev.mk(ev.merge);
  GG_HIP(rec(mg.a));
  const bool defer = defer_ratio && old_nnz && (pend_nnz + dl_nnz) * defer_ratio <= old_nnz;
  if (defer || pend_nnz) {
    if (!pend_nnz) {
      GG_TRY(alloc(pd_ptr, (uint64_t(R) + 1) * 8, err));
      fill_u64(pd_ptr, 0ull, uint64_t(R) + 1);
    }
    GG_TRY(alloc(p2_ptr, (uint64_t(R) + 1) * 8, err));
    const uint64_t p2n = pend_nnz + dl_nnz;
    GG_TRY(grow(p2_col, std::max<uint64_t>(p2n, 1) * 4, err));
    k_add_ptr<<<grid_for(uint64_t(R) + 1), kBlock, 0, stream>>>(
        P<unsigned long long>(p2_ptr), P<unsigned long long>(pd_ptr), P<unsigned long long>(dl_ptr), uint64_t(R) + 1);
    const DCsr pc{P<unsigned long long>(pd_ptr), P<uint32_t>(pd_col)};
    const DCsr dc{P<unsigned long long>(dl_ptr), P<uint32_t>(dl_col)};
    GG_TRY(merge_rows(pc, dc, R, n, C(), DPacked{}, nullptr, P<uint32_t>(p2_col), nullptr, nullptr, err));
    if (pend_nnz)
      k_copy_labels<<<dim3(uint32_t(cus), uint32_t(L)), kBlock, 0, stream>>>(pc, dc, n, C(), DPacked{}, nullptr,
                                                                            P<uint32_t>(p2_col), nullptr, true);
    std::swap(pd_ptr, p2_ptr);
    std::swap(pd_col, p2_col);
    pend_nnz = p2n;
  }
  if (!defer) {
    // pend (which now holds delta) or delta itself goes into old
    if (pend_nnz) {
      GG_TRY(fold_into_old({P<unsigned long long>(pd_ptr), P<uint32_t>(pd_col)}, pend_nnz, nullptr, err));
      pend_nnz = 0;
    } else {
      GG_TRY(fold_into_old({P<unsigned long long>(dl_ptr), P<uint32_t>(dl_col)}, dl_nnz, C(), err));
    }
  }
  GG_HIP(rec(mg.b));
  GG_HIP(hipGetLastError());
  std::swap(dl_ptr, nx_ptr);
  std::swap(dl_col, nx_col);
  std::swap(dpk_len, npk_len);
  std::swap(dpk_ent, npk_ent);
  dpk_ready = npk;
  dl_nnz = N;
  if (on_round) on_round(last_active, R);
  return GG_OK;
}

// ---------------------------------------------------------------------------
// The reference's partitioned schedule (filter != 0 needs it): superstep 0
// closes every partition on its own (precomputation, src/main.cpp:764-804),
// later supersteps sweep the partition pairs i < j (run_computation,
// src/main.cpp:806-858), with the operator sequences of spagpu_s / spagpu_b /
// spagpu (src/kernel.cu:466-780) and mergeAndDiff (src/kernel.cu:943-985).
// Every region (old, delta, tmp of a partition) is a CSR over the
// partition's own rows; an operator is one MODE 2 expand launch plus the
// compaction of its new ids and their merge into tmp.

int Engine::run_r(int filter, const std::vector<std::pair<uint32_t, uint32_t>>& parts, const PartOpts& po,
                  gg_stats* st, std::string& err) {
  if (!has_grammar || !has_graph) {
    err = "grammar and graph must be loaded before gg_closure_run";
    return GG_E_STATE;
  }
  // the reference's schedule runs the rule table as given: with filter != 0
  // its result depends on the rules, not only on their least fixpoint
  use_rules(rules_given);
  ids_pending = false;
  // row ids (label * vertices + vertex) are uint32 in every region and in
  // the assembled closure, which has all n vertices
  if (uint64_t(L) * n >= 0xFFFFFFFFull) {
    err = "num_labels * num_vertices must be < 2^32 (row ids are 32-bit)";
    return GG_E_UNSUPPORTED;
  }
  apply_tuning();
  if (parts.empty()) {
    err = "the partitioned schedule needs at least one vertex interval";
    return GG_E_ARG;
  }
  GG_HIP(hipSetDevice(device));
  struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  struct Region {
    DevBuf ptr, col;
    uint64_t nnz = 0;
    HostBuf hptr, hcol;  // spilled copy (pinned)
  };
  struct Part {
    uint32_t first = 0, np = 0;
    Region r[3];  // old, delta, tmp
    bool has[3] = {false, false, false};
    bool resident = false;
    bool evicted = false;      // has been spilled at least once
    bool in_pending = false;   // resident, its upload still in flight (the compute stream has not waited for it)
    bool out_pending = false;  // spilled, its copy-out in flight: the device buffers are kept until it ends
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
  };
  enum { kO = 0, kD = 1, kT = 2 };
  const int NP = int(parts.size());
  std::vector<Part> pt(static_cast<size_t>(NP));
  uint32_t max_np = 0;
  for (int p = 0; p < NP; ++p) {
    pt[p].first = parts[p].first;
    pt[p].np = parts[p].second - parts[p].first + 1;
    max_np = std::max(max_np, pt[p].np);
  }
  gg_stats S{};
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t copy_stream = nullptr;
  GG_HIP(hipStreamCreate(&copy_stream));
  struct Cleanup {
    std::function<void()> f;
    ~Cleanup() { f(); }
  } cl{[&]() {
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    for (auto& q : pt) {
      for (auto& r : q.r) {
        release(r.ptr);
        release(r.col);
        if (r.hptr.p) (void)hipHostFree(r.hptr.p);
        if (r.hcol.p) (void)hipHostFree(r.hcol.p);
      }
      if (q.ev_in) (void)hipEventDestroy(q.ev_in);
      if (q.ev_out) (void)hipEventDestroy(q.ev_out);
    }
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
  }};
  const uint64_t max_rows = uint64_t(L) * max_np;
  peak = used;  // device_bytes_peak is per run
  GG_TRY(alloc(ctr, sizeof(Counters), err));
  GG_TRY(alloc(seg_off, max_rows * 8, err));
  GG_TRY(alloc(seg_cnt, max_rows * 4, err));
  GG_TRY(alloc(active, max_rows * 4, err));
  GG_TRY(grow(staging, uint64_t(expand_grid()) * kWavesPerBlock * kStageBlock * 2 * 4, err));
  if (!h_ctr) GG_HIP(hipHostMalloc((void**)&h_ctr, sizeof(Counters), hipHostMallocDefault));
  if (!h_u64) GG_HIP(hipHostMalloc((void**)&h_u64, 8, hipHostMallocDefault));
  auto rows_of = [&](const Part& q) { return uint64_t(L) * q.np; };
  auto dcsr = [&](Region& r) { return DCsr{P<unsigned long long>(r.ptr), P<uint32_t>(r.col)}; };
  auto make_empty = [&](Region& r, uint64_t rows) -> int {
    GG_TRY(grow(r.ptr, (rows + 1) * 8, err));
    fill_u64(r.ptr, 0, rows + 1);
    r.nnz = 0;
    return GG_OK;
  };
  // pinned host buffers grow with 25% slack: pinning is slow, and a spilled
  // region usually comes back a little larger
  auto host_alloc = [&](HostBuf& h, size_t bytes) -> int {
    bytes = std::max<size_t>(bytes, 256);
    if (h.bytes >= bytes) return GG_OK;
    if (h.p) {
      // a copy on the copy stream may still read or write the old buffer
      GG_HIP(hipStreamSynchronize(copy_stream));
      (void)hipHostFree(h.p);
    }
    h.p = nullptr;
    h.bytes = 0;
    if (bytes > (1u << 20)) bytes += bytes / 4;
    GG_HIP(hipHostMalloc(&h.p, bytes, hipHostMallocDefault));
    h.bytes = bytes;
    return GG_OK;
  };
  // Residency. A partition is resident (its three regions in HBM) or spilled
  // (in pinned host DRAM). Copies run on the copy stream and are ordered with
  // the compute stream by events only, so they overlap the operators:
  //  * before each step (a pair, or one partition's superstep-0 closure or
  //    merge) the partitions of the next steps of the sweep are uploaded
  //    ahead when they fit the budget next to the step's working memory
  //    (the most any step has needed so far); the compute stream waits for
  //    an upload only when the partition is first used;
  //  * when they do not fit, partitions whose next use is farthest ahead are
  //    copied out ahead, and their device buffers go back once the copy-out
  //    has ended (checked before each step);
  //  * an allocation that still does not fit (the allocator's pressure
  //    callback) first waits for a copy-out in flight, then copies out the
  //    farthest partition the current step does not hold and waits for it.
  // The sweep order is known, so "farthest next use" is Belady's choice.
  // GG_OOC_MAX_RESIDENT (a test hook) also caps the number of resident
  // partitions, so that small graphs go through the spill path.
  hipEvent_t ev = nullptr;
  GG_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  struct EventGuard {
    hipEvent_t& e;
    ~EventGuard() { (void)hipEventDestroy(e); }
  } eg{ev};
  int hold[2] = {-1, -1};
  int max_resident = 0;
  if (const char* mr = std::getenv("GG_OOC_MAX_RESIDENT")) max_resident = std::max(2, std::atoi(mr));
  double copy_ms = 0;  // host time spent issuing copies and waiting for them
  auto make_event = [&](hipEvent_t& e) -> int {
    if (!e) GG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return GG_OK;
  };
  // the device buffers of a copied-out partition go back once its copy has
  // ended (wait: block until it has)
  auto finish_out = [&](int p, bool wait) -> bool {
    Part& q = pt[p];
    if (!q.out_pending) return false;
    if (wait) {
      if (hipEventSynchronize(q.ev_out) != hipSuccess) return false;
    } else if (hipEventQuery(q.ev_out) != hipSuccess) {
      (void)hipGetLastError();  // not ready is not an error
      return false;
    }
    for (auto& r : q.r) {
      release(r.ptr);
      release(r.col);
    }
    q.out_pending = false;
    return true;
  };
  auto reap = [&]() {
    for (int p = 0; p < NP; ++p) finish_out(p, false);
  };
  auto spill = [&](int p, bool wait) -> int {
    Part& q = pt[p];
    if (!po.spill || !q.resident) return GG_OK;
    const auto c0 = std::chrono::steady_clock::now();
    const uint64_t rows = rows_of(q);
    // after the partition's last use on the compute stream
    GG_HIP(hipEventRecord(ev, stream));
    GG_HIP(hipStreamWaitEvent(copy_stream, ev, 0));
    for (auto& r : q.r) {
      if (!r.nnz) continue;  // an empty region comes back as zeros
      GG_TRY(host_alloc(r.hptr, (rows + 1) * 8));
      GG_TRY(host_alloc(r.hcol, r.nnz * 4));
      GG_HIP(hipMemcpyAsync(r.hptr.p, r.ptr.p, (rows + 1) * 8, hipMemcpyDeviceToHost, copy_stream));
      GG_HIP(hipMemcpyAsync(r.hcol.p, r.col.p, r.nnz * 4, hipMemcpyDeviceToHost, copy_stream));
      S.spill_bytes += (rows + 1) * 8 + r.nnz * 4;
    }
    GG_TRY(make_event(q.ev_out));
    GG_HIP(hipEventRecord(q.ev_out, copy_stream));
    q.resident = false;
    q.evicted = true;
    q.in_pending = false;
    q.out_pending = true;
    if (wait) finish_out(p, true);
    copy_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
    return GG_OK;
  };
  // Position of the next step that uses partition q, from the current step:
  // superstep 0 closes partition s at step s; a pair superstep runs pair
  // (i, j) at step i * (NP + 1) + j and merges partition i at step
  // i * (NP + 1) + NP. Partition q is used in row i < q as j and in its own
  // row q. Skipped pairs are ignored.
  int cur_i = 0, cur_j = 0;
  bool in_pairs = false;
  const long kRow = long(NP) + 1, kSweep = long(NP) * (NP + 1);
  auto next_use = [&](int q) -> long {
    if (!in_pairs) return q > cur_i ? long(q) : long(NP) + (q == 0 ? 1 : long(q));
    const long now = long(cur_i) * kRow + cur_j;
    long t;
    if (q > cur_j && q > cur_i) t = long(cur_i) * kRow + q;                       // later in this row
    else if (q > cur_i + 1) t = long(cur_i + 1) * kRow + q;                        // next row, as j
    else if (q == cur_i + 1) t = long(q) * kRow + q + 1;                           // next row, as i
    else t = kSweep + (q == 0 ? 1 : long(q));                                      // next superstep
    return t - now;
  };
  auto evict_one = [&]() -> bool {
    // a copy-out in flight frees its buffers first
    for (int p = 0; p < NP; ++p)
      if (pt[p].out_pending) return finish_out(p, true);
    int best = -1;
    long best_next = -1;
    for (int p = 0; p < NP; ++p) {
      if (!pt[p].resident || p == hold[0] || p == hold[1]) continue;
      const long nu = next_use(p);
      if (nu > best_next) {
        best_next = nu;
        best = p;
      }
    }
    return best >= 0 && spill(best, true) == GG_OK;
  };
  struct PressureGuard {
    Engine* e;
    ~PressureGuard() { e->on_pressure = nullptr; }
  } pg{this};
  if (po.spill) on_pressure = evict_one;
  auto resident_count = [&]() {
    int res = 0;
    for (auto& x : pt) res += x.resident;
    return res;
  };
  // device bytes a partition takes when resident (alloc rounds large blocks
  // to 2 MiB)
  auto part_bytes = [&](const Part& q) {
    const uint64_t rows = rows_of(q);
    uint64_t b = 0;
    for (auto& r : q.r) b += (rows + 1) * 8 + std::max<uint64_t>(r.nnz, 1) * 4 + (4u << 20);
    return b;
  };
  // Makes partition p resident. ahead: an upload ahead of its use (the
  // compute stream waits for it when p is fetched for use).
  auto fetch = [&](int p, bool ahead = false) -> int {
    Part& q = pt[p];
    if (q.out_pending) {
      // needed again while its copy-out is in flight: the device buffers
      // still hold it
      const auto c0 = std::chrono::steady_clock::now();
      GG_HIP(hipEventSynchronize(q.ev_out));
      q.out_pending = false;
      q.resident = true;
      copy_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
      return GG_OK;
    }
    if (q.resident) {
      if (q.in_pending && !ahead) {
        GG_HIP(hipStreamWaitEvent(stream, q.ev_in, 0));
        q.in_pending = false;
      }
      return GG_OK;
    }
    if (po.spill && max_resident) {
      int res = resident_count();
      while (res >= max_resident && evict_one()) res = resident_count();
    }
    const auto c0 = std::chrono::steady_clock::now();
    const uint64_t rows = rows_of(q);
    for (auto& r : q.r) {
      GG_TRY(alloc(r.ptr, (rows + 1) * 8, err));
      GG_TRY(alloc(r.col, std::max<uint64_t>(r.nnz, 1) * 4, err));
      if (!r.nnz) fill_u64(r.ptr, 0, rows + 1);
    }
    // the buffers are allocated in the compute stream's order: the copy
    // stream starts after that point
    GG_HIP(hipEventRecord(ev, stream));
    GG_HIP(hipStreamWaitEvent(copy_stream, ev, 0));
    for (auto& r : q.r) {
      if (!r.nnz) continue;
      GG_HIP(hipMemcpyAsync(r.ptr.p, r.hptr.p, (rows + 1) * 8, hipMemcpyHostToDevice, copy_stream));
      GG_HIP(hipMemcpyAsync(r.col.p, r.hcol.p, r.nnz * 4, hipMemcpyHostToDevice, copy_stream));
      if (q.evicted) S.spill_bytes += (rows + 1) * 8 + r.nnz * 4;  // the first upload is the input
    }
    GG_TRY(make_event(q.ev_in));
    GG_HIP(hipEventRecord(q.ev_in, copy_stream));
    if (!ahead) GG_HIP(hipStreamWaitEvent(stream, q.ev_in, 0));
    q.in_pending = ahead;
    q.resident = true;
    copy_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
    return GG_OK;
  };
  // Working memory of a step beyond its resident partitions: the most any
  // step has needed so far (the device peak within the step above the
  // bytes in use when it started).
  uint64_t step_extra = 0, step_base = 0, run_peak = 0;
  auto step_begin = [&]() {
    run_peak = std::max(run_peak, peak);
    step_base = used;
    peak = used;
  };
  auto step_end = [&]() {
    step_extra = std::max(step_extra, peak > step_base ? peak - step_base : 0);
    run_peak = std::max(run_peak, peak);
  };
  // Before a step: finished copy-outs give their buffers back, then the
  // partitions the next steps will fetch (upcoming, in order) are uploaded
  // ahead while they fit, and the first that does not fit gets room by
  // copying out partitions used later than it.
  auto prepare = [&](const std::vector<int>& upcoming) -> int {
    if (!po.spill) return GG_OK;
    reap();
    for (int p : upcoming) {
      if (pt[p].resident || pt[p].out_pending) continue;
      const uint64_t need = part_bytes(pt[p]);
      int res = resident_count();
      if ((!budget || used + need + step_extra <= budget) && (!max_resident || res < max_resident)) {
        GG_TRY(fetch(p, true));
        S.uploads_ahead++;
        continue;
      }
      // room for p: copy out the partitions used after p, farthest first
      uint64_t pending = 0;
      for (auto& x : pt)
        if (x.out_pending) pending += part_bytes(x);
      const long p_next = next_use(p);
      for (;;) {
        const bool fits_mem = !budget || used - std::min(used, pending) + need + step_extra <= budget;
        const bool fits_count = !max_resident || res < max_resident;
        if (fits_mem && fits_count) break;
        int best = -1;
        long best_next = p_next;
        for (int v = 0; v < NP; ++v) {
          if (!pt[v].resident || v == hold[0] || v == hold[1]) continue;
          const long nu = next_use(v);
          if (nu > best_next) {
            best_next = nu;
            best = v;
          }
        }
        if (best < 0) break;
        pending += part_bytes(pt[best]);
        GG_TRY(spill(best, false));
        S.spills_ahead++;
        --res;
      }
      break;
    }
    return GG_OK;
  };
  auto elapsed_s = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  // initial delta of every partition: its input edges, deduplicated, built
  // on the host by up to 16 threads. The edges are bucketed by partition (a
  // count pass and a scatter pass over the same chunks of the edge list),
  // then each partition is sorted into its rows on its own (counting sort by
  // row, then sort and unique within each row); the pinned copies follow one
  // partition after another. Edges whose source is in no partition are left
  // out, as the rows of the partitions are all there is.
  {
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<bool> oom{false};
    auto par = [&](uint64_t count, auto&& f) {  // f(thread, begin, end) on nt chunks of [0, count)
      std::vector<std::thread> th;
      for (unsigned t = 0; t < nt; ++t) {
        const uint64_t a = count * t / nt, b = count * (t + 1) / nt;
        if (a < b)
          th.emplace_back([&f, &oom, t, a, b]() {
            try {
              f(t, a, b);
            } catch (const std::bad_alloc&) {
              oom = true;
            }
          });
      }
      for (auto& x : th) x.join();
    };
    std::vector<uint32_t> pfirst(static_cast<size_t>(NP));
    for (int p = 0; p < NP; ++p) pfirst[size_t(p)] = pt[p].first;
    // partition of vertex v (intervals sorted by first vertex), -1: none
    auto part_of = [&](uint32_t v) -> int {
      const int p = int(std::upper_bound(pfirst.begin(), pfirst.end(), v) - pfirst.begin()) - 1;
      return p >= 0 && v - pt[p].first < pt[p].np ? p : -1;
    };
    std::vector<uint64_t> pcnt(size_t(nt) * NP, 0);
    par(m, [&](unsigned t, uint64_t a, uint64_t b) {
      uint64_t* c = pcnt.data() + size_t(t) * NP;
      for (uint64_t i = a; i < b; ++i) {
        const int p = part_of(h_src[i]);
        if (p >= 0) c[p]++;
      }
    });
    // bucket offsets: partition-major, then thread
    std::vector<uint64_t> pbeg(size_t(NP) + 1, 0);
    uint64_t tot = 0;
    for (int p = 0; p < NP; ++p) {
      pbeg[size_t(p)] = tot;
      for (unsigned t = 0; t < nt; ++t) {
        const uint64_t c = pcnt[size_t(t) * NP + p];
        pcnt[size_t(t) * NP + p] = tot;
        tot += c;
      }
    }
    pbeg[size_t(NP)] = tot;
    std::vector<uint32_t> bsrc(tot), bdst(tot);
    std::vector<uint8_t> blab(tot);
    par(m, [&](unsigned t, uint64_t a, uint64_t b) {
      uint64_t* o = pcnt.data() + size_t(t) * NP;
      for (uint64_t i = a; i < b; ++i) {
        const int p = part_of(h_src[i]);
        if (p < 0) continue;
        const uint64_t k = o[p]++;
        bsrc[k] = h_src[i];
        bdst[k] = h_dst[i];
        blab[k] = h_lab[i];
      }
    });
    std::vector<std::vector<unsigned long long>> lptrs(static_cast<size_t>(NP));
    std::vector<std::vector<uint32_t>> lcols(static_cast<size_t>(NP));
    std::atomic<int> next_p{0};
    par(nt, [&](unsigned, uint64_t, uint64_t) {
      for (int p; (p = next_p.fetch_add(1)) < NP && !oom;) {
        const Part& q = pt[p];
        const uint64_t rows = rows_of(q);
        const uint64_t b0 = pbeg[size_t(p)], b1 = pbeg[size_t(p) + 1];
        std::vector<uint64_t> cnt(rows + 1, 0);
        for (uint64_t k = b0; k < b1; ++k) cnt[uint64_t(blab[k] - 1) * q.np + (bsrc[k] - q.first) + 1]++;
        for (uint64_t r = 0; r < rows; ++r) cnt[r + 1] += cnt[r];
        std::vector<uint32_t> col(b1 - b0);
        {
          std::vector<uint64_t> fillp(cnt.begin(), cnt.end() - 1);
          for (uint64_t k = b0; k < b1; ++k) col[fillp[uint64_t(blab[k] - 1) * q.np + (bsrc[k] - q.first)]++] = bdst[k];
        }
        std::vector<unsigned long long>& lptr = lptrs[size_t(p)];
        std::vector<uint32_t>& lcol = lcols[size_t(p)];
        lptr.assign(rows + 1, 0);
        lcol.reserve(col.size());
        for (uint64_t r = 0; r < rows; ++r) {
          uint32_t* b = col.data() + cnt[r];
          uint32_t* e = col.data() + cnt[r + 1];
          std::sort(b, e);
          e = std::unique(b, e);
          lcol.insert(lcol.end(), b, e);
          lptr[r + 1] = lcol.size();
        }
      }
    });
    if (oom) {
      err = "out of host memory building the partitions' input rows";
      return GG_E_NOMEM;
    }
    std::vector<uint32_t>().swap(bsrc);
    std::vector<uint32_t>().swap(bdst);
    std::vector<uint8_t>().swap(blab);
    for (int p = 0; p < NP; ++p) {
      Part& q = pt[p];
      const uint64_t rows = rows_of(q);
      Region& d = q.r[kD];
      d.nnz = lcols[size_t(p)].size();
      S.input_unique += d.nnz;
      GG_TRY(host_alloc(d.hptr, (rows + 1) * 8));
      GG_TRY(host_alloc(d.hcol, d.nnz * 4));
      std::memcpy(d.hptr.p, lptrs[size_t(p)].data(), (rows + 1) * 8);
      if (d.nnz) std::memcpy(d.hcol.p, lcols[size_t(p)].data(), d.nnz * 4);
      std::vector<unsigned long long>().swap(lptrs[size_t(p)]);
      std::vector<uint32_t>().swap(lcols[size_t(p)]);
      for (int k : {int(kO), int(kT)}) {
        Region& e2 = q.r[k];
        GG_TRY(host_alloc(e2.hptr, (rows + 1) * 8));
        std::memset(e2.hptr.p, 0, (rows + 1) * 8);
        e2.nnz = 0;
      }
      q.has[kD] = true;
    }
  }
  // scratch regions reused by every operator
  Region F, M;
  struct ScratchGuard {
    Engine* e;
    Region* r[2];
    ~ScratchGuard() {
      for (Region* x : r) {
        e->release(x->ptr);
        e->release(x->col);
      }
    }
  } sg{this, {&F, &M}};
  // M = a U b (disjoint sorted rows), then swap M into a
  auto merge_into = [&](Region& a, Region& b, uint64_t rows, uint32_t np) -> int {
    GG_TRY(grow(M.ptr, (rows + 1) * 8, err));
    GG_TRY(grow(M.col, std::max<uint64_t>(a.nnz + b.nnz, 1) * 4, err));
    k_add_ptr<<<grid_for(rows + 1), kBlock, 0, stream>>>(P<unsigned long long>(M.ptr), P<unsigned long long>(a.ptr),
                                                        P<unsigned long long>(b.ptr), rows + 1);
    GG_TRY(merge_rows(dcsr(a), dcsr(b), uint32_t(rows), np, nullptr, DPacked{}, nullptr, P<uint32_t>(M.col), nullptr,
                      nullptr, err));
    GG_HIP(hipGetLastError());
    M.nnz = a.nnz + b.nnz;
    std::swap(a.ptr, M.ptr);
    std::swap(a.col, M.col);
    std::swap(a.nnz, M.nnz);
    return GG_OK;
  };
  struct Dst {
    Region* r;
    int part;
  };
  // One operator on partition p; its new ids are merged into `seed`
  // (kinds 0-2) or become `into` (kind 3).
  auto op = [&](int p, int kind, Region* key, const std::vector<Dst>& dsts, Region* seed, Region* into,
                uint64_t* nfresh) -> int {
    Part& q = pt[p];
    const uint64_t rows = rows_of(q);
    ExpandArgs a{};
    a.n = n;
    a.g = g;
    a.grab = grab;
    a.active = P<uint32_t>(active);
    a.ctr = C();
    a.seg_off = P<unsigned long long>(seg_off);
    a.seg_cnt = P<uint32_t>(seg_cnt);
    a.ro.kind = kind;
    a.ro.first = q.first;
    a.ro.np = q.np;
    a.ro.seed = dcsr(*seed);
    a.ro.key = key ? dcsr(*key) : DCsr{nullptr, nullptr};
    a.ro.ndst = int(dsts.size());
    for (size_t d = 0; d < dsts.size(); ++d) {
      a.ro.dst[d] = dcsr(*dsts[d].r);
      a.ro.dfirst[d] = pt[dsts[d].part].first;
      a.ro.dn[d] = pt[dsts[d].part].np;
    }
    for (;;) {
      a.staging = P<uint32_t>(staging);
      a.staging_cap = staging.bytes / 4;
      zero_ctr();
      fill_u32(seg_cnt, 0, rows);
      // only rows that can have candidates are expanded: for a join, rows
      // with a key where a destination region has rows
      for (size_t d = 0; d < dsts.size(); ++d) {
        const uint32_t bx = std::max<uint32_t>(1, std::min<uint32_t>(grid_for(a.ro.dn[d]), uint32_t(cus) / L));
        k_region_span<<<dim3(bx, uint32_t(L)), kBlock, 0, stream>>>(a.ro.dst[d], a.ro.dn[d], int(d), C());
      }
      k_region_active<<<std::min<uint32_t>(grid_for(rows), uint32_t(cus)), kBlock, 0, stream>>>(a, uint32_t(rows));
      k_expand<2><<<expand_grid(), kBlock, 0, stream>>>(a);
      GG_HIP(hipGetLastError());
      GG_TRY(read_ctr(err));
      S.expand_launches++;
      if (!h_ctr->overflow) break;
      S.retries++;
      GG_TRY(grow(staging, size_t(h_ctr->cursor + h_ctr->cursor / 4 + 1024) * 4, err));
    }
    const uint64_t N = h_ctr->newe;
    *nfresh = N;
    S.candidates += h_ctr->cand;
    if (!into && N == 0) return GG_OK;
    GG_TRY(grow(F.ptr, (rows + 1) * 8, err));
    GG_TRY(scan(seg_cnt, rows, F.ptr, err));
    GG_TRY(grow(F.col, std::max<uint64_t>(N, 1) * 4, err));
    if (N)
      k_gather<<<gather_grid(), kBlock, 0, stream>>>(P<uint32_t>(active), C(), P<uint32_t>(seg_cnt),
                                                     P<unsigned long long>(seg_off), P<uint32_t>(staging),
                                                     P<unsigned long long>(F.ptr), P<uint32_t>(F.col), n, DPacked{},
                                                     nullptr, nullptr);
    F.nnz = N;
    if (into) {
      std::swap(into->ptr, F.ptr);
      std::swap(into->col, F.col);
      into->nnz = N;
    } else {
      GG_TRY(merge_into(*seed, F, rows, q.np));
    }
    return GG_OK;
  };
  bool changed = false;
  auto join = [&](int p, int key, std::vector<std::pair<int, int>> dsts) -> int {
    if (!pt[p].has[key]) return GG_OK;
    std::vector<Dst> ds;
    for (auto& d : dsts)
      if (pt[d.second].has[d.first]) ds.push_back({&pt[d.second].r[d.first], d.second});
    if (ds.empty()) return GG_OK;
    uint64_t nf = 0;
    GG_TRY(op(p, 0, &pt[p].r[key], ds, &pt[p].r[kT], nullptr, &nf));
    changed |= nf > 0;
    return GG_OK;
  };
  auto unary = [&](int p, int from) -> int {
    if (!pt[p].has[from] || g.nu == 0) return GG_OK;
    uint64_t nf = 0;
    GG_TRY(op(p, 1, &pt[p].r[from], {}, &pt[p].r[kT], nullptr, &nf));
    changed |= nf > 0;
    return GG_OK;
  };
  auto eps = [&](int p) -> int {
    if (g.ne == 0) return GG_OK;
    uint64_t nf = 0;
    GG_TRY(op(p, 2, nullptr, {}, &pt[p].r[kT], nullptr, &nf));
    changed |= nf > 0;
    return GG_OK;
  };
  auto merge_and_diff = [&](int p) -> int {
    Part& q = pt[p];
    const uint64_t rows = rows_of(q);
    if (q.has[kD]) {
      if (!q.has[kO]) {
        std::swap(q.r[kO].ptr, q.r[kD].ptr);
        std::swap(q.r[kO].col, q.r[kD].col);
        std::swap(q.r[kO].nnz, q.r[kD].nnz);
      } else {
        GG_TRY(merge_into(q.r[kO], q.r[kD], rows, q.np));
      }
      GG_TRY(make_empty(q.r[kD], rows));
      q.has[kO] = true;
      q.has[kD] = false;
    }
    if (q.has[kT]) {
      uint64_t nf = 0;
      GG_TRY(op(p, 3, &q.r[kT], {}, &q.r[kO], &q.r[kD], &nf));
      q.has[kD] = nf > 0;
      GG_TRY(make_empty(q.r[kT], rows));
      q.has[kT] = false;
    }
    return GG_OK;
  };
  // cross-partition edge counts of partition p (all three regions), the
  // pair weights of the DDM schedule
  std::vector<std::vector<uint64_t>> ddm(size_t(NP), std::vector<uint64_t>(size_t(NP), 0));
  DevBuf d_bounds, d_hist;
  struct BGuard {
    Engine* e;
    DevBuf* b[2];
    ~BGuard() {
      for (DevBuf* x : b) e->release(*x);
    }
  } bg{this, {&d_bounds, &d_hist}};
  if (po.ddm) {
    std::vector<uint32_t> bounds(size_t(NP) + 1);
    for (int p = 0; p < NP; ++p) bounds[p] = pt[p].first;
    bounds[NP] = n;
    GG_TRY(alloc(d_bounds, bounds.size() * 4, err));
    GG_TRY(alloc(d_hist, size_t(NP) * 8, err));
    GG_HIP(hipMemcpy(d_bounds.p, bounds.data(), bounds.size() * 4, hipMemcpyHostToDevice));
  }
  std::vector<unsigned long long> hist(static_cast<size_t>(NP));
  auto cross = [&](int p) -> int {
    if (!po.ddm) return GG_OK;
    Part& q = pt[p];
    fill_u64(d_hist, 0, uint64_t(NP));
    for (auto& r : q.r)
      if (r.nnz)
        k_cross<<<uint32_t(cus) * 4, kBlock, 0, stream>>>(dcsr(r), uint32_t(rows_of(q)), P<uint32_t>(d_bounds),
                                                         uint32_t(NP), P<unsigned long long>(d_hist));
    GG_HIP(hipGetLastError());
    GG_HIP(hipMemcpyAsync(hist.data(), d_hist.p, size_t(NP) * 8, hipMemcpyDeviceToHost, stream));
    GG_HIP(hipStreamSynchronize(stream));
    for (int j = 0; j < NP; ++j) ddm[p][j] = j == p ? 0 : hist[j];
    return GG_OK;
  };
  if (!po.spill)
    for (int p = 0; p < NP; ++p) GG_TRY(fetch(p));
  if (verbose)
    std::fprintf(stderr, "[gg] partitioned: %d partitions, input on the host after %.1f s\n", NP, elapsed_s());
  // with the DDM schedule a pair runs only when edges cross between the two
  // partitions and one of them has new edges (delta or tmp) to push
  auto dirty = [&](int p) { return pt[p].has[kD] || pt[p].r[kT].nnz > 0; };
  auto pair_runs = [&](int i, int j) { return !po.ddm || (ddm[i][j] + ddm[j][i] != 0 && (dirty(i) || dirty(j))); };
  // the partitions, not resident now, that the next two steps of the sweep
  // fetch, in order (superstep 0 looks no further than its end)
  constexpr int kAhead = 2;
  auto upcoming = [&]() {
    std::vector<int> up;
    int steps = 0;
    auto add = [&](int q) {
      if (!pt[q].resident && q != hold[0] && q != hold[1] && std::find(up.begin(), up.end(), q) == up.end())
        up.push_back(q);
    };
    if (!in_pairs) {
      for (int q = cur_i + 1; q < NP && q <= cur_i + kAhead; ++q) add(q);
      return up;
    }
    for (int i = cur_i; i < NP && steps < kAhead; ++i) {
      for (int j = i == cur_i ? cur_j + 1 : i + 1; j < NP && steps < kAhead; ++j)
        if (pair_runs(i, j)) {
          add(i);
          add(j);
          ++steps;
        }
      if (i > cur_i && steps < kAhead && (pt[i].has[kD] || pt[i].has[kT])) {
        add(i);  // row i's merge step
        ++steps;
      }
    }
    return up;
  };
  // superstep 0 (precomputation)
  for (int p = 0; p < NP; ++p) {
    hold[0] = p;
    hold[1] = -1;
    cur_i = p;
    GG_TRY(fetch(p));
    GG_TRY(prepare(upcoming()));
    step_begin();
    for (;;) {
      Part& q = pt[p];
      q.has[kT] = true;
      if (!q.has[kO]) GG_TRY(eps(p));
      if (filter == 0) GG_TRY(unary(p, kD));
      if (q.has[kO]) {
        if (filter == 0) GG_TRY(join(p, kO, {{kD, p}}));
        GG_TRY(join(p, kD, {{kO, p}, {kD, p}}));
      } else {
        GG_TRY(join(p, kD, {{kD, p}}));
      }
      GG_TRY(merge_and_diff(p));
      S.rounds++;
      if (!q.has[kD]) {
        std::swap(q.r[kD].ptr, q.r[kO].ptr);
        std::swap(q.r[kD].col, q.r[kO].col);
        std::swap(q.r[kD].nnz, q.r[kO].nnz);
        q.has[kD] = q.has[kO];
        q.has[kO] = false;
        break;
      }
    }
    step_end();
    GG_TRY(cross(p));
    if (verbose)
      std::fprintf(stderr,
                   "[gg] superstep 0: %d/%d partitions, %u rounds, %u launches, %u retries, closure %llu, "
                   "%d resident, %.2f GB used, %.2f GB moved (%.1f s), %llu allocs (%.1f s), %.1f s\n",
                   p + 1, NP, S.rounds, S.expand_launches, S.retries,
                   (unsigned long long)(pt[p].r[kD].nnz + pt[p].r[kO].nnz), resident_count(), used / 1e9,
                   S.spill_bytes / 1e9, copy_ms / 1e3, (unsigned long long)alloc_calls, alloc_ms / 1e3, elapsed_s());
  }
  // supersteps 1.. (run_computation)
  uint32_t supersteps = 0;
  in_pairs = true;
  for (;;) {
    ++supersteps;
    for (int i = 0; i < NP; ++i) {
      for (int j = i + 1; j < NP; ++j) {
        if (!pair_runs(i, j)) {
          S.pairs_skipped++;
          continue;
        }
        hold[0] = i;
        hold[1] = j;
        cur_i = i;
        cur_j = j;
        GG_TRY(fetch(i));
        GG_TRY(fetch(j));
        GG_TRY(prepare(upcoming()));
        step_begin();
        S.pairs_run++;
        pt[i].has[kT] = true;
        pt[j].has[kT] = true;
        if (filter == 0) GG_TRY(join(i, kO, {{kD, j}}));
        GG_TRY(join(i, kD, {{kO, j}, {kD, j}}));
        if (filter == 0) GG_TRY(join(j, kO, {{kD, i}}));
        GG_TRY(join(j, kD, {{kO, i}, {kD, i}}));
        for (;;) {
          changed = false;
          for (int side = 0; side < 2; ++side) {
            const int p = side ? j : i, o = side ? i : j;
            if (filter == 0) GG_TRY(join(p, kO, {{kT, i}, {kT, j}}));
            GG_TRY(join(p, kD, {{kT, i}, {kT, j}}));
            GG_TRY(join(p, kT, {{kO, p}, {kD, p}}));
            GG_TRY(join(p, kT, {{kO, o}, {kD, o}}));
            if (filter == 0) GG_TRY(unary(p, kT));
            GG_TRY(join(p, kT, {{kT, i}, {kT, j}}));
          }
          S.rounds++;
          if (!changed) break;
        }
        step_end();
        GG_TRY(cross(i));
        GG_TRY(cross(j));
      }
      // a partition with neither delta nor tmp has nothing to merge
      if (pt[i].has[kD] || pt[i].has[kT]) {
        hold[0] = i;
        hold[1] = -1;
        cur_i = i;
        cur_j = NP;
        GG_TRY(fetch(i));
        GG_TRY(prepare(upcoming()));
        step_begin();
        GG_TRY(merge_and_diff(i));
        step_end();
        GG_TRY(cross(i));
      }
    }
    bool done = true;
    for (auto& q : pt) done &= !q.has[kD];
    if (verbose)
      std::fprintf(stderr,
                   "[gg] superstep %u: pairs run %u skipped %u, %d resident, %.2f GB moved (%.1f s), %u uploads and "
                   "%u copy-outs ahead, %.1f s\n",
                   supersteps, S.pairs_run, S.pairs_skipped, resident_count(), S.spill_bytes / 1e9, copy_ms / 1e3,
                   S.uploads_ahead, S.spills_ahead, elapsed_s());
    if (done) break;
  }
  hold[0] = hold[1] = -1;
  // copies still in flight end before the closure is assembled
  GG_HIP(hipStreamSynchronize(copy_stream));
  for (int p = 0; p < NP; ++p) {
    finish_out(p, true);
    pt[p].in_pending = false;
  }
  peak = std::max(run_peak, peak);
  supersteps_r = supersteps;
  // the closure is every partition's old region: assemble the global CSR
  GG_HIP(hipStreamSynchronize(stream));
  const uint64_t R = rows();
  std::vector<unsigned long long> gptr(R + 1, 0);
  std::vector<std::vector<unsigned long long>> lptr(static_cast<size_t>(NP));
  for (int p = 0; p < NP; ++p) {
    Part& q = pt[p];
    lptr[p].assign(rows_of(q) + 1, 0);
    if (q.r[kO].nnz == 0) continue;
    if (q.resident)
      GG_HIP(hipMemcpy(lptr[p].data(), q.r[kO].ptr.p, lptr[p].size() * 8, hipMemcpyDeviceToHost));
    else
      std::memcpy(lptr[p].data(), q.r[kO].hptr.p, lptr[p].size() * 8);
  }
  for (int l = 0; l < L; ++l)
    for (int p = 0; p < NP; ++p)
      for (uint32_t k = 0; k < pt[p].np; ++k) {
        const uint64_t lr = uint64_t(l) * pt[p].np + k;
        const uint64_t gr = uint64_t(l) * n + pt[p].first + k;
        gptr[gr + 1] = lptr[p][lr + 1] - lptr[p][lr];
      }
  for (uint64_t r = 0; r < R; ++r) gptr[r + 1] += gptr[r];
  const uint64_t total = gptr[R];
  // the per-round scratch is not needed any more
  for (DevBuf* b : {&F.ptr, &F.col, &M.ptr, &M.col, &staging, &seg_off, &seg_cnt, &active}) release(*b);
  GG_TRY(alloc(old_ptr, (R + 1) * 8, err));
  GG_TRY(alloc(old_col, std::max<uint64_t>(total, 1) * 4, err));
  GG_HIP(hipMemcpy(old_ptr.p, gptr.data(), (R + 1) * 8, hipMemcpyHostToDevice));
  for (int l = 0; l < L; ++l)
    for (int p = 0; p < NP; ++p) {
      Part& q = pt[p];
      const uint64_t lb = lptr[p][uint64_t(l) * q.np], le = lptr[p][uint64_t(l + 1) * q.np];
      if (le <= lb) continue;
      uint32_t* dst = static_cast<uint32_t*>(old_col.p) + gptr[uint64_t(l) * n + q.first];
      if (q.resident)
        GG_HIP(hipMemcpy(dst, static_cast<uint32_t*>(q.r[kO].col.p) + lb, (le - lb) * 4, hipMemcpyDeviceToDevice));
      else
        GG_HIP(hipMemcpy(dst, static_cast<uint32_t*>(q.r[kO].hcol.p) + lb, (le - lb) * 4, hipMemcpyHostToDevice));
    }
  GG_HIP(hipStreamSynchronize(stream));
  S.total_edges = total;
  S.fixpoint_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  S.device_bytes_peak = peak;
  old_nnz = total;
  has_closure = true;
  if (st) *st = S;
  return GG_OK;
}

// ---------------------------------------------------------------------------
// engine API

Engine* engine_create(int device, uint64_t budget, int verbose, int& status, std::string& err) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    (void)hipGetLastError();
    status = GG_E_DEVICE;
    err = "no HIP device available";
    return nullptr;
  }
  if (device < 0) {
    int best = -1;
    for (int d = 0; d < count; ++d) {
      hipDeviceProp_t p;
      if (hipGetDeviceProperties(&p, d) == hipSuccess && p.multiProcessorCount > best) {
        best = p.multiProcessorCount;
        device = d;
      }
    }
  }
  if (device < 0 || device >= count) {
    status = GG_E_ARG;
    err = "device ordinal out of range";
    return nullptr;
  }
  auto* en = new Engine();
  en->device = device;
  en->budget = budget;
  en->verbose = verbose;
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreate(&en->stream) != hipSuccess) {
    (void)hipGetLastError();
    delete en;
    status = GG_E_DEVICE;
    err = "cannot initialise HIP device " + std::to_string(device);
    return nullptr;
  }
  en->cus = prop.multiProcessorCount;
  int bpc = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, reinterpret_cast<const void*>(k_expand<0>), kBlock, 0) ==
          hipSuccess &&
      bpc > 0)
    en->expand_blocks_per_cu = en->occ_expand = bpc;
  // the merge and gather kernels hand each wave a fixed share of the rows, so their
  // grids are what fits on the device at once (no second, late wave of blocks)
  bpc = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, reinterpret_cast<const void*>(k_merge), kBlock, 0) ==
          hipSuccess &&
      bpc > 0)
    en->merge_blocks_per_cu = en->occ_merge = bpc;
  bpc = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, reinterpret_cast<const void*>(k_merge_rows), kBlock, 0) ==
          hipSuccess &&
      bpc > 0)
    en->merge_rows_blocks_per_cu = en->occ_merge_rows = bpc;
  bpc = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, reinterpret_cast<const void*>(k_gather), kBlock, 0) ==
          hipSuccess &&
      bpc > 0)
    en->gather_blocks_per_cu = en->occ_gather = bpc;
  (void)hipGetLastError();
  status = GG_OK;
  return en;
}

void engine_destroy(Engine* e) { delete e; }

void engine_trim(Engine* e) {
  if (!e->arena) e->trim_cache();
}

namespace {
// Device form of a rule table: rules by kind, pair-list slots, the labels
// that are the lhs of a unary or binary rule, and the labels kept packed.
int build_rule_set(const std::vector<std::array<int32_t, 3>>& rules, int num_labels, Engine::RuleSet& rs,
                   std::string& err) {
  DevGrammar g{};
  std::vector<bool> is_lhs(size_t(num_labels) + 1, false);
  for (auto& r : rules) {
    if (r[1] == 0) {
      if (g.ne >= kMaxRules) {
        err = "more than 64 eps rules";
        return GG_E_UNSUPPORTED;
      }
      g.e_lhs[g.ne++] = uint8_t(r[0] - 1);
    } else if (r[2] == 0) {
      if (g.nu >= kMaxRules) {
        err = "more than 64 unary rules";
        return GG_E_UNSUPPORTED;
      }
      g.u_lhs[g.nu] = uint8_t(r[0] - 1);
      g.u_rhs[g.nu++] = uint8_t(r[1] - 1);
      is_lhs[r[0]] = true;
    } else {
      if (g.nb >= kMaxRules) {
        err = "more than 64 binary rules";
        return GG_E_UNSUPPORTED;
      }
      g.b_lhs[g.nb] = uint8_t(r[0] - 1);
      g.b_rhs1[g.nb] = uint8_t(r[1] - 1);
      g.b_rhs2[g.nb++] = uint8_t(r[2] - 1);
      is_lhs[r[0]] = true;
    }
  }
  // pair-list slots: binary rules whose C is derived by a unary/binary rule
  for (int i = 0; i < g.nb; ++i) {
    g.b_slot[i] = 0xFF;
    const int C = g.b_rhs2[i] + 1, B = g.b_rhs1[i] + 1;
    if (!is_lhs[C]) continue;
    g.slot_c[g.nslot] = uint8_t(C - 1);
    g.slot_b[g.nslot] = uint8_t(B - 1);
    g.b_slot[i] = uint8_t(g.nslot++);
  }
  for (int i = 0; i < 8; ++i) rs.pack[i] = 0;
  auto mark_pack = [&](int l0) { rs.pack[l0 >> 5] |= 1u << (l0 & 31); };
  for (int i = 0; i < g.nb; ++i) mark_pack(g.b_rhs2[i]);
  for (int l = 1; l <= num_labels; ++l)
    if (is_lhs[l]) mark_pack(l - 1);
  if (const char* np = std::getenv("GG_NO_PACK"))
    if (np[0] == '1')
      for (int i = 0; i < 8; ++i) rs.pack[i] = 0;
  rs.g = g;
  rs.lhs.clear();
  for (int l = 1; l <= num_labels; ++l)
    if (is_lhs[l]) rs.lhs.push_back(uint8_t(l - 1));
  return GG_OK;
}
}  // namespace

int engine_set_grammar(Engine* e, const Grammar& gr, std::string& err) {
  Engine::RuleSet given, planned;
  GG_TRY(build_rule_set(gr.rules, gr.num_labels(), given, err));
  std::vector<int32_t> flat;
  for (auto& r : gr.rules) flat.insert(flat.end(), r.begin(), r.end());
  const std::vector<int32_t> pl = plan_fixpoint_rules(flat.data(), int(gr.rules.size()), kMaxRules);
  std::vector<std::array<int32_t, 3>> prules;
  for (size_t i = 0; i + 2 < pl.size(); i += 3) prules.push_back({pl[i], pl[i + 1], pl[i + 2]});
  GG_TRY(build_rule_set(prules, gr.num_labels(), planned, err));
  e->rules_given = given;
  e->rules_planned = planned;
  e->use_rules(e->rules_planned);
  e->L = gr.num_labels();
  e->has_grammar = true;
  e->has_closure = false;
  return GG_OK;
}

namespace {
int upload_graph(Engine* e, const Graph& gr, std::string& err) {
  GG_HIP(hipSetDevice(e->device));
  e->n = gr.n;
  e->m = gr.m();
  GG_TRY(e->alloc(e->in_src, e->m * 4, err));
  GG_TRY(e->alloc(e->in_dst, e->m * 4, err));
  GG_TRY(e->alloc(e->in_lab, e->m, err));
  if (e->m) {
    GG_HIP(hipMemcpyAsync(e->in_src.p, gr.src.data(), e->m * 4, hipMemcpyHostToDevice, e->stream));
    GG_HIP(hipMemcpyAsync(e->in_dst.p, gr.dst.data(), e->m * 4, hipMemcpyHostToDevice, e->stream));
    GG_HIP(hipMemcpyAsync(e->in_lab.p, gr.lab.data(), e->m, hipMemcpyHostToDevice, e->stream));
  }
  GG_HIP(hipStreamSynchronize(e->stream));
  e->has_graph = true;
  e->has_closure = false;
  return GG_OK;
}
}  // namespace

int engine_set_graph(Engine* e, const Graph& gr, std::string& err) {
  GG_TRY(upload_graph(e, gr, err));
  e->h_src = gr.src;
  e->h_dst = gr.dst;
  e->h_lab = gr.lab;
  return GG_OK;
}

int engine_set_graph(Engine* e, Graph&& gr, std::string& err) {
  GG_TRY(upload_graph(e, gr, err));
  e->h_src = std::move(gr.src);
  e->h_dst = std::move(gr.dst);
  e->h_lab = std::move(gr.lab);
  return GG_OK;
}

int engine_run(Engine* e, int filter, gg_stats* stats, std::string& err) {
  e->has_closure = false;
  const int s = e->run(filter, stats, err);
  if (s == GG_OK && stats) {
    stats->supersteps = 0;
    stats->partitions = 1;
  }
  return s;
}

int engine_run_partitioned(Engine* e, int filter, const std::vector<std::pair<uint32_t, uint32_t>>& parts,
                           const PartOpts& po, gg_stats* stats, std::string& err) {
  e->has_closure = false;
  if (parts.size() > kMaxCrossParts) {
    err = "at most " + std::to_string(kMaxCrossParts) + " partitions";
    return GG_E_UNSUPPORTED;
  }
  gg_stats st{};
  const int s = e->run_r(filter, parts, po, &st, err);
  if (s == GG_OK && stats) {
    *stats = st;
    stats->supersteps = e->supersteps_r;
    stats->partitions = uint32_t(parts.size());
  }
  return s;
}

namespace {
int need_closure(Engine* e, std::string& err) {
  if (!e->has_closure) {
    err = "no closure: call gg_closure_run first";
    return GG_E_STATE;
  }
  return GG_OK;
}

// writes the packed labels' ids that run() left pending
int ensure_ids(Engine* e, std::string& err) {
  if (!e->ids_pending) return GG_OK;
  GG_HIP(hipSetDevice(e->device));
  DPacked plab{};
  for (int i = 0; i < 8; ++i) plab.labels[i] = e->pack_labels[i];
  plab.len = static_cast<uint32_t*>(e->pk_len.p);
  plab.ent = static_cast<unsigned long long*>(e->pk_ent.p);
  k_unpack<<<uint32_t(e->cus) * 8, kBlock, 0, e->stream>>>(static_cast<unsigned long long*>(e->old_ptr.p),
                                                          static_cast<uint32_t*>(e->old_col.p), plab, e->rows(), e->n);
  GG_HIP(hipGetLastError());
  e->ids_pending = false;
  return GG_OK;
}

// need_closure, and the closure's ids written
int need_ids(Engine* e, std::string& err) {
  GG_TRY(need_closure(e, err));
  return ensure_ids(e, err);
}
}  // namespace

int engine_counts(Engine* e, uint64_t* per_label, uint64_t* total, std::string& err) {
  GG_TRY(need_closure(e, err));
  GG_HIP(hipSetDevice(e->device));
  std::vector<unsigned long long> b(size_t(e->L) + 1);
  for (int l = 0; l <= e->L; ++l)
    GG_HIP(hipMemcpy(&b[l], static_cast<unsigned long long*>(e->old_ptr.p) + uint64_t(l) * e->n, 8,
                     hipMemcpyDeviceToHost));
  if (per_label)
    for (int l = 0; l < e->L; ++l) per_label[l] = b[l + 1] - b[l];
  if (total) *total = b[e->L];
  return GG_OK;
}

int engine_degrees(Engine* e, uint64_t* deg, std::string& err) {
  GG_TRY(need_closure(e, err));
  GG_HIP(hipSetDevice(e->device));
  const uint64_t R = e->rows();
  std::vector<unsigned long long> ptr(R + 1);
  GG_HIP(hipMemcpy(ptr.data(), e->old_ptr.p, (R + 1) * 8, hipMemcpyDeviceToHost));
  for (uint32_t u = 0; u < e->n; ++u) {
    uint64_t s = 0;
    for (int l = 0; l < e->L; ++l) {
      const uint64_t r = uint64_t(l) * e->n + u;
      s += ptr[r + 1] - ptr[r];
    }
    deg[u] = s;
  }
  return GG_OK;
}

// Triples of vertices [v0, v1) and combined labels [l0, l0+lg), renumbered to
// the slice (vertex - v0, label - l0 + 1), sorted by (src, dst, label).
static int export_slice(Engine* e, uint32_t v0, uint32_t v1, uint32_t l0, uint32_t lg, uint32_t* buf, uint64_t cap,
                        uint64_t* count, std::string& err) {
  GG_HIP(hipSetDevice(e->device));
  const uint32_t n = e->n, ns = v1 - v0;
  // row pointers of the slice's rows, label by label
  std::vector<unsigned long long> ptr(size_t(lg) * (ns + 1));
  for (uint32_t l = 0; l < lg; ++l)
    GG_HIP(hipMemcpy(ptr.data() + size_t(l) * (ns + 1),
                     static_cast<unsigned long long*>(e->old_ptr.p) + uint64_t(l0 + l) * n + v0, (ns + 1) * 8,
                     hipMemcpyDeviceToHost));
  uint64_t total = 0;
  for (uint32_t l = 0; l < lg; ++l) total += ptr[size_t(l) * (ns + 1) + ns] - ptr[size_t(l) * (ns + 1)];
  if (count) *count = total;
  if (buf == nullptr || cap < total) {
    err = "export buffer too small: need " + std::to_string(total) + " triples";
    return GG_E_ARG;
  }
  // ids of each label's rows (contiguous per label)
  std::vector<std::vector<uint32_t>> col(lg);
  for (uint32_t l = 0; l < lg; ++l) {
    const unsigned long long b = ptr[size_t(l) * (ns + 1)], en = ptr[size_t(l) * (ns + 1) + ns];
    col[l].resize(en - b);
    if (en > b)
      GG_HIP(hipMemcpy(col[l].data(), static_cast<uint32_t*>(e->old_col.p) + b, (en - b) * 4, hipMemcpyDeviceToHost));
  }
  std::vector<uint64_t> off(size_t(ns) + 1, 0);
  for (uint32_t u = 0; u < ns; ++u) {
    uint64_t s = 0;
    for (uint32_t l = 0; l < lg; ++l) s += ptr[size_t(l) * (ns + 1) + u + 1] - ptr[size_t(l) * (ns + 1) + u];
    off[u + 1] = off[u] + s;
  }
  // per vertex: merge the lg sorted label rows by (dst, label)
  auto work = [&](uint32_t u0, uint32_t u1) {
    std::vector<uint64_t> pos(lg), end(lg);
    for (uint32_t u = u0; u < u1; ++u) {
      for (uint32_t l = 0; l < lg; ++l) {
        const unsigned long long lb = ptr[size_t(l) * (ns + 1)];
        pos[l] = ptr[size_t(l) * (ns + 1) + u] - lb;
        end[l] = ptr[size_t(l) * (ns + 1) + u + 1] - lb;
      }
      uint32_t* o = buf + 3 * off[u];
      for (;;) {
        int best = -1;
        uint32_t bd = 0;
        for (uint32_t l = 0; l < lg; ++l) {
          if (pos[l] < end[l] && (best < 0 || col[l][pos[l]] < bd)) {
            best = int(l);
            bd = col[l][pos[l]];
          }
        }
        if (best < 0) break;
        o[0] = u;
        o[1] = bd - v0;
        o[2] = uint32_t(best + 1);
        o += 3;
        ++pos[size_t(best)];
      }
    }
  };
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  const uint32_t per = (ns + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const uint32_t a = std::min<uint64_t>(ns, uint64_t(t) * per), b = std::min<uint64_t>(ns, uint64_t(t + 1) * per);
    if (a < b) th.emplace_back(work, a, b);
  }
  for (auto& t : th) t.join();
  return GG_OK;
}

int engine_export(Engine* e, uint32_t* buf, uint64_t cap, uint64_t* count, std::string& err) {
  GG_TRY(need_ids(e, err));
  return export_slice(e, 0, e->n, 0, uint32_t(e->L), buf, cap, count, err);
}

int engine_slices(Engine* e, const std::vector<uint32_t>& v0, const std::vector<uint32_t>& l0,
                  const std::vector<uint32_t>& lg, std::vector<uint64_t>& counts, std::vector<uint64_t>& sums,
                  std::string& err) {
  GG_TRY(need_closure(e, err));
  GG_HIP(hipSetDevice(e->device));
  const uint32_t G = uint32_t(l0.size());
  std::vector<uint32_t> vg(e->n), coff(G);
  uint32_t nc = 0;
  for (uint32_t g = 0; g < G; ++g) {
    for (uint32_t u = v0[g]; u < v0[g + 1]; ++u) vg[u] = g;
    coff[g] = nc;
    nc += lg[g];
  }
  DevBuf d_v0, d_vg, d_l0, d_lg, d_coff, d_cnt, d_sum;
  struct Guard {
    Engine* e;
    DevBuf* b[7];
    ~Guard() {
      for (DevBuf* x : b) e->release(*x);
    }
  } guard{e, {&d_v0, &d_vg, &d_l0, &d_lg, &d_coff, &d_cnt, &d_sum}};
  GG_TRY(e->alloc(d_v0, (G + 1) * 4, err));
  GG_TRY(e->alloc(d_vg, uint64_t(e->n) * 4, err));
  GG_TRY(e->alloc(d_l0, G * 4, err));
  GG_TRY(e->alloc(d_lg, G * 4, err));
  GG_TRY(e->alloc(d_coff, G * 4, err));
  GG_TRY(e->alloc(d_cnt, uint64_t(nc) * 8, err));
  GG_TRY(e->alloc(d_sum, uint64_t(G) * 8, err));
  GG_HIP(hipMemcpyAsync(d_v0.p, v0.data(), (G + 1) * 4, hipMemcpyHostToDevice, e->stream));
  GG_HIP(hipMemcpyAsync(d_vg.p, vg.data(), uint64_t(e->n) * 4, hipMemcpyHostToDevice, e->stream));
  GG_HIP(hipMemcpyAsync(d_l0.p, l0.data(), G * 4, hipMemcpyHostToDevice, e->stream));
  GG_HIP(hipMemcpyAsync(d_lg.p, lg.data(), G * 4, hipMemcpyHostToDevice, e->stream));
  GG_HIP(hipMemcpyAsync(d_coff.p, coff.data(), G * 4, hipMemcpyHostToDevice, e->stream));
  e->fill_u64(d_sum, 0, G);
  const DSlices sl{static_cast<uint32_t*>(d_v0.p), static_cast<uint32_t*>(d_vg.p), static_cast<uint32_t*>(d_l0.p),
                   static_cast<uint32_t*>(d_lg.p), static_cast<uint32_t*>(d_coff.p), G};
  k_slice_counts<<<e->grid_for(G), kBlock, 0, e->stream>>>(static_cast<unsigned long long*>(e->old_ptr.p), e->n, sl,
                                                           static_cast<unsigned long long*>(d_cnt.p));
  // ids left pending by the run: the packed labels' rows are read as entries
  DPacked plab{};
  for (int i = 0; i < 8; ++i) plab.labels[i] = e->pack_labels[i];
  plab.len = e->ids_pending ? static_cast<uint32_t*>(e->pk_len.p) : nullptr;
  plab.ent = static_cast<unsigned long long*>(e->pk_ent.p);
  k_slice_checksum<<<uint32_t(e->cus) * 8, kBlock, 0, e->stream>>>(
      {static_cast<unsigned long long*>(e->old_ptr.p), static_cast<uint32_t*>(e->old_col.p)}, plab, e->rows(), e->n, sl,
      static_cast<unsigned long long*>(d_sum.p));
  GG_HIP(hipGetLastError());
  counts.assign(nc, 0);
  sums.assign(G, 0);
  GG_HIP(hipMemcpyAsync(counts.data(), d_cnt.p, uint64_t(nc) * 8, hipMemcpyDeviceToHost, e->stream));
  GG_HIP(hipMemcpyAsync(sums.data(), d_sum.p, uint64_t(G) * 8, hipMemcpyDeviceToHost, e->stream));
  GG_HIP(hipStreamSynchronize(e->stream));
  return GG_OK;
}

int engine_export_slice(Engine* e, uint32_t v0, uint32_t v1, uint32_t l0, uint32_t lg, uint32_t* buf, uint64_t cap,
                        uint64_t* count, std::string& err) {
  GG_TRY(need_ids(e, err));
  return export_slice(e, v0, v1, l0, lg, buf, cap, count, err);
}

int engine_checksum(Engine* e, uint64_t* sum, std::string& err) {
  GG_TRY(need_ids(e, err));
  GG_HIP(hipSetDevice(e->device));
  e->zero_ctr();
  k_checksum<<<uint32_t(e->cus) * 8, kBlock, 0, e->stream>>>(
      {static_cast<unsigned long long*>(e->old_ptr.p), static_cast<uint32_t*>(e->old_col.p)}, e->rows(), e->n, e->C());
  GG_HIP(hipGetLastError());
  GG_TRY(e->read_ctr(err));
  *sum = e->h_ctr->checksum;
  return GG_OK;
}

// ---------------------------------------------------------------------------
// Sessions (the serving loop): one running fixpoint that graphs join at round
// boundaries. A joining graph takes the next vertex range: the per-row offset
// arrays are re-laid out for the larger vertex count (rows are label-major,
// so each label's block gains empty rows at its end and no id moves), its
// edges are uploaded and its raw rows are deduplicated inside the next round
// (round_step's inject), which makes them its first delta. A graph whose rows
// get no new edges in a round has reached its fixpoint: its closure is its
// rows of old ∪ pending, which later rounds do not touch. Sessions keep the
// ids of packed labels in old, so results can be read while the fixpoint of
// the other graphs goes on.

namespace {

// rows [rb, re) of one label of one graph: their first vertex u0, the
// graph's vertex base v0, the label in the graph's numbering and the index
// of the (graph, label) item the rows belong to (results are cut into row
// chunks of at most kResultChunk rows, so that large graphs use every CU)
struct SliceItem {
  uint32_t rb, re, u0, v0, lab, item;
};
constexpr uint32_t kResultChunk = 2048;

__global__ void k_remap_ptr(const unsigned long long* in, unsigned long long* out, uint32_t L, uint32_t n0,
                            uint32_t n1) {
  const uint64_t total = uint64_t(L) * n1 + 1;
  for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < total; r += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t l = uint32_t(r / n1), u = uint32_t(r - uint64_t(l) * n1);
    // past the old vertices (and the final entry) a row starts where its
    // label's old block ended
    out[r] = (l < L && u < n0) ? in[uint64_t(l) * n0 + u] : in[uint64_t(l < L ? l + 1 : L) * n0];
  }
}

__global__ void k_remap_len(const uint32_t* in, uint32_t* out, uint32_t L, uint32_t n0, uint32_t n1) {
  const uint64_t total = uint64_t(L) * n1;
  for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < total; r += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t l = uint32_t(r / n1), u = uint32_t(r - uint64_t(l) * n1);
    out[r] = u < n0 ? in[uint64_t(l) * n0 + u] : 0u;
  }
}

__global__ void k_item_counts(const unsigned long long* ptr, const SliceItem* it, uint32_t ni,
                              unsigned long long* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ni; i += gridDim.x * blockDim.x)
    out[i] = ptr[it[i].re] - ptr[it[i].rb];
}

// per chunk of rows: its edge count and checksum in c, added to its item's
// cnt / sums
// plab: rows of packed labels read their packed entries (plab.ent at the
// row's own offset, plab.len[r] of them) when plab.len is set, their ids
// otherwise.
__global__ __launch_bounds__(kBlock) void k_item_results(DCsr c, DPacked plab, uint32_t n, const SliceItem* it,
                                                         uint32_t ni, unsigned long long* cnt,
                                                         unsigned long long* sums) {
  // a wave takes 64 rows of the item at a time, one per lane, and streams
  // their elements together (hash_rows)
  __shared__ unsigned long long s_tab[kWavesPerBlock][kWave], s_tmask[kWavesPerBlock][kPosUnroll],
      s_row[kWavesPerBlock][kWave];
  __shared__ unsigned long long s_bit[32];
  if (threadIdx.x < 32) s_bit[threadIdx.x] = uint64_t(threadIdx.x) * kHashD;
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t wv = threadIdx.x / kWave;
  for (uint32_t i = blockIdx.x; i < ni; i += gridDim.x) {
    const SliceItem x = it[i];
    unsigned long long acc = 0;
    const bool pk = plab.len && x.re > x.rb && packed_label(plab, x.rb / n);
    for (uint32_t r0 = x.rb + wv * kWave; r0 < x.re; r0 += kWavesPerBlock * kWave) {
      const uint32_t r = r0 + uint32_t(lane);
      unsigned long long b = 0;
      uint32_t lp = 0, li = 0;
      if (r < x.re) {
        b = c.ptr[r];
        const unsigned long long e = c.ptr[r + 1];
        if (e > b) {
          if (pk)
            lp = plab.len[r];
          else
            li = uint32_t(e - b);
        }
      }
      const unsigned long long crow = uint64_t(x.u0 + (r - x.rb) - x.v0) * kHashS ^ uint64_t(x.lab) * kHashL;
      hash_rows(plab.ent, c.col, b, lp, li, x.v0, crow, s_tab[wv], s_tmask[wv], s_row[wv], s_bit, lane, acc);
    }
    acc = wave_sum_u64(acc);
    if (lane == 0 && acc) atomicAdd(&sums[x.item], acc);
    if (threadIdx.x == 0) atomicAdd(&cnt[x.item], (unsigned long long)(c.ptr[x.re] - c.ptr[x.rb]));
  }
}

// one item per (graph, label); chunked: each cut into row chunks
std::vector<SliceItem> slice_items(const std::vector<SessionSlice>& sl, uint32_t n, bool chunked = false) {
  std::vector<SliceItem> it;
  uint32_t item = 0;
  for (const SessionSlice& q : sl)
    for (uint32_t i = 0; i < q.lg; ++i, ++item) {
      const uint32_t rb = (q.l0 + i) * n + q.v0, re = rb + q.n;
      if (!chunked) {
        it.push_back({rb, re, q.v0, q.v0, i + 1, item});
        continue;
      }
      for (uint32_t a = rb; a < re; a += kResultChunk)
        it.push_back({a, std::min(re, a + kResultChunk), q.v0 + (a - rb), q.v0, i + 1, item});
    }
  return it;
}

}  // namespace

int engine_session_begin(Engine* e, const Grammar& gr, std::string& err) {
  GG_TRY(engine_set_grammar(e, gr, err));
  if (e->g.ne) {
    err = "a session takes eps rules as explicit self-loop edges";
    return GG_E_ARG;
  }
  GG_HIP(hipSetDevice(e->device));
  e->n = 0;
  e->m = 0;
  e->old_nnz = e->dl_nnz = e->pend_nnz = 0;
  e->pk_ready = e->dpk_ready = e->pairs_ready = false;
  // rows of packed labels keep only their packed entries during the rounds
  // (results read them directly); GG_KEEP_PACKED_IDS=1 keeps their ids too
  e->keep_ids = false;
  if (const char* ki = std::getenv("GG_KEEP_PACKED_IDS")) e->keep_ids = std::atoi(ki) != 0;
  e->apply_tuning();
  e->defer_ratio = e->tune.defer_ratio;
  if (const char* dr = std::getenv("GG_DEFER_RATIO")) e->defer_ratio = std::strtoull(dr, nullptr, 10);
  e->guess_bits = e->tune.guess_min;
  e->last_cand = e->last_active = 0;
  e->last_deferred = Engine::kUnknown;
  e->pr_arena = 0;
  e->sess_rp = RoundPlan{};
  e->sess_S = gg_stats{};
  e->sess_inject = 0;
  e->has_graph = true;
  e->has_closure = false;
  e->peak = e->used;
  GG_TRY(e->alloc(e->ctr, sizeof(Counters), err));
  if (!e->h_ctr) GG_HIP(hipHostMalloc((void**)&e->h_ctr, sizeof(Counters), hipHostMallocDefault));
  if (!e->h_u64) GG_HIP(hipHostMalloc((void**)&e->h_u64, 8, hipHostMallocDefault));
  return GG_OK;
}

int engine_session_admit(Engine* e, uint32_t new_n, const uint32_t* src, const uint32_t* dst, const uint8_t* lab,
                         uint64_t m, std::string& err) {
  GG_HIP(hipSetDevice(e->device));
  const uint32_t n0 = e->n;
  if (new_n < n0) {
    err = "a session's vertex range only grows";
    return GG_E_ARG;
  }
  const uint64_t L = uint64_t(e->L);
  if (L * new_n >= 0xFFFFFFFFull || uint64_t(e->g.nslot) * new_n >= 0xFFFFFFFFull) {
    err = "session too large: labels * vertices (and pair slots * vertices) must stay below 2^32";
    return GG_E_UNSUPPORTED;
  }
  const uint64_t R1 = L * new_n;
  if (n0 == 0) {
    GG_TRY(e->alloc(e->old_ptr, (R1 + 1) * 8, err));
    GG_TRY(e->alloc(e->dl_ptr, (R1 + 1) * 8, err));
    e->fill_u64(e->old_ptr, 0, R1 + 1);
    e->fill_u64(e->dl_ptr, 0, R1 + 1);
    if (e->pack_any_label()) {
      GG_TRY(e->alloc(e->pk_len, R1 * 4, err));
      e->fill_u32(e->pk_len, 0, R1);
    }
  } else if (new_n > n0) {
    auto remap_ptr = [&](DevBuf& b) -> int {
      DevBuf t;
      GG_TRY(e->alloc(t, (R1 + 1) * 8, err));
      k_remap_ptr<<<e->grid_for(R1 + 1), kBlock, 0, e->stream>>>(
          static_cast<unsigned long long*>(b.p), static_cast<unsigned long long*>(t.p), uint32_t(L), n0, new_n);
      std::swap(b, t);
      e->release(t);
      return GG_OK;
    };
    auto remap_len = [&](DevBuf& b) -> int {
      DevBuf t;
      GG_TRY(e->alloc(t, R1 * 4, err));
      k_remap_len<<<e->grid_for(R1), kBlock, 0, e->stream>>>(static_cast<uint32_t*>(b.p), static_cast<uint32_t*>(t.p),
                                                             uint32_t(L), n0, new_n);
      std::swap(b, t);
      e->release(t);
      return GG_OK;
    };
    GG_TRY(remap_ptr(e->old_ptr));
    GG_TRY(remap_ptr(e->dl_ptr));
    if (e->pend_nnz) GG_TRY(remap_ptr(e->pd_ptr));
    if (e->pack_any_label()) GG_TRY(remap_len(e->pk_len));
    if (e->dpk_ready) GG_TRY(remap_len(e->dpk_len));
    GG_HIP(hipGetLastError());
  }
  e->n = new_n;
  e->m = m;
  if (!m) return GG_OK;
  GG_TRY(e->alloc(e->in_src, m * 4, err));
  GG_TRY(e->alloc(e->in_dst, m * 4, err));
  GG_TRY(e->alloc(e->in_lab, m, err));
  GG_HIP(hipMemcpyAsync(e->in_src.p, src, m * 4, hipMemcpyHostToDevice, e->stream));
  GG_HIP(hipMemcpyAsync(e->in_dst.p, dst, m * 4, hipMemcpyHostToDevice, e->stream));
  GG_HIP(hipMemcpyAsync(e->in_lab.p, lab, m, hipMemcpyHostToDevice, e->stream));
  const uint32_t R = e->rows();
  GG_TRY(e->alloc(e->seg_cnt, uint64_t(R) * 4, err));
  GG_TRY(e->alloc(e->raw_ptr, (uint64_t(R) + 1) * 8, err));
  GG_TRY(e->grow(e->raw_col, m * 4, err));
  e->fill_u32(e->seg_cnt, 0, R);
  k_count_raw<<<e->grid_for(m), kBlock, 0, e->stream>>>(static_cast<uint32_t*>(e->in_src.p),
                                                        static_cast<uint8_t*>(e->in_lab.p), m, e->n, e->g,
                                                        static_cast<uint32_t*>(e->seg_cnt.p));
  GG_TRY(e->scan(e->seg_cnt, R, e->raw_ptr, err));
  e->fill_u32(e->seg_cnt, 0, R);
  k_scatter_raw<<<e->grid_for(m), kBlock, 0, e->stream>>>(
      static_cast<uint32_t*>(e->in_src.p), static_cast<uint32_t*>(e->in_dst.p), static_cast<uint8_t*>(e->in_lab.p), m,
      e->n, e->g, static_cast<unsigned long long*>(e->raw_ptr.p), static_cast<uint32_t*>(e->seg_cnt.p),
      static_cast<uint32_t*>(e->raw_col.p));
  GG_HIP(hipGetLastError());
  e->sess_inject = m;
  return GG_OK;
}

int engine_session_round(Engine* e, const std::vector<SessionSlice>& live, std::vector<uint64_t>& delta,
                         std::string& err) {
  GG_HIP(hipSetDevice(e->device));
  delta.assign(live.size(), 0);
  if (e->n == 0 || (e->dl_nnz == 0 && e->sess_inject == 0)) return GG_OK;
  if (e->sess_rp.n != e->n) GG_TRY(e->plan_rounds(e->sess_rp, err));
  RunEvents ev;
  ev.enabled = e->sess_timing;
  GG_TRY(e->round_step(e->sess_rp, e->sess_S, ev, e->sess_inject, err));
  e->sess_inject = 0;
  if (e->sess_timing) {
    GG_HIP(hipStreamSynchronize(e->stream));
    auto sum = [](std::vector<EventPair>& v) {
      double t = 0;
      for (auto& x : v) {
        float ms = 0;
        if (x.a && x.b && hipEventElapsedTime(&ms, x.a, x.b) == hipSuccess) t += ms;
      }
      return t;
    };
    e->sess_S.expand_ms += sum(ev.expand);
    e->sess_S.compact_ms += sum(ev.compact);
    e->sess_S.merge_ms += sum(ev.merge);
    e->sess_S.other_ms += sum(ev.other);
  }
  if (live.empty()) return GG_OK;
  // per graph: its edges in the next delta
  const std::vector<SliceItem> it = slice_items(live, e->n);
  DevBuf d_it, d_out;
  struct Guard {
    Engine* e;
    DevBuf* b[2];
    ~Guard() {
      for (DevBuf* x : b) e->release(*x);
    }
  } guard{e, {&d_it, &d_out}};
  GG_TRY(e->alloc(d_it, it.size() * sizeof(SliceItem), err));
  GG_TRY(e->alloc(d_out, it.size() * 8, err));
  GG_HIP(hipMemcpyAsync(d_it.p, it.data(), it.size() * sizeof(SliceItem), hipMemcpyHostToDevice, e->stream));
  k_item_counts<<<e->grid_for(it.size()), kBlock, 0, e->stream>>>(static_cast<unsigned long long*>(e->dl_ptr.p),
                                                                  static_cast<SliceItem*>(d_it.p), uint32_t(it.size()),
                                                                  static_cast<unsigned long long*>(d_out.p));
  GG_HIP(hipGetLastError());
  std::vector<uint64_t> out(it.size());
  GG_HIP(hipMemcpyAsync(out.data(), d_out.p, it.size() * 8, hipMemcpyDeviceToHost, e->stream));
  GG_HIP(hipStreamSynchronize(e->stream));
  size_t k = 0;
  for (size_t q = 0; q < live.size(); ++q)
    for (uint32_t i = 0; i < live[q].lg; ++i) delta[q] += out[k++];
  return GG_OK;
}

int engine_session_results(Engine* e, const std::vector<SessionSlice>& done, std::vector<uint64_t>& counts,
                           std::vector<uint64_t>& sums, std::string& err) {
  GG_HIP(hipSetDevice(e->device));
  const std::vector<SliceItem> it = slice_items(done, e->n, true);
  uint32_t nitems = 0;
  for (const SessionSlice& q : done) nitems += q.lg;
  counts.assign(nitems, 0);
  sums.assign(done.size(), 0);
  if (it.empty() || e->n == 0) return GG_OK;
  DevBuf d_it, d_cnt, d_sum;
  struct Guard {
    Engine* e;
    DevBuf* b[3];
    ~Guard() {
      for (DevBuf* x : b) e->release(*x);
    }
  } guard{e, {&d_it, &d_cnt, &d_sum}};
  GG_TRY(e->alloc(d_it, it.size() * sizeof(SliceItem), err));
  GG_TRY(e->alloc(d_cnt, uint64_t(nitems) * 8, err));
  GG_TRY(e->alloc(d_sum, uint64_t(nitems) * 8, err));
  GG_HIP(hipMemcpyAsync(d_it.p, it.data(), it.size() * sizeof(SliceItem), hipMemcpyHostToDevice, e->stream));
  e->fill_u64(d_cnt, 0, nitems);
  e->fill_u64(d_sum, 0, nitems);
  const uint32_t grid = uint32_t(std::min<uint64_t>(it.size(), uint64_t(e->cus) * 8));
  // packed labels' rows of old hold only their packed entries unless the
  // ids are kept every round
  DPacked opk{};
  for (int i = 0; i < 8; ++i) opk.labels[i] = e->pack_labels[i];
  opk.len = e->pk_ready && !e->keep_ids ? static_cast<uint32_t*>(e->pk_len.p) : nullptr;
  opk.ent = static_cast<unsigned long long*>(e->pk_ent.p);
  k_item_results<<<grid, kBlock, 0, e->stream>>>({static_cast<unsigned long long*>(e->old_ptr.p),
                                                  static_cast<uint32_t*>(e->old_col.p)},
                                                 opk, e->n, static_cast<SliceItem*>(d_it.p), uint32_t(it.size()),
                                                 static_cast<unsigned long long*>(d_cnt.p),
                                                 static_cast<unsigned long long*>(d_sum.p));
  if (e->pend_nnz)
    k_item_results<<<grid, kBlock, 0, e->stream>>>({static_cast<unsigned long long*>(e->pd_ptr.p),
                                                    static_cast<uint32_t*>(e->pd_col.p)},
                                                   DPacked{}, e->n, static_cast<SliceItem*>(d_it.p),
                                                   uint32_t(it.size()),
                                                   static_cast<unsigned long long*>(d_cnt.p),
                                                   static_cast<unsigned long long*>(d_sum.p));
  GG_HIP(hipGetLastError());
  std::vector<uint64_t> s(nitems);
  GG_HIP(hipMemcpyAsync(counts.data(), d_cnt.p, uint64_t(nitems) * 8, hipMemcpyDeviceToHost, e->stream));
  GG_HIP(hipMemcpyAsync(s.data(), d_sum.p, uint64_t(nitems) * 8, hipMemcpyDeviceToHost, e->stream));
  GG_HIP(hipStreamSynchronize(e->stream));
  size_t k = 0;
  for (size_t q = 0; q < done.size(); ++q)
    for (uint32_t i = 0; i < done[q].lg; ++i) sums[q] += s[k++];
  return GG_OK;
}

void engine_session_end(Engine* e, bool trim) {
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->stream);
  if (trim) {
    engine_release(e);
  } else {
    DevBuf* all[] = {&e->old_ptr, &e->old_col, &e->dl_ptr, &e->dl_col, &e->nx_ptr, &e->nx_col, &e->o2_ptr,
                     &e->o2_col, &e->staging, &e->pk_len, &e->pk_ent, &e->pk2_ent, &e->dpk_len, &e->dpk_ent,
                     &e->npk_len, &e->npk_ent, &e->pd_ptr, &e->pd_col, &e->p2_ptr, &e->p2_col, &e->raw_ptr,
                     &e->raw_col, &e->pr_col};
    for (DevBuf* b : all) e->release(*b);
  }
  e->n = 0;
  e->m = 0;
  e->old_nnz = e->dl_nnz = e->pend_nnz = 0;
  e->pk_ready = e->dpk_ready = false;
  e->sess_rp = RoundPlan{};
  e->sess_inject = 0;
  e->has_graph = false;
  e->has_closure = false;
}

uint64_t engine_device_bytes(Engine* e) { return e->used + e->cached; }

void engine_session_timing(Engine* e, bool on) { e->sess_timing = on; }

std::shared_ptr<DeviceRegion> device_region_create(int device, size_t bytes, std::string& err) {
  if (hipSetDevice(device) != hipSuccess) {
    (void)hipGetLastError();
    err = "cannot use HIP device " + std::to_string(device);
    return nullptr;
  }
  void* p = nullptr;
  const hipError_t he = hipMalloc(&p, bytes);
  if (he != hipSuccess) {
    (void)hipGetLastError();
    err = "hipMalloc(" + std::to_string(bytes) + ") for a device region: " + hipGetErrorString(he);
    return nullptr;
  }
  auto r = std::make_shared<DeviceRegion>();
  r->base = static_cast<char*>(p);
  r->cap = bytes;
  r->owned = true;
  r->free_list.emplace(0, bytes);
  return r;
}

int engine_device(Engine* e) { return e->device; }

size_t region_bytes(const DeviceRegion* r) { return r ? r->cap : 0; }

void engine_on_round(Engine* e, std::function<void(uint64_t, uint64_t)> f) { e->on_round = std::move(f); }

void engine_defer_unpack(Engine* e, bool on) { e->defer_unpack = on; }

double tuned_group_mem_share() { return load_tuning().group_mem_share; }

int engine_use_region(Engine* e, std::shared_ptr<DeviceRegion> r, std::string& err) {
  if (e->arena) return GG_OK;
  if (e->used || !e->cache.empty()) {
    err = "the engine already holds device memory";
    return GG_E_STATE;
  }
  e->region = std::move(r);
  e->arena = e->region->base;
  return GG_OK;
}

gg_stats engine_session_stats(Engine* e, double* alloc_ms) {
  if (alloc_ms) *alloc_ms = e->alloc_ms;
  gg_stats st = e->sess_S;
  st.device_bytes_peak = e->peak;
  return st;
}

// ---------------------------------------------------------------------------
// The allocation-free entry point (gg_closure_run_noalloc). Its buffers come
// from `scratch`, a fixed device region that every GPU the library runs on
// holds for the life of the process. In host code the array's name is not a
// device address: scratch_region() looks the address up once per device, with
// that device current, and only that address is used afterwards.

constexpr size_t kScratchBytes = size_t(256) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

int scratch_region(int device, char** out, std::string& err) {
  if (device < 0 || device >= kMaxScratchDevices) {
    err = "device ordinal " + std::to_string(device) + " out of range for the scratch table";
    return GG_E_DEVICE;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    GG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      err = "hipGetSymbolAddress returned no address for the scratch region";
      return GG_E_DEVICE;
    }
    g_scratch_ptr[device] = p;
  }
  *out = static_cast<char*>(g_scratch_ptr[device]);
  return GG_OK;
}

// Sets ctr->overflow when an input edge names a vertex >= n or a label
// outside 1..L (the caller's device arrays cannot be checked on the host).
__global__ void k_check_edges(const uint32_t* src, const uint32_t* dst, const uint8_t* lab, uint64_t m, uint32_t n,
                              uint32_t L, Counters* ctr) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += uint64_t(gridDim.x) * blockDim.x)
    if (src[i] >= n || dst[i] >= n || lab[i] < 1 || lab[i] > L) ctr->overflow = 1u;
}

}  // namespace

size_t noalloc_scratch_bytes() { return kScratchBytes; }

int engine_run_noalloc(const Grammar& gr, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                       const uint8_t* lab, void* stream, uint64_t* out_ptr, uint32_t* out_col, uint64_t col_cap,
                       uint64_t* total, uint64_t* checksum, uint64_t* per_label, std::string& err) {
  int dev = 0, cus = 0;
  GG_HIP(hipGetDevice(&dev));
  GG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  char* base = nullptr;
  GG_TRY(scratch_region(dev, &base, err));
  Engine e;
  e.device = dev;
  e.cus = cus;
  e.stream = static_cast<hipStream_t>(stream);
  e.owns_stream = false;
  auto reg = std::make_shared<DeviceRegion>();
  reg->base = base;
  reg->cap = kScratchBytes;
  reg->free_list.emplace(0, kScratchBytes);
  e.region = reg;
  e.arena = base;
  e.h_ctr = &e.arena_h_ctr;
  e.h_u64 = &e.arena_h_u64;
  // a small join grid: every wave of it reserves staging space up front
  e.expand_grid_cap = 128;
  GG_TRY(engine_set_grammar(&e, gr, err));
  const uint64_t R = uint64_t(e.L) * n;
  if (R >= 0xFFFFFFFFull) {
    err = "num_labels * num_vertices must be < 2^32";
    return GG_E_UNSUPPORTED;
  }
  // the per-row arrays and the raw input alone: the least the closure needs
  const uint64_t least = R * (8 * 5 + 4 * 4) + (m + uint64_t(e.g.ne) * n) * 8;
  if (least > kScratchBytes) {
    err = "the graph needs at least " + std::to_string(least) + " bytes of device memory; the allocation-free "
          "entry point has a fixed scratch region of " + std::to_string(kScratchBytes) + " bytes";
    return GG_E_NOMEM;
  }
  e.in_src.p = const_cast<uint32_t*>(src);
  e.in_dst.p = const_cast<uint32_t*>(dst);
  e.in_lab.p = const_cast<uint8_t*>(lab);
  e.n = n;
  e.m = m;
  e.has_graph = true;
  GG_TRY(e.alloc(e.ctr, sizeof(Counters), err));
  e.zero_ctr();
  if (m)
    k_check_edges<<<e.grid_for(m), kBlock, 0, e.stream>>>(src, dst, lab, m, n, uint32_t(e.L), e.C());
  GG_HIP(hipGetLastError());
  GG_TRY(e.read_ctr(err));
  if (e.h_ctr->overflow) {
    err = "an input edge names a vertex >= num_vertices or a label outside 1..num_labels";
    return GG_E_ARG;
  }
  gg_stats st{};
  GG_TRY(e.run(0, &st, err));
  const uint64_t tot = st.total_edges;
  if (total) *total = tot;
  if (tot > col_cap) {
    err = "the closure has " + std::to_string(tot) + " edges; col_cap is " + std::to_string(col_cap);
    return GG_E_ARG;
  }
  GG_HIP(hipMemcpyAsync(out_ptr, e.old_ptr.p, (R + 1) * 8, hipMemcpyDeviceToDevice, e.stream));
  if (tot) GG_HIP(hipMemcpyAsync(out_col, e.old_col.p, tot * 4, hipMemcpyDeviceToDevice, e.stream));
  GG_HIP(hipStreamSynchronize(e.stream));
  if (per_label) GG_TRY(engine_counts(&e, per_label, nullptr, err));
  if (checksum) GG_TRY(engine_checksum(&e, checksum, err));
  GG_HIP(hipStreamSynchronize(e.stream));
  return GG_OK;
}

int engine_release(Engine* e) {
  DevBuf* all[] = {&e->old_ptr, &e->old_col, &e->dl_ptr,  &e->dl_col,   &e->nx_ptr, &e->nx_col,   &e->o2_ptr,
                   &e->o2_col,  &e->staging, &e->seg_off, &e->seg_cnt,  &e->active, &e->tiles,
                   &e->fr_bits,
                   &e->pr_ptr,  &e->pr_col,  &e->pk_len, &e->pk_ent, &e->pk2_ent,
                   &e->dpk_len, &e->dpk_ent, &e->npk_len, &e->npk_ent, &e->mlist, &e->mcnt,
                   &e->pd_ptr,  &e->pd_col,  &e->p2_ptr,  &e->p2_col};
  for (DevBuf* b : all) e->release(*b);
  e->ids_pending = false;
  e->pend_nnz = 0;
  e->trim_cache();  // the memory goes back to the device (other contexts may need it)
  e->has_closure = false;
  return GG_OK;
}

int device_query(int* count, int preferred, int* selected, char* name, size_t name_cap, uint64_t* total,
                 uint64_t* free_bytes, std::string& err) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c == 0) {
    (void)hipGetLastError();
    if (count) *count = 0;
    err = "no HIP device available";
    return GG_E_DEVICE;
  }
  int d = preferred;
  if (d < 0) {
    int best = -1;
    for (int i = 0; i < c; ++i) {
      hipDeviceProp_t p;
      if (hipGetDeviceProperties(&p, i) == hipSuccess && p.multiProcessorCount > best) {
        best = p.multiProcessorCount;
        d = i;
      }
    }
  }
  hipDeviceProp_t p;
  GG_HIP(hipGetDeviceProperties(&p, d));
  GG_HIP(hipSetDevice(d));
  size_t fr = 0, tot = 0;
  GG_HIP(hipMemGetInfo(&fr, &tot));
  if (count) *count = c;
  if (selected) *selected = d;
  if (name && name_cap) {
    std::snprintf(name, name_cap, "%s", p.name);
  }
  if (total) *total = tot;
  if (free_bytes) *free_bytes = fr;
  return GG_OK;
}

}  // namespace gg
