// Device-side shared declarations of libparbutterfly_amd (HIP, gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "parbutterfly.h"

#define PBF_WAVE 64

// ------------------------------------------------------------------------------------------
// Context: one HIP stream, a grow-only scratch arena (named slots), timing events.
// ------------------------------------------------------------------------------------------
enum ScratchSlot {
  S_GRAPH = 0,   // uploaded host graph (offV, edgV, offU, edgU packed)
  S_KEY,         // u8 key per vertex (ADEG / ACOCORE level)
  S_ORDER,       // u32 order[n]
  S_RANKOF,      // u32 rank_of[n]
  S_HIST,        // u32 key histogram per block
  S_SCAN,        // scan block sums
  S_ROFF,        // u32 ranked offsets [n+1]
  S_REDG,        // u32 ranked adjacency [2m]
  S_RPOS,        // u32 original CSR position of each ranked slot [2m]
  S_OIDX,        // u32 O-CSR index of each ranked slot [2m]
  S_INPREF,      // u32 in(r) [n]
  S_BOUND,       // u32 endpoint bound [n]
  S_BINS,        // u32 vertex lists per kernel class
  S_COUNTS,      // u64 counts in rank space [n]
  S_MISC,        // small counters / totals (u64)
  S_ETI,         // u32 eti [m]
  S_ITE,         // u32 ite [m]
  S_POOL,        // HBM fallback dense histograms
  S_OUT,         // staging for host outputs
  S_PEEL0, S_PEEL1, S_PEEL2, S_PEEL3, S_PEEL4, S_PEEL5, S_PEEL6, S_PEEL7,
  S_ACORE0, S_ACORE1, S_ACORE2,
  S_HOT,         // replicated hub accumulators (kept zeroed between calls)
  S_BIG,         // long-list classification: list, chunk counts, chunk bases
  S_UB,          // u64 work bound of long lists
  S_RLONG,       // relabel: long-list ids, chunk counts, chunk bases
  S_RHIST,       // relabel: per-chunk key histograms / offsets
  S_SPLIT,       // split endpoints: list, chunk counts, chunk bases
  S_GROWS,       // split endpoints: dense count rows in HBM
  S_LPOS,        // per slot (r -> y): roff[y] and the wedge segment length (uint2)
  S_ROWN,        // per-slot owner rank
  S_LPLAN,       // grouping plan of the light endpoints (bucket counts, starts, word bases)
  S_LREC,        // grouped light endpoints: (i, roff[i], deg, log2 table size)
  S_BSTAGE,      // batch: host graphs staged in HBM
  S_BTAB,        // batch: per-graph table
  S_BUNION,      // batch: union CSR
  S_BFLAG,       // batch: per-vertex side flags
  S_BAUX,        // batch: per-graph |S|, offsets, sides
  S_MLP,         // selector MLP: parameters, Adam moments, samples
  S_BFEAT,       // batch features: per-graph statistics
  S_RLO,         // batch: first rank of each rank's graph block (graph-local entries)
  S_RINFO,       // per rank: (roff, in) packed for the slot-info gathers
  S_SPARSE0,     // sparsification: scanned slot flags, new ids, tile owners
  S_SPARSE1,     // sparsification: the sparsified graph
  S_TEXT0,       // device reader: the file's bytes, tile token counts
  S_TEXT1,       // device reader: list-start flags, tile owners
  S_TEXT2,       // device reader: first-error key
  S_LEARN0,      // learned batch: every part's packed outputs (counts, offsets, totals)
  S_LEARN1,      // learned batch: placement table of the parts' graphs
  S_LEARN2,      // learned batch: final outputs for host callers
  S_WBATCH,      // wedge-bounded counting: wedges per endpoint and their prefix
  S_HLOG,        // per-vertex counting: the log of credits to hub ranks
  S_B32,         // per-edge counting: 32-bit per-slot credits of the wedge pass [2m]
  S_RKEY,        // relabel: (rank, key) per vertex [n]
  S_NUM_SLOTS
};

// Runtime knobs of the counting kernels (defaults; the batch autotuner may change them).
struct CountTune {
  uint32_t light_max = 0;        // distinct-key bound of a light endpoint (0: the mode's default)
  uint32_t light_work = 1u << 14;  // wedge bound of a light endpoint
  uint32_t blocks_per_cu = 4;    // workgroups per CU of the persistent counting kernels
};

struct pbf_context {
  int device = 0;
  hipStream_t stream = nullptr;      // where all work is queued: own_stream or the caller's
  hipStream_t own_stream = nullptr;  // created with the context
  void* buf[S_NUM_SLOTS] = {};
  size_t cap[S_NUM_SLOTS] = {};
  hipEvent_t ev[8] = {};
  pbf_timing timing = {};
  uint64_t* pinned = nullptr;  // small pinned host buffer (readbacks)
  int num_cus = 256;
  bool hot_clear = false;  // S_HOT has been zeroed
  size_t pool_clean = 0;   // bytes of S_POOL known to be zero (the HBM kernel leaves them zero)
  CountTune tune;
  // Memory bounds of the counting pass (pbf_set_limits; 0 = none): endpoints are counted in
  // rank ranges of at most max_wedges wedges, and the HBM-fallback histograms hold at most
  // max_array 32-bit counters (one endpoint's worth at least).
  int64_t max_wedges = 0;
  int64_t max_array = 0;
  void* batch = nullptr;   // batch state (batch.hip)
  // pinned staging for host transfers (pbf_h2d / pbf_d2h): two buffers in turn, so the host
  // copy of one chunk overlaps the DMA of the other
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  bool stage_busy[2] = {false, false};
  // The device's fixed scratch array (the no-allocation entry point), looked up at creation;
  // null when the lookup failed.
  unsigned char* noalloc = nullptr;
  // While fixed.base is set, pbf_scratch carves its slots from [base, base + cap) instead of
  // allocating (the no-allocation entry point), and fails with PBF_ERR_CAPACITY past the end.
  struct {
    unsigned char* base = nullptr;
    size_t cap = 0, used = 0;
  } fixed;
};

// Returns a device buffer of at least `bytes` for the slot (grow-only; contents undefined).
int pbf_scratch(pbf_context* ctx, int slot, size_t bytes, void** out);
// Bytes pbf_scratch can still hand out without allocating: the rest of the fixed region, or
// SIZE_MAX when the arena allocates.
size_t pbf_scratch_room(const pbf_context* ctx);

#define PBF_HIP(call)                                  \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return PBF_ERR_HIP;          \
  } while (0)

#define PBF_CHECK(call)                                \
  do {                                                 \
    int _s = (call);                                   \
    if (_s != PBF_OK) return _s;                       \
  } while (0)

// Device view of the input graph.
struct DevGraph {
  int64_t nv, nu, m;
  const uint32_t* offV;
  const uint32_t* edgV;
  const uint32_t* offU;
  const uint32_t* edgU;
};

// Ranked graph (rankGraph / rankGraphEdges, butterfly_utils.h:579-700) on device.
struct RankedGraph {
  int64_t n, m2;
  uint32_t* order;    // order[r] = combined id
  uint32_t* rank_of;  // rank_of[id] = r
  uint32_t* roff;     // [n+1]
  uint32_t* redg;     // [2m] (rank << 1) | flag(neighbour in S)
  uint32_t* rpos;     // [2m] position in owner's original CSR (EDGE only, else null)
  uint32_t* inpref;   // [n]  neighbours of lower rank
  uint32_t* rown;     // [2m] owner rank of each slot
  int64_t wedges;     // W under this ranking (-1 when not read back)
  unsigned long long* wedges_dev;  // W on the device
  // Union of graphs ranked graph by graph (each graph a contiguous block of ranks): lo[r] is
  // the first rank of r's block and entries hold ranks relative to it.  Null otherwise.
  uint32_t* lo = nullptr;
};

// api.hip: host <-> device copies on the context's stream, always through the context's pinned
// staging buffers (host memory is never handed to the runtime's pageable-memory copy path;
// large transfers run at DMA speed).  pbf_h2d returns once the host source has been read;
// pbf_d2h returns once the host destination is written.
int pbf_h2d(pbf_context* ctx, void* dst, const void* src, size_t bytes);
int pbf_d2h(pbf_context* ctx, void* dst, const void* src, size_t bytes);

// scan.hip
int pbf_scan_excl_u32(pbf_context* ctx, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev);
int pbf_scan_excl_u64(pbf_context* ctx, const uint64_t* in, uint64_t* out, int64_t n, uint64_t* total_dev);
// Fill kernels (used instead of hipMemsetAsync everywhere a caller may capture the stream).
int pbf_fill_u32(pbf_context* ctx, uint32_t* p, uint32_t v, int64_t n);
int pbf_fill_u64(pbf_context* ctx, uint64_t* p, uint64_t v, int64_t n);

// rank.hip
int pbf_dev_side(pbf_context* ctx, const DevGraph& g, int* use_v, int64_t* wedges);
// SIDE ranking needs the count side: use_v, or the per-vertex flags sflag of a batch.
// sort = false: only the keys (S_KEY) are computed; the caller orders the vertices itself.
int pbf_dev_rank(pbf_context* ctx, const DevGraph& g, int rank_type, RankedGraph* R, int use_v = 0,
                 const uint8_t* sflag = nullptr, bool sort = true);
// sflag (optional): per combined id, 1 when the vertex lies on the count side S (a batch of
// graphs with their own sides); else S is U when use_v, V otherwise.  sync: read W back into
// R->wedges (else only R->wedges_dev holds it, and the call never waits for the device).
int pbf_dev_relabel(pbf_context* ctx, const DevGraph& g, int use_v, bool want_pos, RankedGraph* R,
                    const uint8_t* sflag = nullptr, bool sync = true);
int pbf_dev_edge_to_idxs(pbf_context* ctx, const DevGraph& g, int use_v, uint32_t* eti);

// count.hip
enum CountMode { MODE_VERT = 0, MODE_EDGE = 1, MODE_TOTAL = 2 };
// sync: the HBM fallback is launched only when needed, after reading its list length back;
// otherwise it is always launched (it exits at once on an empty list) and the call never
// waits for the device.  rank_space (VERT): counts are left in rank space (*B_out), no
// conversion to ids.
struct CountOpts {
  bool sync = true;
  bool rank_space = false;
  unsigned long long** B_out = nullptr;
  int64_t key_range = 0;  // with graph-local ranks (R.lo): the largest graph's vertex count
};
int pbf_dev_count(pbf_context* ctx, const DevGraph& g, int use_v, const RankedGraph& R, int mode,
                  uint64_t* out_dev /* VERT: n_S, EDGE: m, TOTAL: 1 */, const uint32_t* eti,
                  const CountOpts& opts = CountOpts());

// peel.hip
int pbf_dev_peel_vert(pbf_context* ctx, const DevGraph& g, int use_v, const uint64_t* counts, uint64_t* tip);
int pbf_dev_peel_edge(pbf_context* ctx, const DevGraph& g, int use_v, const uint64_t* counts, uint64_t* wing);

// sparse.hip: the sparsified graph of g (eSparseBipartite / clrSparseBipartite + delZeroDeg)
// into the given device arrays (room for g's sizes); *out gets its sizes and pointers.
int pbf_dev_sparsify(pbf_context* ctx, const DevGraph& g, int sparse_type, int64_t denom, int64_t seed, DevGraph* out,
                     uint32_t* offV2, uint32_t* edgV2, uint32_t* offU2, uint32_t* edgU2);
// The driver's scaling of a sparsified total (BipartiteButterflies.C:213-214).
int pbf_sparse_scale(uint64_t total, int sparse_type, int64_t denom, int64_t* approx);

// ---------------- device helpers ----------------
__device__ __forceinline__ uint32_t pbf_bitlen(uint32_t d) { return d ? 32u - __clz(d) : 0u; }

__device__ __forceinline__ uint64_t pbf_lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Lanes of the wavefront holding the same 6-bit key (inactive lanes excluded via `active`).
__device__ __forceinline__ uint64_t pbf_match_key6(uint32_t k, uint64_t active) {
  uint64_t m = active;
#pragma unroll
  for (int b = 0; b < 6; b++) {
    const uint64_t bal = __ballot((k >> b) & 1u);
    m &= ((k >> b) & 1u) ? bal : ~bal;
  }
  return m;
}

// Makes the compiler wait for x here (an empty asm that reads it): placed after a batch of
// loads and before the atomics that depend on them, so that no wait lands between atomics
// (on this ISA a wait for a load also waits for every atomic issued before it).
__device__ __forceinline__ void pbf_ready(uint32_t x) { asm volatile("" ::"v"(x)); }

template <typename T>
__device__ __forceinline__ T pbf_wave_max(T x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const T y = __shfl_xor(x, d, 64);
    x = y > x ? y : x;
  }
  return x;
}

// Inclusive wave scan (64 lanes).
template <typename T>
__device__ __forceinline__ T pbf_wave_incl_scan(T x) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

template <typename T>
__device__ __forceinline__ T pbf_wave_sum(T x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// wedgeF of cmpWedgeCounts (butterfly_utils.h:1061-1069): C(d,2) with the product d * (d - 1)
// in the 32-bit degree type, i.e. modulo 2^32 (it wraps for d > 65536), then halved.  The
// side choice sums these terms in 64 bits, as the reference does.
__device__ __forceinline__ int64_t pbf_side_term(uint32_t d) { return (int64_t)((uint32_t)(d * (d - 1u)) / 2u); }
