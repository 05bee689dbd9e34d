// Butterfly counting kernels (per vertex, per edge, total) on the ranked graph.
//
// Algorithm (CountWorkEfficientSerial / CountWorkEfficientParallel, apps/hyper/
// butterfly_count.h:1218-1320 and :1402-1497; per edge butterfly_ecount.h:1334-1529; total
// butterfly_count_total.h:662-749, 833-906).  For every ranked endpoint i, the wedges are
// (i, v, u2) with v in N(i) and u2 in N(v), u2 < min(i, v).  Grouping them by u2 gives
// c(i,u2), the number of centres shared with the lowest-ranked vertex u2 of each butterfly:
//   * endpoint credit: B[i] += C(c,2), B[u2] += C(c,2) when i (hence u2) lies on S;
//   * centre credit:   B[v] += c(i,u2) - 1 for every wedge, when v lies on S (i on O);
//   * edge credit:     E[i->v] += c - 1 and E[v->u2] += c - 1 for every wedge with c > 1;
//   * total:           sum of C(c,2) over all groups.
//
// GPU mapping.  One wavefront (light endpoints) or one 256-thread workgroup (heavy ones) owns
// an endpoint i at a time, taken from a device work queue.  The wedges of a chunk of 64
// neighbours are flattened: lane j holds L_j, the length of N(v_j)'s prefix below min(i, v_j),
// a wave scan of L numbers the wedges densely, and windows of 256 consecutive wedge indices
// are handed out, 4 consecutive indices per lane (consecutive lanes read consecutive entries
// of the same N(v): coalesced loads, full lanes whatever the degrees).  The wedge key u2 is
// aggregated in LDS: dense counters indexed by u2 when u2 < i is small, otherwise an
// open-addressing hash.  Counts are scattered with 64-bit atomics; integer adds commute, so
// the results are exact and deterministic.
#include <algorithm>
#include <vector>

#include "pbf_device.h"
#include "team.h"

using namespace pbf;

namespace {
enum TableKind { T_LDS = 0, T_HBM = 1 };

// The lists, their lengths and the group counts a kernel reads (list_len, nblock, ngroups,
// rerun and overflow lists) are written by earlier launches on the stream (classification,
// grouping, the previous counting launch); the launch boundary orders those writes before any
// load, so plain loads read them.  Within a launch, workgroups share only the work queues,
// through device-scope atomics whose returned values carry the claims.
struct CountArgs {
  const uint32_t* roff;
  const uint32_t* redg;
  const uint32_t* inpref;
  const uint2* sinfo;      // per slot (i -> v): roff[v] and L, the length of N(v)'s prefix below min(i, v)
  const uint32_t* bound;   // distinct-u2 bound per endpoint: min(wedges, i)
  // VERT: counts in rank space.  EDGE: counts per ranked slot; an edge's two slots are summed
  // into its O-CSR position at the end (k_edge_out_*).  Slot indexing keeps the credits of a
  // segment, consecutive entries of N(v), on neighbouring addresses.
  unsigned long long* B;
  unsigned long long* total;
  const uint32_t* list;
  const uint32_t* list_len;
  uint32_t* queue;         // work counter
  uint32_t* ovf_list;
  uint32_t* ovf_len;
  uint32_t* pool;          // T_HBM: per-block dense counters [kn] followed by used list [kn]
  int64_t n;
  int64_t kn;              // key range: n, or the largest graph of a union with graph-local ranks
  // fused kernel (k_count2): two work lists and their queues
  const uint32_t* blist;   // endpoints processed by a whole workgroup
  const uint32_t* wlist;   // light endpoints, grouped (see light_group)
  const uint32_t* gstart;  // first wlist index of each light group, and the end
  const uint4* lrec;       // per grouped light endpoint: i, roff[i], deg(i), log2 table size
  const uint32_t* nblock;  // blist length
  const uint32_t* ngroups; // light groups, and those of them on side S
  const uint32_t* ngroups0;
  uint32_t* queues;        // [0] block queue, [1] light-group queue
  uint32_t* rerun_list;    // light endpoints whose group table overflowed (optimistic size)
  uint32_t* rerun_len;
  uint32_t rerun;          // this launch re-runs such endpoints with full-size tables
  unsigned long long* hot; // replicated accumulators for ranks < HOT_N (HOT_R replicas each)
  uint32_t cb;             // count bits of a packed 32-bit hash slot (32 - key bits)
  // A union of graphs with graph-local entries: lo[r] is the first rank of r's graph, every
  // entry of r's list holds (rank - lo[r]) << 1 | flag, and the tables are keyed by the local
  // rank, so each graph behaves as if counted alone.  Null for a single graph (lo = 0).
  const uint32_t* lo;
  // Log of credits to ranks below HOT_N (per-vertex counts): entries (rank << 52) | value,
  // appended by wavefronts in reserved chunks and summed by k_hlog_reduce.  hlog_cap == 0:
  // no log, those credits are replicated atomics.
  unsigned long long* hlog;
  uint32_t* hlog_len;      // entries reserved (may pass hlog_cap; the rest are not stored)
  uint32_t hlog_cap;
  // Per-edge counts: the per-wedge credits of pass 2 go to 32-bit per-slot counters where no
  // sum can reach 2^32 (the output adds them to B).  A slot (v -> u2) receives part of the
  // butterflies of edge (v, u2), at most (deg(v) - 1) (deg(u2) - 1) <= (deg(v) - 1) (maxdeg - 1),
  // so the slots of a centre that keeps that product below 2^32 take 32-bit credits.  Null:
  // every credit is 64-bit.
  uint32_t* B32;
  const uint32_t* maxdeg;  // the largest degree of the ranked graph (written by k_rinfo)
};

__device__ __forceinline__ uint32_t lo_of(const CountArgs& a, uint32_t i) { return a.lo ? a.lo[i] : 0u; }

// HBM fallback: dense counters in HBM (pool slice per block), used list in HBM, one workgroup
// per endpoint.  Only endpoints that no LDS table can hold come here (a workgroup hash that
// overflowed, or an endpoint of degree >= 2^24).
template <int MODE, int TEAM>
__global__ __launch_bounds__(TEAM) void k_count_hbm(CountArgs a) {
  __shared__ TeamShared<TEAM> sh;
  const uint32_t tid = threadIdx.x;
  const uint32_t nlist = *a.list_len;
  uint32_t* dense = a.pool + (size_t)blockIdx.x * (size_t)a.kn * 2;
  uint32_t* dused = dense + a.kn;
  unsigned long long block_total = 0;

  while (true) {
    if (tid == 0) sh.item = atomicAdd(a.queue, 1u);
    team_sync<TEAM>();
    const uint32_t idx = sh.item;
    team_sync<TEAM>();
    if (idx >= nlist) break;
    const uint32_t i = a.list[idx];
    const uint32_t io = a.roff[i], deg = a.roff[i + 1] - io;
    const uint32_t lo = lo_of(a, i), il = i - lo;  // entries and keys are local to i's graph
    const bool i_in_S = (a.redg[io] & 1u) == 0u;  // neighbours of i carry flag 1 iff i is on O
    if (tid == 0) sh.nused = 0;
    team_sync<TEAM>();
    const bool need_pass2 = (MODE == MODE_EDGE) || (MODE == MODE_VERT && !i_in_S);
    for (int pass = 1; pass <= (need_pass2 ? 2 : 1); pass++) {
      for (uint32_t c0 = 0; c0 < deg; c0 += TEAM) {
        const uint32_t j = c0 + tid;
        uint32_t L = 0, vo = 0, v = 0;
        if (j < deg) {
          v = a.redg[io + j] >> 1;
          vo = a.roff[v + lo];
          const uint32_t ip = a.inpref[v + lo];
          L = (v < il) ? ip : rank_lower_bound(a.redg + vo, ip, il);
        }
        uint32_t T;
        const uint32_t incl = team_incl_scan<TEAM>(L, sh, &T);
        sh.P[tid] = incl;
        sh.vs[tid] = vo;
        if (pass == 2) sh.acc[tid] = 0;
        team_sync<TEAM>();
        for (uint32_t w = tid; w < T; w += TEAM) {
          const uint32_t s = seg_of<TEAM>(sh.P, w);
          const uint32_t k = w - (s ? sh.P[s - 1] : 0u);
          const uint32_t p = sh.vs[s] + k;
          const uint32_t u2 = a.redg[p] >> 1;
          if (pass == 1) {
            if (atomicAdd(&dense[u2], 1u) == 0u) dused[atomicAdd(&sh.nused, 1u)] = u2;
          } else {
            const uint32_t c = dense[u2];
            if (c > 1u) {
              atomicAdd(&sh.acc[s], (unsigned long long)(c - 1u));
              if constexpr (MODE == MODE_EDGE) atomicAdd(&a.B[p], (unsigned long long)(c - 1u));
            }
          }
        }
        team_sync<TEAM>();
        if (pass == 2 && j < deg) {
          const unsigned long long s = sh.acc[tid];
          if (s) {
            if constexpr (MODE == MODE_EDGE) atomicAdd(&a.B[io + j], s);
            else atomicAdd(&a.B[v + lo], s);
          }
        }
        team_sync<TEAM>();
      }
      if (pass == 1 && (MODE == MODE_TOTAL || (MODE == MODE_VERT && i_in_S))) {
        unsigned long long sum = 0;
        for (uint32_t q = tid; q < sh.nused; q += TEAM) {
          const uint32_t key = dused[q];
          const uint32_t c = dense[key];
          if (c > 1u) {
            const unsigned long long b = (unsigned long long)c * (c - 1u) / 2u;
            sum += b;
            if constexpr (MODE == MODE_VERT) atomicAdd(&a.B[key + lo], b);
          }
        }
        sum = team_sum<TEAM>(sum, sh);
        if (tid == 0 && sum) {
          if constexpr (MODE == MODE_VERT) atomicAdd(&a.B[i], sum);
          else block_total += sum;
        }
      }
    }
    team_sync<TEAM>();
    // leave the whole slice zero, the used list included: a later call with another vertex
    // count lays its dense counters over this storage
    for (uint32_t q = tid; q < sh.nused; q += TEAM) {
      dense[dused[q]] = 0;
      dused[q] = 0;
    }
    team_sync<TEAM>();
  }
  if constexpr (MODE == MODE_TOTAL) {
    if (tid == 0 && block_total) atomicAdd(a.total, block_total);
  }
}

// ------------------------------------------------------------------------------------------
// Fused counting kernel (one launch): a workgroup queue of heavy endpoints, then a wavefront
// queue of light ones.  Tables:
//   * dense u32 counters indexed by u2 when i < D32_MAX (u2 < i),
//   * dense u16 counters (two per word) when i < D16_MAX and deg(i) < 65536 (so c < 65536),
//   * otherwise open addressing: packed 32-bit slots (key << cb | count) when the counts fit
//     in cb bits, else 64-bit slots (key << 32 | count).
// An endpoint is light when its distinct-key bound min(wedges, i) fits a wavefront's private
// 2048-slot table at load <= 3/4 and its wedges are few; such a table cannot overflow.  A
// workgroup hash that fills up sends its endpoint to the HBM kernel.  Credits to ranks <
// HOT_N (the hubs, which receive most credits) go to HOT_R replicated accumulators to spread
// atomic contention; k_hot_reduce folds them into the counts.
// ------------------------------------------------------------------------------------------
constexpr int NWAVE = 4;
constexpr int TPB = NWAVE * 64;
constexpr uint32_t POOL_WORDS = 8192;             // 32 KB of LDS per workgroup
constexpr uint32_t D32_MAX = POOL_WORDS;          // dense u32: i < 8192
constexpr uint32_t D16_MAX = 2 * POOL_WORDS;      // dense u16: i < 16384
constexpr uint32_t LT_WORDS = POOL_WORDS / NWAVE; // wavefront table: 2048 packed slots
constexpr uint32_t LG_CLAIM = 2;                  // light groups a wavefront claims at once
constexpr uint32_t HB_CLAIM = 1;                  // heavy endpoints a workgroup claims at once
constexpr uint32_t LIGHT_MAX = LT_WORDS / 4 * 3 - 1;  // light endpoint: <= 1535 distinct keys ...
constexpr uint32_t LIGHT_WORK = 1u << 14;         // ... and at most 16384 wedges
static_assert(LIGHT_MAX + LIGHT_MAX / 3 + 1 <= LT_WORDS, "a light table must fit its wavefront's slots");
static_assert(NWAVE * LT_WORDS <= POOL_WORDS, "the light tables of a workgroup live in its pool");
constexpr uint32_t WIDE_DEG = 1u << 24;           // per-window u32 sums need deg(i) < 2^24
constexpr uint32_t PROBE_MAX = 64;                // probes per insert into a light sub-table
constexpr uint32_t WIN = 4;                       // wedge indices per lane per window
constexpr uint32_t WIN_SPAN = 64 * WIN;           // window of 256 consecutive wedge indices
constexpr uint32_t HOT_N = 4096, HOT_R = 16;
constexpr uint32_t EPOCH_MAX = 1u << 26;          // window tags are epoch << 6 | segment

enum TabMode { TM_D32 = 0, TM_D16 = 1, TM_HASH = 2, TM_HASH32 = 3, TM_G32 = 4 };

// LDS pointers carry address space 3 so that every access compiles to ds_* instructions
// (through a generic pointer the compiler emits flat accesses, which also wait for all
// outstanding global loads).
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) unsigned long long lds_u64;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;

__device__ __forceinline__ uint32_t lds_load(lds_u32* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_store(lds_u32* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t lds_add(lds_u32* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ unsigned long long lds_add(lds_u64* p, unsigned long long v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t lds_cas(lds_u32* p, uint32_t expect, uint32_t desired) {
  __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
  return expect;
}
__device__ __forceinline__ unsigned long long lds_cas(lds_u64* p, unsigned long long expect, unsigned long long desired) {
  __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
  return expect;
}

// Orders this wavefront's LDS accesses around cross-lane exchanges through LDS.  LDS
// operations of one wavefront are performed in issue order, so only the compiler needs
// fencing here.
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// ---------------- wavefront scans on DPP (no LDS traffic) ----------------
// row_shr:d (0x110 + d) within rows of 16 lanes, row_bcast:15 (0x142) into rows 1 and 3,
// row_bcast:31 (0x143) into rows 2 and 3, wave_shr:1 (0x138), wave_shl:1 (0x130).  Lanes
// without a source read 0 (bound_ctrl), rows outside row_mask keep 0.
template <int CTRL, int ROWM = 0xf>
__device__ __forceinline__ uint32_t dpp0(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWM, 0xf, true);
}

__device__ __forceinline__ uint32_t wave_incl_add(uint32_t x) {
  x += dpp0<0x111>(x);
  x += dpp0<0x112>(x);
  x += dpp0<0x114>(x);
  x += dpp0<0x118>(x);
  x += dpp0<0x142, 0xa>(x);
  x += dpp0<0x143, 0xc>(x);
  return x;
}

__device__ __forceinline__ uint32_t wave_incl_max(uint32_t x) {
  x = max(x, dpp0<0x111>(x));
  x = max(x, dpp0<0x112>(x));
  x = max(x, dpp0<0x114>(x));
  x = max(x, dpp0<0x118>(x));
  x = max(x, dpp0<0x142, 0xa>(x));
  x = max(x, dpp0<0x143, 0xc>(x));
  return x;
}

// Segmented inclusive sum: lanes with equal keys s form runs (s non-decreasing over lanes).
__device__ __forceinline__ uint32_t wave_seg_incl_add(uint32_t x, uint32_t s) {
#define PBF_SEG_STEP(CTRL, ROWM)                  \
  {                                               \
    const uint32_t y = dpp0<CTRL, ROWM>(x);       \
    const uint32_t t = dpp0<CTRL, ROWM>(s);       \
    x += (t == s) ? y : 0u;                       \
  }
  PBF_SEG_STEP(0x111, 0xf)
  PBF_SEG_STEP(0x112, 0xf)
  PBF_SEG_STEP(0x114, 0xf)
  PBF_SEG_STEP(0x118, 0xf)
  PBF_SEG_STEP(0x142, 0xa)
  PBF_SEG_STEP(0x143, 0xc)
#undef PBF_SEG_STEP
  return x;
}

__device__ __forceinline__ uint32_t wave_shr1(uint32_t x) { return dpp0<0x138>(x); }
__device__ __forceinline__ uint32_t wave_shl1_or(uint32_t x, uint32_t fill) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)x, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t lane_read(uint32_t x, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, (int)lane);
}

// ---------------- tables ----------------
struct Tab {
  lds_u32* w;        // LDS words
  uint32_t mask, shift, cb;
  const uint32_t* g; // TM_G32: dense counts in HBM (lookups only)
  uint32_t plim;     // inserts: probes beyond the first before giving up
};

// Inserts key u2: returns 1 when the key is new, 0 when it was present, 2 when the probe
// sequence found no room (the table is full).
template <int TM>
__device__ __forceinline__ uint32_t tab_ins(const Tab& t, uint32_t u2) {
  if constexpr (TM == TM_D32) {
    lds_add(t.w + u2, 1u);
    return 0;
  } else if constexpr (TM == TM_D16) {
    lds_add(t.w + (u2 >> 1), 1u << ((u2 & 1u) << 4));
    return 0;
  } else if constexpr (TM == TM_HASH32) {
    uint32_t h = hash_slot(u2, t.shift);
    const uint32_t mine = (u2 << t.cb) | 1u;
    for (uint32_t probe = 0; probe <= t.plim; probe++) {
      const uint32_t old = lds_cas(t.w + h, EMPTY, mine);
      if (old == EMPTY) return 1;
      if ((old >> t.cb) == u2) { lds_add(t.w + h, 1u); return 0; }
      h = (h + 1) & t.mask;
    }
    return 2;
  } else {
    lds_u64* s = (lds_u64*)t.w;
    uint32_t h = hash_slot(u2, t.shift);
    const unsigned long long mine = ((unsigned long long)u2 << 32) | 1ull;
    for (uint32_t probe = 0; probe <= t.plim; probe++) {
      const unsigned long long old = lds_cas(s + h, ~0ull, mine);
      if (old == ~0ull) return 1;
      if ((uint32_t)(old >> 32) == u2) { lds_add(s + h, 1ull); return 0; }
      h = (h + 1) & t.mask;
    }
    return 2;
  }
}

template <int TM>
__device__ __forceinline__ uint32_t tab_get(const Tab& t, uint32_t u2) {
  if constexpr (TM == TM_D32) {
    return t.w[u2];
  } else if constexpr (TM == TM_G32) {
    return t.g[u2];
  } else if constexpr (TM == TM_D16) {
    return (t.w[u2 >> 1] >> ((u2 & 1u) << 4)) & 0xFFFFu;
  } else if constexpr (TM == TM_HASH32) {
    uint32_t h = hash_slot(u2, t.shift);
    for (uint32_t probe = 0; probe <= t.mask; probe++) {
      const uint32_t cur = t.w[h];
      if ((cur >> t.cb) == u2) return cur & ((1u << t.cb) - 1u);  // EMPTY never matches a key
      if (cur == EMPTY) return 0u;
      h = (h + 1) & t.mask;
    }
    return 0u;
  } else {
    const lds_u64* s = (const lds_u64*)t.w;
    uint32_t h = hash_slot(u2, t.shift);
    for (uint32_t probe = 0; probe <= t.mask; probe++) {
      const unsigned long long cur = s[h];
      if ((uint32_t)(cur >> 32) == u2 && cur != ~0ull) return (uint32_t)cur;
      if (cur == ~0ull) return 0u;
      h = (h + 1) & t.mask;
    }
    return 0u;
  }
}

// Continuations of a packed-hash insert / lookup whose first probe (slot h - 1) collided.
__device__ __forceinline__ uint32_t tab_ins_from(const Tab& t, uint32_t u2, uint32_t h) {
  const uint32_t mine = (u2 << t.cb) | 1u;
  for (uint32_t probe = 1; probe <= t.plim; probe++) {
    const uint32_t old = lds_cas(t.w + h, EMPTY, mine);
    if (old == EMPTY) return 1;
    if ((old >> t.cb) == u2) { lds_add(t.w + h, 1u); return 0; }
    h = (h + 1) & t.mask;
  }
  return 2;
}
__device__ __forceinline__ uint32_t tab_get_from(const Tab& t, uint32_t u2, uint32_t h) {
  for (uint32_t probe = 1; probe <= t.mask; probe++) {
    const uint32_t cur = t.w[h];
    if ((cur >> t.cb) == u2) return cur & ((1u << t.cb) - 1u);
    if (cur == EMPTY) return 0u;
    h = (h + 1) & t.mask;
  }
  return 0u;
}

// Credit to rank x: hub ranks go to replicated accumulators.
__device__ __forceinline__ void credit(const CountArgs& a, uint32_t x, unsigned long long val) {
  if (x < HOT_N) atomicAdd(&a.hot[x * HOT_R + (blockIdx.x & (HOT_R - 1))], val);
  else atomicAdd(&a.B[x], val);
}

// ---------------- the hub-credit log ----------------
// Most endpoint credits go to the hubs (ranks below HOT_N): on a power-law graph over 90 % of
// them.  Scattered 64-bit device atomics run at a small fraction of the rate of coalesced
// ones, so a wavefront appends its hub credits to an HBM log instead: consecutive entries of
// a chunk of HLOG_CHUNK it reserved with one counter atomic (coalesced stores);
// k_hlog_reduce sums each workgroup's share of the log per rank in LDS into a row of partial
// sums and k_hlog_sum adds the rows' columns to B.  Entries are (rank << 52) | value; a value of
// 2^52 or more, an entry past the log's capacity, and every credit to another rank take the
// direct atomic.  An abandoned chunk's unused entries are zeroed (credits of 0).
constexpr uint32_t HLOG_CHUNK = 1024;
constexpr unsigned long long HLOG_VMAX = 1ull << 52;

struct HotLog {
  uint32_t base = 0;           // the wavefront's chunk (wave-uniform)
  uint32_t fill = HLOG_CHUNK;  // entries of it used; HLOG_CHUNK: no chunk yet
};

// Zeroes the unused entries of the wavefront's chunk.
__device__ __forceinline__ void hlog_pad(const CountArgs& a, const HotLog& hl) {
  for (uint32_t q = hl.fill + __lane_id(); q < HLOG_CHUNK; q += 64) {
    const uint64_t p = (uint64_t)hl.base + q;
    if (p < a.hlog_cap) a.hlog[p] = 0ull;
  }
}

// Credits val to rank x on the lanes with act; every lane of the wavefront calls it together.
__device__ __forceinline__ void credit_wave(const CountArgs& a, HotLog& hl, bool act, uint32_t x,
                                            unsigned long long val) {
  const bool logged = act && x < HOT_N && val < HLOG_VMAX && a.hlog_cap != 0u;
  if (act && !logged) credit(a, x, val);
  const uint64_t lm = __ballot(logged);
  if (lm == 0ull) return;
  const uint32_t cnt = (uint32_t)__popcll(lm);
  if (hl.fill + cnt > HLOG_CHUNK) {
    hlog_pad(a, hl);
    uint32_t b = 0;
    if (__lane_id() == (uint32_t)__ffsll((long long)__ballot(1)) - 1u) b = atomicAdd(a.hlog_len, HLOG_CHUNK);
    hl.base = (uint32_t)__builtin_amdgcn_readfirstlane((int)b);
    hl.fill = 0;
  }
  if (logged) {
    const uint64_t p = (uint64_t)hl.base + hl.fill + (uint32_t)__popcll(lm & pbf_lanemask_lt());
    if (p < a.hlog_cap) a.hlog[p] = ((unsigned long long)x << 52) | val;
    else credit(a, x, val);
  }
  hl.fill += cnt;
}

// Sums the log per rank in LDS and adds each rank's sum to its count (coalesced atomics).
__global__ __launch_bounds__(256) void k_hlog_reduce(const unsigned long long* __restrict__ hlog,
                                                     const uint32_t* __restrict__ len, uint32_t cap,
                                                     unsigned long long* __restrict__ part) {
  __shared__ unsigned long long acc[HOT_N];
  for (uint32_t x = threadIdx.x; x < HOT_N; x += blockDim.x) acc[x] = 0;
  __syncthreads();
  const uint32_t L = min(*len, cap);
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < L; p += gridDim.x * blockDim.x) {
    const unsigned long long e = hlog[p];
    if (e) atomicAdd(&acc[e >> 52], e & (HLOG_VMAX - 1ull));
  }
  __syncthreads();
  // row blockIdx.x of the partial sums (plain stores: the ranks' atomics from every workgroup
  // would pile up on the same HOT_N words)
  for (uint32_t x = threadIdx.x; x < HOT_N; x += blockDim.x) part[(size_t)blockIdx.x * HOT_N + x] = acc[x];
}

// Rank x of the hub log: the sum of column x of the partial sums over rows [HLOG_ROWS_PER * by,
// + HLOG_ROWS_PER), added to B[x].
constexpr uint32_t HLOG_ROWS_PER = 64;
__global__ __launch_bounds__(256) void k_hlog_sum(const unsigned long long* __restrict__ part, uint32_t rows,
                                                  unsigned long long* __restrict__ B, int64_t n) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t r0 = blockIdx.y * HLOG_ROWS_PER, r1 = min(r0 + HLOG_ROWS_PER, rows);
  if (x >= HOT_N || (int64_t)x >= n) return;
  unsigned long long s = 0;
#pragma unroll 8
  for (uint32_t r = r0; r < r1; r++) s += part[(size_t)r * HOT_N + x];
  if (s) atomicAdd(&B[x], s);
}

// ---------------- windowed wedge enumeration ----------------
// Per wavefront LDS: the window's segment-start flags and the per-segment partial sums.
struct WaveCtx {
  lds_u32* flags;   // WIN_SPAN words, 16-byte aligned
  lds_u64* acc;     // 64 partial sums (pass 2), kept zero between chunks
  uint32_t epoch;   // tag generation of the flags
  HotLog hl;        // the wavefront's chunk of the hub-credit log
};

// Overflow tracking of a shared hash table (one workgroup).
struct Track {
  lds_u32* nused;
  lds_u32* ovf;
  uint32_t maxfill;
  bool on;
};

// A chunk: up to 64 segments (endpoint i, centre v), one per lane.  Segment s covers the
// wedge indices [P, P + L) of the chunk, i.e. the entries N(v)[0, L) below min(i, v).
struct Chunk {
  uint32_t v;      // centre rank (global)
  uint32_t slot;   // ranked slot of (i -> v): the edge credited in EDGE mode
  uint32_t L, P;   // segment length and first wedge index
  uint32_t T;      // wedges of the chunk (wave-uniform)
  uint32_t base;   // slot of wedge index x of this segment: base + x
  uint32_t td;     // grouped tables: the endpoint's sub-table (word offset | log2 size << 16)
};

// A chunk from its lanes' slot records (si, the slot's entry e), loaded beforehand.
__device__ __forceinline__ Chunk make_chunk(uint2 si, uint32_t e, uint32_t slot, bool valid, uint32_t td, uint32_t lo) {
  Chunk ch;
  ch.v = 0; ch.L = 0; ch.slot = slot; ch.td = td;
  uint32_t vo = 0;
  if (valid) {
    ch.v = (e >> 1) + lo;  // global rank of the centre
    vo = si.x;
    ch.L = si.y;
  }
  const uint32_t incl = wave_incl_add(ch.L);
  ch.T = lane_read(incl, 63);
  ch.P = incl - ch.L;
  ch.base = vo - ch.P;
  return ch;
}

__device__ __forceinline__ Chunk load_chunk(const CountArgs& a, uint32_t slot, bool valid, uint32_t td, uint32_t lo) {
  Chunk ch;
  ch.v = 0; ch.L = 0; ch.slot = slot; ch.td = td;
  uint32_t vo = 0;
  if (valid) {
    const uint2 si = a.sinfo[slot];
    ch.v = (a.redg[slot] >> 1) + lo;  // global rank of the centre
    vo = si.x;
    ch.L = si.y;
  }
  const uint32_t incl = wave_incl_add(ch.L);
  ch.T = lane_read(incl, 63);
  ch.P = incl - ch.L;
  ch.base = vo - ch.P;
  return ch;
}

// One window: lane l holds wedge indices b0 + 4l + k (blocked layout) or b0 + 64k + l
// (interleaved layout), k < nval.  In the interleaved layout the k-th wedges of consecutive
// lanes are consecutive entries of N(v): a per-wedge credit to the slot of each is one
// coalesced atomic per k (64-bit atomics reach about 3.6x the rate of the blocked layout's
// 4-word lane stride on this part).
struct Win {
  uint32_t ent[WIN];  // ranked entry (rank << 1 | flag) of u2
  uint32_t pos[WIN];  // its slot in the ranked adjacency
  uint32_t seg[WIN];  // segment (lane) within the chunk
  uint32_t td[WIN];   // grouped tables: sub-table of the segment
  uint32_t nval;
};

__device__ __forceinline__ uint32_t next_tag(WaveCtx& wc) {
  if (++wc.epoch == EPOCH_MAX) {
    // tags would wrap: forget every stale start
    wave_lds_fence();
    *(lds_u32x4*)(wc.flags + 4 * __lane_id()) = u32x4{0u, 0u, 0u, 0u};
    wave_lds_fence();
    wc.epoch = 1;
  }
  return wc.epoch << 6;
}

// Maps the window at b0 to (segment, slot) per wedge index and issues its loads.  A segment
// s with L_s > 0 starting inside the window writes its tag (epoch << 6 | s) at its start; an
// inclusive max-scan over the window then yields, per index, the last segment started at or
// before it.  The segment covering b0 itself (it may have started in an earlier window) is
// the highest lane with L > 0 and P <= b0.  Stale flags from earlier windows of the chunk
// hold smaller segments, those of earlier chunks smaller epochs: the max ignores both.
template <bool MULTI, bool IL>
__device__ __forceinline__ void win_map(const CountArgs& a, lds_u32* flags, uint32_t b0, const Chunk& ch, uint32_t tag,
                                        Win& wv) {
  const uint32_t lane = __lane_id();
  if (ch.L != 0u && ch.P >= b0 && ch.P - b0 < WIN_SPAN) flags[ch.P - b0] = tag | lane;
  const uint64_t cm = __ballot(ch.L != 0u && ch.P <= b0);
  const uint32_t carry = tag | (63u - (uint32_t)__clzll(cm));
  wave_lds_fence();
  if constexpr (IL) {
    // row k = indices b0 + 64k .. b0 + 64k + 63, one per lane: a max-scan per row, each row
    // starting from the maximum of everything before it
    uint32_t f[WIN];
#pragma unroll
    for (uint32_t k = 0; k < WIN; k++) f[k] = flags[64 * k + lane];
    wave_lds_fence();
    const uint32_t q0 = b0 + lane;
    wv.nval = q0 < ch.T ? min((ch.T - q0 + 63u) / 64u, WIN) : 0u;
    uint32_t run = carry;
#pragma unroll
    for (uint32_t k = 0; k < WIN; k++) {
      const uint32_t m = max(wave_incl_max(f[k]), run);
      run = lane_read(m, 63);
      const uint32_t sg = m & 63u;
      wv.seg[k] = sg;
      wv.pos[k] = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(sg << 2), (int)ch.base) + q0 + 64u * k;
      if constexpr (MULTI) wv.td[k] = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(sg << 2), (int)ch.td);
    }
#pragma unroll
    for (uint32_t k = 0; k < WIN; k++) wv.ent[k] = a.redg[k < wv.nval ? wv.pos[k] : 0u];
    return;
  }
  const u32x4 f = *(lds_u32x4*)(flags + 4 * lane);
  wave_lds_fence();
  const uint32_t m0 = f.x, m1 = max(m0, f.y), m2 = max(m1, f.z), m3 = max(m2, f.w);
  const uint32_t cin = max(wave_shr1(wave_incl_max(m3)), carry);
  const uint32_t q0 = b0 + WIN * lane;
  wv.nval = q0 < ch.T ? min(ch.T - q0, WIN) : 0u;
  const uint32_t mm[WIN] = {m0, m1, m2, m3};
#pragma unroll
  for (uint32_t k = 0; k < WIN; k++) {
    const uint32_t s = max(cin, mm[k]) & 63u;
    wv.seg[k] = s;
    wv.pos[k] = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(s << 2), (int)ch.base) + q0 + k;
    if constexpr (MULTI) wv.td[k] = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(s << 2), (int)ch.td);
  }
#pragma unroll
  // unconditional loads (lanes past the end read slot 0, and their entries are never used): a
  // fixed count of loads in flight lets the consumer wait for this window's loads only, not for
  // the next window's, which are issued before this one is aggregated
  for (uint32_t k = 0; k < WIN; k++) wv.ent[k] = a.redg[k < wv.nval ? wv.pos[k] : 0u];
}

// The sub-table of a grouped endpoint.
__device__ __forceinline__ Tab sub_tab(const Tab& t, uint32_t td) {
  const uint32_t lg = td >> 16;
  const uint32_t mask = (1u << lg) - 1u;
  return Tab{t.w + (td & 0xFFFFu), mask, 32u - lg, t.cb, nullptr, mask < PROBE_MAX ? mask : PROBE_MAX - 1u};
}

// Pass 1 on one window: returns the lane's number of new keys, plus 256 on a failed insert.
// Packed hash tables take the first probe of all WIN wedges at once (their CAS round trips
// overlap); only collisions continue probing one wedge at a time.
template <int TM, bool MULTI>
__device__ __forceinline__ uint32_t win_insert(const Tab& t, const Win& wv) {
  uint32_t r = 0;
  if constexpr (TM == TM_HASH32) {
    uint32_t h[WIN], old[WIN];
#pragma unroll
    for (uint32_t k = 0; k < WIN; k++) {
      const Tab tk = MULTI ? sub_tab(t, wv.td[k]) : t;
      const uint32_t u2 = wv.ent[k] >> 1;
      h[k] = hash_slot(u2, tk.shift);
      old[k] = k < wv.nval ? lds_cas(tk.w + h[k], EMPTY, (u2 << t.cb) | 1u) : EMPTY;
    }
    uint32_t slow = 0;
#pragma unroll
    for (uint32_t k = 0; k < WIN; k++) {
      if (k < wv.nval) {
        const uint32_t u2 = wv.ent[k] >> 1;
        if (old[k] == EMPTY) r += 1u;
        else if ((old[k] >> t.cb) == u2) lds_add((MULTI ? sub_tab(t, wv.td[k]) : t).w + h[k], 1u);
        else slow |= 1u << k;
      }
    }
    if (__ballot(slow != 0u)) {
#pragma unroll
      for (uint32_t k = 0; k < WIN; k++) {
        if (slow & (1u << k)) {
          const Tab tk = MULTI ? sub_tab(t, wv.td[k]) : t;
          const uint32_t s = tab_ins_from(tk, wv.ent[k] >> 1, (h[k] + 1u) & tk.mask);
          r += (s == 2u) ? 256u : s;
        }
      }
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < WIN; k++) {
      if (k < wv.nval) {
        const uint32_t s = tab_ins<TM>(MULTI ? sub_tab(t, wv.td[k]) : t, wv.ent[k] >> 1);
        r += (s == 2u) ? 256u : s;
      }
    }
  }
  return r;
}

// Counts of the WIN wedges of a lane; packed hash tables read all first probes at once.
template <int TM, bool MULTI>
__device__ __forceinline__ void win_get(const Tab& t, const Win& wv, uint32_t* c) {
  if constexpr (TM == TM_HASH32) {
    uint32_t h[WIN], cur[WIN];
#pragma unroll
    for (uint32_t k = 0; k < WIN; k++) {
      const Tab tk = MULTI ? sub_tab(t, wv.td[k]) : t;
      h[k] = hash_slot(wv.ent[k] >> 1, tk.shift);
      cur[k] = k < wv.nval ? tk.w[h[k]] : EMPTY;
    }
    const uint32_t cmask = (1u << t.cb) - 1u;
    uint32_t slow = 0;
#pragma unroll
    for (uint32_t k = 0; k < WIN; k++) {
      c[k] = 0;
      if (k < wv.nval) {
        if ((cur[k] >> t.cb) == (wv.ent[k] >> 1)) c[k] = cur[k] & cmask;
        else if (cur[k] != EMPTY) slow |= 1u << k;
      }
    }
    if (__ballot(slow != 0u)) {
#pragma unroll
      for (uint32_t k = 0; k < WIN; k++) {
        if (slow & (1u << k)) {
          const Tab tk = MULTI ? sub_tab(t, wv.td[k]) : t;
          c[k] = tab_get_from(tk, wv.ent[k] >> 1, (h[k] + 1u) & tk.mask);
        }
      }
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < WIN; k++) c[k] = k < wv.nval ? tab_get<TM>(MULTI ? sub_tab(t, wv.td[k]) : t, wv.ent[k] >> 1) : 0u;
  }
}

// Pass 2 on one window: c - 1 per wedge with c > 1, credited to the edge (v -> u2) (EDGE) and
// summed per segment into acc.  A lane folds its own wedges of its last segment; the others
// (a segment boundary inside the lane) go straight to acc.  Lane sums are then reduced per
// segment with a segmented scan; the last lane of each run adds the run's sum.  The u32 sums
// hold at most 256 values below deg(i) < 2^24.
template <int MODE, int TM, bool MULTI, bool IL>
__device__ __forceinline__ void win_credit(const CountArgs& a, const Tab& t, const Win& wv, lds_u64* acc,
                                           uint64_t wide) {
  uint32_t val[WIN], cnt[WIN];
  uint32_t last = wv.seg[0];
  win_get<TM, MULTI>(t, wv, cnt);
  if constexpr (IL) {
    // row by row: the credits of a row go to consecutive slots, and the row's segment sums
    // come from a segmented scan over its lanes (segments are non-decreasing along a row)
#pragma unroll
    for (uint32_t k = 0; k < WIN; k++) {
      const uint32_t c = cnt[k];
      const uint32_t v = (k < wv.nval && c > 1u) ? c - 1u : 0u;
      if constexpr (MODE == MODE_EDGE) {
        if (v) {
          // wide: the segments whose centre's slots need 64-bit credits
          if (wide != 0ull && ((wide >> wv.seg[k]) & 1ull)) atomicAdd(&a.B[wv.pos[k]], (unsigned long long)v);
          else atomicAdd(a.B32 + wv.pos[k], v);
        }
      }
      if (__ballot(v != 0u) == 0ull) continue;
      const uint32_t key = k < wv.nval ? wv.seg[k] : 64u;
      const uint32_t x = wave_seg_incl_add(v, key);
      const uint32_t knext = wave_shl1_or(key, ~0u);
      if (key < 64u && x && knext != key) lds_add(acc + key, (unsigned long long)x);
    }
    return;
  }
#pragma unroll
  for (uint32_t k = 0; k < WIN; k++) {
    val[k] = 0;
    if (k < wv.nval) {
      last = wv.seg[k];
      const uint32_t c = cnt[k];
      if (c > 1u) {
        val[k] = c - 1u;
        if constexpr (MODE == MODE_EDGE) atomicAdd(&a.B[wv.pos[k]], (unsigned long long)(c - 1u));
      }
    }
  }
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < WIN; k++) {
    if (wv.seg[k] == last) sum += val[k];
    else if (val[k]) lds_add(acc + wv.seg[k], (unsigned long long)val[k]);
  }
  if (__ballot(sum != 0u) == 0ull) return;
  const uint32_t key = wv.nval ? last : 64u;  // lanes past the end form a trailing run
  const uint32_t x = wave_seg_incl_add(sum, key);
  // cross-lane reads run with every lane active: a lane disabled by the branch below would
  // not supply its key to its left neighbour
  const uint32_t knext = wave_shl1_or(key, ~0u);
  if (key < 64u && x && knext != key) lds_add(acc + key, (unsigned long long)x);
}

// Runs windows wstart, wstart + wstep, ... of a chunk.  The next window is mapped and its
// loads issued before the current one is aggregated.  Pass 2 ends with the per-segment
// credits (centre v, or edge i -> v).  Returns false when a table overflowed in pass 1.
// Window layout of a pass: interleaved for the per-edge credit pass (a coalesced atomic per
// wedge row), blocked otherwise (a lane's wedges mostly share its segment).
template <int MODE, bool PASS2>
constexpr bool win_interleaved() { return MODE == MODE_EDGE && PASS2; }

template <int MODE, bool PASS2, int TM, bool MULTI>
__device__ __forceinline__ bool chunk_windows(const CountArgs& a, const Tab& t, const Chunk& ch, uint32_t wstart,
                                              uint32_t wstep, WaveCtx& wc, const Track& tr) {
  constexpr bool IL = win_interleaved<MODE, PASS2>();
  const uint32_t lane = __lane_id();
  if (ch.T <= wstart) return true;
  const uint32_t tag = next_tag(wc);
  uint64_t wide = ~0ull;  // per segment: its centre's slots take 64-bit credits
  if constexpr (PASS2 && MODE == MODE_EDGE) {
    if (a.B32 != nullptr) {
      bool w = false;
      if (ch.L != 0u) {
        const uint64_t d = a.roff[ch.v + 1] - a.roff[ch.v];
        const uint32_t md = *a.maxdeg;
        w = (d - 1u) * (uint64_t)(md > 0u ? md - 1u : 0u) >= (1ull << 32);
      }
      wide = __ballot(w);
    }
  }
  Win cur, nxt;
  uint32_t b0 = wstart;
  win_map<MULTI, IL>(a, wc.flags, b0, ch, tag, cur);
  // the first window's loads and lane exchanges complete before the loop: with nothing of
  // `cur` pending on entry, the waits inside the loop cover only what they must (otherwise
  // they also wait for the next window's loads and exchanges, issued just before, and the
  // pipelining is lost).  0x0070: vmcnt(0) lgkmcnt(0).
  __builtin_amdgcn_s_waitcnt(0x0070);
  while (true) {
    const uint32_t nb = b0 + wstep;
    const bool more = nb < ch.T;
    if (more) win_map<MULTI, IL>(a, wc.flags, nb, ch, tag, nxt);
    if constexpr (!PASS2) {
      const uint32_t r = win_insert<TM, MULTI>(t, cur);
      const bool failed = __ballot(r >= 256u) != 0ull;
      if (tr.on) {
        const uint32_t nnew = lane_read(wave_incl_add(r & 255u), 63);
        if (lane == 0 && (failed || lds_add(tr.nused, nnew) + nnew >= tr.maxfill)) lds_store(tr.ovf, 1u);
        wave_lds_fence();
        if (__builtin_amdgcn_readfirstlane(lds_load(tr.ovf))) return false;
      } else if (failed) {
        if (lane == 0) lds_store(tr.ovf, 1u);
        return false;
      }
    } else {
      win_credit<MODE, TM, MULTI, IL>(a, t, cur, wc.acc, wide);
    }
    if (!more) break;
    cur = nxt;
    b0 = nb;
  }
  if constexpr (PASS2) {
    wave_lds_fence();
    const unsigned long long mine = wc.acc[lane];
    if (mine) wc.acc[lane] = 0;
    if constexpr (MODE == MODE_EDGE) {
      if (mine) atomicAdd(&a.B[ch.slot], mine);
    } else {
      credit_wave(a, wc.hl, mine != 0ull, ch.v, mine);
    }
    wave_lds_fence();
  }
  return true;
}

// Enumerates the wedges of endpoint i over neighbours [jbeg, jend) with wave w of nw.  For
// long lists (several waves) the waves take different 64-neighbour chunks; otherwise they
// share the windows of each chunk.  Returns false when a tracked table overflowed.
template <int MODE, bool PASS2, int TM>
__device__ __forceinline__ bool enum_pass(const CountArgs& a, const Tab& t, uint32_t i, uint32_t io, uint32_t lo,
                                          uint32_t jbeg, uint32_t jend, uint32_t w, uint32_t nw, WaveCtx& wc,
                                          const Track& tr, const Chunk* first = nullptr) {
  const bool split = nw > 1 && jend - jbeg >= 8 * 64;
  const uint32_t cstart = split ? w * 64 : 0u, cstep = split ? nw * 64 : 64u;
  const uint32_t wstart = split ? 0u : w * WIN_SPAN, wstep = split ? WIN_SPAN : nw * WIN_SPAN;
  for (uint32_t c0 = jbeg + cstart; c0 < jend; c0 += cstep) {
    const uint32_t j = c0 + __lane_id();
    // `first`: the chunk at jbeg, already loaded
    const Chunk ch = (first != nullptr && c0 == jbeg) ? *first : load_chunk(a, io + j, j < jend, 0u, lo);
    if (!chunk_windows<MODE, PASS2, TM, false>(a, t, ch, wstart, wstep, wc, tr)) return false;
  }
  return true;
}

// Endpoint credits from a finished table: C(c,2) to every key (VERT) and the sum (returned).
// Cells q = gt, gt + gn, ... of `cells` (gt - lane is the wavefront's first).
template <int MODE, int TM>
__device__ __forceinline__ unsigned long long sweep(const CountArgs& a, const Tab& t, uint32_t cells, uint32_t gt,
                                                    uint32_t gn, uint32_t lo, HotLog& hl) {
  unsigned long long sum = 0;
  if constexpr (TM == TM_HASH32) {
    // packed hash tables (a power of two of at least 64 cells): 4 cells per lane per step,
    // and a step whose cells hold no count above 1 costs one ballot
    const uint32_t cmask = (1u << t.cb) - 1u;
    for (uint32_t qb = 4 * (gt - __lane_id()); qb < cells; qb += 4 * gn) {
      const uint32_t q = qb + 4 * __lane_id();
      u32x4 f = u32x4{EMPTY, EMPTY, EMPTY, EMPTY};
      if (q < cells) f = *(lds_u32x4*)(t.w + q);
      const uint32_t cur[4] = {f.x, f.y, f.z, f.w};
      uint32_t c[4];
      bool any = false;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        c[k] = cur[k] == EMPTY ? 0u : (cur[k] & cmask);
        any |= c[k] > 1u;
      }
      if (__ballot(any) == 0ull) continue;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const unsigned long long b = c[k] > 1u ? (unsigned long long)c[k] * (c[k] - 1u) / 2u : 0ull;
        sum += b;
        if constexpr (MODE == MODE_VERT) credit_wave(a, hl, c[k] > 1u, (cur[k] >> t.cb) + lo, b);
      }
    }
    return sum;
  }
  // every lane of a wavefront runs the same iterations (the credits are wavefront-wide)
  for (uint32_t qb = gt - __lane_id(); qb < cells; qb += gn) {
    const uint32_t q = qb + __lane_id();
    uint32_t key = 0, c = 0;
    if (q < cells) {
      if constexpr (TM == TM_D32) {
        key = q; c = t.w[q];
      } else if constexpr (TM == TM_D16) {
        key = q; c = (t.w[q >> 1] >> ((q & 1u) << 4)) & 0xFFFFu;
      } else if constexpr (TM == TM_HASH32) {
        const uint32_t cur = t.w[q];
        key = cur >> t.cb; c = cur == EMPTY ? 0u : (cur & ((1u << t.cb) - 1u));
      } else {
        const unsigned long long cur = ((const lds_u64*)t.w)[q];
        key = (uint32_t)(cur >> 32); c = cur == ~0ull ? 0u : (uint32_t)cur;
      }
    }
    const unsigned long long b = c > 1u ? (unsigned long long)c * (c - 1u) / 2u : 0ull;
    sum += b;
    if constexpr (MODE == MODE_VERT) credit_wave(a, hl, c > 1u, key + lo, b);
  }
  return sum;
}

// A heavy endpoint's record: the endpoint, roff[i], deg(i), its distinct-key bound, its first
// entry (whose flag says whether i lies on S) and its graph's first rank.
struct HRec {
  uint32_t i, io, deg, bnd, e0, lo;
};

__device__ __forceinline__ void hrec_lists(const CountArgs& a, HRec& r) {
  r.io = a.roff[r.i];
  r.deg = a.roff[r.i + 1] - r.io;
  r.bnd = a.bound[r.i];
  r.lo = lo_of(a, r.i);
}

// Heavy endpoint, whole workgroup.  Returns false when its hash table overflowed (nothing was
// credited).  after_pass1(0) runs after pass 1, after_pass1(1) after the barrier that follows,
// at_end before the last barrier (all three on every path): the item pipeline's loads.
template <int MODE, int TM, typename FA, typename FB>
__device__ bool block_item_tm(const CountArgs& a, const Tab& t, uint32_t i, uint32_t io, uint32_t lo, uint32_t deg,
                              bool i_in_S, uint32_t cells, WaveCtx& wc, const Track& tr, unsigned long long* red,
                              unsigned long long* total_acc, const Chunk& one, FA&& after_pass1, FB&& at_end) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // the list's first chunk was loaded before the table was cleared, and serves both passes
  enum_pass<MODE, false, TM>(a, t, i, io, lo, 0, deg, w, NWAVE, wc, tr, &one);
  after_pass1(0);
  __syncthreads();
  after_pass1(1);
  if (lds_load(tr.ovf)) {
    at_end();
    return false;
  }
  if (MODE == MODE_TOTAL || (MODE == MODE_VERT && i_in_S)) {
    unsigned long long sum = pbf_wave_sum(sweep<MODE, TM>(a, t, cells, tid, TPB, lo, wc.hl));
    if (lane == 0) red[w] = sum;
    __syncthreads();
    if (tid == 0) {
      unsigned long long tot = 0;
      for (int q = 0; q < NWAVE; q++) tot += red[q];
      if (tot) {
        if constexpr (MODE == MODE_VERT) credit(a, i, tot);
        else *total_acc += tot;
      }
    }
  }
  if (MODE == MODE_EDGE || (MODE == MODE_VERT && !i_in_S))
    enum_pass<MODE, true, TM>(a, t, i, io, lo, 0, deg, w, NWAVE, wc, tr, &one);
  at_end();
  __syncthreads();
  return true;
}

template <int MODE, typename FA, typename FB>
__device__ bool block_item(const CountArgs& a, const HRec& r, lds_u32* pool, WaveCtx& wc, lds_u32* nused, lds_u32* ovf,
                           unsigned long long* red, unsigned long long* total_acc, FA&& after_pass1, FB&& at_end) {
  const uint32_t tid = threadIdx.x;
  const uint32_t i = r.i, io = r.io, deg = r.deg;
  const bool i_in_S = (r.e0 & 1u) == 0u;
  const uint32_t lo = r.lo, il = i - lo;  // keys are local ranks below il
  // the list's first chunk (its first 64 slots): loaded now, in flight while the table is cleared
  const uint32_t lane = tid & 63u;
  const bool in1 = lane < deg;
  uint2 si1 = make_uint2(0u, 0u);
  uint32_t e1 = 0;
  if (in1) {
    si1 = a.sinfo[io + lane];
    e1 = a.redg[io + lane];
  }
  Tab t{pool, 0, 0, a.cb, nullptr, 0};
  Track tr{nused, ovf, 0xFFFFFFFFu, false};
  int tm;
  uint32_t words, fill = 0, cells = il;
  if (il < D32_MAX) {
    tm = TM_D32; words = il;
  } else if (il < D16_MAX && deg < 65536u) {
    tm = TM_D16; words = (il + 1) >> 1;
  } else {
    const uint32_t bnd = r.bnd;
    const bool packed = a.cb < 32u && deg < (1u << a.cb);
    const uint32_t capmax = packed ? POOL_WORDS : POOL_WORDS / 2;
    const uint64_t want = (uint64_t)bnd + bnd / 3 + 1;
    uint32_t cap = a.rerun ? capmax : 64u;  // a re-run endpoint's bound was optimistic
    while (cap < want && cap < capmax) cap <<= 1;
    tm = packed ? TM_HASH32 : TM_HASH;
    t.mask = cap - 1;
    t.plim = cap - 1;
    t.shift = 32 - (31 - __clz(cap));
    if (cap == capmax) {
      // the bound does not fit: count keys; every wave may insert one more window after
      // the flag is raised, so occupancy stays below 3/4
      tr.on = true;
      tr.maxfill = (capmax / 4) * 3 - NWAVE * WIN_SPAN;
    }
    words = packed ? cap : 2 * cap;
    fill = EMPTY;
    cells = cap;
  }
  // cleared 4 words per thread per step (the pool has room for words rounded up to 4)
  for (uint32_t q = 4 * tid; q < words; q += 4 * TPB) *(lds_u32x4*)(pool + q) = u32x4{fill, fill, fill, fill};
  if (tid == 0) { *nused = 0; *ovf = 0; }
  __syncthreads();
  const Chunk one = make_chunk(si1, e1, io + lane, in1, 0u, lo);
  switch (tm) {
    case TM_D32:
      return block_item_tm<MODE, TM_D32>(a, t, i, io, lo, deg, i_in_S, cells, wc, tr, red, total_acc, one, after_pass1,
                                         at_end);
    case TM_D16:
      return block_item_tm<MODE, TM_D16>(a, t, i, io, lo, deg, i_in_S, cells, wc, tr, red, total_acc, one, after_pass1,
                                         at_end);
    case TM_HASH32:
      return block_item_tm<MODE, TM_HASH32>(a, t, i, io, lo, deg, i_in_S, cells, wc, tr, red, total_acc, one, after_pass1,
                                            at_end);
    default:
      return block_item_tm<MODE, TM_HASH>(a, t, i, io, lo, deg, i_in_S, cells, wc, tr, red, total_acc, one, after_pass1,
                                          at_end);
  }
}

// Table size of a light endpoint with distinct-key bound b (load <= 3/4), at least 16 slots.
__device__ __forceinline__ uint32_t light_cap(uint32_t b) {
  uint32_t cap = 16;
  while (cap < b + b / 3 + 1) cap <<= 1;
  return cap;
}

// Largest k < nk with key_k <= x, keys held by lanes 0..nk-1 in increasing order (nk <= 32).
__device__ __forceinline__ uint32_t lane_upper(uint32_t key, uint32_t nk, uint32_t x) {
  uint32_t s = 0;
#pragma unroll
  for (uint32_t bit = 16; bit; bit >>= 1) {
    const uint32_t cand = s + bit;
    const uint32_t kc = (uint32_t)__builtin_amdgcn_ds_bpermute((int)((cand & 63u) << 2), (int)key);
    if (cand < nk && kc <= x) s = cand;
  }
  return s;
}

// A group of light endpoints (one wavefront): up to 32 endpoints whose tables, sorted by
// decreasing power-of-two size, tile the wavefront's LDS slots exactly (buddy-aligned).  Their
// adjacency lists are concatenated into one flattened slot range, taken 64 slots per chunk;
// every wedge is aggregated in its endpoint's sub-table.  One enumeration thus serves several
// endpoints, and the dependent loads of an endpoint's setup are paid once per group.  All
// endpoints of a group lie on the same side.  Lane k < nk holds endpoint k's record rc =
// (i, roff[i], deg(i), log2 table size).  at_b runs after pass 1 and at_c at the end (the
// caller's prefetches of the next groups).  Returns false if a table overflowed (then nothing
// was credited).
template <int MODE, typename FB, typename FC>
__device__ bool light_group(const CountArgs& a, uint32_t nk, const uint4& rc, bool o_side, lds_u32* tw, uint32_t& wm,
                            WaveCtx& wc, lds_u64* iacc, lds_u32* ovf, unsigned long long* total_acc, FB at_b, FC at_c) {
  const uint32_t lane = __lane_id();
  const bool own = lane < nk;
  const uint32_t i_k = own ? rc.x : 0u, io_k = own ? rc.y : 0u, deg_k = own ? rc.z : 0u;
  const uint32_t lg_k = own ? rc.w : 0u, cap_k = own ? 1u << rc.w : 0u;
  const uint32_t lo_k = own ? lo_of(a, i_k) : 0u;
  const uint32_t cinc = wave_incl_add(cap_k);
  const uint32_t words = lane_read(cinc, 63);  // <= LT_WORDS by construction of the groups
  const uint32_t off_k = cinc - cap_k;
  const uint32_t dinc = wave_incl_add(deg_k);
  const uint32_t DS = lane_read(dinc, 63);
  const uint32_t dex_k = dinc - deg_k;
  const uint32_t td_k = off_k | (lg_k << 16);
  const uint32_t dirty = min(wm, words);
  for (uint32_t q = WIN * lane; q < dirty; q += WIN_SPAN) *(lds_u32x4*)(tw + q) = u32x4{EMPTY, EMPTY, EMPTY, EMPTY};
  if (lane == 0) lds_store(ovf, 0u);
  wm = max(wm, words);
  wave_lds_fence();
  const Tab t{tw, 0, 0, a.cb, nullptr, 0};
  const Track tr{nullptr, ovf, 0xFFFFFFFFu, false};
  const bool need_pass2 = MODE == MODE_EDGE || (MODE == MODE_VERT && o_side);
  // chunk c0 of the concatenated lists: lane -> endpoint (largest k with dex_k <= f; every
  // light endpoint has a neighbour, so dex_k is strictly increasing)
  auto chunk_at = [&](uint32_t c0) {
    const uint32_t f = c0 + lane;
    const uint32_t k = lane_upper(dex_k, nk, f);
    const int kb = (int)(k << 2);
    const uint32_t slot = (uint32_t)__builtin_amdgcn_ds_bpermute(kb, (int)io_k) + f -
                          (uint32_t)__builtin_amdgcn_ds_bpermute(kb, (int)dex_k);
    const uint32_t td = (uint32_t)__builtin_amdgcn_ds_bpermute(kb, (int)td_k);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(kb, (int)lo_k);
    return load_chunk(a, slot, f < DS, td, lo);
  };
  bool ok = true;
  if (DS <= 64) {
    // one chunk: its setup serves both passes
    const Chunk ch = chunk_at(0);
    ok = chunk_windows<MODE, false, TM_HASH32, true>(a, t, ch, 0, WIN_SPAN, wc, tr);
    wave_lds_fence();
    at_b();
    if (ok && need_pass2) chunk_windows<MODE, true, TM_HASH32, true>(a, t, ch, 0, WIN_SPAN, wc, tr);
  } else {
    for (uint32_t c0 = 0; c0 < DS && ok; c0 += 64)
      ok = chunk_windows<MODE, false, TM_HASH32, true>(a, t, chunk_at(c0), 0, WIN_SPAN, wc, tr);
    wave_lds_fence();
    at_b();
    if (ok && need_pass2)
      for (uint32_t c0 = 0; c0 < DS; c0 += 64) chunk_windows<MODE, true, TM_HASH32, true>(a, t, chunk_at(c0), 0, WIN_SPAN, wc, tr);
  }
  if (ok && (MODE == MODE_TOTAL || (MODE == MODE_VERT && !o_side))) {
    // endpoint credits: each lane sweeps 4 words of one sub-table per step (sub-tables are
    // multiples of 16 words) and clears them; per-endpoint sums are reduced with a segmented
    // scan (an endpoint's sum of C(c,2) is at most T^2 / 2 < 2^32) into iacc
    const uint32_t cmask = (1u << a.cb) - 1u;
    for (uint32_t q0 = 0; q0 < words; q0 += WIN_SPAN) {
      const uint32_t q = q0 + WIN * lane;
      const bool in = q < words;
      u32x4 f = u32x4{EMPTY, EMPTY, EMPTY, EMPTY};
      if (in) {
        f = *(lds_u32x4*)(tw + q);
        *(lds_u32x4*)(tw + q) = u32x4{EMPTY, EMPTY, EMPTY, EMPTY};
      }
      const uint32_t cur[4] = {f.x, f.y, f.z, f.w};
      // the 4 words lie in one sub-table (sub-tables are multiples of 16 words); every lane
      // takes part in the cross-lane steps (see win_credit)
      const uint32_t kq = lane_upper(off_k, nk, q);
      const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(kq << 2), (int)lo_k);
      uint32_t sum = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint32_t c = cur[k] == EMPTY ? 0u : (cur[k] & cmask);
        const uint32_t b = c > 1u ? c * (c - 1u) / 2u : 0u;
        sum += b;
        if constexpr (MODE == MODE_VERT) credit_wave(a, wc.hl, c > 1u, (cur[k] >> a.cb) + lo, b);
      }
      const uint32_t k = in ? kq : 64u;
      const uint32_t x = wave_seg_incl_add(sum, k);
      const uint32_t knext = wave_shl1_or(k, ~0u);
      if (k < 64u && x && knext != k) lds_add(iacc + k, (unsigned long long)x);
    }
    if (words >= wm) wm = 0;  // swept clean
    wave_lds_fence();
    const unsigned long long s = lane < nk ? iacc[lane] : 0ull;
    if (s) {
      iacc[lane] = 0;
      if constexpr (MODE != MODE_VERT) *total_acc += s;
    }
    if constexpr (MODE == MODE_VERT) credit_wave(a, wc.hl, s != 0ull, i_k, s);
    wave_lds_fence();
  }
  at_c();
  return ok;
}

// PH: 1 = heavy endpoints (workgroup queue), 2 = light groups (wavefront queue), 3 = both.
template <int MODE, int PH>
__global__ __launch_bounds__(TPB) void k_count2(CountArgs a) {
  __shared__ __align__(16) uint32_t pool[POOL_WORDS];
  __shared__ __align__(16) uint32_t s_flags[NWAVE][WIN_SPAN];
  __shared__ unsigned long long s_acc[NWAVE][64];
  __shared__ unsigned long long s_iacc[NWAVE][32];
  __shared__ unsigned long long s_red[NWAVE];
  __shared__ uint32_t s_item, s_nused, s_ovf, s_wovf[NWAVE];
  __shared__ HRec s_hrec;
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned long long total_acc = 0;
  *(lds_u32x4*)((lds_u32*)s_flags[wid] + 4 * lane) = u32x4{0u, 0u, 0u, 0u};
  s_acc[wid][lane] = 0;
  if (lane < 32) s_iacc[wid][lane] = 0;
  __syncthreads();
  WaveCtx wc{(lds_u32*)s_flags[wid], (lds_u64*)s_acc[wid], 0};
  // phase 1: heavy endpoints, whole workgroup.  Software pipeline over items k, k+1, k+2 (thread
  // 0): while item k runs, item k+2 is claimed and item k+1's record is loaded -- its endpoint
  // at the start, its list bounds and key bound after pass 1, its first entry at the end -- so
  // an item starts from its record in LDS, not from three dependent loads.
  if constexpr ((PH & 1) != 0) {
  const uint32_t nb = *a.nblock;
  uint32_t n1 = 0, n2 = 0;  // thread 0: the next item and the one after it
  HRec r1{};                // thread 0: the next item's record
  // thread 0 claims items HB_CLAIM at a time, the next batch as soon as the current one is
  // taken up (one claim per item would serialise on the queue's counter, about 12 ns each)
  uint32_t cb = 0, cleft = 0, pend = 0;
  auto take = [&]() -> uint32_t {
    if (cleft == 0) {
      cb = pend;
      cleft = HB_CLAIM;
      pend = atomicAdd(&a.queues[0], HB_CLAIM);
    }
    cleft--;
    return cb++;
  };
  if (threadIdx.x == 0) {
    pend = atomicAdd(&a.queues[0], HB_CLAIM);
    const uint32_t q0 = take();
    n1 = take();
    HRec r{};
    if (q0 < nb) {
      r.i = a.blist[q0];
      hrec_lists(a, r);
      r.e0 = a.redg[r.io];  // a listed endpoint has wedges, so a non-empty list
    }
    s_hrec = r;
    s_item = q0;
  }
  __syncthreads();
  while (true) {
    const uint32_t idx = s_item;
    const HRec r = s_hrec;
    __syncthreads();  // every thread holds the record before thread 0 stores the next one
    if (idx >= nb) break;
    if (threadIdx.x == 0) {
      n2 = take();  // a workgroup's items increase, so after the first past nb all are
      if (n1 < nb) r1.i = a.blist[n1];
    }
    auto after_pass1 = [&](int stage) {
      if (threadIdx.x == 0 && n1 < nb) {
        if (stage == 0) hrec_lists(a, r1);
        else r1.e0 = a.redg[r1.io];
      }
    };
    auto at_end = [&]() {
      if (threadIdx.x == 0) {
        s_hrec = r1;
        s_item = n1;
        n1 = n2;
      }
    };
    if (!block_item<MODE>(a, r, (lds_u32*)pool, wc, (lds_u32*)&s_nused, (lds_u32*)&s_ovf, s_red, &total_acc, after_pass1,
                          at_end)) {
      if (threadIdx.x == 0) a.ovf_list[atomicAdd(a.ovf_len, 1u)] = r.i;
    }
    __syncthreads();
  }
  }
  // phase 2: groups of light endpoints, one per wavefront, private quarter of the pool
  if constexpr ((PH & 2) != 0) {
  const uint32_t ng = *a.ngroups, ng0 = *a.ngroups0;
  lds_u32* tw = (lds_u32*)(pool + wid * LT_WORDS);
  uint32_t wm = LT_WORDS;  // phase 1 left arbitrary words
  // Software pipeline over groups k, k+1, k+2: while group k runs, the records of k+1 are
  // loaded (after its pass 1) and the bounds of k+2 (at its end), whose claim was issued at
  // its start; the dependent loads of a group's setup thus leave the critical path.
  // Groups are claimed LG_CLAIM at a time, the next batch claimed as soon as the current one is
  // taken up (its atomic has LG_CLAIM groups to return): same-address atomics serialise at
  // about 12 ns each, so one claim per group would bound the launch at that rate.
  uint32_t cb = 0, cleft = LG_CLAIM, pend = 0;
  if (lane == 0) { cb = atomicAdd(&a.queues[1], LG_CLAIM); pend = atomicAdd(&a.queues[1], LG_CLAIM); }
  cb = (uint32_t)__shfl((int)cb, 0, 64);
  auto take = [&]() -> uint32_t {  // wave-uniform
    if (cleft == 0) {
      cb = (uint32_t)__shfl((int)pend, 0, 64);
      cleft = LG_CLAIM;
      if (lane == 0) pend = atomicAdd(&a.queues[1], LG_CLAIM);
    }
    cleft--;
    return cb++;
  };
  uint32_t g = take(), gn = take();
  uint32_t gs = 0, ge = 0, ns = 0, ne = 0;
  uint4 rc = make_uint4(0u, 0u, 0u, 0u);
  if (g < ng) {
    gs = a.gstart[g]; ge = a.gstart[g + 1];
    if (lane < ge - gs) rc = a.lrec[gs + lane];
  }
  if (gn < ng) { ns = a.gstart[gn]; ne = a.gstart[gn + 1]; }
  while (g < ng) {
    const uint32_t gnn = take();  // a wave's groups increase, so after the first past ng all are
    uint4 nrc = make_uint4(0u, 0u, 0u, 0u);
    uint32_t nns = 0, nne = 0;
    auto at_b = [&]() {
      if (gn < ng && lane < ne - ns) nrc = a.lrec[ns + lane];
    };
    auto at_c = [&]() {
      if (gnn < ng) { nns = a.gstart[gnn]; nne = a.gstart[gnn + 1]; }
    };
    const uint32_t nk = ge - gs;
    if (!light_group<MODE>(a, nk, rc, g >= ng0, tw, wm, wc, (lds_u64*)s_iacc[wid], (lds_u32*)&s_wovf[wid], &total_acc,
                           at_b, at_c)) {
      uint32_t o = 0;
      if (lane == 0) o = atomicAdd(a.rerun_len, nk);
      o = __shfl(o, 0, 64);
      if (lane < nk) a.rerun_list[o + lane] = rc.x;
    }
    g = gn; gs = ns; ge = ne; rc = nrc;
    gn = gnn; ns = nns; ne = nne;
  }
  }
  if constexpr (MODE == MODE_VERT) hlog_pad(a, wc.hl);
  if constexpr (MODE == MODE_TOTAL) {
    total_acc = pbf_wave_sum(total_acc);
    if (lane == 0 && total_acc) atomicAdd(a.total, total_acc);
  }
}

// Folds the replicated hub accumulators into B and clears them.
__global__ void k_hot_reduce(unsigned long long* __restrict__ hot, unsigned long long* __restrict__ B, int64_t n) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < (int64_t)HOT_N && x < n;
       x += (int64_t)gridDim.x * blockDim.x) {
    unsigned long long s = 0;
#pragma unroll
    for (uint32_t r = 0; r < HOT_R; r++) { s += hot[x * HOT_R + r]; hot[x * HOT_R + r] = 0; }
    if (s) B[x] += s;
  }
}

// Work classification by the exact wedge count of each endpoint, T_i = sum over v in N(i) of
// L_v (in(v) when v < i, else the position of i inside N(v)); min(T_i, i) bounds its
// distinct keys.  Lists of at most 64 entries are summed by their own lane; longer lists are
// cut into chunks of CL_CHUNK entries summed by separate wavefronts (k_wedge_chunks), so a
// hub's list does not serialise one wavefront.  Appends to the work lists are aggregated per
// wavefront (one atomic per class per wave).
constexpr uint32_t CL_CHUNK = 2048;

__device__ __forceinline__ uint32_t seg_len(const uint2* __restrict__ sinfo, uint32_t slot) { return sinfo[slot].y; }

// lens: [0] light, [1] heavy, [11] HBM fallback.  A light endpoint's table is sized by its
// distinct-key bound min(T, i); when that exceeds a wavefront table but T itself is small, the
// table is sized optimistically for T / 2 keys (the endpoint is re-run with a full-size table
// if it overflows).
struct ClassOut {
  uint32_t* bound;
  uint32_t* wlist;
  uint32_t* blist;
  uint32_t* olist;
  uint32_t* lens;
  uint32_t cb;
  uint32_t light_max;   // <= LIGHT_MAX
  uint32_t light_work;
  const uint32_t* lo;   // graph-local ranks (CountArgs::lo), or null
  uint32_t rlo, rhi;    // only endpoints with ranks in [rlo, rhi) are listed (a batch)
};

// Classifies endpoint i (a thread of a 256-thread workgroup, every thread calling) and appends
// it to its list.  The appends of the workgroup take one counter atomic per list (lists[w][c]:
// wavefront w's count of class c; per-wavefront atomics on the three counters serialise), then
// every lane places its entry.
__device__ __forceinline__ void classify_append(uint32_t i, bool valid, uint64_t T, uint32_t deg, const ClassOut& co,
                                                uint32_t (*cnt)[4], uint32_t* base) {
  uint32_t* __restrict__ bound = co.bound;
  uint32_t* __restrict__ wlist = co.wlist;
  uint32_t* __restrict__ blist = co.blist;
  uint32_t* __restrict__ olist = co.olist;
  uint32_t* __restrict__ lens = co.lens;
  const uint32_t cb = co.cb;
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t cls = 0;  // 0 none, 1 wavefront, 2 workgroup, 3 HBM
  valid = valid && i >= co.rlo && i < co.rhi;
  if (valid && T) {
    const uint32_t il = co.lo ? i - co.lo[i] : i;  // distinct keys are local ranks below il
    const uint32_t b = (uint32_t)(T < (uint64_t)il ? T : (uint64_t)il);
    if (b) {
      const uint32_t est = (b > co.light_max && T <= co.light_work) ? min(b, (uint32_t)(T / 2)) : b;
      const bool light = est <= co.light_max && T <= co.light_work && cb < 32u && deg < (1u << cb);
      cls = deg >= WIDE_DEG ? 3u : light ? 1u : 2u;
      bound[i] = light ? est : b;
    } else {
      bound[i] = 0;
    }
  }
  const uint64_t m1 = __ballot(cls == 1), m2 = __ballot(cls == 2), m3 = __ballot(cls == 3);
  if (lane < 3) cnt[wid][lane] = (uint32_t)__popcll(lane == 0 ? m1 : lane == 1 ? m2 : m3);
  __syncthreads();
  if (threadIdx.x < 3) {
    const uint32_t c = threadIdx.x;
    const uint32_t tot = cnt[0][c] + cnt[1][c] + cnt[2][c] + cnt[3][c];
    base[c] = tot ? atomicAdd(&lens[c == 0 ? 0 : c == 1 ? 1 : 11], tot) : 0u;
  }
  __syncthreads();
  if (cls) {
    const uint32_t c = cls - 1u;
    uint32_t pos = base[c];
    for (uint32_t w = 0; w < wid; w++) pos += cnt[w][c];
    pos += (uint32_t)__popcll((c == 0 ? m1 : c == 1 ? m2 : m3) & pbf_lanemask_lt());
    (c == 0 ? wlist : c == 1 ? blist : olist)[pos] = i;
  }
  __syncthreads();  // cnt and base are reused by the next call
}

// Short lists: classified here.  Long lists: listed in `big` with their chunk counts.
__global__ __launch_bounds__(256) void k_classify(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ redg,
                                                  const uint32_t* __restrict__ inpref, const uint2* __restrict__ sinfo,
                                                  int64_t n, ClassOut co, uint32_t* __restrict__ big,
                                                  uint32_t* __restrict__ big_chunks, unsigned long long* __restrict__ w64) {
  __shared__ unsigned long long s_T[4][64];
  __shared__ uint32_t s_cnt[4][4], s_base[4];
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // every thread of a workgroup runs the same steps (the appends are per workgroup)
  for (int64_t b0 = (int64_t)blockIdx.x * 256; b0 < n; b0 += (int64_t)gridDim.x * 256) {
    const int64_t base = b0 + wid * 64;
    const int64_t i = base + lane;
    uint32_t io = 0, deg = 0;
    if (i < n) { io = roff[i]; deg = roff[i + 1] - io; }
    const bool small = deg <= 64;
    // T of the short lists, summed by the whole wavefront over their concatenated slots (at most
    // 4096): lane l reads slot q = q0 + l of the concatenation, whose owner k is found among the
    // lanes' starts, so the reads are coalesced and every lane stays busy
    const uint32_t d = (i < n && small) ? deg : 0u;
    const uint32_t cinc = wave_incl_add(d);
    const uint32_t C = lane_read(cinc, 63), c = cinc - d;
    s_T[wid][lane] = 0;
    wave_lds_fence();
    for (uint32_t q0 = 0; q0 < C; q0 += 64) {
      const uint32_t q = q0 + lane;
      uint32_t k = 0;
#pragma unroll
      for (uint32_t bit = 32; bit; bit >>= 1) {  // largest k with c_k <= q
        const uint32_t cand = k + bit;
        const uint32_t cc = (uint32_t)__builtin_amdgcn_ds_bpermute((int)((cand & 63u) << 2), (int)c);
        if (cand < 64u && cc <= q) k = cand;
      }
      const uint32_t ck = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(k << 2), (int)c);
      const uint32_t iok = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(k << 2), (int)io);
      if (q < C) atomicAdd(&s_T[wid][k], (unsigned long long)seg_len(sinfo, iok + (q - ck)));
    }
    wave_lds_fence();
    const uint64_t T = s_T[wid][lane];
    classify_append((uint32_t)i, i < n && small, T, deg, co, s_cnt, s_base);
    const bool big_item = i < n && !small && i >= (int64_t)co.rlo && i < (int64_t)co.rhi;
    const uint64_t mb = __ballot(big_item);
    uint32_t bb = 0;
    if (lane == 0 && mb) bb = atomicAdd(&co.lens[2], (uint32_t)__popcll(mb));
    bb = __shfl(bb, 0, 64);
    if (big_item) {
      const uint32_t slot = bb + __popcll(mb & pbf_lanemask_lt());
      big[slot] = (uint32_t)i;
      big_chunks[slot] = (deg + CL_CHUNK - 1) / CL_CHUNK;
      w64[i] = 0;
    }
  }
}

// One wavefront per chunk of a long list (chunk c of big vertex b, located by a binary search
// over the scanned chunk counts); partial wedge counts are added into w64.
__global__ __launch_bounds__(256) void k_wedge_chunks(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ redg,
                                                      const uint32_t* __restrict__ inpref, const uint2* __restrict__ sinfo,
                                                      const uint32_t* __restrict__ big, const uint32_t* __restrict__ chunk_base,
                                                      const uint32_t* __restrict__ lens, const uint32_t* __restrict__ total_chunks,
                                                      unsigned long long* __restrict__ w64) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t nbig = lens[2], tot = *total_chunks;
  const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t item = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); item < tot; item += nwaves) {
    uint32_t lo = 0, hi = nbig;  // last b with chunk_base[b] <= item
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (chunk_base[mid] <= item) lo = mid; else hi = mid;
    }
    const uint32_t i = big[lo], c = item - chunk_base[lo];
    const uint32_t io = roff[i], deg = roff[i + 1] - io;
    const uint32_t b0 = c * CL_CHUNK, b1 = min(deg, b0 + CL_CHUNK);
    uint64_t part = 0;
    uint32_t j = b0 + lane;
    for (; j + 192 < b1; j += 256)
      part += (uint64_t)seg_len(sinfo, io + j) + seg_len(sinfo, io + j + 64) + seg_len(sinfo, io + j + 128) +
              seg_len(sinfo, io + j + 192);
    for (; j < b1; j += 64) part += seg_len(sinfo, io + j);
    part = pbf_wave_sum(part);
    if (lane == 0 && part) atomicAdd(&w64[i], part);
  }
}

// Endpoints with more than SPLIT_DEG neighbours and i < SPLIT_IMAX (few keys per chunk of
// neighbours: latency-bound in one workgroup) are split
// across workgroups (k_split_*); the other long ones are classified as usual.
constexpr uint32_t SPLIT_DEG = 1u << 12, SPLIT_CHUNK = 1u << 11, SPLIT_IMAX = 2048;

__global__ __launch_bounds__(256) void k_classify_big(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ big,
                                                      const uint32_t* __restrict__ lens_in,
                                                      const unsigned long long* __restrict__ w64, ClassOut co,
                                                      uint32_t* __restrict__ split, uint32_t* __restrict__ split_chunks,
                                                      uint32_t split_cap) {
  __shared__ uint32_t s_cnt[4][4], s_base[4];
  const uint32_t nbig = lens_in[2];
  const uint32_t lane = threadIdx.x & 63;
  // every thread of a workgroup runs the same iterations (the appends are per workgroup)
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < nbig; b0 += stride) {
    const uint32_t t = b0 + threadIdx.x;
    const bool valid = t < nbig;
    const uint32_t i = valid ? big[t] : 0u;
    const uint32_t deg = valid ? roff[i + 1] - roff[i] : 0u;
    const unsigned long long T = valid ? w64[i] : 0ull;
    const uint32_t il = (valid && co.lo) ? i - co.lo[i] : i;
    const bool cand = valid && T && il > 0 && deg > SPLIT_DEG && deg < WIDE_DEG && il < SPLIT_IMAX;
    const uint64_t mc = __ballot(cand);
    uint32_t b = 0;
    if (lane == 0 && mc) b = atomicAdd(&co.lens[8], (uint32_t)__popcll(mc));
    b = __shfl(b, 0, 64);
    const uint32_t slot = b + __popcll(mc & pbf_lanemask_lt());
    // past the capacity (several graphs of a union, each with its own low ranks) an endpoint
    // is classified as an ordinary heavy one; the consumers clamp the count to split_cap
    const bool sp = cand && slot < split_cap;
    classify_append(i, valid && !sp, T, deg, co, s_cnt, s_base);
    if (sp) {
      split[slot] = i;
      split_chunks[slot] = (deg + SPLIT_CHUNK - 1) / SPLIT_CHUNK;
    }
  }
}

// ---------------- grouping of the light endpoints ----------------
// Light endpoints are sorted by (side, table size descending) with a counting sort over 16
// buckets: bucket = side * 8 + (11 - log2 cap) for caps 2048 .. 16, side 1 when a VERT endpoint
// lies on O (its groups run pass 2 instead of the sweep).  Within a side the tables then tile
// consecutive LT_WORDS-word groups: grouping counts every table as max(cap, 64) words (so a
// group holds at most 32 endpoints, one record per lane of its first half), and since these
// sizes are powers of two in decreasing order, the prefix of every endpoint is a multiple of its
// own size and a group starts wherever the prefix is a multiple of LT_WORDS.  Inside a group
// the tables are laid out by their real sizes (light_group), which only packs them tighter.
// plan (u32): [0, 16) bucket counts, [16, 32) bucket starts, [32, 34) first group of each side;
// wbase (u64): word prefix of each bucket within its side.
constexpr uint32_t LB_SIDE = 8, LBUCKETS = 2 * LB_SIDE;
constexpr uint32_t PLAN_START = LBUCKETS, PLAN_GROUP0 = 2 * LBUCKETS;
// grouping words of a table of bucket r (2^(11 - r) slots): never below 64
__device__ __forceinline__ uint32_t group_words(uint32_t r) { return 2048u >> (r < 5u ? r : 5u); }

template <int MODE>
__device__ __forceinline__ uint32_t light_bucket(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ redg,
                                                 const uint32_t* __restrict__ bound, uint32_t i) {
  const uint32_t side = (MODE == MODE_VERT && (redg[roff[i]] & 1u)) ? 1u : 0u;
  return side * LB_SIDE + (11u - (31u - (uint32_t)__clz(light_cap(bound[i]))));
}

// Two-level counting sort over LG_BLOCKS workgroups (the same items visit the same workgroup in
// the histogram and the scatter): per-workgroup LDS histograms, one scan per bucket over the
// workgroups, then a scatter whose positions come from LDS counters seeded with the
// workgroup's offsets.  Order within a bucket is arbitrary; counts do not depend on it.
constexpr uint32_t LG_BLOCKS = 256;

template <int MODE>
__global__ __launch_bounds__(256) void k_light_hist(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ redg,
                                                    const uint32_t* __restrict__ bound, const uint32_t* __restrict__ wl,
                                                    const uint32_t* __restrict__ lens, uint32_t* __restrict__ bh) {
  __shared__ uint32_t h[LBUCKETS];
  if (threadIdx.x < LBUCKETS) h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t nl = lens[0];
  for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < nl; t += LG_BLOCKS * 256)
    atomicAdd(&h[light_bucket<MODE>(roff, redg, bound, wl[t])], 1u);
  __syncthreads();
  if (threadIdx.x < LBUCKETS) bh[blockIdx.x * LBUCKETS + threadIdx.x] = h[threadIdx.x];
}

// One workgroup of LG_BLOCKS threads: bucket totals, starts and word bases, the groups of each
// side, and every workgroup's starting position per bucket (bo).
__global__ __launch_bounds__(LG_BLOCKS) void k_light_plan(const uint32_t* __restrict__ bh, uint32_t* __restrict__ bo,
                                                          uint32_t* __restrict__ plan, unsigned long long* __restrict__ wbase,
                                                          uint32_t* __restrict__ ctl, uint32_t* __restrict__ gstart) {
  __shared__ uint32_t wsum[LG_BLOCKS / 64];
  __shared__ uint32_t tot[LBUCKETS];
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t excl[LBUCKETS];
  for (uint32_t q = 0; q < LBUCKETS; q++) {
    const uint32_t x = bh[t * LBUCKETS + q];
    const uint32_t inc = wave_incl_add(x);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t k = 0; k < LG_BLOCKS / 64; k++) {
      if (k < w) before += wsum[k];
      all += wsum[k];
    }
    excl[q] = before + inc - x;
    if (t == 0) tot[q] = all;
    __syncthreads();
  }
  if (t == 0) {
    uint32_t start = 0, groups = 0;
    for (uint32_t side = 0; side < 2; side++) {
      unsigned long long words = 0;
      for (uint32_t r = 0; r < LB_SIDE; r++) {
        const uint32_t b = side * LB_SIDE + r;
        plan[b] = tot[b];
        plan[PLAN_START + b] = start;
        start += tot[b];
        wbase[b] = words;
        words += (unsigned long long)tot[b] * group_words(r);
      }
      plan[PLAN_GROUP0 + side] = groups;
      groups += (uint32_t)((words + LT_WORDS - 1) / LT_WORDS);
      if (side == 0) ctl[10] = groups;
    }
    ctl[5] = groups;
    gstart[groups] = start;
  }
  __syncthreads();
  for (uint32_t q = 0; q < LBUCKETS; q++) bo[t * LBUCKETS + q] = plan[PLAN_START + q] + excl[q];
}

template <int MODE>
__global__ __launch_bounds__(256) void k_light_scatter(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ redg,
                                                       const uint32_t* __restrict__ bound, const uint32_t* __restrict__ wl,
                                                       const uint32_t* __restrict__ lens, const uint32_t* __restrict__ bo,
                                                       const uint32_t* __restrict__ plan,
                                                       const unsigned long long* __restrict__ wbase, uint32_t* __restrict__ out,
                                                       uint4* __restrict__ rec, uint32_t* __restrict__ gstart) {
  __shared__ uint32_t c[LBUCKETS];
  if (threadIdx.x < LBUCKETS) c[threadIdx.x] = bo[blockIdx.x * LBUCKETS + threadIdx.x];
  __syncthreads();
  const uint32_t nl = lens[0];
  for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < nl; t += LG_BLOCKS * 256) {
    const uint32_t i = wl[t];
    const uint32_t b = light_bucket<MODE>(roff, redg, bound, i);
    const uint32_t pos = atomicAdd(&c[b], 1u);
    out[pos] = i;
    const uint32_t io = roff[i];
    rec[pos] = make_uint4(i, io, roff[i + 1] - io, 11u - b % LB_SIDE);
    const unsigned long long prefix =
        wbase[b] + (unsigned long long)(pos - plan[PLAN_START + b]) * group_words(b % LB_SIDE);
    if ((prefix & (LT_WORDS - 1)) == 0) gstart[plan[PLAN_GROUP0 + b / LB_SIDE] + (uint32_t)(prefix / LT_WORDS)] = pos;
  }
}

// Split endpoints, pass 1: one workgroup per (endpoint, slice of SPLIT_CHUNK neighbours) counts
// its wedges in a dense LDS table and adds the non-zero counts into the endpoint's dense HBM
// array G (row of SPLIT_IMAX words).
template <int MODE>
__global__ __launch_bounds__(TPB) void k_split_pass1(CountArgs a, const uint32_t* __restrict__ split,
                                                     const uint32_t* __restrict__ chunk_base, const uint32_t* __restrict__ ctl,
                                                     uint32_t* __restrict__ G, uint32_t split_cap) {
  __shared__ __align__(16) uint32_t pool[POOL_WORDS];
  __shared__ __align__(16) uint32_t s_flags[NWAVE][WIN_SPAN];
  __shared__ unsigned long long s_acc[NWAVE][64];
  __shared__ uint32_t s_ovf;
  const uint32_t nsplit = min(ctl[8], split_cap), tot = ctl[9];
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  *(lds_u32x4*)((lds_u32*)s_flags[wid] + 4 * lane) = u32x4{0u, 0u, 0u, 0u};
  s_acc[wid][lane] = 0;
  if (threadIdx.x == 0) s_ovf = 0;
  WaveCtx wc{(lds_u32*)s_flags[wid], (lds_u64*)s_acc[wid], 0};
  const Tab t{(lds_u32*)pool, 0, 0, a.cb, nullptr, 0};
  const Track tr{nullptr, (lds_u32*)&s_ovf, 0xFFFFFFFFu, false};
  for (uint32_t item = blockIdx.x; item < tot; item += gridDim.x) {
    uint32_t lo = 0, hi = nsplit;
    while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1; if (chunk_base[mid] <= item) lo = mid; else hi = mid; }
    const uint32_t i = split[lo], c = item - chunk_base[lo];
    const uint32_t io = a.roff[i], deg = a.roff[i + 1] - io;
    const uint32_t glo = lo_of(a, i), il = i - glo;
    for (uint32_t q = threadIdx.x; q < il; q += TPB) pool[q] = 0u;
    __syncthreads();
    const uint32_t j1 = min(deg, (c + 1) * SPLIT_CHUNK);
    enum_pass<MODE, false, TM_D32>(a, t, i, io, glo, c * SPLIT_CHUNK, j1, wid, NWAVE, wc, tr);
    __syncthreads();
    uint32_t* row = G + (size_t)lo * SPLIT_IMAX;
    for (uint32_t q = threadIdx.x; q < il; q += TPB)
      if (pool[q]) atomicAdd(&row[q], pool[q]);
    __syncthreads();
  }
}

// Split endpoints: endpoint credits (VERT, endpoint on S) or the total, one workgroup each.
template <int MODE>
__global__ __launch_bounds__(TPB) void k_split_credit(CountArgs a, const uint32_t* __restrict__ split,
                                                      const uint32_t* __restrict__ ctl, const uint32_t* __restrict__ G,
                                                      uint32_t split_cap) {
  __shared__ unsigned long long red[NWAVE];
  const uint32_t nsplit = min(ctl[8], split_cap);
  HotLog hl;
  for (uint32_t b = blockIdx.x; b < nsplit; b += gridDim.x) {
    const uint32_t i = split[b];
    const bool i_in_S = (a.redg[a.roff[i]] & 1u) == 0u;
    if (!(MODE == MODE_TOTAL || (MODE == MODE_VERT && i_in_S))) continue;
    const uint32_t* row = G + (size_t)b * SPLIT_IMAX;
    const uint32_t glo = lo_of(a, i), il = i - glo;
    unsigned long long sum = 0;
    // every lane of a wavefront runs the same iterations (the credits are wavefront-wide)
    for (uint32_t qb = threadIdx.x & ~63u; qb < il; qb += TPB) {
      const uint32_t q = qb + (threadIdx.x & 63u);
      const uint32_t c = q < il ? row[q] : 0u;
      const unsigned long long v = c > 1u ? (unsigned long long)c * (c - 1u) / 2u : 0ull;
      sum += v;
      if constexpr (MODE == MODE_VERT) credit_wave(a, hl, c > 1u, q + glo, v);
    }
    sum = pbf_wave_sum(sum);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long tot = 0;
      for (int q = 0; q < NWAVE; q++) tot += red[q];
      if (tot) {
        if constexpr (MODE == MODE_VERT) credit(a, i, tot);
        else atomicAdd(a.total, tot);
      }
    }
    __syncthreads();
  }
  if constexpr (MODE == MODE_VERT) hlog_pad(a, hl);
}

// Split endpoints, pass 2 (VERT with the endpoint on O, or EDGE): counts looked up in G.
template <int MODE>
__global__ __launch_bounds__(TPB) void k_split_pass2(CountArgs a, const uint32_t* __restrict__ split,
                                                     const uint32_t* __restrict__ chunk_base, const uint32_t* __restrict__ ctl,
                                                     const uint32_t* __restrict__ G, uint32_t split_cap) {
  __shared__ __align__(16) uint32_t s_flags[NWAVE][WIN_SPAN];
  __shared__ unsigned long long s_acc[NWAVE][64];
  __shared__ uint32_t s_ovf;
  const uint32_t nsplit = min(ctl[8], split_cap), tot = ctl[9];
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  *(lds_u32x4*)((lds_u32*)s_flags[wid] + 4 * lane) = u32x4{0u, 0u, 0u, 0u};
  s_acc[wid][lane] = 0;
  if (threadIdx.x == 0) s_ovf = 0;
  __syncthreads();
  WaveCtx wc{(lds_u32*)s_flags[wid], (lds_u64*)s_acc[wid], 0};
  const Track tr{nullptr, (lds_u32*)&s_ovf, 0xFFFFFFFFu, false};
  for (uint32_t item = blockIdx.x; item < tot; item += gridDim.x) {
    uint32_t lo = 0, hi = nsplit;
    while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1; if (chunk_base[mid] <= item) lo = mid; else hi = mid; }
    const uint32_t i = split[lo], c = item - chunk_base[lo];
    const uint32_t io = a.roff[i], deg = a.roff[i + 1] - io;
    const bool i_in_S = (a.redg[io] & 1u) == 0u;
    if (!(MODE == MODE_EDGE || (MODE == MODE_VERT && !i_in_S))) continue;
    const Tab t{nullptr, 0, 0, a.cb, G + (size_t)lo * SPLIT_IMAX, 0};
    const uint32_t j1 = min(deg, (c + 1) * SPLIT_CHUNK);
    enum_pass<MODE, true, TM_G32>(a, t, i, io, lo_of(a, i), c * SPLIT_CHUNK, j1, wid, NWAVE, wc, tr);
  }
  if constexpr (MODE == MODE_VERT) hlog_pad(a, wc.hl);
}

// O-CSR index of every ranked slot: the payload is an O-CSR position when the slot's owner
// is on O (flag 1: neighbour on S), else an S-CSR position mapped through eti
// (CountERank's conversion, butterfly_ecount.h:1723-1741).
// Per-edge output from the per-slot counts: every edge has one slot owned by its S end (flag 0)
// and one owned by its O end (flag 1); the first pass stores, the second adds, each writing
// every edge exactly once (no atomics).
// A slot's count is its 64-bit counter plus its 32-bit one.
__global__ void k_edge_out_s(const uint32_t* __restrict__ redg, const uint32_t* __restrict__ oidx,
                             const unsigned long long* __restrict__ Bs, const uint32_t* __restrict__ B32, int64_t m2,
                             unsigned long long* __restrict__ out) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m2; p += (int64_t)gridDim.x * blockDim.x)
    if ((redg[p] & 1u) == 0u) out[oidx[p]] = Bs[p] + B32[p];
}
__global__ void k_edge_out_o(const uint32_t* __restrict__ redg, const uint32_t* __restrict__ oidx,
                             const unsigned long long* __restrict__ Bs, const uint32_t* __restrict__ B32, int64_t m2,
                             unsigned long long* __restrict__ out) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m2; p += (int64_t)gridDim.x * blockDim.x)
    if (redg[p] & 1u) out[oidx[p]] += Bs[p] + B32[p];
}

__global__ void k_oidx(const uint32_t* __restrict__ redg, const uint32_t* __restrict__ rpos, const uint32_t* __restrict__ eti,
                       int64_t m2, uint32_t* __restrict__ oidx) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m2; p += (int64_t)gridDim.x * blockDim.x)
    oidx[p] = (redg[p] & 1u) ? rpos[p] : eti[rpos[p]];
}

// For every slot (r -> y): roff[y] and the segment length of the wedges (r, y, u2), i.e. the
// number of entries of N(y) ranked below min(r, y): in(y) when y < r, else the position of r
// inside N(y) (a binary search over the in-prefix of N(y)).  The counting kernels read both
// beside the slot, which keeps two dependent loads and the search off their critical path.
// rinfo[r] = (roff[r], in(r)): one 8-byte gather per slot instead of two 4-byte ones.
// With maxdeg (cleared beforehand), also the largest degree.
__global__ __launch_bounds__(256) void k_rinfo(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ inpref, int64_t n,
                        uint2* __restrict__ rinfo, uint32_t* __restrict__ maxdeg) {
  uint32_t md = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t o = roff[r];
    rinfo[r] = make_uint2(o, inpref[r]);
    md = max(md, roff[r + 1] - o);
  }
  if (maxdeg) {
    // one atomic per workgroup: atomics of every wavefront on one address serialise
    __shared__ uint32_t part[4];
    md = wave_incl_max(md);
    if (__lane_id() == 63u) part[threadIdx.x >> 6] = md;
    __syncthreads();
    if (threadIdx.x == 0) {
      md = max(max(part[0], part[1]), max(part[2], part[3]));
      if (md) atomicMax(maxdeg, md);
    }
  }
}

__global__ void k_slotinfo(const uint2* __restrict__ rinfo, const uint32_t* __restrict__ redg,
                           const uint32_t* __restrict__ rown, const uint32_t* __restrict__ lo, int64_t m2,
                           uint2* __restrict__ sinfo) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m2; p += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t rg = rown[p], l = lo ? lo[rg] : 0u;
    const uint32_t r = rg - l, y = redg[p] >> 1;  // both local to the graph of rg
    const uint2 yi = rinfo[y + l];
    const uint32_t yo = yi.x, ip = yi.y;
    sinfo[p] = make_uint2(yo, y < r ? ip : rank_lower_bound(redg + yo, ip < r ? ip : r, r));
  }
}

// CountRank's conversion (butterfly_count.h:1600-1610): out[x] = B[rank(x)] for x in S.
__global__ void k_convert_vert(const unsigned long long* __restrict__ B, const uint32_t* __restrict__ rank_of, int64_t base,
                               int64_t ns, unsigned long long* __restrict__ out) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < ns; x += (int64_t)gridDim.x * blockDim.x)
    out[x] = B[rank_of[base + x]];
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) p[x] = v;
}
__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, int64_t n) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) p[x] = v;
}

// Upper bound on the number of vertices with more than d of the m2 list entries.
inline int64_t g_m2_over(int64_t m2, uint32_t d) { return m2 / (int64_t)d + 1; }

inline unsigned grid_for(int64_t n, int threads, int max_blocks) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)b;
}

constexpr int HBM_TEAM = 256, HBM_BLOCKS = 32, HBM_BLOCKS_NOSYNC = 4;

// Wedges of every endpoint (the sum of its slots' segment lengths), a wavefront per endpoint.
__global__ __launch_bounds__(256) void k_endpoint_wedges(const uint32_t* __restrict__ roff, const uint2* __restrict__ sinfo,
                                                         int64_t n, unsigned long long* __restrict__ T) {
  const uint32_t lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += nw) {
    const uint32_t b = roff[i], e = roff[i + 1];
    unsigned long long t = 0;
    for (uint32_t p = b + lane; p < e; p += 64) t += sinfo[p].y;
    t = pbf_wave_sum(t);
    if (lane == 0) T[i] = t;
  }
}

// Rank ranges of at most max_wedges wedges each, consecutive and greedy (getNextWedgeIdx,
// butterfly_utils.h:1700-1714, 1948-1959: the binary search over the wedge prefix for the
// first endpoint that no longer fits).  An endpoint with more wedges than max_wedges on its own
// is a capacity error (the reference exits: "Space must accomodate seagulls originating from
// one vertex").
int wedge_batches(pbf_context* ctx, const RankedGraph& R, const uint2* sinfo, int64_t max_wedges,
                  std::vector<uint32_t>* bounds) {
  const int64_t n = R.n;
  void* tp;
  PBF_CHECK(pbf_scratch(ctx, S_WBATCH, 16 * (size_t)(n + 1), &tp));
  unsigned long long* T = static_cast<unsigned long long*>(tp);
  unsigned long long* P = T + n + 1;
  hipLaunchKernelGGL(k_endpoint_wedges, dim3(grid_for(n, 4, 8192)), dim3(256), 0, ctx->stream, R.roff, sinfo, n, T);
  PBF_CHECK(pbf_scan_excl_u64(ctx, reinterpret_cast<const uint64_t*>(T), reinterpret_cast<uint64_t*>(P), n,
                              reinterpret_cast<uint64_t*>(P + n)));
  std::vector<uint64_t> pre((size_t)n + 1);
  PBF_CHECK(pbf_d2h(ctx, pre.data(), P, 8 * (size_t)(n + 1)));
  bounds->assign(1, 0u);
  int64_t lo = 0;
  while (lo < n) {
    // largest hi with pre[hi] - pre[lo] <= max_wedges
    const uint64_t lim = pre[(size_t)lo] + (uint64_t)max_wedges;
    const int64_t hi = (int64_t)(std::upper_bound(pre.begin() + lo, pre.end(), lim) - pre.begin()) - 1;
    if (hi <= lo) return PBF_ERR_CAPACITY;
    bounds->push_back((uint32_t)hi);
    lo = hi;
  }
  return PBF_OK;
}

// Runs the counting phase: classification, the fused kernel, the hub fold and (only if some
// endpoint needs it) the HBM fallback.  lens/queues are cleared with a fill kernel.  With a
// wedge bound (ctx->max_wedges) the endpoints are counted rank range by rank range.
template <int MODE>
int run_count(pbf_context* ctx, const RankedGraph& R, CountArgs a0, uint32_t* lists, uint32_t* ctl, bool sync) {
  const int64_t n = R.n;
  // ctl: [0] light len, [1] heavy len, [2] big len, [3] heavy queue, [4] light-group queue,
  //      [5] light groups, [6] re-run len, [7] total big chunks, [8] split len,
  //      [9] total split items, [10] light groups on side S, [11] HBM-fallback len,
  //      [13] HBM queue, [14] zero, [16], [17] re-run queues
  // lists: [0, n) light (classification order), [n, 2n) heavy, [2n, 3n) re-run,
  //        [3n, 4n) light grouped, [4n, 5n + 1) group starts, [5n + 1, 6n + 1) HBM fallback
  void *bp2, *up, *pl;
  PBF_CHECK(pbf_scratch(ctx, S_BIG, 4 * (size_t)(3 * n + 64), &bp2));
  PBF_CHECK(pbf_scratch(ctx, S_UB, 8 * (size_t)(n ? n : 1), &up));
  uint32_t* big = static_cast<uint32_t*>(bp2);
  uint32_t* big_chunks = big + n;
  uint32_t* chunk_base = big + 2 * n;
  unsigned long long* w64 = static_cast<unsigned long long*>(up);
  {
    void* rip;
    PBF_CHECK(pbf_scratch(ctx, S_RINFO, 8 * (size_t)(n ? n : 1), &rip));
    uint2* rinfo = static_cast<uint2*>(rip);
    hipLaunchKernelGGL(k_rinfo, dim3(grid_for(n, 256, 16384)), dim3(256), 0, ctx->stream, R.roff, R.inpref, n, rinfo,
                       const_cast<uint32_t*>(a0.maxdeg));
    hipLaunchKernelGGL(k_slotinfo, dim3(grid_for(R.m2, 256, 16384)), dim3(256), 0, ctx->stream, rinfo, R.redg, R.rown, R.lo,
                       R.m2, const_cast<uint2*>(a0.sinfo));
  }
  PBF_CHECK(pbf_scratch(ctx, S_LPLAN, 4 * (64 + 2 * LG_BLOCKS * LBUCKETS) + 8 * LBUCKETS, &pl));
  uint32_t* plan = static_cast<uint32_t*>(pl);
  uint32_t* bh = plan + 64;
  void* lrp;
  PBF_CHECK(pbf_scratch(ctx, S_LREC, 16 * (size_t)(n ? n : 1), &lrp));
  uint4* lrec = static_cast<uint4*>(lrp);
  uint32_t* bo = bh + LG_BLOCKS * LBUCKETS;
  unsigned long long* wbase = reinterpret_cast<unsigned long long*>(bo + LG_BLOCKS * LBUCKETS);
  uint32_t* hbm_list = lists + 5 * n + 1;
  // split endpoints: list, chunk counts, chunk bases, dense rows
  void *sp, *gp;
  // a single graph has at most SPLIT_IMAX split endpoints (distinct ranks below it); a union
  // with graph-local ranks may have more, and the capacity then bounds them
  const int64_t max_split = std::min<int64_t>(R.lo ? 4 * SPLIT_IMAX : SPLIT_IMAX, 2 * g_m2_over(R.m2, SPLIT_DEG) + 8);
  PBF_CHECK(pbf_scratch(ctx, S_SPLIT, 4 * (size_t)(3 * max_split + 64), &sp));
  PBF_CHECK(pbf_scratch(ctx, S_GROWS, 4 * (size_t)max_split * SPLIT_IMAX, &gp));
  uint32_t* split = static_cast<uint32_t*>(sp);
  uint32_t* split_chunks = split + max_split;
  uint32_t* split_base = split + 2 * max_split;
  uint32_t* G = static_cast<uint32_t*>(gp);
  // light-table bound: the context's, or the mode's default.  Per-vertex counting: 767 keys
  // (1024-slot sub-tables; 13.2 ms against 14.6 ms at 1535 on the Chung-Lu 2^20 x 2^18 graph).
  // Per-edge counting: 1535 (2048-slot sub-tables), which its second pass over every endpoint
  // favours on the R-MAT 2^22 x 2^22 graph (82 ms against 94 ms at 767) more than the Chung-Lu
  // graph favours 767 (18.0 against 19.2 ms).
  const uint32_t want = ctx->tune.light_max ? ctx->tune.light_max : (MODE == MODE_EDGE ? LIGHT_MAX : 767u);
  const uint32_t lmax = want < LIGHT_MAX ? want : LIGHT_MAX;
  const unsigned grid = (unsigned)(ctx->num_cus * (ctx->tune.blocks_per_cu ? ctx->tune.blocks_per_cu : 4));
  std::vector<uint32_t> bounds{0u, (uint32_t)n};
  if (ctx->max_wedges > 0 && n > 0) PBF_CHECK(wedge_batches(ctx, R, a0.sinfo, ctx->max_wedges, &bounds));
  ctx->timing.batches = (int64_t)bounds.size() - 1;
  ctx->timing.hub_vertices = sync ? 0 : -1;
  for (size_t bi = 0; bi + 1 < bounds.size(); bi++) {
    CountArgs a = a0;
    a.hlog_len = ctl + 20;  // cleared with ctl below
    const ClassOut co{const_cast<uint32_t*>(a.bound), lists, lists + n, hbm_list, ctl, a.cb, lmax, ctx->tune.light_work,
                      R.lo, bounds[bi], bounds[bi + 1]};
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctl, 0u, (int64_t)64);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, plan, 0u, (int64_t)64);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, ctx->stream, big_chunks, 0u, n);
    hipLaunchKernelGGL(k_classify, dim3(grid_for(n, 256, 8192)), dim3(256), 0, ctx->stream, R.roff, R.redg, R.inpref, a.sinfo,
                       n, co, big, big_chunks, w64);
    PBF_CHECK(pbf_scan_excl_u32(ctx, big_chunks, chunk_base, n, ctl + 7));
    hipLaunchKernelGGL(k_wedge_chunks, dim3(ctx->num_cus * 8), dim3(256), 0, ctx->stream, R.roff, R.redg, R.inpref, a.sinfo,
                       big, chunk_base, ctl, ctl + 7, w64);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(max_split, 256, 64)), dim3(256), 0, ctx->stream, split_chunks, 0u, max_split);
    hipLaunchKernelGGL(k_classify_big, dim3(64), dim3(256), 0, ctx->stream, R.roff, big, ctl, w64, co, split, split_chunks,
                       (uint32_t)max_split);
    PBF_CHECK(pbf_scan_excl_u32(ctx, split_chunks, split_base, max_split, ctl + 9));
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(max_split * SPLIT_IMAX, 256, 4096)), dim3(256), 0, ctx->stream, G, 0u,
                       max_split * (int64_t)SPLIT_IMAX);
    hipLaunchKernelGGL(k_split_pass1<MODE>, dim3(ctx->num_cus * 2), dim3(TPB), 0, ctx->stream, a, split, split_base, ctl, G,
                       (uint32_t)max_split);
    hipLaunchKernelGGL(k_split_credit<MODE>, dim3(256), dim3(TPB), 0, ctx->stream, a, split, ctl, G, (uint32_t)max_split);
    if (MODE != MODE_TOTAL)
      hipLaunchKernelGGL(k_split_pass2<MODE>, dim3(ctx->num_cus * 2), dim3(TPB), 0, ctx->stream, a, split, split_base, ctl, G,
                         (uint32_t)max_split);
    // light endpoints: sort into groups
    hipLaunchKernelGGL(k_light_hist<MODE>, dim3(LG_BLOCKS), dim3(256), 0, ctx->stream, R.roff, R.redg, a.bound, lists, ctl,
                       bh);
    hipLaunchKernelGGL(k_light_plan, dim3(1), dim3(LG_BLOCKS), 0, ctx->stream, bh, bo, plan, wbase, ctl, lists + 4 * n);
    hipLaunchKernelGGL(k_light_scatter<MODE>, dim3(LG_BLOCKS), dim3(256), 0, ctx->stream, R.roff, R.redg, a.bound, lists, ctl,
                       bo, plan, wbase, lists + 3 * n, lrec, lists + 4 * n);
    a.wlist = lists + 3 * n;
    a.gstart = lists + 4 * n;
    a.lrec = lrec;
    a.blist = lists + n;
    a.nblock = ctl + 1;
    a.ngroups = ctl + 5;
    a.ngroups0 = ctl + 10;
    a.queues = ctl + 3;
    a.ovf_list = hbm_list;
    a.ovf_len = ctl + 11;
    a.rerun_list = lists + 2 * n;
    a.rerun_len = ctl + 6;
    a.rerun = 0;
    hipLaunchKernelGGL((k_count2<MODE, 1>), dim3(grid), dim3(TPB), 0, ctx->stream, a);
    hipLaunchKernelGGL((k_count2<MODE, 2>), dim3(grid), dim3(TPB), 0, ctx->stream, a);
    // light endpoints whose optimistic table overflowed: again, one workgroup each, full tables
    a.blist = lists + 2 * n;
    a.nblock = ctl + 6;
    a.ngroups = ctl + 14;
    a.ngroups0 = ctl + 14;
    a.queues = ctl + 16;
    a.rerun = 1;
    hipLaunchKernelGGL((k_count2<MODE, 1>), dim3(grid), dim3(TPB), 0, ctx->stream, a);
    // the log's reducer: four workgroups per CU (32 KB of accumulators each)
    if (MODE == MODE_VERT && a.hlog_cap) {
      const uint32_t rows = (uint32_t)ctx->num_cus * 4;
      unsigned long long* part = a.hlog + a.hlog_cap;  // rows x HOT_N partial sums after the log
      hipLaunchKernelGGL(k_hlog_reduce, dim3(rows), dim3(256), 0, ctx->stream, a.hlog, a.hlog_len, a.hlog_cap, part);
      hipLaunchKernelGGL(k_hlog_sum, dim3(HOT_N / 256, (rows + HLOG_ROWS_PER - 1) / HLOG_ROWS_PER), dim3(256), 0, ctx->stream,
                         part, rows, a.B, n);
    }
    PBF_HIP(hipGetLastError());
    uint32_t novf = 0;
    // each block needs 2 kn words of dense counters: with small key ranges (a union of small
    // graphs) a 256 MB pool holds many more blocks
    const size_t per_block = sizeof(uint32_t) * 2 * (size_t)a.kn;
    const int64_t fit = std::min<int64_t>((int64_t)ctx->num_cus * 4, ((int64_t)256 << 20) / (int64_t)per_block);
    int blocks = (int)std::max<int64_t>(HBM_BLOCKS_NOSYNC, fit);
    if (sync) {
      PBF_HIP(hipMemcpyAsync(ctx->pinned + 16, ctl + 11, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
      PBF_HIP(hipStreamSynchronize(ctx->stream));
      novf = (uint32_t)(ctx->pinned[16] & 0xFFFFFFFFu);
      ctx->timing.hub_vertices += novf;
      const int64_t most = std::max<int64_t>(HBM_BLOCKS, fit);
      blocks = (int)std::min<int64_t>((int64_t)novf, most);
      // a fixed scratch region holds as many blocks' counters as fit (the work queue serves any
      // number of blocks); not even one is a capacity error
      const size_t room = pbf_scratch_room(ctx);
      if (novf && room != SIZE_MAX) {
        const size_t fit_room = room > 512 ? (room - 512) / per_block : 0;
        if (fit_room == 0) return PBF_ERR_CAPACITY;
        blocks = (int)std::min<size_t>((size_t)blocks, fit_room);
      }
    }
    // an array bound (-a) caps the histograms: as many blocks as it holds, at least one
    if (ctx->max_array > 0)
      blocks = (int)std::max<int64_t>(1, std::min<int64_t>(blocks, ctx->max_array / (2 * std::max<int64_t>(a.kn, 1))));
    if (novf || !sync) {
      // endpoints no LDS table could hold: dense HBM histograms (every counter and used-list
      // entry is restored to zero after its endpoint, so the pool is cleared only when first
      // allocated)
      void* pp;
      const size_t pool_bytes = per_block * blocks;
      PBF_CHECK(pbf_scratch(ctx, S_POOL, pool_bytes, &pp));
      if (ctx->pool_clean < pool_bytes) {
        PBF_CHECK(pbf_fill_u32(ctx, static_cast<uint32_t*>(pp), 0u, (int64_t)(pool_bytes / 4)));
        ctx->pool_clean = pool_bytes;
      }
      a.pool = static_cast<uint32_t*>(pp);
      a.list = hbm_list; a.list_len = ctl + 11; a.queue = ctl + 13;
      hipLaunchKernelGGL((k_count_hbm<MODE, HBM_TEAM>), dim3(blocks), dim3(HBM_TEAM), 0, ctx->stream, a);
      PBF_HIP(hipGetLastError());
    }
  }
  if (MODE != MODE_EDGE && MODE != MODE_TOTAL)
    hipLaunchKernelGGL(k_hot_reduce, dim3(grid_for(HOT_N, 256, 64)), dim3(256), 0, ctx->stream, a0.hot, a0.B, n);
  return PBF_OK;
}
}  // namespace

int pbf_dev_count(pbf_context* ctx, const DevGraph& g, int use_v, const RankedGraph& R, int mode, uint64_t* out_dev,
                  const uint32_t* eti, const CountOpts& opts) {
  const int64_t n = R.n, m2 = R.m2;
  void *bp, *lp, *cp, *mp, *hp, *op = nullptr;
  PBF_CHECK(pbf_scratch(ctx, S_BOUND, 4 * (size_t)(n ? n : 1), &bp));
  PBF_CHECK(pbf_scratch(ctx, S_BINS, 4 * (size_t)(6 * n + 128), &lp));
  PBF_CHECK(pbf_scratch(ctx, S_MISC, 64 * sizeof(uint64_t), &mp));
  PBF_CHECK(pbf_scratch(ctx, S_HOT, 8 * (size_t)HOT_N * HOT_R, &hp));
  void* lpp;
  PBF_CHECK(pbf_scratch(ctx, S_LPOS, 8 * (size_t)(m2 ? m2 : 1), &lpp));
  uint32_t* lists = static_cast<uint32_t*>(lp);
  uint32_t* ctl = lists + 6 * n + 64;
  CountArgs a{};
  a.roff = R.roff; a.redg = R.redg; a.inpref = R.inpref; a.bound = static_cast<uint32_t*>(bp);
  a.sinfo = static_cast<uint2*>(lpp);
  a.n = n;
  a.total = static_cast<unsigned long long*>(mp) + 12;
  a.hot = static_cast<unsigned long long*>(hp);
  a.lo = R.lo;
  {
    // key bits: keys (local ranks) are below key_range <= 2^kb - 1, so a packed key is never
    // all ones
    const int64_t key_range = (R.lo && opts.key_range > 0) ? opts.key_range : n;
    a.kn = key_range > 0 ? key_range : 1;
    uint32_t kb = 0;
    while (kb < 32 && ((uint64_t)1 << kb) - 1 < (uint64_t)key_range) kb++;
    a.cb = 32 - kb;
  }
  if (!ctx->hot_clear) {
    // the accumulators are left zeroed by every fold; clear them once per context
    hipLaunchKernelGGL(k_fill_u64, dim3(64), dim3(256), 0, ctx->stream, a.hot, 0ull, (int64_t)HOT_N * HOT_R);
    ctx->hot_clear = true;
  }
  if (mode == MODE_VERT) {
    PBF_CHECK(pbf_scratch(ctx, S_COUNTS, 8 * (size_t)(n ? n : 1), &cp));
    a.B = static_cast<unsigned long long*>(cp);
    // the hub-credit log: up to 16 entries per slot (at most 2^27), none in a fixed scratch
    // region (there the hub credits stay replicated atomics)
    if (pbf_scratch_room(ctx) == SIZE_MAX && n > 0) {
      const int64_t cap = std::min<int64_t>(int64_t(1) << 27, 16 * m2 + (1 << 16));
      void* hl;
      // the log, then the reducer's rows of partial sums (one per workgroup, HOT_N each)
      PBF_CHECK(pbf_scratch(ctx, S_HLOG, 8 * ((size_t)cap + (size_t)ctx->num_cus * 4 * HOT_N), &hl));
      a.hlog = static_cast<unsigned long long*>(hl);
      a.hlog_cap = (uint32_t)cap;
    }
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256, 8192)), dim3(256), 0, ctx->stream, a.B, 0ull, n);
    PBF_CHECK(run_count<MODE_VERT>(ctx, R, a, lists, ctl, opts.sync));
    if (opts.B_out) *opts.B_out = a.B;
    const int64_t ns = use_v ? g.nu : g.nv, base = use_v ? g.nv : 0;
    if (ns && !opts.rank_space)
      hipLaunchKernelGGL(k_convert_vert, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, ctx->stream, a.B, R.rank_of, base, ns,
                         reinterpret_cast<unsigned long long*>(out_dev));
  } else if (mode == MODE_EDGE) {
    PBF_CHECK(pbf_scratch(ctx, S_OIDX, 4 * (size_t)(m2 ? m2 : 1), &op));
    uint32_t* oidx = static_cast<uint32_t*>(op);
    if (m2) hipLaunchKernelGGL(k_oidx, dim3(grid_for(m2, 256, 16384)), dim3(256), 0, ctx->stream, R.redg, R.rpos, eti, m2, oidx);
    PBF_CHECK(pbf_scratch(ctx, S_COUNTS, 8 * (size_t)(m2 ? m2 : 1), &cp));
    a.B = static_cast<unsigned long long*>(cp);
    void* b32;
    PBF_CHECK(pbf_scratch(ctx, S_B32, 4 * (size_t)(m2 ? m2 : 1), &b32));
    a.B32 = static_cast<uint32_t*>(b32);
    uint32_t* md = static_cast<uint32_t*>(mp) + 28;
    a.maxdeg = md;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, md, 0u, (int64_t)1);
    if (m2) hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(m2, 256, 8192)), dim3(256), 0, ctx->stream, a.B, 0ull, m2);
    if (m2) hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(m2, 256, 8192)), dim3(256), 0, ctx->stream, a.B32, 0u, m2);
    PBF_CHECK(run_count<MODE_EDGE>(ctx, R, a, lists, ctl, opts.sync));
    if (m2) {
      unsigned long long* out = reinterpret_cast<unsigned long long*>(out_dev);
      hipLaunchKernelGGL(k_edge_out_s, dim3(grid_for(m2, 256, 16384)), dim3(256), 0, ctx->stream, R.redg, oidx, a.B, a.B32, m2,
                         out);
      hipLaunchKernelGGL(k_edge_out_o, dim3(grid_for(m2, 256, 16384)), dim3(256), 0, ctx->stream, R.redg, oidx, a.B, a.B32, m2,
                         out);
    }
  } else {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, a.total, 0ull, (int64_t)1);
    PBF_CHECK(run_count<MODE_TOTAL>(ctx, R, a, lists, ctl, opts.sync));
    PBF_HIP(hipMemcpyAsync(out_dev, a.total, 8, hipMemcpyDeviceToDevice, ctx->stream));
  }
  PBF_HIP(hipGetLastError());
  return PBF_OK;
}
