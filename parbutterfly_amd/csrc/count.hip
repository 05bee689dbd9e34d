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
// GPU mapping.  A "team" (one wavefront, or one 256-thread workgroup for heavy endpoints)
// owns one endpoint i at a time, taken from a device work queue.  The team enumerates i's
// wedges flattened: for a chunk of TEAM neighbours v, each thread computes its segment length
// L_v (in(v) when v < i, else the position of i inside N(v) by binary search), a team scan of
// L gives a dense wedge index space, and each thread takes wedges w, w+TEAM, ... locating its
// segment by binary search over the scanned lengths (so consecutive threads read consecutive
// entries of the same N(v): coalesced loads, full lanes whatever the degrees).  Wedges are
// aggregated by u2 in an open-addressing hash table in LDS (atomic CAS on the key, atomic add
// on the count).  Heavy endpoints whose distinct-u2 bound exceeds the LDS table are moved to a
// second kernel that uses a dense per-workgroup histogram in HBM.  Counts are scattered with
// 64-bit atomics; integer atomic adds commute, so results are exact and deterministic.
#include "pbf_device.h"
#include "team.h"

using namespace pbf;

namespace {
enum TableKind { T_LDS = 0, T_HBM = 1 };

struct CountArgs {
  const uint32_t* roff;
  const uint32_t* redg;
  const uint32_t* inpref;
  const uint32_t* lpos;    // per slot (i -> v) with v > i: position of i inside N(v)
  const uint32_t* bound;   // distinct-u2 bound per endpoint
  const uint32_t* oidx;    // EDGE: O-CSR index of each ranked slot
  unsigned long long* B;   // VERT: counts in rank space; EDGE: counts in O-CSR order
  unsigned long long* total;
  const uint32_t* list;
  const uint32_t* list_len;
  uint32_t* queue;         // work counter
  uint32_t* ovf_list;
  uint32_t* ovf_len;
  uint32_t* pool;          // T_HBM: per-block dense counters [n] followed by used list [n]
  int64_t n;
  // fused kernel (k_count2): two work lists and their queues
  const uint32_t* blist;   // endpoints processed by a whole workgroup
  const uint32_t* wlist;   // endpoints processed by one wavefront
  const uint32_t* lens;    // [0] wlist length, [1] blist length
  uint32_t* queues;        // [0] block queue, [1] wave queue
  unsigned long long* hot; // replicated accumulators for ranks < HOT_N (HOT_R replicas each)
  uint32_t cb;             // count bits of a packed 32-bit hash slot (32 - key bits)
};

// Core: process the endpoints of one list with one team per endpoint.
//   KIND == T_LDS: LDS hash table of capacity CAP (USED: keep a used-slot list so that the
//                  table is cleared lazily; otherwise clear the first `cap` slots per vertex).
//   KIND == T_HBM: dense counters in HBM (pool slice per block), used list in HBM.
template <int MODE, int TEAM, int CAP, bool USED, int KIND>
__global__ __launch_bounds__(TEAM) void k_count(CountArgs a) {
  __shared__ TeamShared<TEAM> sh;
  __shared__ uint32_t lkeys[KIND == T_LDS ? CAP : 1];
  __shared__ uint32_t lcnts[KIND == T_LDS ? CAP : 1];
  __shared__ uint32_t lused[(KIND == T_LDS && USED) ? CAP : 1];
  const uint32_t tid = threadIdx.x;
  const uint32_t nlist = *a.list_len;

  uint32_t* dense = nullptr;
  uint32_t* dused = nullptr;
  if constexpr (KIND == T_HBM) {
    dense = a.pool + (size_t)blockIdx.x * (size_t)a.n * 2;
    dused = dense + a.n;
  }
  if constexpr (KIND == T_LDS && USED) {
    for (uint32_t s = tid; s < CAP; s += TEAM) { lkeys[s] = EMPTY; lcnts[s] = 0; }
  }
  unsigned long long block_total = 0;

  while (true) {
    if (tid == 0) sh.item = atomicAdd(a.queue, 1u);
    team_sync<TEAM>();
    const uint32_t idx = sh.item;
    team_sync<TEAM>();
    if (idx >= nlist) break;
    const uint32_t i = a.list[idx];
    const uint32_t io = a.roff[i], deg = a.roff[i + 1] - io;
    const bool i_in_S = (a.redg[io] & 1u) == 0u;  // neighbours of i carry flag 1 iff i is on O

    LdsTable t;
    if constexpr (KIND == T_LDS) {
      uint32_t bnd = a.bound[i];
      uint32_t cap = CAP;
      if constexpr (!USED) {
        uint64_t want = (uint64_t)bnd + bnd / 3 + 1;
        cap = 256;
        while (cap < want && cap < CAP) cap <<= 1;
      }
      uint32_t lg = 31 - __clz(cap);
      t.keys = lkeys; t.cnts = lcnts; t.used = USED ? lused : nullptr;
      t.shift = 32 - lg; t.mask = cap - 1;
      // Only a full-size table can overflow: a smaller one was sized from a bound on the
      // number of distinct keys.  After the flag is raised every thread completes at most its
      // current insert, so occupancy stays below maxfill + TEAM < CAP and probing ends.
      t.maxfill = (cap == CAP) ? (CAP / 4) * 3 - TEAM : 0xFFFFFFFFu;
      if constexpr (!USED) {
        for (uint32_t s = tid; s < cap; s += TEAM) { lkeys[s] = EMPTY; lcnts[s] = 0; }
      }
    }
    if (tid == 0) { sh.nused = 0; sh.ovf = 0; }
    team_sync<TEAM>();

    const bool need_pass2 = (MODE == MODE_EDGE) || (MODE == MODE_VERT && !i_in_S);
    for (int pass = 1; pass <= (need_pass2 ? 2 : 1); pass++) {
      if (pass == 2) {
        // endpoint groups are complete: abandon on overflow (never reached when bound fits)
        if (sh.ovf) break;
      }
      for (uint32_t c0 = 0; c0 < deg; c0 += TEAM) {
        const uint32_t j = c0 + tid;
        uint32_t L = 0, vo = 0, v = 0;
        if (j < deg) {
          v = a.redg[io + j] >> 1;
          vo = a.roff[v];
          const uint32_t ip = a.inpref[v];
          L = (v < i) ? ip : rank_lower_bound(a.redg + vo, ip, i);
        }
        uint32_t T;
        const uint32_t incl = team_incl_scan<TEAM>(L, sh, &T);
        sh.P[tid] = incl;
        sh.vs[tid] = vo;
        sh.vv[tid] = v;
        if (pass == 2) sh.acc[tid] = 0;
        team_sync<TEAM>();
        for (uint32_t w = tid; w < T; w += TEAM) {
          const uint32_t s = seg_of<TEAM>(sh.P, w);
          const uint32_t k = w - (s ? sh.P[s - 1] : 0u);
          const uint32_t p = sh.vs[s] + k;
          const uint32_t u2 = a.redg[p] >> 1;
          if (pass == 1) {
            if constexpr (KIND == T_LDS) {
              if (!sh.ovf) lds_insert<TEAM>(t, sh, u2);
            } else {
              if (atomicAdd(&dense[u2], 1u) == 0u) dused[atomicAdd(&sh.nused, 1u)] = u2;
            }
          } else {
            uint32_t c;
            if constexpr (KIND == T_LDS) c = lds_lookup(t, u2);
            else c = dense[u2];
            if (c > 1u) {
              atomicAdd(&sh.acc[s], (unsigned long long)(c - 1u));
              if constexpr (MODE == MODE_EDGE) atomicAdd(&a.B[a.oidx[p]], (unsigned long long)(c - 1u));
            }
          }
        }
        team_sync<TEAM>();
        if (pass == 2 && j < deg) {
          const unsigned long long s = sh.acc[tid];
          if (s) {
            if constexpr (MODE == MODE_EDGE) atomicAdd(&a.B[a.oidx[io + j]], s);
            else atomicAdd(&a.B[v], s);
          }
        }
        team_sync<TEAM>();
      }
      if (pass == 1) {
        team_sync<TEAM>();
        if (sh.ovf) break;
        if (MODE == MODE_TOTAL || (MODE == MODE_VERT && i_in_S)) {
          // endpoint credit: C(c,2) per group
          unsigned long long sum = 0;
          const uint32_t extent = (KIND == T_HBM || USED) ? sh.nused : (t.mask + 1);
          for (uint32_t q = tid; q < extent; q += TEAM) {
            uint32_t key, c;
            if constexpr (KIND == T_LDS) {
              const uint32_t slot = USED ? lused[q] : q;
              key = lkeys[slot];
              c = (key == EMPTY) ? 0u : lcnts[slot];
            } else {
              key = dused[q];
              c = dense[key];
            }
            if (c > 1u) {
              const unsigned long long b = (unsigned long long)c * (c - 1u) / 2u;
              sum += b;
              if constexpr (MODE == MODE_VERT) atomicAdd(&a.B[key], b);
            }
          }
          sum = team_sum<TEAM>(sum, sh);
          if (tid == 0 && sum) {
            if constexpr (MODE == MODE_VERT) atomicAdd(&a.B[i], sum);
            else block_total += sum;
          }
        }
      }
    }
    team_sync<TEAM>();
    const bool overflowed = sh.ovf != 0;
    if (overflowed && tid == 0) a.ovf_list[atomicAdd(a.ovf_len, 1u)] = i;
    // restore the table for the next endpoint
    if constexpr (KIND == T_LDS && USED) {
      const uint32_t nu_ = sh.nused < CAP ? sh.nused : CAP;
      for (uint32_t q = tid; q < nu_; q += TEAM) { const uint32_t slot = lused[q]; lkeys[slot] = EMPTY; lcnts[slot] = 0; }
    }
    if constexpr (KIND == T_HBM) {
      for (uint32_t q = tid; q < sh.nused; q += TEAM) dense[dused[q]] = 0;
    }
    team_sync<TEAM>();
  }
  if constexpr (MODE == MODE_TOTAL) {
    if (tid == 0 && block_total) atomicAdd(a.total, block_total);
  }
}

// ------------------------------------------------------------------------------------------
// Fused counting kernel (one launch): a workgroup queue of heavy endpoints, then a wavefront
// queue of light ones.
//
// Per endpoint i the wedges are enumerated segment by segment (segment = N(v)[0, L_v) for a
// neighbour v).  Segments of at least LONG_SEG entries are streamed by a whole wavefront
// (64 consecutive entries per step: coalesced, no per-wedge search); shorter segments are
// walked by the lane that owns them.  The wedge key u2 is aggregated in LDS:
//   * dense u32 counters indexed by u2 when i < D32_MAX (u2 < i),
//   * dense u16 counters (two per word) when i < D16_MAX and deg(i) < 65536 (so c < 65536),
//   * otherwise an open-addressing hash of 64-bit slots (u2 << 32 | count).
// Light endpoints (bound <= WAVE_BOUND distinct keys, bounded work) get a private 1024-slot
// hash per wavefront; heavy ones share a workgroup table.  A workgroup hash that fills up
// sends its endpoint to the HBM dense-histogram kernel.  Credits to ranks < HOT_N (the hubs,
// which receive most credits) go to HOT_R replicated accumulators to spread atomic
// contention; k_hot_reduce folds them into the counts.
// ------------------------------------------------------------------------------------------
constexpr int NWAVE = 4;
constexpr int TPB = NWAVE * 64;
constexpr uint32_t POOL_WORDS = 8192;          // 32 KB of LDS per workgroup
constexpr uint32_t D32_MAX = POOL_WORDS;       // dense u32: i < 8192
constexpr uint32_t D16_MAX = 2 * POOL_WORDS;   // dense u16: i < 16384
constexpr uint32_t HB_CAP = POOL_WORDS / 2;    // workgroup hash: 4096 64-bit slots
constexpr uint32_t HW_CAP = POOL_WORDS / 2 / NWAVE;  // wavefront hash: 1024 64-bit slots
constexpr uint32_t WAVE_BOUND = (HW_CAP / 4) * 3 - 64;  // light endpoint: <= 704 distinct keys
constexpr uint32_t WAVE_WORK = 1u << 14;             // ... and at most 16384 wedges (bound)
constexpr uint32_t LONG_SEG = 32;
constexpr uint32_t HOT_N = 4096, HOT_R = 16;
constexpr unsigned long long EMPTY64 = 0xFFFFFFFF00000000ull;

enum TabMode { TM_D32 = 0, TM_D16 = 1, TM_HASH = 2, TM_HASH32 = 3, TM_G32 = 4 };

// TM_HASH32 packs a slot into 32 bits, key << cb | count, usable when every count of the
// endpoint fits in cb bits (c <= deg(i) < 2^cb) and every key fits in 32 - cb bits; the all-ones
// word is the empty slot.  TM_HASH uses 64-bit slots, key << 32 | count.
// LDS pointers carry address space 3 so that every access compiles to ds_* instructions
// (through a generic pointer the compiler emits flat accesses, which also wait for all
// outstanding global loads).
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) unsigned long long lds_u64;

__device__ __forceinline__ uint32_t lds_load(lds_u32* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ unsigned long long lds_load(lds_u64* p) {
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

struct Tab {
  lds_u32* w;                    // pool words
  lds_u64* s;                    // same memory as 64-bit slots
  uint32_t mode, shift, mask, maxfill, cb;
  lds_u32* nused;
  lds_u32* ovf;
  const uint32_t* g;             // TM_G32: dense counts in HBM (lookups only)
};

__device__ __forceinline__ void tab_insert(const Tab& t, uint32_t u2) {
  if (t.mode == TM_D32) {
    lds_add(t.w + u2, 1u);
  } else if (t.mode == TM_D16) {
    lds_add(t.w + (u2 >> 1), 1u << ((u2 & 1u) << 4));
  } else if (t.mode == TM_HASH32) {
    uint32_t h = hash_slot(u2, t.shift);
    const uint32_t mine = (u2 << t.cb) | 1u;
    for (uint32_t probe = 0; probe <= t.mask; probe++) {
      const uint32_t cur = lds_load(t.w + h);
      if (cur != EMPTY && (cur >> t.cb) == u2) { lds_add(t.w + h, 1u); return; }
      if (cur == EMPTY) {
        const uint32_t old = lds_cas(t.w + h, EMPTY, mine);
        if (old == EMPTY) {
          if (lds_add(t.nused, 1u) >= t.maxfill) lds_store(t.ovf, 1u);
          return;
        }
        if ((old >> t.cb) == u2) { lds_add(t.w + h, 1u); return; }
      }
      h = (h + 1) & t.mask;
    }
    lds_store(t.ovf, 1u);
  } else {
    uint32_t h = hash_slot(u2, t.shift);
    const unsigned long long mine = ((unsigned long long)u2 << 32) | 1ull;
    for (uint32_t probe = 0; probe <= t.mask; probe++) {
      const unsigned long long cur = lds_load(t.s + h);
      if ((uint32_t)(cur >> 32) == u2) { lds_add(t.s + h, 1ull); return; }
      if (cur == EMPTY64) {
        const unsigned long long old = lds_cas(t.s + h, EMPTY64, mine);
        if (old == EMPTY64) {
          if (lds_add(t.nused, 1u) >= t.maxfill) lds_store(t.ovf, 1u);
          return;
        }
        if ((uint32_t)(old >> 32) == u2) { lds_add(t.s + h, 1ull); return; }
      }
      h = (h + 1) & t.mask;
    }
    lds_store(t.ovf, 1u);
  }
}

__device__ __forceinline__ uint32_t tab_lookup(const Tab& t, uint32_t u2) {
  if (t.mode == TM_D32) return t.w[u2];
  if (t.mode == TM_G32) return t.g[u2];
  if (t.mode == TM_D16) return (t.w[u2 >> 1] >> ((u2 & 1u) << 4)) & 0xFFFFu;
  uint32_t h = hash_slot(u2, t.shift);
  if (t.mode == TM_HASH32) {
    for (uint32_t probe = 0; probe <= t.mask; probe++) {
      const uint32_t cur = t.w[h];
      if (cur == EMPTY) return 0u;
      if ((cur >> t.cb) == u2) return cur & ((1u << t.cb) - 1u);
      h = (h + 1) & t.mask;
    }
    return 0u;
  }
  for (uint32_t probe = 0; probe <= t.mask; probe++) {
    const unsigned long long cur = t.s[h];
    if ((uint32_t)(cur >> 32) == u2) return (uint32_t)cur;
    if (cur == EMPTY64) return 0u;
    h = (h + 1) & t.mask;
  }
  return 0u;
}

// Number of table cells and the (key, count) of cell q.
__device__ __forceinline__ uint32_t tab_cells(const Tab& t, uint32_t i) {
  return (t.mode == TM_HASH || t.mode == TM_HASH32) ? t.mask + 1 : i;
}
__device__ __forceinline__ uint32_t tab_cell(const Tab& t, uint32_t q, uint32_t* key) {
  if (t.mode == TM_D32) { *key = q; return t.w[q]; }
  if (t.mode == TM_D16) { *key = q; return (t.w[q >> 1] >> ((q & 1u) << 4)) & 0xFFFFu; }
  if (t.mode == TM_HASH32) {
    const uint32_t cur = t.w[q];
    *key = cur >> t.cb;
    return cur == EMPTY ? 0u : (cur & ((1u << t.cb) - 1u));
  }
  const unsigned long long cur = t.s[q];
  *key = (uint32_t)(cur >> 32);
  return cur == EMPTY64 ? 0u : (uint32_t)cur;
}

// Credit to rank x: hub ranks go to replicated accumulators.
__device__ __forceinline__ void credit(const CountArgs& a, uint32_t x, unsigned long long val) {
  if (x < HOT_N) atomicAdd(&a.hot[x * HOT_R + (blockIdx.x & (HOT_R - 1))], val);
  else atomicAdd(&a.B[x], val);
}

// Enumerates the wedges of endpoint i: the neighbours are taken 64 at a time; their segment
// lengths L_v are scanned into a flattened wedge range [0, T), and the wavefronts of the group
// share that range (wave w of nw takes indices w*64 + lane, then every 64*nw).  Each index is
// mapped to its segment by a binary search over the scanned lengths held in lanes (six
// shuffles), so any mix of short and long segments keeps every lane busy, and U indices are
// resolved before their loads are issued so that U loads per lane are in flight.  PASS2 looks
// the counts up and credits centres (VERT) or edges (EDGE): per-segment sums are reduced over
// the contiguous lanes of a segment and accumulated in the wave's LDS slots `acc`.
constexpr int UNROLL = 4;

template <int MODE, bool PASS2>
__device__ __forceinline__ void enum_wedges(const CountArgs& a, const Tab& t, uint32_t i, uint32_t io, uint32_t jbeg,
                                            uint32_t deg, uint32_t w, uint32_t nw, lds_u64* acc) {
  const uint32_t lane = threadIdx.x & 63;
  // high-degree endpoints: the waves take different neighbour chunks; otherwise they share the
  // wedge range of each chunk
  const bool split_chunks = nw > 1 && deg - jbeg >= 8 * 64;
  const uint32_t cstart = split_chunks ? w * 64 : 0u, cstep = split_chunks ? nw * 64 : 64u;
  const uint32_t wstart = split_chunks ? 0u : w * 64, wstep = split_chunks ? 64u : nw * 64;
  for (uint32_t c0 = jbeg + cstart; c0 < deg; c0 += cstep) {
    const uint32_t j = c0 + lane;
    uint32_t v = 0, vo = 0, L = 0;
    if (j < deg) {
      v = a.redg[io + j] >> 1;
      vo = a.roff[v];
      const uint32_t ip = a.inpref[v];
      // at most i entries of N(v) rank below i
      // the prefix of N(v) below min(i, v): all of in(v) when v < i, else up to i's own slot
      L = (v < i) ? ip : a.lpos[io + j];
    }
    const uint32_t incl = pbf_wave_incl_scan(L);
    const uint32_t T = __shfl(incl, 63, 64);
    if (T == 0) continue;
    const uint32_t P = incl - L;  // first flattened index of this lane's segment
    if (PASS2) {
      acc[lane] = 0;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
    }
    for (uint32_t b0 = wstart; b0 < T; b0 += wstep * UNROLL) {
      uint32_t ent[UNROLL], seg[UNROLL], pos[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; u++) {
        const uint32_t widx = b0 + (uint32_t)u * wstep + lane;
        // largest s with P_s <= widx (zero-length segments share a start with the next one,
        // and the larger index wins)
        uint32_t s = 0;
#pragma unroll
        for (uint32_t bit = 32; bit; bit >>= 1) {
          const uint32_t cand = s + bit;
          const uint32_t pc = __shfl(P, (int)(cand & 63u), 64);
          if (cand < 64u && pc <= widx) s = cand;
        }
        const uint32_t ps = __shfl(P, (int)s, 64), vs = __shfl(vo, (int)s, 64);
        pos[u] = vs + (widx - ps);
        seg[u] = widx < T ? s : 64u;
        ent[u] = widx < T ? a.redg[pos[u]] : 0u;
      }
#pragma unroll
      for (int u = 0; u < UNROLL; u++) {
        if (!PASS2) {
          if (seg[u] < 64u && !lds_load(t.ovf)) tab_insert(t, ent[u] >> 1);
        } else {
          unsigned long long val = 0;
          if (seg[u] < 64u) {
            const uint32_t c = tab_lookup(t, ent[u] >> 1);
            if (c > 1u) {
              val = c - 1u;
              if constexpr (MODE == MODE_EDGE) atomicAdd(&a.B[a.oidx[pos[u]]], val);
            }
          }
          // segmented inclusive scan over the contiguous lanes of each segment
          const uint32_t sg = seg[u];
#pragma unroll
          for (int d = 1; d < 64; d <<= 1) {
            const unsigned long long y = __shfl_up(val, d, 64);
            const uint32_t sy = __shfl_up(sg, d, 64);
            if ((int)lane >= d && sy == sg) val += y;
          }
          const uint32_t snext = __shfl_down(sg, 1, 64);
          const bool tail = lane == 63 || snext != sg;
          if (tail && sg < 64u && val) lds_add(acc + sg, val);
        }
      }
    }
    if (PASS2) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      const unsigned long long mine = acc[lane];
      if (j < deg && mine) {
        if constexpr (MODE == MODE_EDGE) atomicAdd(&a.B[a.oidx[io + j]], mine);
        else credit(a, v, mine);
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
}

template <bool BLOCK>
__device__ __forceinline__ void group_sync() {
  if constexpr (BLOCK) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
}

// Processes endpoint i with a group of waves (the workgroup, or one wavefront) on table t.
// Returns false when the table overflowed (nothing was credited).
template <int MODE, bool BLOCK>
__device__ bool process_endpoint(const CountArgs& a, Tab& t, uint32_t i, uint32_t bound, unsigned long long* red,
                                 unsigned long long* total_acc, lds_u64* acc) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t w = BLOCK ? (threadIdx.x >> 6) : 0u, nw = BLOCK ? NWAVE : 1u;
  const uint32_t gt = BLOCK ? threadIdx.x : lane, gn = BLOCK ? TPB : 64u;
  const uint32_t io = a.roff[i], deg = a.roff[i + 1] - io;
  const bool i_in_S = (a.redg[io] & 1u) == 0u;
  // choose and clear the table
  uint32_t words;
  if (BLOCK && i < D32_MAX) { t.mode = TM_D32; words = i; }
  else if (BLOCK && i < D16_MAX && deg < 65536u) { t.mode = TM_D16; words = (i + 1) >> 1; }
  else {
    // packed 32-bit slots when the counts (<= deg) and the keys (< n) both fit
    const bool packed = t.cb < 32u && deg < (1u << t.cb);
    const uint32_t capmax = (BLOCK ? POOL_WORDS : POOL_WORDS / NWAVE) / (packed ? 1u : 2u);
    const uint64_t want = (uint64_t)bound + bound / 3 + 1;
    uint32_t cap = 64;
    while (cap < want && cap < capmax) cap <<= 1;
    t.mode = packed ? TM_HASH32 : TM_HASH;
    t.mask = cap - 1;
    t.shift = 32 - (31 - __clz(cap));
    t.maxfill = (cap == capmax) ? (capmax / 4) * 3 - gn : 0xFFFFFFFFu;
    words = packed ? cap : 2 * cap;
  }
  if (t.mode == TM_HASH) {
    for (uint32_t q = gt; q <= t.mask; q += gn) t.s[q] = EMPTY64;
  } else if (t.mode == TM_HASH32) {
    for (uint32_t q = gt; q <= t.mask; q += gn) t.w[q] = EMPTY;
  } else {
    for (uint32_t q = gt; q < words; q += gn) t.w[q] = 0u;
  }
  if (gt == 0) { t.nused[0] = 0; t.ovf[0] = 0; }
  group_sync<BLOCK>();
  enum_wedges<MODE, false>(a, t, i, io, 0, deg, w, nw, acc);
  group_sync<BLOCK>();
  if (lds_load(t.ovf)) return false;
  if (MODE == MODE_TOTAL || (MODE == MODE_VERT && i_in_S)) {
    unsigned long long sum = 0;
    const uint32_t cells = tab_cells(t, i);
    for (uint32_t q = gt; q < cells; q += gn) {
      uint32_t key;
      const uint32_t c = tab_cell(t, q, &key);
      if (c > 1u) {
        const unsigned long long b = (unsigned long long)c * (c - 1u) / 2u;
        sum += b;
        if constexpr (MODE == MODE_VERT) credit(a, key, b);
      }
    }
    sum = pbf_wave_sum(sum);
    if constexpr (BLOCK) {
      if (lane == 0) red[w] = sum;
      __syncthreads();
      if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int q = 0; q < NWAVE; q++) tot += red[q];
        if (tot) {
          if constexpr (MODE == MODE_VERT) credit(a, i, tot);
          else *total_acc += tot;
        }
      }
    } else {
      if (lane == 0 && sum) {
        if constexpr (MODE == MODE_VERT) credit(a, i, sum);
        else *total_acc += sum;
      }
    }
  }
  if (MODE == MODE_EDGE || (MODE == MODE_VERT && !i_in_S)) enum_wedges<MODE, true>(a, t, i, io, 0, deg, w, nw, acc);
  group_sync<BLOCK>();
  return true;
}

template <int MODE>
__global__ __launch_bounds__(TPB) void k_count2(CountArgs a) {
  __shared__ __align__(16) uint32_t pool[POOL_WORDS];
  __shared__ uint32_t s_item, s_nused[NWAVE + 1], s_ovf[NWAVE + 1];
  __shared__ unsigned long long s_red[NWAVE];
  __shared__ unsigned long long s_acc[NWAVE][64];
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned long long total_acc = 0;
  Tab t;
  t.cb = a.cb;
  // phase 1: heavy endpoints, whole workgroup
  const uint32_t nb = a.lens[1];
  t.w = (lds_u32*)pool;
  t.s = (lds_u64*)pool;
  t.nused = (lds_u32*)&s_nused[NWAVE];
  t.ovf = (lds_u32*)&s_ovf[NWAVE];
  while (true) {
    if (threadIdx.x == 0) s_item = atomicAdd(&a.queues[0], 1u);
    __syncthreads();
    const uint32_t idx = s_item;
    __syncthreads();
    if (idx >= nb) break;
    const uint32_t i = a.blist[idx];
    if (!process_endpoint<MODE, true>(a, t, i, a.bound[i], s_red, &total_acc, (lds_u64*)s_acc[wid])) {
      if (threadIdx.x == 0) a.ovf_list[atomicAdd(a.ovf_len, 1u)] = i;
    }
  }
  // phase 2: light endpoints, one per wavefront, private quarter of the pool
  const uint32_t nwl = a.lens[0];
  t.w = (lds_u32*)(pool + wid * (POOL_WORDS / NWAVE));
  t.s = (lds_u64*)(pool + wid * (POOL_WORDS / NWAVE));
  t.nused = (lds_u32*)&s_nused[wid];
  t.ovf = (lds_u32*)&s_ovf[wid];
  // the next item is claimed before the current one is processed, hiding the queue latency
  uint32_t next = 0;
  if (lane == 0) next = atomicAdd(&a.queues[1], 1u);
  next = __shfl(next, 0, 64);
  while (next < nwl) {
    const uint32_t idx = next;
    if (lane == 0) next = atomicAdd(&a.queues[1], 1u);
    const uint32_t i = a.wlist[idx];
    if (!process_endpoint<MODE, false>(a, t, i, a.bound[i], nullptr, &total_acc, (lds_u64*)s_acc[wid])) {
      if (lane == 0) a.ovf_list[atomicAdd(a.ovf_len, 1u)] = i;
    }
    next = __shfl(next, 0, 64);
  }
  if constexpr (MODE == MODE_TOTAL) {
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

// Work classification: UB_i = sum_{v in N(i)} in(v) bounds both the wedges and (with i) the
// distinct keys of endpoint i.  Lists of at most 64 entries are summed by their own lane;
// longer lists are cut into chunks of CL_CHUNK entries summed by separate wavefronts
// (k_ub_chunks), so a hub's list does not serialise one wavefront.  Appends to the work lists
// are aggregated per wavefront (one atomic per class per wave).
constexpr uint32_t CL_CHUNK = 2048;

__device__ __forceinline__ void classify_append(uint32_t i, bool valid, uint64_t ub, uint32_t* __restrict__ bound,
                                                uint32_t* __restrict__ wlist, uint32_t* __restrict__ blist,
                                                uint32_t* __restrict__ lens) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t cls = 0;  // 0 none, 1 wavefront, 2 workgroup
  if (valid && ub) {
    const uint32_t b = (uint32_t)(ub < (uint64_t)i ? ub : (uint64_t)i);
    // a wavefront takes every endpoint of bounded work (optimistically: its table may
    // overflow, and the endpoint is then redone by a workgroup); the rest go to workgroups
    if (b) cls = ub <= WAVE_WORK ? 1u : 2u;
    bound[i] = b;
  }
  const uint64_t m1 = __ballot(cls == 1), m2 = __ballot(cls == 2);
  uint32_t base1 = 0, base2 = 0;
  if (lane == 0) {
    if (m1) base1 = atomicAdd(&lens[0], (uint32_t)__popcll(m1));
    if (m2) base2 = atomicAdd(&lens[1], (uint32_t)__popcll(m2));
  }
  base1 = __shfl(base1, 0, 64);
  base2 = __shfl(base2, 0, 64);
  const uint64_t lt = pbf_lanemask_lt();
  if (cls == 1) wlist[base1 + __popcll(m1 & lt)] = i;
  if (cls == 2) blist[base2 + __popcll(m2 & lt)] = i;
}

// Short lists: classified here.  Long lists: listed in `big` with their chunk counts.
__global__ __launch_bounds__(256) void k_classify(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ redg,
                                                  const uint32_t* __restrict__ inpref, int64_t n, uint32_t* __restrict__ bound,
                                                  uint32_t* __restrict__ wlist, uint32_t* __restrict__ blist,
                                                  uint32_t* __restrict__ lens, uint32_t* __restrict__ big,
                                                  uint32_t* __restrict__ big_chunks, unsigned long long* __restrict__ ub64) {
  const uint32_t lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64; base < n; base += nwaves * 64) {
    const int64_t i = base + lane;
    uint32_t io = 0, deg = 0;
    if (i < n) { io = roff[i]; deg = roff[i + 1] - io; }
    uint64_t ub = 0;
    const bool small = deg <= 64;
    if (small)
      for (uint32_t j = 0; j < deg; j++) ub += inpref[redg[io + j] >> 1];
    classify_append((uint32_t)i, i < n && small, ub, bound, wlist, blist, lens);
    const uint64_t mb = __ballot(!small);
    uint32_t bb = 0;
    if (lane == 0 && mb) bb = atomicAdd(&lens[2], (uint32_t)__popcll(mb));
    bb = __shfl(bb, 0, 64);
    if (!small) {
      const uint32_t slot = bb + __popcll(mb & pbf_lanemask_lt());
      big[slot] = (uint32_t)i;
      big_chunks[slot] = (deg + CL_CHUNK - 1) / CL_CHUNK;
      ub64[i] = 0;
    }
  }
}

// One wavefront per chunk of a long list (chunk c of big vertex b, located by a binary search
// over the scanned chunk counts); partial sums are added into ub64.
__global__ __launch_bounds__(256) void k_ub_chunks(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ redg,
                                                   const uint32_t* __restrict__ inpref, const uint32_t* __restrict__ big,
                                                   const uint32_t* __restrict__ chunk_base, const uint32_t* __restrict__ lens,
                                                   const uint32_t* __restrict__ total_chunks,
                                                   unsigned long long* __restrict__ ub64) {
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
      part += (uint64_t)inpref[redg[io + j] >> 1] + inpref[redg[io + j + 64] >> 1] + inpref[redg[io + j + 128] >> 1] +
              inpref[redg[io + j + 192] >> 1];
    for (; j < b1; j += 64) part += inpref[redg[io + j] >> 1];
    part = pbf_wave_sum(part);
    if (lane == 0 && part) atomicAdd(&ub64[i], part);
  }
}

// Endpoints with more than SPLIT_DEG neighbours and a dense table (i < D32_MAX) are split
// across workgroups (k_split_*); the other long ones are classified as usual.
constexpr uint32_t SPLIT_DEG = 1u << 15, SPLIT_CHUNK = 1u << 13;

__global__ __launch_bounds__(256) void k_classify_big(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ big,
                                                      const uint32_t* __restrict__ lens_in,
                                                      const unsigned long long* __restrict__ ub64, uint32_t* __restrict__ bound,
                                                      uint32_t* __restrict__ wlist, uint32_t* __restrict__ blist,
                                                      uint32_t* __restrict__ lens, uint32_t* __restrict__ split,
                                                      uint32_t* __restrict__ split_chunks) {
  const uint32_t nbig = lens_in[2];
  const uint32_t lane = threadIdx.x & 63;
  // every wavefront runs the same number of iterations (the ballots inside need the whole wave)
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < nbig; base += stride) {
    const uint32_t t = base + lane;
    const bool valid = t < nbig;
    const uint32_t i = valid ? big[t] : 0u;
    const uint32_t deg = valid ? roff[i + 1] - roff[i] : 0u;
    const unsigned long long ub = valid ? ub64[i] : 0ull;
    const bool sp = valid && ub && i > 0 && deg > SPLIT_DEG && i < D32_MAX;
    classify_append(i, valid && !sp, ub, bound, wlist, blist, lens);
    const uint64_t ms = __ballot(sp);
    uint32_t b = 0;
    if (lane == 0 && ms) b = atomicAdd(&lens[8], (uint32_t)__popcll(ms));
    b = __shfl(b, 0, 64);
    if (sp) {
      const uint32_t slot = b + __popcll(ms & pbf_lanemask_lt());
      split[slot] = i;
      split_chunks[slot] = (deg + SPLIT_CHUNK - 1) / SPLIT_CHUNK;
    }
  }
}

// Split endpoints, pass 1: one workgroup per (endpoint, slice of SPLIT_CHUNK neighbours) counts
// its wedges in a dense LDS table and adds the non-zero counts into the endpoint's dense HBM
// array G (row of D32_MAX words).
template <int MODE>
__global__ __launch_bounds__(TPB) void k_split_pass1(CountArgs a, const uint32_t* __restrict__ split,
                                                     const uint32_t* __restrict__ chunk_base, const uint32_t* __restrict__ ctl,
                                                     uint32_t* __restrict__ G) {
  __shared__ __align__(16) uint32_t pool[POOL_WORDS];
  __shared__ uint32_t s_nused, s_ovf;
  __shared__ unsigned long long s_acc[NWAVE][64];
  const uint32_t nsplit = ctl[8], tot = ctl[9];
  const uint32_t wid = threadIdx.x >> 6;
  Tab t;
  t.w = (lds_u32*)pool; t.s = nullptr; t.mode = TM_D32; t.nused = (lds_u32*)&s_nused; t.ovf = (lds_u32*)&s_ovf;
  t.g = nullptr; t.cb = a.cb;
  for (uint32_t item = blockIdx.x; item < tot; item += gridDim.x) {
    uint32_t lo = 0, hi = nsplit;
    while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1; if (chunk_base[mid] <= item) lo = mid; else hi = mid; }
    const uint32_t i = split[lo], c = item - chunk_base[lo];
    const uint32_t io = a.roff[i], deg = a.roff[i + 1] - io;
    for (uint32_t q = threadIdx.x; q < i; q += TPB) pool[q] = 0u;
    if (threadIdx.x == 0) { s_nused = 0; s_ovf = 0; }
    __syncthreads();
    const uint32_t j1 = min(deg, (c + 1) * SPLIT_CHUNK);
    enum_wedges<MODE, false>(a, t, i, io, c * SPLIT_CHUNK, j1, wid, NWAVE, (lds_u64*)s_acc[wid]);
    __syncthreads();
    uint32_t* row = G + (size_t)lo * D32_MAX;
    for (uint32_t q = threadIdx.x; q < i; q += TPB)
      if (pool[q]) atomicAdd(&row[q], pool[q]);
    __syncthreads();
  }
}

// Split endpoints: endpoint credits (VERT, endpoint on S) or the total, one workgroup each.
template <int MODE>
__global__ __launch_bounds__(TPB) void k_split_credit(CountArgs a, const uint32_t* __restrict__ split,
                                                      const uint32_t* __restrict__ ctl, const uint32_t* __restrict__ G) {
  __shared__ unsigned long long red[NWAVE];
  const uint32_t nsplit = ctl[8];
  for (uint32_t b = blockIdx.x; b < nsplit; b += gridDim.x) {
    const uint32_t i = split[b];
    const bool i_in_S = (a.redg[a.roff[i]] & 1u) == 0u;
    if (!(MODE == MODE_TOTAL || (MODE == MODE_VERT && i_in_S))) continue;
    const uint32_t* row = G + (size_t)b * D32_MAX;
    unsigned long long sum = 0;
    for (uint32_t q = threadIdx.x; q < i; q += TPB) {
      const uint32_t c = row[q];
      if (c > 1u) {
        const unsigned long long v = (unsigned long long)c * (c - 1u) / 2u;
        sum += v;
        if constexpr (MODE == MODE_VERT) credit(a, q, v);
      }
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
}

// Split endpoints, pass 2 (VERT with the endpoint on O, or EDGE): counts looked up in G.
template <int MODE>
__global__ __launch_bounds__(TPB) void k_split_pass2(CountArgs a, const uint32_t* __restrict__ split,
                                                     const uint32_t* __restrict__ chunk_base, const uint32_t* __restrict__ ctl,
                                                     const uint32_t* __restrict__ G) {
  __shared__ unsigned long long s_acc[NWAVE][64];
  __shared__ uint32_t s_ovf;
  const uint32_t nsplit = ctl[8], tot = ctl[9];
  const uint32_t wid = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_ovf = 0;
  __syncthreads();
  for (uint32_t item = blockIdx.x; item < tot; item += gridDim.x) {
    uint32_t lo = 0, hi = nsplit;
    while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1; if (chunk_base[mid] <= item) lo = mid; else hi = mid; }
    const uint32_t i = split[lo], c = item - chunk_base[lo];
    const uint32_t io = a.roff[i], deg = a.roff[i + 1] - io;
    const bool i_in_S = (a.redg[io] & 1u) == 0u;
    if (!(MODE == MODE_EDGE || (MODE == MODE_VERT && !i_in_S))) continue;
    Tab t;
    t.w = nullptr; t.s = nullptr; t.mode = TM_G32; t.g = G + (size_t)lo * D32_MAX; t.ovf = (lds_u32*)&s_ovf;
    t.nused = nullptr;
    t.cb = a.cb;
    const uint32_t j1 = min(deg, (c + 1) * SPLIT_CHUNK);
    enum_wedges<MODE, true>(a, t, i, io, c * SPLIT_CHUNK, j1, wid, NWAVE, (lds_u64*)s_acc[wid]);
  }
}

// O-CSR index of every ranked slot: the payload is an O-CSR position when the slot's owner
// is on O (flag 1: neighbour on S), else an S-CSR position mapped through eti
// (CountERank's conversion, butterfly_ecount.h:1723-1741).
__global__ void k_oidx(const uint32_t* __restrict__ redg, const uint32_t* __restrict__ rpos, const uint32_t* __restrict__ eti,
                       int64_t m2, uint32_t* __restrict__ oidx) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m2; p += (int64_t)gridDim.x * blockDim.x)
    oidx[p] = (redg[p] & 1u) ? rpos[p] : eti[rpos[p]];
}

// For every slot (r -> y) with y > r: the position of r inside N(y), i.e. the number of
// entries of N(y) ranked below r (a binary search over the in-prefix of N(y)).  The counting
// kernel reads it beside the slot, keeping the search off its dependent-load chain.
__global__ void k_lpos(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ redg,
                       const uint32_t* __restrict__ rown, const uint32_t* __restrict__ inpref, int64_t m2,
                       uint32_t* __restrict__ lpos) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m2; p += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t r = rown[p], y = redg[p] >> 1;
    if (y > r) {
      const uint32_t ip = inpref[y];
      lpos[p] = rank_lower_bound(redg + roff[y], ip < r ? ip : r, r);
    }
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

constexpr int HBM_TEAM = 256, HBM_BLOCKS = 32;

// Runs the counting phase: classification, the fused kernel, the hub fold and (only if some
// workgroup table overflowed) the HBM fallback.  lens/queues are cleared with a fill kernel.
template <int MODE>
int run_count(pbf_context* ctx, const RankedGraph& R, CountArgs a, uint32_t* lists, uint32_t* ctl) {
  const int64_t n = R.n;
  // ctl: [0] wlist len, [1] blist len, [2] big len, [3] q block, [4] q wave, [5] q ovf,
  //      [6] ovf len, [7] total big chunks, [8] split len, [9] total split items,
  //      [10], [11] queues of the overflow pass, [12] its overflow len, [13] HBM queue
  void *bp2, *up;
  PBF_CHECK(pbf_scratch(ctx, S_BIG, 4 * (size_t)(3 * n + 64), &bp2));
  PBF_CHECK(pbf_scratch(ctx, S_UB, 8 * (size_t)(n ? n : 1), &up));
  uint32_t* big = static_cast<uint32_t*>(bp2);
  uint32_t* big_chunks = big + n;
  uint32_t* chunk_base = big + 2 * n;
  unsigned long long* ub64 = static_cast<unsigned long long*>(up);
  hipLaunchKernelGGL(k_lpos, dim3(grid_for(R.m2, 256, 16384)), dim3(256), 0, ctx->stream, R.roff, R.redg, R.rown,
                     R.inpref, R.m2, const_cast<uint32_t*>(a.lpos));
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctl, 0u, (int64_t)16);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, ctx->stream, big_chunks, 0u, n);
  hipLaunchKernelGGL(k_classify, dim3(grid_for(n, 256, 8192)), dim3(256), 0, ctx->stream, R.roff, R.redg, R.inpref, n,
                     const_cast<uint32_t*>(a.bound), lists, lists + n, ctl, big, big_chunks, ub64);
  PBF_CHECK(pbf_scan_excl_u32(ctx, big_chunks, chunk_base, n, ctl + 7));
  hipLaunchKernelGGL(k_ub_chunks, dim3(ctx->num_cus * 8), dim3(256), 0, ctx->stream, R.roff, R.redg, R.inpref, big,
                     chunk_base, ctl, ctl + 7, ub64);
  // split endpoints: list, chunk counts, chunk bases, dense rows
  void *sp, *gp;
  const int64_t max_split = 2 * g_m2_over(R.m2, SPLIT_DEG) + 8;
  PBF_CHECK(pbf_scratch(ctx, S_SPLIT, 4 * (size_t)(3 * max_split + 64), &sp));
  PBF_CHECK(pbf_scratch(ctx, S_GROWS, 4 * (size_t)max_split * D32_MAX, &gp));
  uint32_t* split = static_cast<uint32_t*>(sp);
  uint32_t* split_chunks = split + max_split;
  uint32_t* split_base = split + 2 * max_split;
  uint32_t* G = static_cast<uint32_t*>(gp);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(max_split, 256, 64)), dim3(256), 0, ctx->stream, split_chunks, 0u, max_split);
  hipLaunchKernelGGL(k_classify_big, dim3(64), dim3(256), 0, ctx->stream, R.roff, big, ctl, ub64,
                     const_cast<uint32_t*>(a.bound), lists, lists + n, ctl, split, split_chunks);
  PBF_CHECK(pbf_scan_excl_u32(ctx, split_chunks, split_base, max_split, ctl + 9));
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(max_split * D32_MAX, 256, 4096)), dim3(256), 0, ctx->stream, G, 0u,
                     max_split * (int64_t)D32_MAX);
  hipLaunchKernelGGL(k_split_pass1<MODE>, dim3(ctx->num_cus * 2), dim3(TPB), 0, ctx->stream, a, split, split_base, ctl, G);
  hipLaunchKernelGGL(k_split_credit<MODE>, dim3(256), dim3(TPB), 0, ctx->stream, a, split, ctl, G);
  if (MODE != MODE_TOTAL)
    hipLaunchKernelGGL(k_split_pass2<MODE>, dim3(ctx->num_cus * 2), dim3(TPB), 0, ctx->stream, a, split, split_base, ctl, G);
  a.wlist = lists;
  a.blist = lists + n;
  a.lens = ctl;
  a.queues = ctl + 3;
  a.ovf_list = lists + 2 * n;
  a.ovf_len = ctl + 6;
  hipLaunchKernelGGL(k_count2<MODE>, dim3(ctx->num_cus * 4), dim3(TPB), 0, ctx->stream, a);
  // endpoints whose wavefront table overflowed: again, one workgroup each (lens = {ctl[5] = 0,
  // ctl[6] = overflow count})
  a.blist = lists + 2 * n;
  a.lens = ctl + 5;
  a.queues = ctl + 10;
  a.ovf_list = lists + 3 * n;
  a.ovf_len = ctl + 12;
  hipLaunchKernelGGL(k_count2<MODE>, dim3(ctx->num_cus * 4), dim3(TPB), 0, ctx->stream, a);
  PBF_HIP(hipGetLastError());
  if (MODE != MODE_EDGE && MODE != MODE_TOTAL)
    hipLaunchKernelGGL(k_hot_reduce, dim3(grid_for(HOT_N, 256, 64)), dim3(256), 0, ctx->stream, a.hot, a.B, n);
  PBF_HIP(hipMemcpyAsync(ctx->pinned + 16, ctl + 12, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  const uint32_t novf = (uint32_t)(ctx->pinned[16] & 0xFFFFFFFFu);
  ctx->timing.hub_vertices = novf;
  if (novf) {
    // workgroup tables overflowed too: dense HBM histograms
    void* pp;
    const int blocks = (int)(novf < (uint32_t)HBM_BLOCKS ? novf : (uint32_t)HBM_BLOCKS);
    PBF_CHECK(pbf_scratch(ctx, S_POOL, sizeof(uint32_t) * 2 * (size_t)n * blocks, &pp));
    PBF_HIP(hipMemsetAsync(pp, 0, sizeof(uint32_t) * 2 * (size_t)n * blocks, ctx->stream));
    a.pool = static_cast<uint32_t*>(pp);
    a.list = lists + 3 * n; a.list_len = ctl + 12; a.queue = ctl + 13;
    a.ovf_list = lists; a.ovf_len = ctl + 14;  // never written by the HBM kernel
    hipLaunchKernelGGL((k_count<MODE, HBM_TEAM, 1, true, T_HBM>), dim3(blocks), dim3(HBM_TEAM), 0, ctx->stream, a);
    PBF_HIP(hipGetLastError());
  }
  return PBF_OK;
}
}  // namespace

int pbf_dev_count(pbf_context* ctx, const DevGraph& g, int use_v, const RankedGraph& R, int mode, uint64_t* out_dev,
                  const uint32_t* eti) {
  const int64_t n = R.n, m2 = R.m2;
  void *bp, *lp, *cp, *mp, *hp, *op = nullptr;
  PBF_CHECK(pbf_scratch(ctx, S_BOUND, 4 * (size_t)(n ? n : 1), &bp));
  PBF_CHECK(pbf_scratch(ctx, S_BINS, 4 * (size_t)(4 * n + 64), &lp));
  PBF_CHECK(pbf_scratch(ctx, S_MISC, 64 * sizeof(uint64_t), &mp));
  PBF_CHECK(pbf_scratch(ctx, S_HOT, 8 * (size_t)HOT_N * HOT_R, &hp));
  void* lpp;
  PBF_CHECK(pbf_scratch(ctx, S_LPOS, 4 * (size_t)(m2 ? m2 : 1), &lpp));
  uint32_t* lists = static_cast<uint32_t*>(lp);
  uint32_t* ctl = lists + 4 * n;
  CountArgs a{};
  a.roff = R.roff; a.redg = R.redg; a.inpref = R.inpref; a.bound = static_cast<uint32_t*>(bp);
  a.lpos = static_cast<uint32_t*>(lpp);
  a.n = n;
  a.total = static_cast<unsigned long long*>(mp) + 12;
  a.hot = static_cast<unsigned long long*>(hp);
  {
    uint32_t kb = 0;  // key bits: n <= 2^kb - 1, so a packed key is never all ones
    while (kb < 32 && ((uint64_t)1 << kb) - 1 < (uint64_t)n) kb++;
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
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256, 8192)), dim3(256), 0, ctx->stream, a.B, 0ull, n);
    PBF_CHECK(run_count<MODE_VERT>(ctx, R, a, lists, ctl));
    const int64_t ns = use_v ? g.nu : g.nv, base = use_v ? g.nv : 0;
    if (ns)
      hipLaunchKernelGGL(k_convert_vert, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, ctx->stream, a.B, R.rank_of, base, ns,
                         reinterpret_cast<unsigned long long*>(out_dev));
  } else if (mode == MODE_EDGE) {
    PBF_CHECK(pbf_scratch(ctx, S_OIDX, 4 * (size_t)(m2 ? m2 : 1), &op));
    uint32_t* oidx = static_cast<uint32_t*>(op);
    if (m2) hipLaunchKernelGGL(k_oidx, dim3(grid_for(m2, 256, 16384)), dim3(256), 0, ctx->stream, R.redg, R.rpos, eti, m2, oidx);
    a.oidx = oidx;
    a.B = reinterpret_cast<unsigned long long*>(out_dev);
    if (g.m)
      hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(g.m, 256, 8192)), dim3(256), 0, ctx->stream, a.B, 0ull, (int64_t)g.m);
    PBF_CHECK(run_count<MODE_EDGE>(ctx, R, a, lists, ctl));
  } else {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, a.total, 0ull, (int64_t)1);
    PBF_CHECK(run_count<MODE_TOTAL>(ctx, R, a, lists, ctl));
    PBF_HIP(hipMemcpyAsync(out_dev, a.total, 8, hipMemcpyDeviceToDevice, ctx->stream));
  }
  PBF_HIP(hipGetLastError());
  return PBF_OK;
}
