// Butterfly peeling on the GPU: tip numbers (vertices of S) and wing numbers (edges).
//
// Semantics follow the reference's bucketed peeling (Peel / PeelOrigParallel, apps/hyper/
// butterfly_peel.h:647-711, 896-934; PeelE / PeelEOrigParallel, butterfly_epeel.h:339-443,
// 519-569; updateBuckets, butterfly_putils.h:1027-1115; increasing buckets of ligra/bucket.h):
//   repeat: k = min D over live items; every live item with D == k is peeled in this round;
//   the peeled items' butterflies are subtracted from the live items they share them with;
//   every updated live item gets D = max(B, k).  The tip / wing number is D at removal.
// The reference's bucket structure is replaced by min-count frontiers: per round a min
// reduction, a selection (stream compaction of D == k), a two-hop walk of the frontier with
// 64-bit atomic decrements, and an update of the touched items.  Rounds run in batches on the
// device; the host reads the live count once per batch.
//
// Vertex peel walk (PeelOrigParallel :670-702): for a peeled u, every u2 != u reached through a
// centre v loses C(c(u,u2), 2), c counted in an LDS hash (same team enumeration as counting).
// Edge peel walk (PeelEOrigParallel :364-429): for a peeled edge e = (v,u) and every live
// u2 in N(v), the butterflies {u,u2} x {v,v2} with v2 in N(u2) & N(u) and all four edges live
// are charged once, by their minimum-index peeled edge; each of the other three edges loses one.
#include "pbf_device.h"
#include "team.h"

using namespace pbf;

namespace {
constexpr unsigned long long KNONE = ~0ull;
// work[x] of a peeled item: its walk cost; bit 31 is the side an edge walk starts from
constexpr uint32_t WORK_MASK = 0x7FFFFFFFu, WORK_SSIDE = 0x80000000u;

// Control words, list entries and walk targets are read with relaxed agent-scope atomic loads
// (vector loads; a plain load of a uniform address would be a scalar-cache load).  Every such
// value is written by an earlier launch on the stream: no kernel here reads what another
// workgroup of the same launch stored, since within a launch workgroups share data only through
// device-scope atomics whose returned values carry it (queue claims, list reservations).  The
// end of the writing launch and the start of the reading one already order those writes before
// any load, plain or atomic; the atomic form is the one these kernels have been tested with.
__device__ __forceinline__ uint32_t dev_ld(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long dev_ld(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

inline unsigned grid_for(int64_t n, int threads, int max_blocks) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)b;
}

// ---------------- shared round kernels ----------------
// ctl layout (u64): [0] k, [1] live count, [2] k of the previous round; u32 view from word 8:
// list lengths, then queues.
struct Ctl {
  unsigned long long* k;
  unsigned long long* live;
  unsigned long long* kprev;
  uint32_t* len;    // [0] small, [1] large, [2] ovf, [3] ntouched, [4] previous ntouched, [5] found,
                    // [6] split (reservations; the list holds at most its capacity)
  uint32_t* queue;  // [0] small, [1] large, [2] ovf, [3..5] split walk counters; all reset per round
};
constexpr int VS_MAX_ROWS = 32, CTL_QUEUES = 8;

// The split class of the vertex peel: frontier items with at least min_paths two-hop paths, or
// more possible distinct two-hop neighbours (work) than max_work, are walked by many
// workgroups at once (k_vsplit_*), up to cap of them per round.  paths == null: no split
// class (edge peel).
struct SplitSel {
  const uint32_t* paths;
  uint32_t min_paths, max_work, cap;
  uint32_t* list;
};

__global__ void k_ctl_init(Ctl c) {
  if (threadIdx.x == 0) { *c.k = KNONE; *c.kprev = KNONE; }
  if (threadIdx.x < 8) c.len[threadIdx.x] = 0;
}

// The previous round's touched list stays readable until this round's walk appends to it.
__global__ void k_round_begin(Ctl c) {
  if (threadIdx.x == 0) {
    *c.kprev = dev_ld(c.k);
    c.len[4] = dev_ld(c.len + 3);
    c.len[5] = 0;
    *c.k = KNONE;
    for (int q = 0; q < 4; q++) c.len[q] = 0;
    c.len[6] = 0;
    for (int q = 0; q < CTL_QUEUES; q++) c.queue[q] = 0;
  }
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    unsigned long long y = __shfl_xor(x, d, 64);
    x = y < x ? y : x;
  }
  return x;
}

// The round kernels visit items ids[0, n) (every item that may still be live), or 0..n-1
// when ids is null.
// Appends x to a list when flag is set, one counter atomic per wavefront (the lanes active at
// the call take part; the lowest of them reserves the room).  A per-lane atomic on one shared
// counter would serialise every append of the launch.
__device__ __forceinline__ void wave_append(bool flag, uint32_t x, uint32_t* list, uint32_t* len) {
  const uint64_t fm = __ballot(flag);
  if (fm == 0ull) return;
  const uint32_t leader = (uint32_t)__ffsll((long long)__ballot(1)) - 1u;
  uint32_t base = 0;
  if (__lane_id() == leader) base = atomicAdd(len, (uint32_t)__popcll(fm));
  base = (uint32_t)__shfl((int)base, (int)leader, 64);
  if (flag) list[base + __popcll(fm & pbf_lanemask_lt())] = x;
}

// wave_append into a list of capacity cap: true for the lanes whose x found room (the counter
// still counts every request).
__device__ __forceinline__ bool wave_append_cap(bool flag, uint32_t x, uint32_t* list, uint32_t* len, uint32_t cap) {
  const uint64_t fm = __ballot(flag);
  if (fm == 0ull) return false;
  const uint32_t leader = (uint32_t)__ffsll((long long)__ballot(1)) - 1u;
  uint32_t base = 0;
  if (__lane_id() == leader) base = atomicAdd(len, (uint32_t)__popcll(fm));
  base = (uint32_t)__shfl((int)base, (int)leader, 64);
  if (!flag) return false;
  const uint32_t pos = base + __popcll(fm & pbf_lanemask_lt());
  if (pos >= cap) return false;
  list[pos] = x;
  return true;
}

// Frontier from the previous round's touched items.  After a round at level k every live item
// has D >= k, and the live items with D == k are all touched ones (an untouched item keeps its
// D, which was not k or it would have been peeled).  So when a touched item has D == k the new
// minimum is k again and the frontier is exactly those items; else the full scan below runs.
__global__ void k_select_touched(const unsigned long long* __restrict__ D, const unsigned long long* __restrict__ B,
                                 uint8_t* __restrict__ state,
                                 const uint32_t* __restrict__ touched, const uint32_t* __restrict__ work,
                                 uint32_t small_max, uint32_t* __restrict__ small_list, uint32_t* __restrict__ large_list,
                                 SplitSel ss, Ctl c) {
  // control words, state and D are loaded before any of them is tested
  const unsigned long long k = dev_ld(c.kprev);
  const uint32_t nt = dev_ld(c.len + 4);
  if (k == KNONE) return;
  uint32_t nsel = 0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    const uint32_t x = touched[t];
    const uint8_t st = state[x];
    const unsigned long long d = D[x];
    const bool sel = st == 1 && d == k;
    uint32_t w = 0;
    if (sel) {
      state[x] = 2;
      nsel++;
      // an item in no remaining butterfly (B == 0) has nothing to decrement: no walk
      w = B[x] ? (work[x] & WORK_MASK) : 0u;
    }
    bool sp = false;
    if (ss.paths)
      sp = wave_append_cap(sel && w != 0 && (w > ss.max_work || ss.paths[x] >= ss.min_paths), x, ss.list, &c.len[6], ss.cap);
    wave_append(sel && !sp && w != 0 && w <= small_max, x, small_list, &c.len[0]);
    wave_append(sel && !sp && w > small_max, x, large_list, &c.len[1]);
  }
  nsel = pbf_wave_sum(nsel);
  if ((threadIdx.x & 63) == 0 && nsel) {
    atomicAdd(c.live, (unsigned long long)(-(long long)nsel));
    c.len[5] = 1u;
    *c.k = k;
  }
}

// k = min D over live items (state != 0), unless the touched items gave the frontier.
__global__ __launch_bounds__(256) void k_min_live(const unsigned long long* __restrict__ D, const uint8_t* __restrict__ state,
                                                  const uint32_t* __restrict__ ids, int64_t n, Ctl c) {
  __shared__ unsigned long long s_min[4];
  if (dev_ld(c.len + 5)) return;
  unsigned long long mn = KNONE;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t x = ids ? ids[p] : (uint32_t)p;
    const uint8_t st = state[x];
    const unsigned long long d = D[x];  // with the state, not after it
    if (st) mn = d < mn ? d : mn;
  }
  // one atomic per workgroup: per-wavefront atomics on the one word serialise the launch
  mn = wave_min_u64(mn);
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = mn;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) mn = s_min[w] < mn ? s_min[w] : mn;
    if (mn != KNONE) atomicMin(c.k, mn);
  }
}

// Frontier selection: live items with D == k become current (state 2) and are listed by work
// class (small / large by `work`); `live` is decremented by the number selected.
__global__ void k_select(const unsigned long long* __restrict__ D, const unsigned long long* __restrict__ B,
                         uint8_t* __restrict__ state,
                         const uint32_t* __restrict__ ids, int64_t n, const uint32_t* __restrict__ work,
                         uint32_t small_max, uint32_t* __restrict__ small_list, uint32_t* __restrict__ large_list,
                         SplitSel ss, Ctl c) {
  const unsigned long long k = dev_ld(c.k);
  const uint32_t found = dev_ld(c.len + 5);
  if (k == KNONE || found) return;
  uint32_t nsel = 0;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t x = ids ? ids[p] : (uint32_t)p;
    const uint8_t st = state[x];
    const unsigned long long d = D[x];  // with the state, not after it
    const bool sel = st == 1 && d == k;
    uint32_t w = 0;
    if (sel) {
      state[x] = 2;
      nsel++;
      // an item in no remaining butterfly (B == 0) has nothing to decrement: no walk
      w = B[x] ? (work[x] & WORK_MASK) : 0u;
    }
    bool sp = false;
    if (ss.paths)
      sp = wave_append_cap(sel && w != 0 && (w > ss.max_work || ss.paths[x] >= ss.min_paths), x, ss.list, &c.len[6], ss.cap);
    wave_append(sel && !sp && w != 0 && w <= small_max, x, small_list, &c.len[0]);
    wave_append(sel && !sp && w > small_max, x, large_list, &c.len[1]);
  }
  nsel = pbf_wave_sum(nsel);
  if ((threadIdx.x & 63) == 0 && nsel) atomicAdd(c.live, (unsigned long long)(-(long long)nsel));
}

// Touched live items: D = max(B, k) (updateBuckets, butterfly_putils.h:1036-1043).
// End of a round, one launch for two independent passes: touched items get D[x] = max(B[x], k)
// (and lose their flag), current items (state 2) die.  A touched item is live, never current.
__global__ void k_update_retire(const uint32_t* __restrict__ touched, uint32_t* __restrict__ tflag,
                                const unsigned long long* __restrict__ B, unsigned long long* __restrict__ D,
                                uint8_t* __restrict__ state, const uint32_t* __restrict__ ids, int64_t n, Ctl c) {
  const unsigned long long k = dev_ld(c.k);
  const uint32_t nt = dev_ld(c.len + 3);
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    const uint32_t x = touched[t];
    tflag[x] = 0;
    const unsigned long long b = B[x];
    D[x] = b > k ? b : k;
  }
  if (k == KNONE) return;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t x = ids ? ids[p] : (uint32_t)p;
    if (state[x] == 2) state[x] = 0;
  }
}

// ---------------- vertex peel: minimum of D by blocks of items ----------------
// The S items are cut into nb <= 256 blocks of 2^bs_log; vmin[b] is the exact minimum of D over
// the live items of block b at the start of every round.  A round flags the blocks whose items
// it peeled or decremented (plain stores of 1: an atomic min per decremented item on 256 words
// serialised), and the next round's k_vblocks_prepare scans only those blocks again: it retires
// the peeled items and sets D = max(B, k) (updateBuckets) on the live ones.  So the new k is the minimum of the nb block minima, only the blocks
// holding it are searched for the frontier, and late rounds, which touch few blocks, scan little.
struct VBlocks {
  unsigned long long* vmin;  // [nb]
  uint32_t* flag;            // [nb] 1: the block had frontier or decremented items (or needs its first scan)
  uint32_t nb;
  uint32_t bs_log;
};

// Round start: counters reset (workgroup 0), then every flagged block retires its current items
// and recomputes its minimum.
__global__ __launch_bounds__(256) void k_vblocks_prepare(VBlocks vb, uint8_t* __restrict__ state,
                                                         const unsigned long long* __restrict__ B,
                                                         unsigned long long* __restrict__ D, int64_t n, Ctl c) {
  __shared__ unsigned long long s_min[4];
  const unsigned long long k = dev_ld(c.k);  // the previous round's level (KNONE before the first)
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int q = 0; q < 7; q++) c.len[q] = 0;
    for (int q = 0; q < CTL_QUEUES; q++) c.queue[q] = 0;
  }
  const uint32_t b = blockIdx.x;
  if (dev_ld(vb.flag + b) == 0u) return;
  const int64_t lo = (int64_t)b << vb.bs_log;
  const int64_t hi = lo + (int64_t(1) << vb.bs_log) < n ? lo + (int64_t(1) << vb.bs_log) : n;
  unsigned long long mn = KNONE;
  for (int64_t x = lo + threadIdx.x; x < hi; x += blockDim.x) {
    uint8_t st = state[x];
    unsigned long long d = D[x];
    const unsigned long long bx = B[x];
    if (st == 2) { state[x] = 0; st = 0; }
    if (st == 1) {
      // D = max(B, k) holds for every live item, decremented this round or not: an untouched
      // one with B < k has D == k already (D >= k, or it would have been peeled; D was set to
      // max(B, k') with k' <= k)
      if (k != KNONE) {
        const unsigned long long nd = bx > k ? bx : k;
        if (nd != d) { D[x] = nd; d = nd; }
      }
      mn = d < mn ? d : mn;
    }
  }
  mn = wave_min_u64(mn);
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = mn;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) mn = s_min[w] < mn ? s_min[w] : mn;
    __hip_atomic_store(vb.vmin + b, mn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(vb.flag + b, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Frontier selection: every workgroup reduces the block minima to k (workgroup 0 publishes it);
// workgroup b then lists the items of block b with D == k when its minimum is k, as k_select
// does, and flags the block for the next round's scan.
__global__ __launch_bounds__(256) void k_vblocks_select(VBlocks vb, const unsigned long long* __restrict__ D,
                                                        const unsigned long long* __restrict__ B, uint8_t* __restrict__ state,
                                                        int64_t n, const uint32_t* __restrict__ work, uint32_t small_max,
                                                        uint32_t* __restrict__ small_list, uint32_t* __restrict__ large_list,
                                                        SplitSel ss, Ctl c) {
  __shared__ unsigned long long s_min[4];
  __shared__ unsigned long long s_own;
  // nb <= 256 = the workgroup size: thread i holds block i's minimum
  const uint32_t b = blockIdx.x;
  const unsigned long long v = threadIdx.x < vb.nb ? dev_ld(vb.vmin + threadIdx.x) : KNONE;
  if (threadIdx.x == b) s_own = v;
  const unsigned long long mn = wave_min_u64(v);
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = mn;
  __syncthreads();
  unsigned long long k = s_min[0];
  for (int w = 1; w < 4; w++) k = s_min[w] < k ? s_min[w] : k;
  if (b == 0 && threadIdx.x == 0) __hip_atomic_store(c.k, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (k == KNONE || s_own != k) return;
  const int64_t lo = (int64_t)b << vb.bs_log;
  const int64_t hi = lo + (int64_t(1) << vb.bs_log) < n ? lo + (int64_t(1) << vb.bs_log) : n;
  uint32_t nsel = 0;
  // every thread runs the same number of iterations: the appends are per wavefront
  for (int64_t x0 = lo; x0 < hi; x0 += blockDim.x) {
    const int64_t xx = x0 + threadIdx.x;
    const uint32_t x = (uint32_t)xx;
    bool sel = false;
    uint32_t w = 0;
    if (xx < hi) {
      const uint8_t st = state[x];
      const unsigned long long d = D[x];
      sel = st == 1 && d == k;
      if (sel) {
        state[x] = 2;
        nsel++;
        // an item in no remaining butterfly (B == 0) has nothing to decrement: no walk
        w = B[x] ? (work[x] & WORK_MASK) : 0u;
      }
    }
    bool sp = false;
    if (ss.paths)
      sp = wave_append_cap(sel && w != 0 && (w > ss.max_work || ss.paths[x] >= ss.min_paths), x, ss.list, &c.len[6], ss.cap);
    wave_append(sel && !sp && w != 0 && w <= small_max, x, small_list, &c.len[0]);
    wave_append(sel && !sp && w > small_max, x, large_list, &c.len[1]);
  }
  if (threadIdx.x == 0) __hip_atomic_store(vb.flag + b, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  nsel = pbf_wave_sum(nsel);
  if ((threadIdx.x & 63) == 0 && nsel) atomicAdd(c.live, (unsigned long long)(-(long long)nsel));
}

__device__ __forceinline__ void touch(uint32_t x, uint32_t* tflag, uint32_t* touched, uint32_t* ntouched) {
  wave_append(atomicExch(&tflag[x], 1u) == 0u, x, touched, ntouched);
}

// ---------------- vertex peel walk ----------------
enum TableKind { T_LDS = 0, T_HBM = 1 };
// Vertex walk teams: wavefronts with 1024-slot tables for small items, workgroups with
// 8192-slot tables for large ones, 32 workgroups with HBM rows for the rest.
constexpr int V_SMALL_TEAM = 64, V_SMALL_CAP = 1024, V_LARGE_TEAM = 256, V_LARGE_CAP = 8192, V_HBM_BLOCKS = 32;

struct PeelVArgs {
  const uint32_t* offS;
  const uint32_t* edgS;
  const uint32_t* offO;
  const uint32_t* edgO;
  int64_t ns;
  unsigned long long* B;
  const unsigned long long* D;
  const uint32_t* bound;
  const uint32_t* list;
  const uint32_t* list_len;
  uint32_t* queue;
  uint32_t* ovf_list;
  uint32_t* ovf_len;
  uint32_t* tflag;
  uint32_t* touched;
  uint32_t* ntouched;
  uint32_t* pool;
  const unsigned long long* k;
  uint32_t* bflag;          // block flags of the block minima (see VBlocks)
  uint32_t bs_log;
};

// A walk team's LDS: its shared scratch and, for LDS tables, keys, counts and used slots.
template <int TEAM, int CAP, bool USED, int KIND>
struct WalkLds {
  TeamShared<TEAM> sh;
  uint32_t keys[KIND == T_LDS ? CAP : 1];
  uint32_t cnts[KIND == T_LDS ? CAP : 1];
  uint32_t used[(KIND == T_LDS && USED) ? CAP : 1];
};

// The barrier of a walk team: the workgroup, or for a one-wavefront team inside a larger
// workgroup the wavefront (an LDS fence and a wave barrier).
template <int TEAM>
__device__ __forceinline__ void walk_barrier() {
  if constexpr (TEAM == 64) {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
}

// Frontier walks of one team (team index `team` of its role, thread `tid` of TEAM): items are
// drained from the list's queue.  A TEAM of 256 is a whole workgroup; a TEAM of 64 is one
// wavefront and synchronises only itself.
template <int TEAM, int CAP, bool USED, int KIND>
__device__ void peelv_walk(const PeelVArgs& a, WalkLds<TEAM, CAP, USED, KIND>& L, uint32_t team, uint32_t tid) {
  TeamShared<TEAM>& sh = L.sh;
  uint32_t* lkeys = L.keys;
  uint32_t* lcnts = L.cnts;
  uint32_t* lused = L.used;
  const unsigned long long k = dev_ld(a.k);
  if (k == KNONE) return;
  const uint32_t nlist = dev_ld(a.list_len);
  // the first nlist teams drain the queue: the others leave without touching its counter
  if (team >= nlist) return;
  uint32_t* dense = nullptr;
  uint32_t* dused = nullptr;
  if constexpr (KIND == T_HBM) {
    dense = a.pool + (size_t)team * (size_t)a.ns * 2;
    dused = dense + a.ns;
  }
  if constexpr (KIND == T_LDS && USED) {
    for (uint32_t s = tid; s < CAP; s += TEAM) { lkeys[s] = EMPTY; lcnts[s] = 0; }
  }
  while (true) {
    if (tid == 0) sh.item = atomicAdd(a.queue, 1u);
    walk_barrier<TEAM>();
    const uint32_t idx = sh.item;
    walk_barrier<TEAM>();
    if (idx >= nlist) break;
    const uint32_t u = dev_ld(a.list + idx);
    const uint32_t uo = a.offS[u], deg = a.offS[u + 1] - uo;
    LdsTable t;
    if constexpr (KIND == T_LDS) {
      const uint32_t bnd = dev_ld(a.bound + u);
      uint32_t cap = CAP;
      if constexpr (!USED) {
        uint64_t want = (uint64_t)bnd + bnd / 3 + 1;
        cap = 256;
        while (cap < want && cap < CAP) cap <<= 1;
      }
      const uint32_t lg = 31 - __clz(cap);
      t.keys = lkeys; t.cnts = lcnts; t.used = USED ? lused : nullptr;
      t.shift = 32 - lg; t.mask = cap - 1;
      t.maxfill = (cap == CAP) ? (CAP / 4) * 3 - TEAM : 0xFFFFFFFFu;
      if constexpr (!USED) {
        for (uint32_t s = tid; s < cap; s += TEAM) { lkeys[s] = EMPTY; lcnts[s] = 0; }
      }
    }
    if (tid == 0) { sh.nused = 0; sh.ovf = 0; }
    walk_barrier<TEAM>();
    for (uint32_t c0 = 0; c0 < deg; c0 += TEAM) {
      const uint32_t j = c0 + tid;
      uint32_t L = 0, vo = 0;
      if (j < deg) {
        const uint32_t v = a.edgS[uo + j];
        vo = a.offO[v];
        L = a.offO[v + 1] - vo;
      }
      uint32_t T;
      const uint32_t incl = team_incl_scan<TEAM>(L, sh, &T);
      sh.P[tid] = incl;
      sh.vs[tid] = vo;
      walk_barrier<TEAM>();
      for (uint32_t w = tid; w < T; w += TEAM) {
        const uint32_t s = seg_of<TEAM>(sh.P, w);
        const uint32_t kk = w - (s ? sh.P[s - 1] : 0u);
        const uint32_t u2 = a.edgO[sh.vs[s] + kk];
        if (u2 == u) continue;
        if constexpr (KIND == T_LDS) {
          if (!sh.ovf) lds_insert<TEAM>(t, sh, u2);
        } else {
          if (atomicAdd(&dense[u2], 1u) == 0u) dused[atomicAdd(&sh.nused, 1u)] = u2;
        }
      }
      walk_barrier<TEAM>();
    }
    const bool overflowed = sh.ovf != 0;
    if (!overflowed) {
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
        if (c > 1u && a.D[key] > k) {
          atomicAdd(&a.B[key], (unsigned long long)(-(long long)((unsigned long long)c * (c - 1u) / 2u)));
          a.bflag[key >> a.bs_log] = 1u;  // the next prepare pass updates D of the block
        }
      }
    } else if (tid == 0) {
      a.ovf_list[atomicAdd(a.ovf_len, 1u)] = u;
    }
    walk_barrier<TEAM>();
    if constexpr (KIND == T_LDS && USED) {
      const uint32_t nu_ = sh.nused < CAP ? sh.nused : CAP;
      for (uint32_t q = tid; q < nu_; q += TEAM) { const uint32_t slot = lused[q]; lkeys[slot] = EMPTY; lcnts[slot] = 0; }
    }
    if constexpr (KIND == T_HBM) {
      for (uint32_t q = tid; q < sh.nused; q += TEAM) dense[dused[q]] = 0;
    }
    walk_barrier<TEAM>();
  }
}

// bound[u] = min(sum_{v in N(u)} deg(v), ns): distinct two-hop neighbours of u; paths[u] = the
// sum itself (two-hop paths, saturating at 2^31 - 1).
__global__ __launch_bounds__(256) void k_peelv_bound(const uint32_t* __restrict__ offS, const uint32_t* __restrict__ edgS,
                                                     const uint32_t* __restrict__ offO, int64_t ns, uint32_t* __restrict__ bound,
                                                     uint32_t* __restrict__ paths) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int64_t u = (int64_t)blockIdx.x * 4 + wid; u < ns; u += (int64_t)gridDim.x * 4) {
    uint64_t s = 0;
    for (uint32_t p = offS[u] + lane; p < offS[u + 1]; p += 64) {
      const uint32_t v = edgS[p];
      s += offO[v + 1] - offO[v];
    }
    s = pbf_wave_sum(s);
    if (lane == 0) {
      bound[u] = (uint32_t)(s < (uint64_t)ns ? s : (uint64_t)ns);
      paths[u] = (uint32_t)(s < (uint64_t)WORK_MASK ? s : (uint64_t)WORK_MASK);
    }
  }
}

// ---------------- vertex peel: walks split over many workgroups ----------------
// A frontier vertex u with many two-hop paths (the late rounds of a power-law graph peel one
// such vertex at a time) is walked by the whole device: its centres v in N(u) are cut into
// chunks of VS_CHUNK, a chunk's paths into units of VS_PCH, and workgroups take units; the
// counts c(u, u2) go to a dense row of u in HBM (global atomics whose results are not waited
// for), and every counted key sets the flag of its block of 64 keys; k_vsplit_credit then
// sweeps the flagged blocks, subtracts C(c, 2) from every live u2 and clears rows and flags.
constexpr uint32_t VS_CHUNK = 256, VS_PCH = 256, VS_MIN_PATHS = 1u << 15;

struct VSplitArgs {
  const uint32_t* offS;
  const uint32_t* edgS;
  const uint32_t* offO;
  const uint32_t* edgO;
  int64_t ns;
  const uint32_t* list;     // split items, row r = list[r]
  const uint32_t* len;      // reservations of the list (rows = min(*len, cap))
  uint32_t cap;
  uint32_t* rows;           // per row: dense counters [ns], zero between rounds
  uint32_t* flags;          // per row: one word per block of 64 keys [nblk], zero between rounds
  uint32_t nblk;
  unsigned long long* sub;  // units after each chunk's first: (row << 59) | (global chunk << 27) | unit,
                            // the units of one chunk consecutive
  uint32_t sub_cap;
  uint2* pre;               // per global chunk g < pre_cap: the chunk's (prefix, O-list start) pairs
  uint32_t pre_cap;
  uint32_t* ctr;            // [1] units listed
  unsigned long long* B;
  const unsigned long long* D;
  uint32_t* bflag;          // block flags of the block minima (see VBlocks)
  uint32_t bs_log;
  const unsigned long long* k;
};

__device__ __forceinline__ uint32_t vs_nrows(const VSplitArgs& a) {
  const uint32_t n = dev_ld(a.len);
  return n < a.cap ? n : a.cap;
}

// Row r's chunk starts: every thread computes the prefix of chunk counts over the rows (at
// most cap rows, read from L2).  Returns the row holding global chunk g and the chunk in it.
__device__ __forceinline__ void vs_locate(const VSplitArgs& a, uint32_t nrows, uint32_t g, uint32_t* row, uint32_t* chunk) {
  uint32_t before = 0;
  for (uint32_t r = 0; r < nrows; r++) {
    const uint32_t u = dev_ld(a.list + r);
    const uint32_t nch = (a.offS[u + 1] - a.offS[u] + VS_CHUNK - 1) / VS_CHUNK;
    if (g < before + nch) { *row = r; *chunk = g - before; return; }
    before += nch;
  }
  *row = nrows; *chunk = 0;
}

// Chunk c of u: the inclusive prefix of its centres' path counts into sh.P and their O-list
// starts into sh.vs (one centre per thread).  Returns the chunk's path total, which is below
// 2^32: the centres are distinct, so their paths are distinct edges.
__device__ uint32_t vs_chunk(const VSplitArgs& a, TeamShared<256>& sh, uint32_t u, uint32_t c, uint2* keep) {
  const uint32_t tid = threadIdx.x;
  const uint32_t uo = a.offS[u], deg = a.offS[u + 1] - uo;
  const uint32_t j = c * VS_CHUNK + tid;
  uint32_t L = 0, vo = 0;
  if (j < deg) {
    const uint32_t v = a.edgS[uo + j];
    vo = a.offO[v];
    L = a.offO[v + 1] - vo;
  }
  uint32_t T;
  const uint32_t incl = team_incl_scan<256>(L, sh, &T);
  sh.P[tid] = incl;
  sh.vs[tid] = vo;
  if (keep) keep[tid] = make_uint2(incl, vo);
  __syncthreads();
  return T;
}

// The stored prefix of global chunk g (written by phase 1, vsplit_chunks) into sh; returns its total.
__device__ uint32_t vs_chunk_stored(const VSplitArgs& a, TeamShared<256>& sh, uint32_t g) {
  const uint32_t tid = threadIdx.x;
  const uint32_t* q = reinterpret_cast<const uint32_t*>(a.pre + (size_t)g * VS_CHUNK + tid);
  const uint32_t incl = dev_ld(q), vo = dev_ld(q + 1);
  __syncthreads();  // every thread is done with the previous chunk
  sh.P[tid] = incl;
  sh.vs[tid] = vo;
  __syncthreads();
  return sh.P[VS_CHUNK - 1];
}

// Paths [s * VS_PCH, min((s + 1) * VS_PCH, T)) of the chunk in sh, counted into row r of u:
// VS_PCH / 256 paths per thread, all key loads issued before the counter updates, which
// return nothing (a key's block flag is set by a plain store of 1 instead).
__device__ __forceinline__ void vs_count(const VSplitArgs& a, const TeamShared<256>& sh, uint32_t r, uint32_t u,
                                         uint32_t s, uint32_t T) {
  const uint32_t tid = threadIdx.x;
  uint32_t* row = a.rows + (size_t)r * (size_t)a.ns;
  uint32_t* flag = a.flags + (size_t)r * a.nblk;
  const uint64_t w0 = (uint64_t)s * VS_PCH;
  const uint32_t wend = (uint32_t)(w0 + VS_PCH < T ? w0 + VS_PCH : T);
  constexpr uint32_t PER = VS_PCH / 256;
  uint32_t key[PER];
#pragma unroll
  for (uint32_t q = 0; q < PER; q++) {
    const uint32_t w = (uint32_t)w0 + q * 256 + tid;
    key[q] = u;
    if (w < wend) {
      const uint32_t sg = seg_of<256>(sh.P, w);
      const uint32_t kk = w - (sg ? sh.P[sg - 1] : 0u);
      key[q] = a.edgO[sh.vs[sg] + kk];
    }
  }
#pragma unroll
  for (uint32_t q = 0; q < PER; q++) {
    if (key[q] != u) {
      atomicAdd(&row[key[q]], 1u);
      flag[key[q] >> 6] = 1u;
    }
  }
}

// Phase 1: workgroups take whole chunks (chunk g, g + nb, ... for workgroup bid of nb) and
// count their first unit; a chunk with more than VS_PCH paths reserves consecutive room for its
// other units and lists them for phase 2.  The list has room for every unit of a round: a row
// has at most m / VS_PCH + (its chunks) units, as its paths u -> v -> u2 are distinct edges.
__device__ void vsplit_chunks(const VSplitArgs& a, TeamShared<256>& sh, uint32_t bid, uint32_t nb) {
  const unsigned long long k = dev_ld(a.k);
  const uint32_t nrows = vs_nrows(a);
  if (k == KNONE || nrows == 0) return;
  uint32_t total = 0, u0 = 0, nch0 = 0;
  for (uint32_t r = 0; r < nrows; r++) {
    const uint32_t u = dev_ld(a.list + r);
    const uint32_t nch = (a.offS[u + 1] - a.offS[u] + VS_CHUNK - 1) / VS_CHUNK;
    if (r == 0) { u0 = u; nch0 = nch; }
    total += nch;
  }
  for (uint32_t g = bid; g < total; g += nb) {
    // the chunks of row 0 (most rounds have one row) need no search
    uint32_t r = 0, c = g, u = u0;
    if (g >= nch0) {
      vs_locate(a, nrows, g, &r, &c);
      u = dev_ld(a.list + r);
    }
    const uint32_t T = vs_chunk(a, sh, u, c, g < a.pre_cap ? a.pre + (size_t)g * VS_CHUNK : nullptr);
    const uint32_t nunits = (uint32_t)(((uint64_t)T + VS_PCH - 1) / VS_PCH);
    if (nunits > 1 && threadIdx.x == 0) sh.item = atomicAdd(&a.ctr[1], nunits - 1);
    vs_count(a, sh, r, u, 0, T);
    __syncthreads();
    const uint32_t base = sh.item;
    for (uint32_t s = 1 + threadIdx.x; s < nunits; s += 256) {
      const uint32_t pos = base + s - 1;
      if (pos < a.sub_cap) a.sub[pos] = ((unsigned long long)r << 59) | ((unsigned long long)g << 27) | s;
    }
  }
}

// Phase 2: workgroup b takes the listed units [b * per, (b + 1) * per); consecutive units of
// one chunk reuse the chunk's prefix already in LDS, which is read from the copy phase 1 kept
// (the first pre_cap chunks of a round) or else scanned again; the next list entry is loaded
// while the current unit runs.
__device__ void vsplit_units(const VSplitArgs& a, TeamShared<256>& sh, uint32_t bid, uint32_t nb) {
  const unsigned long long k = dev_ld(a.k);
  const uint32_t listed = dev_ld(a.ctr + 1);
  if (k == KNONE) return;
  const uint32_t n = listed < a.sub_cap ? listed : a.sub_cap;
  const uint32_t per = (uint32_t)(((uint64_t)n + nb - 1) / nb);
  const uint64_t q0 = (uint64_t)bid * per;
  const uint64_t q1 = q0 + per < n ? q0 + per : n;
  if (q0 >= q1) return;
  uint32_t cr = ~0u, cg = ~0u, u = 0, T = 0;
  unsigned long long e = dev_ld(a.sub + q0);
  for (uint64_t q = q0; q < q1; q++) {
    const unsigned long long en = q + 1 < q1 ? dev_ld(a.sub + q + 1) : 0ull;
    const uint32_t r = (uint32_t)(e >> 59), g = (uint32_t)(e >> 27) & 0xFFFFFFFFu;
    if (r != cr || g != cg) {
      u = dev_ld(a.list + r);
      if (g < a.pre_cap) {
        T = vs_chunk_stored(a, sh, g);
      } else {
        uint32_t r2, c;
        vs_locate(a, vs_nrows(a), g, &r2, &c);
        T = vs_chunk(a, sh, u, c, nullptr);
      }
      cr = r;
      cg = g;
    }
    vs_count(a, sh, r, u, (uint32_t)e & 0x7FFFFFFu, T);
    e = en;
  }
}

// A round's walks run in two launches.  Launch 1: phase 1 of the split walks in workgroups
// [0, chunk_blocks), wavefront teams (4 per workgroup) on the small list in the next
// small_blocks, workgroup teams on the large list in the rest.
__global__ __launch_bounds__(256) void k_peelv_walks(PeelVArgs small, PeelVArgs large, VSplitArgs vs, uint32_t chunk_blocks,
                                                     uint32_t small_blocks) {
  __shared__ union {
    TeamShared<256> sh;
    WalkLds<V_SMALL_TEAM, V_SMALL_CAP, true, T_LDS> s[4];
    WalkLds<V_LARGE_TEAM, V_LARGE_CAP, false, T_LDS> l;
  } lds;
  if (blockIdx.x < chunk_blocks) {
    vsplit_chunks(vs, lds.sh, blockIdx.x, chunk_blocks);
  } else if (blockIdx.x < chunk_blocks + small_blocks) {
    const uint32_t w = threadIdx.x >> 6;
    peelv_walk<V_SMALL_TEAM, V_SMALL_CAP, true, T_LDS>(small, lds.s[w], (blockIdx.x - chunk_blocks) * 4 + w,
                                                        threadIdx.x & 63);
  } else {
    peelv_walk<V_LARGE_TEAM, V_LARGE_CAP, false, T_LDS>(large, lds.l, blockIdx.x - chunk_blocks - small_blocks,
                                                         threadIdx.x);
  }
}

// Launch 2: the HBM-row walks of the items no LDS table held (their list is complete, launch 1
// having ended) in workgroups [0, V_HBM_BLOCKS), phase 2 of the split walks in the others.
__global__ __launch_bounds__(256) void k_vsplit_units(VSplitArgs a, PeelVArgs hbm) {
  __shared__ union {
    TeamShared<256> sh;
    WalkLds<V_LARGE_TEAM, 1, true, T_HBM> h;
  } lds;
  if (blockIdx.x < (uint32_t)V_HBM_BLOCKS)
    peelv_walk<V_LARGE_TEAM, 1, true, T_HBM>(hbm, lds.h, blockIdx.x, threadIdx.x);
  else
    vsplit_units(a, lds.sh, blockIdx.x - V_HBM_BLOCKS, gridDim.x - V_HBM_BLOCKS);
}

// One wavefront per block of 64 keys, over all rows at once: lane r < cap reads row r's flag
// of the block (rows past the round's count are never flagged), every lane its key's D, both
// before either is used; then C(c_r, 2) summed over the flagged rows comes off B of every live
// key, the counters and flags are cleared, and a block with a decremented key is flagged for
// the next round's prepare pass (which sets D = max(B, k) over the block).
constexpr uint32_t VS_CREDIT_WAVES = 16;
__global__ __launch_bounds__(64 * VS_CREDIT_WAVES) void k_vsplit_credit(VSplitArgs a) {
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  const unsigned long long k = dev_ld(a.k);
  for (uint64_t b0 = (uint64_t)blockIdx.x * VS_CREDIT_WAVES; b0 < a.nblk; b0 += (uint64_t)gridDim.x * VS_CREDIT_WAVES) {
    const uint64_t b = b0 + wid;
    if (b >= a.nblk) continue;
    const uint32_t key = (uint32_t)b * 64u + lane;
    uint32_t* fl = a.flags + (size_t)lane * a.nblk + b;
    const uint32_t f = lane < a.cap ? dev_ld(fl) : 0u;
    const bool in = key < (uint64_t)a.ns;
    const unsigned long long d = in ? a.D[key] : 0ull;
    uint64_t fm = __ballot(f != 0u);
    if (fm == 0ull) continue;
    if (f != 0u) __hip_atomic_store(fl, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long sub = 0;
    while (fm) {
      const uint32_t r = (uint32_t)__ffsll((long long)fm) - 1u;
      fm &= fm - 1ull;
      uint32_t* row = a.rows + (size_t)r * (size_t)a.ns;
      const uint32_t c = in ? dev_ld(row + key) : 0u;
      if (c != 0u) __hip_atomic_store(row + key, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      sub += (unsigned long long)c * (c - (c != 0u)) / 2u;
    }
    const bool hit = k != KNONE && sub != 0ull && d > k;
    if (hit) atomicAdd(&a.B[key], (unsigned long long)(-(long long)sub));
    // the 64 keys lie in one block of the block minima: one flag for the wavefront
    if (__ballot(hit) != 0ull && lane == 0) a.bflag[key >> a.bs_log] = 1u;
  }
}

// ---------------- edge peel walk ----------------
// The walk reads a compacted adjacency: every entry is (neighbour, edge id), edge id = the
// edge's O-CSR index; entries of dead edges are dropped whenever the live edges fall to half
// of the entries (the first view is the input CSR itself).
struct PeelEArgs {
  const uint32_t* offS;     // S side: offsets, neighbours (O vertices), edge ids
  const uint32_t* nbrS;
  const uint32_t* idS;
  const uint32_t* offO;     // O side: offsets, neighbours (S vertices), edge ids
  const uint32_t* nbrO;
  const uint32_t* idO;
  const uint32_t* u_of;     // S vertex of each edge id (the input O-CSR)
  const uint32_t* owner_o;  // O vertex of each edge id
  const uint32_t* work;     // bit 31: walk from the S end (see k_peele_work)
  const uint8_t* state;     // 0 dead, 1 live, 2 peeled this round
  unsigned long long* B;
  const uint32_t* list;
  const uint32_t* list_len;
  uint32_t* queue;
  uint32_t* tflag;
  uint32_t* touched;
  uint32_t* ntouched;
  const unsigned long long* k;
};

// Edge f is charged by peeled edge e when it is live and, if peeled too, has a larger index.
__device__ __forceinline__ bool usable(const uint8_t* state, uint32_t f, uint32_t e) {
  const uint8_t s = state[f];
  return s == 1 || (s == 2 && f > e);
}

// Walk of a peeled edge e = (v, u).  Every butterfly through e is {v, v2} x {u, u2}; it is found
// either from v (u2 in N(v), v2 in N(u2), test v2 in N(u)) or, symmetrically, from u (v2 in
// N(u), u2 in N(v2), test u2 in N(v)), whichever end has the smaller two-hop work.  Below, a is
// the starting end, b the other one: c runs over N(a) \ {b}, d over N(c) \ {a}, and d must lie in
// N(b) (staged in LDS).  The edge (a, c) loses the butterflies of its segment at once; (c, d)
// and (b, d) lose one each.
template <int TEAM, int UCAP>
__global__ __launch_bounds__(TEAM) void k_peele_walk(PeelEArgs a) {
  __shared__ TeamShared<TEAM> sh;
  __shared__ uint32_t nb_v[UCAP];    // N(b), ascending
  __shared__ uint32_t nb_f[UCAP];    // edge ids of N(b)
  // membership filter of N(b): one bit per hashed id (no false negatives), so most tested
  // vertices are rejected by one LDS read before any search or state load
  constexpr int BM_LOG = UCAP >= 2048 ? 13 : 11;
  __shared__ uint32_t nb_bits[(1 << BM_LOG) / 32];
  const uint32_t tid = threadIdx.x;
  if (dev_ld(a.k) == KNONE) return;
  const uint32_t nlist = dev_ld(a.list_len);
  // the first nlist workgroups drain the queue: the others leave without touching its counter
  if (blockIdx.x >= nlist) return;
  while (true) {
    if (tid == 0) sh.item = atomicAdd(a.queue, 1u);
    __syncthreads();
    const uint32_t idx = sh.item;
    __syncthreads();
    if (idx >= nlist) break;
    const uint32_t e = dev_ld(a.list + idx);
    const bool from_s = (a.work[e] & WORK_SSIDE) != 0u;
    const uint32_t v = a.owner_o[e], u = a.u_of[e];
    const uint32_t va = from_s ? u : v, vb = from_s ? v : u;
    // X: lists of a's side; Y: lists of the other side (those of c's and of b)
    const uint32_t* offX = from_s ? a.offS : a.offO;
    const uint32_t* nbrX = from_s ? a.nbrS : a.nbrO;
    const uint32_t* idX = from_s ? a.idS : a.idO;
    const uint32_t* offY = from_s ? a.offO : a.offS;
    const uint32_t* nbrY = from_s ? a.nbrO : a.nbrS;
    const uint32_t* idY = from_s ? a.idO : a.idS;
    const uint32_t ao = offX[va], adeg = offX[va + 1] - ao;
    const uint32_t bo = offY[vb], bdeg = offY[vb + 1] - bo;
    const bool cached = bdeg <= (uint32_t)UCAP;
    if (cached) {
      for (uint32_t q = tid; q < (1u << BM_LOG) / 32; q += TEAM) nb_bits[q] = 0;
      __syncthreads();
      for (uint32_t q = tid; q < bdeg; q += TEAM) {
        const uint32_t y = nbrY[bo + q];
        nb_v[q] = y;
        nb_f[q] = idY[bo + q];
        const uint32_t h = (y * 0x9E3779B1u) >> (32 - BM_LOG);
        atomicOr(&nb_bits[h >> 5], 1u << (h & 31));
      }
    }
    // B[e] counts exactly the butterflies through e whose four edges are live or peeled in
    // this round: once the walk has met that many, nothing is left to find
    const unsigned long long target = dev_ld(a.B + e);
    volatile uint32_t* found = &sh.nused;
    if (tid == 0) *found = 0;
    __syncthreads();
    for (uint32_t c0 = 0; c0 < adeg; c0 += TEAM) {
      const uint32_t j = c0 + tid;
      uint32_t L = 0, so = 0, fc = 0;
      if (j < adeg) {
        const uint32_t q = ao + j;  // edge (a, c)
        const uint32_t c = nbrX[q];
        fc = idX[q];
        if (c != vb && usable(a.state, fc, e)) {
          so = offY[c];
          L = offY[c + 1] - so;
        }
      }
      uint32_t T;
      const uint32_t incl = team_incl_scan<TEAM>(L, sh, &T);
      sh.P[tid] = incl;
      sh.vs[tid] = so;
      sh.acc[tid] = 0;
      __syncthreads();
      for (uint32_t w = tid; w < T; w += TEAM) {
        if (*found >= target) break;
        const uint32_t s = seg_of<TEAM>(sh.P, w);
        const uint32_t kk = w - (s ? sh.P[s - 1] : 0u);
        const uint32_t p2 = sh.vs[s] + kk;
        const uint32_t d = nbrY[p2];
        if (d == va) continue;
        // edge (b, d): d in N(b)?  And edge (c, d) still there?
        uint32_t lo = 0, hi = bdeg, fd;
        uint8_t sd;
        if (cached) {
          const uint32_t h = (d * 0x9E3779B1u) >> (32 - BM_LOG);
          if (!((nb_bits[h >> 5] >> (h & 31)) & 1u)) continue;
          while (lo < hi) { const uint32_t mid = (lo + hi) >> 1; if (nb_v[mid] < d) lo = mid + 1; else hi = mid; }
          if (lo == bdeg || nb_v[lo] != d) continue;
          fd = idY[p2];
          sd = a.state[fd];
          if (sd == 0) continue;
        } else {
          fd = idY[p2];
          sd = a.state[fd];
          if (sd == 0) continue;
          while (lo < hi) { const uint32_t mid = (lo + hi) >> 1; if (nbrY[bo + mid] < d) lo = mid + 1; else hi = mid; }
          if (lo == bdeg || nbrY[bo + lo] != d) continue;
        }
        const uint32_t fb = cached ? nb_f[lo] : idY[bo + lo];
        const uint8_t sb = a.state[fb];
        if (sb == 0) continue;
        atomicAdd((uint32_t*)found, 1u);  // a butterfly through e
        if (!(sd == 1 || fd > e) || !(sb == 1 || fb > e)) continue;  // charged by a smaller peeled edge
        atomicAdd(&sh.acc[s], 1ull);
        if (a.state[fd] == 1) { atomicAdd(&a.B[fd], ~0ull); touch(fd, a.tflag, a.touched, a.ntouched); }
        if (a.state[fb] == 1) { atomicAdd(&a.B[fb], ~0ull); touch(fb, a.tflag, a.touched, a.ntouched); }
      }
      __syncthreads();
      if (j < adeg) {
        const unsigned long long cnt = sh.acc[tid];
        if (cnt && a.state[fc] == 1) {
          atomicAdd(&a.B[fc], (unsigned long long)(-(long long)cnt));
          touch(fc, a.tflag, a.touched, a.ntouched);
        }
      }
      __syncthreads();
      if (*found >= target) break;  // uniform: read by every thread after the barrier
    }
    __syncthreads();  // the next item resets the counter
  }
}

// Two-hop work of every vertex over the current adjacency: sum of the neighbours' degrees.
__global__ __launch_bounds__(256) void k_hop_work(const uint32_t* __restrict__ off, const uint32_t* __restrict__ nbr,
                                                  const uint32_t* __restrict__ off_other, int64_t nvert,
                                                  uint32_t* __restrict__ hop) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int64_t x = (int64_t)blockIdx.x * 4 + wid; x < nvert; x += (int64_t)gridDim.x * 4) {
    uint64_t s = 0;
    const uint32_t b = off[x], e = off[x + 1];
    for (uint32_t q = b + lane; q < e; q += 64) {
      const uint32_t y = nbr[q];
      s += off_other[y + 1] - off_other[y];
    }
    s = pbf_wave_sum(s);
    if (lane == 0) hop[x] = (uint32_t)(s < WORK_MASK ? s : WORK_MASK);
  }
}

// work[e] for e = (v, u): the smaller of the two ends' two-hop work (saturating), bit 31 set
// when the walk starts from u; owner_o (when given) receives each edge's O vertex.
__global__ __launch_bounds__(256) void k_peele_work(const uint32_t* __restrict__ offO, const uint32_t* __restrict__ nbrO,
                                                    const uint32_t* __restrict__ idO, int64_t no,
                                                    const uint32_t* __restrict__ hopO, const uint32_t* __restrict__ hopS,
                                                    uint32_t* __restrict__ work, uint32_t* __restrict__ owner_o) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int64_t v = (int64_t)blockIdx.x * 4 + wid; v < no; v += (int64_t)gridDim.x * 4) {
    const uint32_t b = offO[v], e = offO[v + 1], hv = hopO[v];
    for (uint32_t q = b + lane; q < e; q += 64) {
      const uint32_t hu = hopS[nbrO[q]], id = idO[q];
      work[id] = hu < hv ? (hu | WORK_SSIDE) : hv;
      if (owner_o) owner_o[id] = (uint32_t)v;
    }
  }
}

__global__ void k_iota(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x)
    p[x] = (uint32_t)x;
}

// Adjacency compaction: flag[p] = entry p's edge is live (flag[n] = 0), then, from the
// exclusive scan pos of the flags, the live entries move to pos[p] (order kept) and every
// offset off[x] becomes pos[off[x]].
__global__ void k_live_flags(const uint32_t* __restrict__ id, int64_t n, const uint8_t* __restrict__ state,
                             uint32_t* __restrict__ flag) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p <= n; p += (int64_t)gridDim.x * blockDim.x)
    flag[p] = (p < n && state[id[p]] == 1) ? 1u : 0u;
}
__global__ void k_compact(const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ id,
                          const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, int64_t n,
                          uint32_t* __restrict__ nbr2, uint32_t* __restrict__ id2) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
    if (flag[p]) { nbr2[pos[p]] = nbr[p]; id2[pos[p]] = id[p]; }
}
__global__ void k_new_off(const uint32_t* __restrict__ off, int64_t nvert, const uint32_t* __restrict__ pos,
                          uint32_t* __restrict__ off2) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x <= nvert; x += (int64_t)gridDim.x * blockDim.x)
    off2[x] = pos[off[x]];
}

__global__ void k_init_peel(const unsigned long long* __restrict__ counts, unsigned long long* __restrict__ B,
                            unsigned long long* __restrict__ D, uint8_t* __restrict__ state, uint32_t* __restrict__ tflag,
                            int64_t n) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long c = counts[x];
    B[x] = c;
    D[x] = c;
    state[x] = 1;
    tflag[x] = 0;
  }
}

// Edge walks are latency bound: the wavefront teams keep a small LDS copy of N(b) (longer lists
// are searched in place) so that 8 teams fit per SIMD.
constexpr int E_SMALL_TEAM = 64, E_SMALL_UCAP = 256, E_SMALL_PER_CU = 32;
constexpr int E_LARGE_TEAM = 256, E_UCAP = 2048, E_LARGE_PER_CU = 7;
constexpr int ROUND_BATCH = 16;

// Peel-state buffers common to both peels.
struct PeelBufs {
  unsigned long long *B, *D;
  uint8_t* state;
  uint32_t *tflag, *touched, *work, *lists, *owner_o;
  Ctl c;
};

int peel_bufs(pbf_context* ctx, int64_t n, PeelBufs* pb) {
  void *p0, *p1, *p2, *p3, *p4, *p5, *mp;
  const size_t nn = (size_t)(n ? n : 1);
  PBF_CHECK(pbf_scratch(ctx, S_PEEL0, 8 * nn, &p0));
  PBF_CHECK(pbf_scratch(ctx, S_PEEL1, 8 * nn, &p1));
  PBF_CHECK(pbf_scratch(ctx, S_PEEL2, nn, &p2));
  PBF_CHECK(pbf_scratch(ctx, S_PEEL3, 4 * nn * 2, &p3));  // tflag | touched
  PBF_CHECK(pbf_scratch(ctx, S_PEEL4, 4 * nn * 2, &p4));  // work | owner_o
  PBF_CHECK(pbf_scratch(ctx, S_PEEL5, 4 * nn * 3, &p5));  // lists: small | large | ovf
  PBF_CHECK(pbf_scratch(ctx, S_PEEL6, 256, &mp));
  pb->B = static_cast<unsigned long long*>(p0);
  pb->D = static_cast<unsigned long long*>(p1);
  pb->state = static_cast<uint8_t*>(p2);
  pb->tflag = static_cast<uint32_t*>(p3);
  pb->touched = pb->tflag + nn;
  pb->work = static_cast<uint32_t*>(p4);
  pb->owner_o = pb->work + nn;
  pb->lists = static_cast<uint32_t*>(p5);
  unsigned long long* m64 = static_cast<unsigned long long*>(mp);
  pb->c.k = m64;
  pb->c.live = m64 + 1;
  pb->c.kprev = m64 + 2;
  pb->c.len = reinterpret_cast<uint32_t*>(m64 + 4);
  pb->c.queue = reinterpret_cast<uint32_t*>(m64 + 8);
  return PBF_OK;
}

// Runs rounds in batches until no live item remains; returns the number of rounds up to the
// batch after which none remained.  One batch is queued ahead of the one whose live count the
// host waits for (each batch copies the count to its own pinned word and records its own
// event), so the device does not idle while the host checks; rounds queued after the last
// item has gone find no minimum and do nothing.
template <class RoundF>
int run_rounds(pbf_context* ctx, PeelBufs& pb, int64_t n, RoundF round, int64_t* rounds_out) {
  ctx->pinned[33] = (uint64_t)n;
  PBF_HIP(hipMemcpyAsync(pb.c.live, ctx->pinned + 33, 8, hipMemcpyHostToDevice, ctx->stream));
  // every round peels at least one live item, so n rounds always suffice
  const int64_t max_batches = n / ROUND_BATCH + 1;
  auto issue = [&](int64_t b) -> int {
    for (int r = 0; r < ROUND_BATCH; r++) PBF_CHECK(round());
    PBF_HIP(hipMemcpyAsync(ctx->pinned + 40 + (b & 1), pb.c.live, 8, hipMemcpyDeviceToHost, ctx->stream));
    PBF_HIP(hipEventRecord(ctx->ev[6 + (b & 1)], ctx->stream));
    return PBF_OK;
  };
  PBF_CHECK(issue(0));
  int64_t issued = 1;
  int st = PBF_ERR_DOMAIN;
  for (int64_t b = 0; b < max_batches; b++) {
    if (issued == b + 1 && issued < max_batches) PBF_CHECK(issue(issued++));
    PBF_HIP(hipEventSynchronize(ctx->ev[6 + (b & 1)]));
    if (ctx->pinned[40 + (b & 1)] == 0) {
      *rounds_out = (b + 1) * ROUND_BATCH;
      st = PBF_OK;
      break;
    }
  }
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  if (st != PBF_OK) *rounds_out = issued * ROUND_BATCH;
  return st;
}
}  // namespace

int pbf_dev_peel_vert(pbf_context* ctx, const DevGraph& g, int use_v, const uint64_t* counts, uint64_t* tip) {
  const int64_t ns = use_v ? g.nu : g.nv;
  if (ns == 0) return PBF_OK;
  PeelBufs pb;
  PBF_CHECK(peel_bufs(ctx, ns, &pb));
  PeelVArgs a{};
  a.offS = use_v ? g.offU : g.offV;
  a.edgS = use_v ? g.edgU : g.edgV;
  a.offO = use_v ? g.offV : g.offU;
  a.edgO = use_v ? g.edgV : g.edgU;
  a.ns = ns;
  a.B = pb.B;
  a.D = pb.D;
  a.bound = pb.work;
  a.tflag = pb.tflag;
  a.touched = pb.touched;
  a.ntouched = pb.c.len + 3;
  a.k = pb.c.k;
  hipLaunchKernelGGL(k_init_peel, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<const unsigned long long*>(counts), pb.B, pb.D, pb.state, pb.tflag, ns);
  hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, ctx->stream, pb.c);
  uint32_t* paths = pb.owner_o;  // the vertex peel has no edge owners
  hipLaunchKernelGGL(k_peelv_bound, dim3(grid_for(ns, 4, 16384)), dim3(256), 0, ctx->stream, a.offS, a.edgS, a.offO, ns,
                     pb.work, paths);
  PBF_HIP(hipGetLastError());
  // HBM fallback pool, only when some vertex could overflow the largest LDS table
  const uint32_t small_max = (V_SMALL_CAP / 4) * 3 - V_SMALL_TEAM;
  void* pp;
  PBF_CHECK(pbf_scratch(ctx, S_POOL, sizeof(uint32_t) * 2 * (size_t)ns * V_HBM_BLOCKS, &pp));
  PBF_HIP(hipMemsetAsync(pp, 0, sizeof(uint32_t) * 2 * (size_t)ns * V_HBM_BLOCKS, ctx->stream));
  ctx->pool_clean = 0;  // the counting kernels re-clear the pool before their next use
  a.pool = static_cast<uint32_t*>(pp);
  uint32_t* small_list = pb.lists;
  uint32_t* large_list = pb.lists + ns;
  uint32_t* ovf_list = pb.lists + 2 * ns;
  // split walks: up to vrows rows per round, each a dense row of ns words and its block flags
  const int64_t no = use_v ? g.nv : g.nu;
  const uint32_t vrows = (uint32_t)std::max<int64_t>(1, std::min<int64_t>(VS_MAX_ROWS, (int64_t(1) << 31) / (8 * ns)));
  const uint32_t sub_cap = vrows * (uint32_t)(g.m / VS_PCH + no / VS_CHUNK + 2);
  const uint32_t nblk = (uint32_t)((ns + 63) / 64);
  // kept chunk prefixes: every chunk of a round up to 2^16 of them (2 KB each)
  const uint32_t pre_cap = (uint32_t)std::min<int64_t>((int64_t)vrows * (no / VS_CHUNK + 2), int64_t(1) << 16);
  VSplitArgs vs{};
  VBlocks vb{};
  {
    const size_t words = (size_t)vrows * ((size_t)ns + nblk) + (size_t)vrows;
    void* sp;
    PBF_CHECK(pbf_scratch(ctx, S_PEEL7, 4 * words + 8 * (size_t)sub_cap + 8 * (size_t)pre_cap * VS_CHUNK + 12 * 256 + 64, &sp));
    uint32_t* w = static_cast<uint32_t*>(sp);
    vs.rows = w;
    vs.flags = w + (size_t)vrows * ns;
    uint32_t* vlist = vs.flags + (size_t)vrows * nblk;
    vs.sub = reinterpret_cast<unsigned long long*>(reinterpret_cast<uintptr_t>(vlist + vrows + 2) & ~uintptr_t(7));
    // rows and flags start at zero; every round's credit pass leaves them so
    PBF_CHECK(pbf_fill_u32(ctx, vs.rows, 0u, (int64_t)vrows * ((int64_t)ns + nblk)));
    vs.list = vlist;
    vs.nblk = nblk;
    vs.pre = reinterpret_cast<uint2*>(vs.sub + sub_cap);
    vs.pre_cap = pre_cap;
    // the block minima of D: at most 256 blocks of at least 1024 items, all scanned in round 1
    vb.bs_log = 10;
    while (((ns - 1) >> vb.bs_log) + 1 > 256) vb.bs_log++;
    vb.nb = (uint32_t)(((ns - 1) >> vb.bs_log) + 1);
    vb.vmin = reinterpret_cast<unsigned long long*>(vs.pre + (size_t)pre_cap * VS_CHUNK);
    vb.flag = reinterpret_cast<uint32_t*>(vb.vmin + 256);
    PBF_CHECK(pbf_fill_u32(ctx, vb.flag, 1u, vb.nb));
  }
  const unsigned credit_grid = grid_for(nblk, VS_CREDIT_WAVES, ctx->num_cus * 2);  // a wavefront per block
  vs.offS = a.offS; vs.edgS = a.edgS; vs.offO = a.offO; vs.edgO = a.edgO; vs.ns = ns;
  vs.len = pb.c.len + 6;
  vs.cap = vrows;
  vs.sub_cap = sub_cap;
  vs.ctr = pb.c.queue + 3;
  vs.B = pb.B; vs.D = pb.D; vs.k = pb.c.k;
  vs.bflag = vb.flag; vs.bs_log = vb.bs_log;
  a.bflag = vb.flag; a.bs_log = vb.bs_log;
  // beyond the large LDS table's fill limit a walk would overflow into the one-workgroup HBM path
  const SplitSel ss{paths, VS_MIN_PATHS, (V_LARGE_CAP / 4) * 3 - V_LARGE_TEAM, vrows, const_cast<uint32_t*>(vs.list)};
  auto round = [&]() -> int {
    hipLaunchKernelGGL(k_vblocks_prepare, dim3(vb.nb), dim3(256), 0, ctx->stream, vb, pb.state, pb.B, pb.D, ns, pb.c);
    hipLaunchKernelGGL(k_vblocks_select, dim3(vb.nb), dim3(256), 0, ctx->stream, vb, pb.D, pb.B, pb.state, ns, pb.work,
                       small_max, small_list, large_list, ss, pb.c);
    // walks: phase 1 of the split walks (2 * num_cus workgroups), wavefront teams (3 * num_cus
    // workgroups of 4) and workgroup teams (2 * num_cus) in one launch; then the HBM-row walks of
    // what overflowed together with phase 2 of the split walks (8 * num_cus workgroups)
    PeelVArgs s = a;
    s.list = small_list; s.list_len = pb.c.len + 0; s.queue = pb.c.queue + 0;
    s.ovf_list = ovf_list; s.ovf_len = pb.c.len + 2;
    PeelVArgs l = s;
    l.list = large_list; l.list_len = pb.c.len + 1; l.queue = pb.c.queue + 1;
    PeelVArgs h = s;
    h.list = ovf_list; h.list_len = pb.c.len + 2; h.queue = pb.c.queue + 2;
    const uint32_t chunk_blocks = (uint32_t)ctx->num_cus * 2, small_blocks = (uint32_t)ctx->num_cus * 3;
    hipLaunchKernelGGL(k_peelv_walks, dim3(chunk_blocks + small_blocks + ctx->num_cus * 2), dim3(256), 0, ctx->stream, s, l,
                       vs, chunk_blocks, small_blocks);
    hipLaunchKernelGGL(k_vsplit_units, dim3(V_HBM_BLOCKS + ctx->num_cus * 8), dim3(256), 0, ctx->stream, vs, h);
    hipLaunchKernelGGL(k_vsplit_credit, dim3(credit_grid), dim3(64 * VS_CREDIT_WAVES), 0, ctx->stream, vs);
    PBF_HIP(hipGetLastError());
    return PBF_OK;
  };
  int64_t rounds = 0;
  PBF_CHECK(run_rounds(ctx, pb, ns, round, &rounds));
  ctx->timing.rounds = rounds;
  PBF_HIP(hipMemcpyAsync(tip, pb.D, 8 * (size_t)ns, hipMemcpyDeviceToDevice, ctx->stream));
  return PBF_OK;
}

int pbf_dev_peel_edge(pbf_context* ctx, const DevGraph& g, int use_v, const uint64_t* counts, uint64_t* wing) {
  const int64_t m = g.m;
  if (m == 0) return PBF_OK;
  PeelBufs pb;
  PBF_CHECK(peel_bufs(ctx, m, &pb));
  void* ep;
  PBF_CHECK(pbf_scratch(ctx, S_ETI, 4 * (size_t)m, &ep));
  uint32_t* eti = static_cast<uint32_t*>(ep);
  PBF_CHECK(pbf_dev_edge_to_idxs(ctx, g, use_v, eti));
  const int64_t no = use_v ? g.nv : g.nu, ns = use_v ? g.nu : g.nv;
  // compaction buffers: two sets of (offO, nbrO, idO, offS, nbrS, idS), plus flags and positions
  const size_t set_words = size_t(no + 1) + size_t(ns + 1) + 4 * size_t(m);
  void* cp;
  PBF_CHECK(pbf_scratch(ctx, S_PEEL7, 4 * (2 * set_words + 2 * size_t(m + 1) + size_t(no + ns)), &cp));
  struct View { uint32_t *offO, *nbrO, *idO, *offS, *nbrS, *idS; };
  View set[2];
  uint32_t* w = static_cast<uint32_t*>(cp);
  for (int k = 0; k < 2; k++) {
    set[k].offO = w; w += no + 1;
    set[k].nbrO = w; w += m;
    set[k].idO = w; w += m;
    set[k].offS = w; w += ns + 1;
    set[k].nbrS = w; w += m;
    set[k].idS = w; w += m;
  }
  uint32_t* flag = w;
  uint32_t* pos = w + m + 1;
  uint32_t* hopO = pos + m + 1;  // two-hop work per O vertex, then per S vertex
  uint32_t* hopS = hopO + no;
  // the first view is the input CSR, edge ids = O-CSR positions (O side) / eti (S side)
  View cur{const_cast<uint32_t*>(use_v ? g.offV : g.offU), const_cast<uint32_t*>(use_v ? g.edgV : g.edgU), set[0].idO,
           const_cast<uint32_t*>(use_v ? g.offU : g.offV), const_cast<uint32_t*>(use_v ? g.edgU : g.edgV), eti};
  int next = 1;
  hipLaunchKernelGGL(k_iota, dim3(grid_for(m, 256, 16384)), dim3(256), 0, ctx->stream, set[0].idO, m);
  PeelEArgs a{};
  a.u_of = use_v ? g.edgV : g.edgU;
  a.owner_o = pb.owner_o;
  a.work = pb.work;
  a.state = pb.state;
  a.B = pb.B;
  a.tflag = pb.tflag;
  a.touched = pb.touched;
  a.ntouched = pb.c.len + 3;
  a.k = pb.c.k;
  auto use_view = [&](const View& v) {
    a.offO = v.offO; a.nbrO = v.nbrO; a.idO = v.idO;
    a.offS = v.offS; a.nbrS = v.nbrS; a.idS = v.idS;
  };
  use_view(cur);
  hipLaunchKernelGGL(k_init_peel, dim3(grid_for(m, 256, 16384)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<const unsigned long long*>(counts), pb.B, pb.D, pb.state, pb.tflag, m);
  hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, ctx->stream, pb.c);
  // per-edge walk cost and starting end over the current view
  auto estimate = [&](uint32_t* owner) -> int {
    hipLaunchKernelGGL(k_hop_work, dim3(grid_for(no, 4, 16384)), dim3(256), 0, ctx->stream, cur.offO, cur.nbrO, cur.offS,
                       no, hopO);
    hipLaunchKernelGGL(k_hop_work, dim3(grid_for(ns, 4, 16384)), dim3(256), 0, ctx->stream, cur.offS, cur.nbrS, cur.offO,
                       ns, hopS);
    hipLaunchKernelGGL(k_peele_work, dim3(grid_for(no, 4, 16384)), dim3(256), 0, ctx->stream, cur.offO, cur.nbrO, cur.idO,
                       no, hopO, hopS, pb.work, owner);
    PBF_HIP(hipGetLastError());
    return PBF_OK;
  };
  PBF_CHECK(estimate(pb.owner_o));
  const uint32_t small_max = 1024;  // walk work up to which one wavefront handles the edge
  uint32_t* small_list = pb.lists;
  uint32_t* large_list = pb.lists + m;
  int64_t entries = m;  // entries of the current view: every edge that may still be live
  auto round = [&]() -> int {
    const unsigned sel_grid = grid_for(entries, 256, 8192);
    hipLaunchKernelGGL(k_round_begin, dim3(1), dim3(64), 0, ctx->stream, pb.c);
    hipLaunchKernelGGL(k_select_touched, dim3(ctx->num_cus * 4), dim3(256), 0, ctx->stream, pb.D, pb.B, pb.state, pb.touched,
                       pb.work, small_max, small_list, large_list, SplitSel{}, pb.c);
    hipLaunchKernelGGL(k_min_live, dim3(grid_for(entries, 1024, 2048)), dim3(256), 0, ctx->stream, pb.D, pb.state, cur.idO,
                       entries, pb.c);
    hipLaunchKernelGGL(k_select, dim3(sel_grid), dim3(256), 0, ctx->stream, pb.D, pb.B, pb.state, cur.idO, entries, pb.work,
                       small_max, small_list, large_list, SplitSel{}, pb.c);
    PeelEArgs s = a;
    s.list = small_list; s.list_len = pb.c.len + 0; s.queue = pb.c.queue + 0;
    hipLaunchKernelGGL((k_peele_walk<E_SMALL_TEAM, E_SMALL_UCAP>), dim3(ctx->num_cus * E_SMALL_PER_CU), dim3(E_SMALL_TEAM), 0,
                       ctx->stream, s);
    s.list = large_list; s.list_len = pb.c.len + 1; s.queue = pb.c.queue + 1;
    hipLaunchKernelGGL((k_peele_walk<E_LARGE_TEAM, E_UCAP>), dim3(ctx->num_cus * E_LARGE_PER_CU), dim3(E_LARGE_TEAM), 0,
                       ctx->stream, s);
    hipLaunchKernelGGL(k_update_retire, dim3(std::max(grid_for(entries, 256, 4096), sel_grid)), dim3(256), 0,
                       ctx->stream, pb.touched, pb.tflag, pb.B, pb.D, pb.state, cur.idO, entries, pb.c);
    PBF_HIP(hipGetLastError());
    return PBF_OK;
  };
  // drops the dead entries of both sides into the other buffer set, then re-estimates the work
  auto compact = [&]() -> int {
    View& nx = set[next];
    const int64_t sides[2] = {no, ns};
    for (int side = 0; side < 2; side++) {
      const uint32_t* off = side ? cur.offS : cur.offO;
      const uint32_t* nbr = side ? cur.nbrS : cur.nbrO;
      const uint32_t* id = side ? cur.idS : cur.idO;
      const unsigned grid = grid_for(entries + 1, 256, 16384);
      hipLaunchKernelGGL(k_live_flags, dim3(grid), dim3(256), 0, ctx->stream, id, entries, pb.state, flag);
      PBF_CHECK(pbf_scan_excl_u32(ctx, flag, pos, entries + 1, nullptr));
      hipLaunchKernelGGL(k_compact, dim3(grid), dim3(256), 0, ctx->stream, nbr, id, flag, pos, entries,
                         side ? nx.nbrS : nx.nbrO, side ? nx.idS : nx.idO);
      hipLaunchKernelGGL(k_new_off, dim3(grid_for(sides[side] + 1, 256, 16384)), dim3(256), 0, ctx->stream, off,
                         sides[side], pos, side ? nx.offS : nx.offO);
    }
    PBF_HIP(hipGetLastError());
    cur = nx;
    next ^= 1;
    use_view(cur);
    return estimate(nullptr);
  };
  ctx->pinned[33] = (uint64_t)m;
  PBF_HIP(hipMemcpyAsync(pb.c.live, ctx->pinned + 33, 8, hipMemcpyHostToDevice, ctx->stream));
  int64_t rounds = 0;
  uint64_t live = (uint64_t)m;
  // every round peels at least one live edge, so m rounds always suffice
  while (rounds <= m) {
    for (int r = 0; r < ROUND_BATCH; r++) PBF_CHECK(round());
    rounds += ROUND_BATCH;
    PBF_HIP(hipMemcpyAsync(ctx->pinned + 32, pb.c.live, 8, hipMemcpyDeviceToHost, ctx->stream));
    PBF_HIP(hipStreamSynchronize(ctx->stream));
    live = ctx->pinned[32];
    if (live == 0) break;
    // the live edges are exactly the state-1 entries after a round; compact at half occupancy
    if (2 * (int64_t)live < entries) {
      PBF_CHECK(compact());
      entries = (int64_t)live;
    }
  }
  ctx->timing.rounds = rounds;
  if (live != 0) return PBF_ERR_DOMAIN;
  PBF_HIP(hipMemcpyAsync(wing, pb.D, 8 * (size_t)m, hipMemcpyDeviceToDevice, ctx->stream));
  return PBF_OK;
}
