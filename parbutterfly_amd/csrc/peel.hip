// Butterfly peeling on the GPU: tip numbers (vertices of S) and wing numbers (edges).
//
// Semantics follow the reference's bucketed peeling (Peel / PeelOrigParallel, apps/hyper/
// butterfly_peel.h:647-711, 896-934; PeelE / PeelEOrigParallel, butterfly_epeel.h:339-443,
// 519-569; updateBuckets, butterfly_putils.h:1027-1115; increasing buckets of ligra/bucket.h):
//   repeat: k = min D over live items; every live item with D == k is peeled in this round;
//   the peeled items' butterflies are subtracted from the live items they share them with;
//   every updated live item gets D = max(B, k).  The tip / wing number is D at removal.
// The reference's bucket structure is replaced by min-count frontiers: per round a min
// reduction, a selection of the items with D == k, a walk of the frontier with 64-bit atomic
// decrements, and an update of the touched items.  Rounds run in batches on the device; the
// host reads the live count once per batch.
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

// Every control word, list entry and walk target a kernel here reads was written by an earlier
// launch on the stream: no kernel reads what another workgroup of the same launch stored, since
// within a launch workgroups share data only through device-scope atomics whose returned values
// carry it (queue claims, list reservations).  The end of the writing launch and the start of
// the reading one order those writes before any load, plain or atomic.  The edge peel reads
// them with relaxed agent-scope atomic loads (vector loads; a plain load of a uniform address
// would be a scalar-cache load), the form it has been tested with.
__device__ __forceinline__ uint32_t dev_ld(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long dev_ld(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The tip peel's kernels read what earlier launches wrote with ordinary vector loads, which may
// hit in L2 (an agent-scope atomic load bypasses it); a control word at a uniform address is
// read as a workgroup-scope atomic, so that it stays a vector load too rather than a
// scalar-cache load.
__device__ __forceinline__ uint32_t ctl_ld(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ unsigned long long ctl_ld(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

inline unsigned grid_for(int64_t n, int threads, int max_blocks) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)b;
}

// ---------------- shared round kernels ----------------
// ctl layout (u64): [0] k, [1] live count, [2] k of the previous round; u32 view from word 8:
// list lengths, then queues.  The vertex peel uses len[0..2] its own way (k_vpeel_*).
struct Ctl {
  unsigned long long* k;
  unsigned long long* live;
  unsigned long long* kprev;
  uint32_t* len;    // edge peel: [0] small, [1] large, [3] ntouched, [4] previous ntouched, [5] found
  uint32_t* queue;  // edge peel: [0] small, [1] large; reset per round
};
constexpr int VS_MAX_ROWS = 32, CTL_QUEUES = 8;

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

// Appends the x[i] whose f[i] is set, all N of every lane with one counter atomic per
// wavefront.  Every lane of the wavefront calls it together.
template <uint32_t N>
__device__ __forceinline__ void wave_append_n(const uint32_t (&x)[N], const bool (&f)[N], uint32_t* list, uint32_t* len) {
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t i = 0; i < N; i++) cnt += f[i] ? 1u : 0u;
  const uint32_t incl = pbf_wave_incl_scan(cnt);
  const uint32_t tot = (uint32_t)__shfl((int)incl, 63, 64);
  if (tot == 0u) return;
  uint32_t base = 0;
  if (__lane_id() == 0) base = atomicAdd(len, tot);
  base = (uint32_t)__shfl((int)base, 0, 64) + incl - cnt;
#pragma unroll
  for (uint32_t i = 0; i < N; i++)
    if (f[i]) list[base++] = x[i];
}

// The round kernels visit items ids[0, n) (every item that may still be live), or 0..n-1
// when ids is null.  The selection kernels below take SEL_PER items per thread at a time (items p0 + 256 i + tid
// of a workgroup step), every load of a stage in flight together; a step's loop bound is the
// same for the whole workgroup, so every lane takes part in the wavefront appends.
constexpr uint32_t SEL_PER = 4;

// Marks the selected items current (state 2) and lists them by work class; returns how many.
__device__ __forceinline__ uint32_t select_items(const uint32_t (&x)[SEL_PER], const bool (&in)[SEL_PER],
                                                 unsigned long long k, const unsigned long long* __restrict__ D,
                                                 const unsigned long long* __restrict__ B, uint8_t* __restrict__ state,
                                                 const uint32_t* __restrict__ work, uint32_t small_max,
                                                 uint32_t* __restrict__ small_list, uint32_t* __restrict__ large_list,
                                                 uint32_t* len) {
  uint8_t st[SEL_PER];
  unsigned long long d[SEL_PER], bx[SEL_PER];
  uint32_t wk[SEL_PER];
#pragma unroll
  for (uint32_t i = 0; i < SEL_PER; i++) {
    st[i] = state[x[i]];
    d[i] = D[x[i]];
    bx[i] = B[x[i]];
    wk[i] = work[x[i]];
  }
  uint32_t nsel = 0;
  bool fs[SEL_PER], fl[SEL_PER];
#pragma unroll
  for (uint32_t i = 0; i < SEL_PER; i++) {
    const bool sel = in[i] & (st[i] == 1) & (d[i] == k);
    if (sel) { state[x[i]] = 2; nsel++; }
    // an item in no remaining butterfly (B == 0) has nothing to decrement: no walk
    const uint32_t w = sel && bx[i] ? (wk[i] & WORK_MASK) : 0u;
    fs[i] = w != 0u && w <= small_max;
    fl[i] = w > small_max;
  }
  wave_append_n<SEL_PER>(x, fs, small_list, len + 0);
  wave_append_n<SEL_PER>(x, fl, large_list, len + 1);
  return nsel;
}

// Frontier from the previous round's touched items.  After a round at level k every live item
// has D >= k, and the live items with D == k are all touched ones (an untouched item keeps its
// D, which was not k or it would have been peeled).  So when a touched item has D == k the new
// minimum is k again and the frontier is exactly those items; else the full scan below runs.
__global__ __launch_bounds__(256) void k_select_touched(const unsigned long long* __restrict__ D,
                                                        const unsigned long long* __restrict__ B, uint8_t* __restrict__ state,
                                                        const uint32_t* __restrict__ touched, const uint32_t* __restrict__ work,
                                                        uint32_t small_max, uint32_t* __restrict__ small_list,
                                                        uint32_t* __restrict__ large_list, Ctl c) {
  const unsigned long long k = dev_ld(c.kprev);
  const uint32_t nt = dev_ld(c.len + 4);
  if (k == KNONE) return;
  uint32_t nsel = 0;
  for (uint32_t t0 = blockIdx.x * 256u * SEL_PER; t0 < nt; t0 += gridDim.x * 256u * SEL_PER) {
    uint32_t x[SEL_PER];
    bool in[SEL_PER];
#pragma unroll
    for (uint32_t i = 0; i < SEL_PER; i++) {
      const uint32_t t = t0 + 256u * i + threadIdx.x;
      in[i] = t < nt;
      x[i] = touched[t < nt ? t : nt - 1u];
    }
    nsel += select_items(x, in, k, D, B, state, work, small_max, small_list, large_list, c.len);
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
  for (int64_t p0 = (int64_t)blockIdx.x * 256 * SEL_PER; p0 < n; p0 += (int64_t)gridDim.x * 256 * SEL_PER) {
    uint32_t x[SEL_PER];
#pragma unroll
    for (uint32_t i = 0; i < SEL_PER; i++) {
      const int64_t p = p0 + 256 * i + threadIdx.x;
      const int64_t pc = p < n ? p : n - 1;
      x[i] = ids ? ids[pc] : (uint32_t)pc;
    }
#pragma unroll
    for (uint32_t i = 0; i < SEL_PER; i++) {
      const uint8_t st = state[x[i]];
      const unsigned long long d = D[x[i]];
      if (st) mn = d < mn ? d : mn;  // an index past n repeats item n - 1
    }
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
__global__ __launch_bounds__(256) void k_select(const unsigned long long* __restrict__ D,
                                                const unsigned long long* __restrict__ B, uint8_t* __restrict__ state,
                                                const uint32_t* __restrict__ ids, int64_t n,
                                                const uint32_t* __restrict__ work, uint32_t small_max,
                                                uint32_t* __restrict__ small_list, uint32_t* __restrict__ large_list, Ctl c) {
  const unsigned long long k = dev_ld(c.k);
  const uint32_t found = dev_ld(c.len + 5);
  if (k == KNONE || found) return;
  uint32_t nsel = 0;
  for (int64_t p0 = (int64_t)blockIdx.x * 256 * SEL_PER; p0 < n; p0 += (int64_t)gridDim.x * 256 * SEL_PER) {
    uint32_t x[SEL_PER];
    bool in[SEL_PER];
#pragma unroll
    for (uint32_t i = 0; i < SEL_PER; i++) {
      const int64_t p = p0 + 256 * i + threadIdx.x;
      in[i] = p < n;
      const int64_t pc = p < n ? p : n - 1;
      x[i] = ids ? ids[pc] : (uint32_t)pc;
    }
    nsel += select_items(x, in, k, D, B, state, work, small_max, small_list, large_list, c.len);
  }
  nsel = pbf_wave_sum(nsel);
  if ((threadIdx.x & 63) == 0 && nsel) atomicAdd(c.live, (unsigned long long)(-(long long)nsel));
}

// Touched live items: D = max(B, k) (updateBuckets, butterfly_putils.h:1036-1043).
// End of a round, one launch for two independent passes: touched items get D[x] = max(B[x], k)
// (and lose their flag), current items (state 2) die.  A touched item is live, never current.
__global__ __launch_bounds__(256) void k_update_retire(const uint32_t* __restrict__ touched, uint32_t* __restrict__ tflag,
                                                       const unsigned long long* __restrict__ B,
                                                       unsigned long long* __restrict__ D, uint8_t* __restrict__ state,
                                                       const uint32_t* __restrict__ ids, int64_t n, Ctl c) {
  const unsigned long long k = dev_ld(c.k);
  const uint32_t nt = dev_ld(c.len + 3);
  for (uint32_t t0 = blockIdx.x * 256u * SEL_PER; t0 < nt; t0 += gridDim.x * 256u * SEL_PER) {
    uint32_t x[SEL_PER];
#pragma unroll
    for (uint32_t i = 0; i < SEL_PER; i++) {
      const uint32_t t = t0 + 256u * i + threadIdx.x;
      x[i] = touched[t < nt ? t : nt - 1u];
    }
    unsigned long long bx[SEL_PER];
#pragma unroll
    for (uint32_t i = 0; i < SEL_PER; i++) bx[i] = B[x[i]];
#pragma unroll
    for (uint32_t i = 0; i < SEL_PER; i++) {
      if (t0 + 256u * i + threadIdx.x < nt) {
        tflag[x[i]] = 0;
        D[x[i]] = bx[i] > k ? bx[i] : k;
      }
    }
  }
  if (k == KNONE) return;
  for (int64_t p0 = (int64_t)blockIdx.x * 256 * SEL_PER; p0 < n; p0 += (int64_t)gridDim.x * 256 * SEL_PER) {
    uint32_t x[SEL_PER];
#pragma unroll
    for (uint32_t i = 0; i < SEL_PER; i++) {
      const int64_t p = p0 + 256 * i + threadIdx.x;
      const int64_t pc = p < n ? p : n - 1;
      x[i] = ids ? ids[pc] : (uint32_t)pc;
    }
    uint8_t st[SEL_PER];
#pragma unroll
    for (uint32_t i = 0; i < SEL_PER; i++) st[i] = state[x[i]];
#pragma unroll
    for (uint32_t i = 0; i < SEL_PER; i++)
      if (st[i] == 2) state[x[i]] = 0;  // repeated items (past n) write the same 0
  }
}

// ---------------- vertex (tip) peel ----------------
// A round is two launches.
//   k_vpeel_prepare, one workgroup per block of S items (at most 256 blocks of at least 1024):
//     retires the previous round's frontier, subtracts C(c, 2) for the counts the previous
//     round's split walks left in their rows, sets D = max(B, k) on the live items of every
//     block that changed (updateBuckets, butterfly_putils.h:1036-1043), and keeps per block the
//     exact minimum of D over its live items, how many items reach it and the first NARG of them.
//   k_vpeel_round, every workgroup: reduces the block minima to the level k, builds the
//     frontier (live items with D == k: at most FR_CAP of them, at most vrows of the split
//     class) from the blocks holding k, the same list in every workgroup, and walks it.
// An item of the split class (many two-hop paths, or more possible two-hop neighbours than an
// LDS table holds) is walked by many workgroups: its centres are cut into chunks of 256 and the
// paths of a chunk into units of at most 256, numbered by static tables (sv, cu); the counts
// c(u, x) go to a dense row of the item in HBM and reach B in the next prepare pass.  Small and
// large items are counted in LDS tables by one wavefront or one workgroup, which subtract
// C(c, 2) from B directly.
// Items at level k that a round leaves out (the caps) keep D == k and are taken by the next
// rounds, still at level k.  A butterfly has exactly two S vertices, so removing the level's
// items in parts subtracts from every other item x the same C(c(u, x), 2) per removed u as
// removing them at once, and no item with D <= k is ever decremented: the tip numbers are
// those of PeelOrigParallel.
// Both launches are short chains of dependent loads (a round is one level, and most levels of a
// power-law graph peel one vertex), so the code keeps those chains short: static per-item
// records, block scans with every load of a step issued together, no flags on the split rows.
constexpr int V_SMALL_TEAM = 64, V_SMALL_CAP = 1024, V_LARGE_TEAM = 256, V_LARGE_CAP = 8192;
constexpr uint32_t FR_CAP = 1024;     // frontier items per round (4 per thread of a workgroup)
constexpr uint32_t NARG = 8;          // items at a block's minimum that the prepare pass lists
constexpr uint32_t VS_PCH = 256;      // paths per split unit, one per thread
constexpr uint32_t VS_MIN_PATHS = 1u << 15;
constexpr uint32_t V_ROUND_PER_CU = 2;
constexpr uint32_t PREP_PER = 4;      // items per thread in one step of the prepare pass
constexpr uint32_t WALK_PER = 8;      // paths per thread in flight in an LDS walk
constexpr uint32_t VS_BATCH = 8;      // paths per thread in flight in a split walk
constexpr uint32_t V_SMALL_MAX = (V_SMALL_CAP / 4) * 3 - V_SMALL_TEAM;  // the tables' fill limits
constexpr uint32_t V_LARGE_MAX = (V_LARGE_CAP / 4) * 3 - V_LARGE_TEAM;
enum FrontierClass { FC_NONE = 0, FC_SMALL = 1, FC_LARGE = 2, FC_SPLIT = 3 };

struct VPeel {
  const uint32_t* offS;
  const uint32_t* edgS;
  const uint32_t* offO;
  const uint32_t* edgO;
  const uint2* sv;           // per S slot (u -> v_j): (two-hop paths of u through v_0 .. v_{j-1}, offO[v_j])
  const uint32_t* cu;        // per chunk of 256 centres of an item: the item's units before it
  const uint4* rec;          // per S item u: offS[u], deg(u), two-hop paths T(u), class by work
  const uint2* rec2;         // per S item u: its first chunk in cu, its units
  int64_t ns;
  unsigned long long* B;
  unsigned long long* D;
  uint8_t* state;            // 1 live, 0 retired
  unsigned long long* vmin;  // [nb] minimum of D over the block's live items (KNONE: none)
  uint32_t* vcnt;            // [nb] items at that minimum, at most NARG + 1
  uint32_t* varg;            // [nb * NARG] the first of them, ascending
  uint32_t* vflag;           // [nb] 1: the block has decremented items (or needs its first scan)
  uint32_t nb, bs_log;
  uint32_t* fr;              // [FR_CAP] the round's frontier (~0u: left for a later round)
  uint32_t* ctl;             // [0] frontier entries, [1] split rows of the round, [2] error flag
  unsigned long long* k;     // level of the last round (KNONE before the first)
  unsigned long long* live;
  uint32_t* rows;            // [vrows][ns] split-walk counts, zero outside a round
  uint32_t vrows;
};

// Retire, credit, update, block minimum.  A block of at most 256 * PREP_PER items is one step:
// everything the pass needs (control words, the previous frontier, the items' state, D and B,
// the first split row's counts) is loaded at once, before anything is known, so the pass waits
// for one round trip to memory; its items at the minimum are then listed from registers.
// Larger blocks take the general path below, a step at a time.
__global__ __launch_bounds__(256) void k_vpeel_prepare(VPeel a) {
  __shared__ unsigned long long s_min[4];
  __shared__ uint32_t s_wc[PREP_PER][4];
  __shared__ uint32_t s_ret[PREP_PER * 8];  // retired items of a one-step block, a bit each
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
  const uint32_t b = blockIdx.x;
  const int64_t lo = (int64_t)b << a.bs_log;
  const int64_t hi = lo + (int64_t(1) << a.bs_log) < a.ns ? lo + (int64_t(1) << a.bs_log) : a.ns;
  const bool one = hi - lo <= 256 * (int64_t)PREP_PER;
  // the block's first 256 * PREP_PER items (state, D, B, row 0's count, whether the last round
  // had split rows or not), at clamped indices (hi > lo) and before the control words, so that
  // every load is in flight at once; they are masked only once all are issued
  uint8_t st[PREP_PER];
  unsigned long long d[PREP_PER], bx[PREP_PER], sub[PREP_PER];
  uint32_t c0[PREP_PER];
#pragma unroll
  for (uint32_t i = 0; i < PREP_PER; i++) {
    const int64_t xx = lo + 256 * i + tid;
    const int64_t xc = xx < hi ? xx : hi - 1;
    st[i] = a.state[xc];
    d[i] = a.D[xc];
    bx[i] = a.B[xc];
    c0[i] = a.rows[xc];
  }
  if (tid < PREP_PER * 8) s_ret[tid] = 0u;
  __syncthreads();
  const unsigned long long k = ctl_ld(a.k);
  const uint32_t nfr = ctl_ld(a.ctl + 0), nrows = ctl_ld(a.ctl + 1), flag = ctl_ld(a.vflag + b);
  const uint32_t f0 = a.fr[tid];  // FR_CAP >= 256 entries
  if (one) {
#pragma unroll
    for (uint32_t i = 0; i < PREP_PER; i++) {
      if (lo + 256 * i + tid >= hi) { st[i] = 0; d[i] = 0ull; bx[i] = 0ull; c0[i] = 0u; }
    }
    // the previous round's frontier items of this block retire (their D is their tip number)
    uint32_t ret = 0;
    for (uint32_t j = tid; j < nfr; j += 256) {
      const uint32_t x = j < 256 ? f0 : a.fr[j];
      if ((int64_t)x >= lo && (int64_t)x < hi) {
        a.state[x] = 0;
        atomicOr(&s_ret[(x - (uint32_t)lo) >> 5], 1u << (x & 31u));
        ret++;
      }
    }
    ret = pbf_wave_sum(ret);
    if (lane == 0 && ret) atomicAdd(a.live, (unsigned long long)(-(long long)ret));
    const bool need = __syncthreads_or(ret != 0 || flag != 0 || nrows != 0);
    unsigned long long mn = KNONE;
    if (need) {
#pragma unroll
      for (uint32_t i = 0; i < PREP_PER; i++) {
        const uint32_t r = 256 * i + tid;
        if ((s_ret[r >> 5] >> (r & 31u)) & 1u) st[i] = 0;
        sub[i] = 0;
        if (nrows && c0[i]) {
          a.rows[lo + r] = 0u;
          sub[i] = (unsigned long long)c0[i] * (c0[i] - 1u) / 2u;
        }
      }
      for (uint32_t rr = 1; rr < nrows; rr++) {
        uint32_t* row = a.rows + (size_t)rr * (size_t)a.ns;
        uint32_t c[PREP_PER];
#pragma unroll
        for (uint32_t i = 0; i < PREP_PER; i++) {
          const int64_t xx = lo + 256 * i + tid;
          const uint32_t cv = row[xx < hi ? xx : hi - 1];
          c[i] = xx < hi ? cv : 0u;
        }
#pragma unroll
        for (uint32_t i = 0; i < PREP_PER; i++) {
          if (c[i]) {
            row[lo + 256 * i + tid] = 0u;
            sub[i] += (unsigned long long)c[i] * (c[i] - 1u) / 2u;
          }
        }
      }
#pragma unroll
      for (uint32_t i = 0; i < PREP_PER; i++) {
        if (st[i] != 1) continue;
        const int64_t xx = lo + 256 * i + tid;
        // as the LDS walks: items at or below the level (the rest of the level) keep B
        if (sub[i] && d[i] > k) { bx[i] -= sub[i]; a.B[xx] = bx[i]; }
        if (k != KNONE) {
          const unsigned long long nd = bx[i] > k ? bx[i] : k;
          if (nd != d[i]) { a.D[xx] = nd; d[i] = nd; }
        }
        mn = d[i] < mn ? d[i] : mn;
      }
    }
    if (!need) return;
    mn = wave_min_u64(mn);
    if (lane == 0) s_min[wid] = mn;
    __syncthreads();
    mn = s_min[0];
    for (int w = 1; w < 4; w++) mn = s_min[w] < mn ? s_min[w] : mn;
    uint32_t found = 0;
    if (mn != KNONE) {
      uint64_t bm[PREP_PER];
#pragma unroll
      for (uint32_t i = 0; i < PREP_PER; i++) {
        bm[i] = __ballot(st[i] == 1 && d[i] == mn);
        if (lane == 0) s_wc[i][wid] = (uint32_t)__popcll(bm[i]);
      }
      __syncthreads();
#pragma unroll
      for (uint32_t i = 0; i < PREP_PER; i++) {
        uint32_t before = found;
        for (uint32_t w = 0; w < 4; w++) {
          if (w < wid) before += s_wc[i][w];
          found += s_wc[i][w];
        }
        if ((bm[i] >> lane) & 1ull) {
          const uint32_t pos = before + (uint32_t)__popcll(bm[i] & pbf_lanemask_lt());
          if (pos < NARG) a.varg[(size_t)b * NARG + pos] = (uint32_t)(lo + 256 * i + tid);
        }
      }
    }
    if (tid == 0) {
      __hip_atomic_store(a.vmin + b, mn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(a.vcnt + b, found < NARG + 1 ? found : NARG + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(a.vflag + b, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return;
  }
  // general path (blocks of more than 256 * PREP_PER items): retire, then a step at a time
  uint32_t ret = 0;
  for (uint32_t j = tid; j < nfr; j += 256) {
    const uint32_t x = j < 256 ? f0 : a.fr[j];
    if ((int64_t)x >= lo && (int64_t)x < hi) { a.state[x] = 0; ret++; }
  }
  ret = pbf_wave_sum(ret);
  if (lane == 0 && ret) atomicAdd(a.live, (unsigned long long)(-(long long)ret));
  if (!__syncthreads_or(ret != 0 || flag != 0 || nrows != 0)) return;
  unsigned long long mn = KNONE;
  for (int64_t x0 = lo; x0 < hi; x0 += 256 * PREP_PER) {
#pragma unroll
    for (uint32_t i = 0; i < PREP_PER; i++) {
      const int64_t xx = x0 + 256 * i + tid;
      const bool in = xx < hi;
      const int64_t xc = in ? xx : hi - 1;
      const uint8_t s0 = a.state[xc];
      const unsigned long long d0 = a.D[xc], b0 = a.B[xc];
      st[i] = in ? s0 : 0;
      d[i] = in ? d0 : 0ull;
      bx[i] = in ? b0 : 0ull;
      sub[i] = 0;
    }
    for (uint32_t r = 0; r < nrows; r++) {
      uint32_t* row = a.rows + (size_t)r * (size_t)a.ns;
      uint32_t c[PREP_PER];
#pragma unroll
      for (uint32_t i = 0; i < PREP_PER; i++) {
        const int64_t xx = x0 + 256 * i + tid;
        const uint32_t cv = row[xx < hi ? xx : hi - 1];
        c[i] = xx < hi ? cv : 0u;
      }
#pragma unroll
      for (uint32_t i = 0; i < PREP_PER; i++) {
        if (c[i]) {
          row[x0 + 256 * i + tid] = 0u;
          sub[i] += (unsigned long long)c[i] * (c[i] - 1u) / 2u;
        }
      }
    }
#pragma unroll
    for (uint32_t i = 0; i < PREP_PER; i++) {
      if (st[i] != 1) continue;
      const int64_t xx = x0 + 256 * i + tid;
      if (sub[i] && d[i] > k) { bx[i] -= sub[i]; a.B[xx] = bx[i]; }
      if (k != KNONE) {
        const unsigned long long nd = bx[i] > k ? bx[i] : k;
        if (nd != d[i]) { a.D[xx] = nd; d[i] = nd; }
      }
      mn = d[i] < mn ? d[i] : mn;
    }
  }
  mn = wave_min_u64(mn);
  if (lane == 0) s_min[wid] = mn;
  __syncthreads();
  mn = s_min[0];
  for (int w = 1; w < 4; w++) mn = s_min[w] < mn ? s_min[w] : mn;
  // the first NARG live items at the minimum, ascending; counting stops past NARG
  uint32_t found = 0;
  if (mn != KNONE) {
    for (int64_t x0 = lo; x0 < hi && found <= NARG; x0 += 256) {
      const int64_t xx = x0 + tid;
      const int64_t xc = xx < hi ? xx : hi - 1;
      const uint8_t s0 = a.state[xc];
      const unsigned long long d0 = a.D[xc];
      const bool sel = (xx < hi) & (s0 == 1) & (d0 == mn);  // no short circuit: both loads in flight
      const uint64_t bm = __ballot(sel);
      __syncthreads();  // s_wc of the previous step has been read
      if (lane == 0) s_wc[0][wid] = (uint32_t)__popcll(bm);
      __syncthreads();
      uint32_t before = found, tot = 0;
      for (uint32_t w = 0; w < 4; w++) {
        if (w < wid) before += s_wc[0][w];
        tot += s_wc[0][w];
      }
      if (sel) {
        const uint32_t pos = before + (uint32_t)__popcll(bm & pbf_lanemask_lt());
        if (pos < NARG) a.varg[(size_t)b * NARG + pos] = (uint32_t)xx;
      }
      found += tot;
    }
  }
  if (tid == 0) {
    __hip_atomic_store(a.vmin + b, mn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(a.vcnt + b, found < NARG + 1 ? found : NARG + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(a.vflag + b, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// A walk team's LDS: its shared scratch and the table's keys, counts and used slots.
template <int TEAM, int CAP, bool USED>
struct WalkLds {
  TeamShared<TEAM> sh;
  uint32_t keys[CAP];
  uint32_t cnts[CAP];
  uint32_t used[USED ? CAP : 1];
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

// The keys of PER paths of a thread, w + i * stride for i < PER, among the paths [0, lim) of a
// step whose N centres start at the paths P[0 .. N) (nondecreasing, P[0] = 0, ~0u past the last
// centre) with centre j's path p at edgO[vs[j] + p]; u where the path is not below lim.  The
// centre is found by a search of fixed depth, one level for all PER paths at a time, and a path
// past lim is searched as the last one, so that every LDS read and every key load is
// unconditional and each level's reads (and all the key loads) are in flight together.
template <uint32_t N, uint32_t PER>
__device__ __forceinline__ void path_keys(const uint32_t* P, const uint32_t* vs, const uint32_t* __restrict__ edgO,
                                          uint32_t w, uint32_t stride, uint32_t lim, uint32_t u, uint32_t (&key)[PER]) {
  uint32_t wq[PER], l[PER];
#pragma unroll
  for (uint32_t i = 0; i < PER; i++) {
    const uint32_t wi = w + i * stride;
    wq[i] = wi < lim ? wi : lim - 1u;
    l[i] = 0;
  }
#pragma unroll
  for (uint32_t step = N / 2; step; step >>= 1) {
#pragma unroll
    for (uint32_t i = 0; i < PER; i++) l[i] = P[l[i] + step] <= wq[i] ? l[i] + step : l[i];
  }
#pragma unroll
  for (uint32_t i = 0; i < PER; i++) {
    const uint32_t kv = edgO[vs[l[i]] + wq[i]];
    key[i] = w + i * stride < lim ? kv : u;
  }
}

// The walk of one frontier item u (its record r) by one team (thread tid of TEAM): every
// two-hop path u -> v -> u2 (u2 != u) counted in an LDS hash (PeelOrigParallel,
// butterfly_peel.h:670-702), then C(c, 2) comes off B[u2] for every u2, and u2's block is flagged
// for the next prepare pass.  The reference decrements only the u2 with D[u2] > k; the others
// need no test here: a live u2 at the level keeps D = max(B, k) = k however B falls (B stays the
// exact count of its butterflies, never below zero), and a retired u2's B is never read again,
// so the sweep issues its atomics without first waiting for a load of D.  USED tables keep a list of their used slots; the
// others are sized from the item's bound on distinct keys, min(T(u), ns), and swept whole.  The
// classes keep every item's distinct keys below the table's fill limit, so a table cannot fill
// up; if one did, the round's error flag is set and the peel fails.  A step takes TEAM centres,
// whose paths are numbered by the slot records: the centre j of the step starts at sv[j].x.
template <int TEAM, int CAP, bool USED>
__device__ void vwalk_item(const VPeel& a, WalkLds<TEAM, CAP, USED>& L, uint32_t u, uint4 r, uint32_t tid) {
  TeamShared<TEAM>& sh = L.sh;
  const uint32_t uo = r.x, deg = r.y, T = r.z;
  uint32_t cap = CAP;
  if constexpr (!USED) {
    const uint32_t bnd = (int64_t)T < a.ns ? T : (uint32_t)a.ns;
    const uint64_t want = (uint64_t)bnd + bnd / 3 + 1;
    cap = 256;
    while (cap < want && cap < CAP) cap <<= 1;
  }
  const uint32_t lg = 31 - __clz(cap);
  LdsTable t;
  t.keys = L.keys; t.cnts = L.cnts; t.used = USED ? L.used : nullptr;
  t.shift = 32 - lg; t.mask = cap - 1;
  t.maxfill = (cap == CAP) ? (CAP / 4) * 3 - TEAM : 0xFFFFFFFFu;
  for (uint32_t s = tid; s < cap; s += TEAM) { L.keys[s] = EMPTY; L.cnts[s] = 0; }
  if (tid == 0) { sh.nused = 0; sh.ovf = 0; }
  walk_barrier<TEAM>();
  for (uint32_t c0 = 0; c0 < deg; c0 += TEAM) {
    // the step's paths [cs, ce) and, per centre, its first path and list start
    const uint32_t cs = a.sv[uo + c0].x;
    const uint32_t ce = c0 + TEAM < deg ? a.sv[uo + c0 + TEAM].x : T;
    const uint32_t j = c0 + tid;
    uint2 e = make_uint2(0xFFFFFFFFu, 0u);
    if (j < deg) e = a.sv[uo + j];
    // path p of centre j is N(v_j)[p - e.x]; with w = p - cs that is (e.y - (e.x - cs)) + w
    sh.P[tid] = e.x == 0xFFFFFFFFu ? e.x : e.x - cs;
    sh.vs[tid] = e.y - (e.x - cs);
    walk_barrier<TEAM>();
    // WALK_PER paths per thread at a time, every key loaded before any is inserted
    for (uint32_t w0 = 0; w0 < ce - cs; w0 += TEAM * WALK_PER) {
      uint32_t key[WALK_PER];
      path_keys<TEAM, WALK_PER>(sh.P, sh.vs, a.edgO, w0 + tid, TEAM, ce - cs, u, key);
#pragma unroll
      for (uint32_t i = 0; i < WALK_PER; i++)
        if (key[i] != u && !sh.ovf) lds_insert<TEAM>(t, sh, key[i]);
    }
    walk_barrier<TEAM>();
  }
  if (sh.ovf) {
    if (tid == 0) a.ctl[2] = 1u;
    walk_barrier<TEAM>();
    return;
  }
  const uint32_t extent = USED ? sh.nused : cap;
  for (uint32_t q = tid; q < extent; q += TEAM) {
    const uint32_t slot = USED ? L.used[q] : q;
    const uint32_t key = L.keys[slot];
    const uint32_t c = key == EMPTY ? 0u : L.cnts[slot];
    if (c > 1u) {
      atomicAdd(&a.B[key], (unsigned long long)(-(long long)((unsigned long long)c * (c - 1u) / 2u)));
      a.vflag[key >> a.bs_log] = 1u;  // the next prepare pass updates D over the block
    }
  }
  walk_barrier<TEAM>();
}

// A split task's LDS: the chunk layout and a table that sums the task's counts per key.
constexpr uint32_t AGG_CAP = 4096, AGG_PROBES = 16;
struct SplitLds {
  TeamShared<256> sh;
  uint32_t keys[AGG_CAP];
  uint32_t cnts[AGG_CAP];
  uint32_t used[AGG_CAP];
};

// Counts key in the task's table (bounded probing); false when no slot was found.
__device__ __forceinline__ bool agg_insert(SplitLds& L, uint32_t key) {
  uint32_t h = hash_slot(key, 32u - 12u);  // AGG_CAP = 2^12 slots
  for (uint32_t probe = 0; probe < AGG_PROBES; probe++) {
    const uint32_t old = atomicCAS(&L.keys[h], EMPTY, key);
    if (old == EMPTY) {
      L.used[atomicAdd(&L.sh.nused, 1u)] = h;
      atomicAdd(&L.cnts[h], 1u);
      return true;
    }
    if (old == key) {
      atomicAdd(&L.cnts[h], 1u);
      return true;
    }
    h = (h + 1u) & (AGG_CAP - 1u);
  }
  return false;
}

// Paths [wa, wb) of the chunk starting at centre c0 of the item (record r4): the chunk's
// centres are laid out in LDS (one per thread) and every thread then takes VS_BATCH paths at a
// time, all their keys loaded before any is counted.  Keys of retired items are dropped (they
// take no credit); the others are summed in the task's table, or counted into row r directly
// when the table has no room for them.
__device__ void vsplit_paths(const VPeel& a, SplitLds& L, uint32_t r, uint32_t u, uint4 r4, uint32_t c0,
                             uint32_t wa, uint32_t wb) {
  TeamShared<256>& sh = L.sh;
  const uint32_t tid = threadIdx.x;
  const uint32_t uo = r4.x, deg = r4.y;
  const uint32_t cs = a.sv[uo + c0].x;
  const uint32_t j = c0 + tid;
  uint2 e = make_uint2(0xFFFFFFFFu, 0u);
  if (j < deg) e = a.sv[uo + j];
  __syncthreads();  // the previous chunk's readers are done with sh
  // path p of centre j is N(v_j)[p - e.x]; with w = p - cs that is (e.y - (e.x - cs)) + w
  sh.P[tid] = e.x == 0xFFFFFFFFu ? e.x : e.x - cs;
  sh.vs[tid] = e.y - (e.x - cs);
  __syncthreads();
  uint32_t* row = a.rows + (size_t)r * (size_t)a.ns;
  for (uint32_t w0 = wa; w0 < wb; w0 += 256u * VS_BATCH) {
    uint32_t key[VS_BATCH];
    path_keys<256, VS_BATCH>(sh.P, sh.vs, a.edgO, w0 + tid, 256u, wb, u, key);
    // unconditional loads (key[i] is u where there is no path): the VS_BATCH loads are in
    // flight together, where a guarded load would wait for each in turn
    uint8_t live[VS_BATCH];
#pragma unroll
    for (uint32_t i = 0; i < VS_BATCH; i++) live[i] = a.state[key[i]];
#pragma unroll
    for (uint32_t i = 0; i < VS_BATCH; i++)
      if (key[i] != u && live[i] == 1 && !agg_insert(L, key[i])) atomicAdd(row + key[i], 1u);
  }
}

// Units [q0, q1) of row r's item u (record r4, chunks from r2.x, r2.y units in all): unit q of
// chunk c (the last chunk with cu[c] <= q) covers the chunk's paths [256 (q - cu[c]), + 256).
// The first chunk is found by a 256-wide probe of cu (a 256-ary search beyond 256 chunks);
// the units of one chunk are walked together.  The task's table starts empty and is added to
// the row at the end, one atomic per key.
__device__ void vsplit_units(const VPeel& a, SplitLds& L, uint32_t r, uint32_t u, uint4 r4, uint2 r2,
                             uint32_t q0, uint32_t q1) {
  TeamShared<256>& sh = L.sh;
  const uint32_t tid = threadIdx.x;
  const uint32_t deg = r4.y, T = r4.z;
  const uint32_t nch = (deg + 255u) / 256u;
  for (uint32_t s = tid; s < AGG_CAP; s += 256) { L.keys[s] = EMPTY; L.cnts[s] = 0; }
  uint32_t lo = 0, len = nch;
  if (tid == 0) { sh.item = 0; sh.nused = 0; }  // cu[lo]
  while (len > 1) {
    const uint32_t step = (len + 255u) / 256u;
    const uint32_t j = lo + tid * step;
    uint32_t cv = 0;
    const bool le = j < lo + len && (cv = a.cu[r2.x + j]) <= q0;
    const uint32_t c = (uint32_t)__syncthreads_count(le);  // probes 0 .. c-1 lie at or before q0
    if (le && tid == c - 1u) sh.item = cv;
    const uint32_t nlo = lo + (c - 1u) * step;
    len = (c * step < len ? c * step : len) - (c - 1u) * step;
    lo = nlo;
  }
  __syncthreads();
  uint32_t cfirst = sh.item;  // units before chunk lo
  for (uint32_t c = lo; c < nch && cfirst < q1; c++) {
    // the chunk's units [cfirst, cnext) and paths [0, ce - cs) from its first
    const uint32_t c0 = c * 256u;
    const uint32_t cnext = c + 1 < nch ? a.cu[r2.x + c + 1] : r2.y;
    const uint32_t cs = a.sv[r4.x + c0].x;
    const uint32_t ce = c0 + 256u < deg ? a.sv[r4.x + c0 + 256u].x : T;
    const uint32_t qa = q0 > cfirst ? q0 : cfirst, qb = q1 < cnext ? q1 : cnext;
    const uint32_t wa = (qa - cfirst) * 256u;
    const uint32_t wb = (qb - cfirst) * 256u < ce - cs ? (qb - cfirst) * 256u : ce - cs;
    vsplit_paths(a, L, r, u, r4, c0, wa, wb);
    cfirst = cnext;
  }
  __syncthreads();
  uint32_t* row = a.rows + (size_t)r * (size_t)a.ns;
  for (uint32_t q = tid; q < sh.nused; q += 256) {
    const uint32_t slot = L.used[q];
    atomicAdd(row + L.keys[slot], L.cnts[slot]);
  }
}

// The round: level, frontier, walks.  Tasks of the round, in order: one per large item (a
// workgroup team), split-unit ranges (at most one per workgroup of the grid), one per 4 small
// items (a wavefront team each); workgroup b runs tasks b, b + grid, ...
__global__ __launch_bounds__(256) void k_vpeel_round(VPeel a) {
  __shared__ union {
    SplitLds sp;
    WalkLds<V_LARGE_TEAM, V_LARGE_CAP, false> l;
    WalkLds<V_SMALL_TEAM, V_SMALL_CAP, true> s[4];
  } lds;
  __shared__ uint32_t f_id[FR_CAP];
  __shared__ uint32_t f_list[FR_CAP];   // small items from the front, large ones from the end
  __shared__ uint4 s_row[VS_MAX_ROWS];  // split rows: the item's record
  __shared__ uint2 s_row2[VS_MAX_ROWS];
  __shared__ uint32_t s_rowu[VS_MAX_ROWS];  // the item
  __shared__ uint32_t s_upfx[VS_MAX_ROWS + 1];  // split units before row r
  __shared__ unsigned long long s_min[4];
  __shared__ uint64_t s_kb[4], s_big[4];
  __shared__ uint32_t s_wc[4], s_cc[4][3];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
  const bool lead = blockIdx.x == 0 && tid == 0;
  // the level: the least block minimum (nb <= 256 = the workgroup size), with every block's
  // count and first item at its minimum loaded before the level is known
  const unsigned long long v = tid < a.nb ? a.vmin[tid] : KNONE;
  const uint32_t vc = tid < a.nb ? a.vcnt[tid] : 0u;
  const uint32_t v0 = tid < a.nb ? a.varg[(size_t)tid * NARG] : 0u;
  unsigned long long k = wave_min_u64(v);
  if (lane == 0) s_min[wid] = k;
  __syncthreads();
  k = s_min[0];
  for (int w = 1; w < 4; w++) k = s_min[w] < k ? s_min[w] : k;
  if (lead) __hip_atomic_store(a.k, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (k == KNONE) {
    if (lead) { a.ctl[0] = 0u; a.ctl[1] = 0u; }
    return;
  }
  // candidates: the live items with D == k of the blocks holding k, in block order
  const bool isk = v == k;
  const uint32_t cnt = isk ? vc : 0u;
  const uint32_t incl = pbf_wave_incl_scan(cnt);
  {
    const uint64_t km = __ballot(isk), bg = __ballot(cnt > NARG);
    if (lane == 63) s_wc[wid] = incl;
    if (lane == 0) { s_kb[wid] = km; s_big[wid] = bg; }
  }
  __syncthreads();
  uint32_t n = 0, before = 0;
  for (uint32_t w = 0; w < 4; w++) {
    if (w < wid) before += s_wc[w];
    n += s_wc[w];
  }
  if ((s_big[0] | s_big[1] | s_big[2] | s_big[3]) == 0ull) {
    // every block holding k listed all its items at k
    const uint32_t pos = before + incl - cnt;
    for (uint32_t j = 0; j < cnt && pos + j < FR_CAP; j++) f_id[pos + j] = j ? a.varg[(size_t)tid * NARG + j] : v0;
  } else {
    n = 0;
    for (uint32_t w4 = 0; w4 < 4 && n < FR_CAP; w4++) {
      uint64_t m = s_kb[w4];
      while (m && n < FR_CAP) {
        const uint32_t bb = w4 * 64u + (uint32_t)__ffsll((long long)m) - 1u;
        m &= m - 1ull;
        const uint32_t c = ctl_ld(a.vcnt + bb);
        if (c <= NARG) {
          if (tid < c && n + tid < FR_CAP) f_id[n + tid] = a.varg[(size_t)bb * NARG + tid];
          n += c;
          continue;
        }
        const int64_t lo = (int64_t)bb << a.bs_log;
        const int64_t hi = lo + (int64_t(1) << a.bs_log) < a.ns ? lo + (int64_t(1) << a.bs_log) : a.ns;
        for (int64_t x0 = lo; x0 < hi && n < FR_CAP; x0 += 256) {
          const int64_t xx = x0 + tid;
          const int64_t xc = xx < hi ? xx : hi - 1;
          const uint8_t s0 = a.state[xc];
          const unsigned long long d0 = a.D[xc];
          const bool sel = (xx < hi) & (s0 == 1) & (d0 == k);
          const uint64_t bm = __ballot(sel);
          __syncthreads();
          if (lane == 0) s_wc[wid] = (uint32_t)__popcll(bm);
          __syncthreads();
          uint32_t bf = n, tot = 0;
          for (uint32_t w = 0; w < 4; w++) {
            if (w < wid) bf += s_wc[w];
            tot += s_wc[w];
          }
          const uint32_t pos = bf + (uint32_t)__popcll(bm & pbf_lanemask_lt());
          if (sel && pos < FR_CAP) f_id[pos] = (uint32_t)xx;
          n += tot;
        }
      }
    }
  }
  n = n < FR_CAP ? n : FR_CAP;
  __syncthreads();
  // classes, placed by ballots; split items beyond the round's rows stay for a later round
  // (frontier entry ~0u)
  uint32_t nsmall = 0, nlarge = 0, nsplit_all = 0;
  const uint64_t lt = pbf_lanemask_lt();
  for (uint32_t j0 = 0; j0 < n; j0 += 256) {
    const uint32_t j = j0 + tid;
    uint32_t cls = FC_NONE, x = 0;
    uint4 rc = make_uint4(0u, 0u, 0u, 0u);
    uint2 rc2 = make_uint2(0u, 0u);
    if (j < n) {
      x = f_id[j];
      rc = a.rec[x];
      rc2 = a.rec2[x];
      // an item in no remaining butterfly (B == 0) has nothing to decrement: no walk
      cls = a.B[x] ? rc.w : (uint32_t)FC_NONE;
    }
    const uint64_t bs = __ballot(cls == FC_SMALL), bl = __ballot(cls == FC_LARGE), bp = __ballot(cls == FC_SPLIT);
    __syncthreads();  // s_cc of the previous step has been read
    if (lane == 0) {
      s_cc[wid][0] = (uint32_t)__popcll(bs);
      s_cc[wid][1] = (uint32_t)__popcll(bl);
      s_cc[wid][2] = (uint32_t)__popcll(bp);
    }
    __syncthreads();
    uint32_t ps = nsmall, pl = nlarge, pp = nsplit_all;
    for (uint32_t w = 0; w < 4; w++) {
      if (w < wid) { ps += s_cc[w][0]; pl += s_cc[w][1]; pp += s_cc[w][2]; }
      nsmall += s_cc[w][0];
      nlarge += s_cc[w][1];
      nsplit_all += s_cc[w][2];
    }
    bool keep = true;
    if (cls == FC_SMALL) {
      f_list[ps + (uint32_t)__popcll(bs & lt)] = x;
    } else if (cls == FC_LARGE) {
      f_list[FR_CAP - 1u - (pl + (uint32_t)__popcll(bl & lt))] = x;  // large items from the end
    } else if (cls == FC_SPLIT) {
      const uint32_t r = pp + (uint32_t)__popcll(bp & lt);
      if (r < a.vrows) { s_row[r] = rc; s_row2[r] = rc2; s_rowu[r] = x; }
      else keep = false;
    }
    if (blockIdx.x == 0 && j < n) a.fr[j] = keep ? x : ~0u;
  }
  const uint32_t nsplit = nsplit_all < a.vrows ? nsplit_all : a.vrows;
  if (lead) { a.ctl[0] = n; a.ctl[1] = nsplit; }
  __syncthreads();
  // split rows and their units
  if (tid < 64) {
    const uint32_t un = tid < nsplit ? s_row2[tid].y : 0u;
    const uint32_t ui = pbf_wave_incl_scan(un);
    if (tid < nsplit) s_upfx[tid + 1] = ui;
    if (tid == 0) s_upfx[0] = 0;
  }
  __syncthreads();
  const uint32_t nunits = s_upfx[nsplit];
  const uint32_t nut = nunits < gridDim.x ? nunits : gridDim.x;  // unit ranges
  const uint32_t ntasks = nlarge + nut + (nsmall + 3u) / 4u;
  for (uint32_t task = blockIdx.x; task < ntasks; task += gridDim.x) {
    __syncthreads();  // LDS of the previous task
    if (task < nlarge) {
      const uint32_t u = f_list[FR_CAP - 1u - task];
      vwalk_item<V_LARGE_TEAM, V_LARGE_CAP, false>(a, lds.l, u, a.rec[u], tid);
    } else if (task < nlarge + nut) {
      const uint32_t t = task - nlarge;
      uint32_t q0 = (uint32_t)((uint64_t)nunits * t / nut), q1 = (uint32_t)((uint64_t)nunits * (t + 1) / nut);
      while (q0 < q1) {
        uint32_t r = 0;
        while (s_upfx[r + 1] <= q0) r++;
        const uint32_t e = q1 < s_upfx[r + 1] ? q1 : s_upfx[r + 1];
        vsplit_units(a, lds.sp, r, s_rowu[r], s_row[r], s_row2[r], q0 - s_upfx[r], e - s_upfx[r]);
        q0 = e;
        __syncthreads();
      }
    } else {
      const uint32_t i = (task - nlarge - nut) * 4u + wid;
      if (i < nsmall) {
        const uint32_t u = f_list[i];
        vwalk_item<V_SMALL_TEAM, V_SMALL_CAP, true>(a, lds.s[wid], u, a.rec[u], lane);
      }
    }
  }
}

// The static tables of the tip peel, one wavefront per S vertex u, a chunk of 256 slots per
// step (4 consecutive slots per lane): sv[e] = (paths of u through its centres before the slot's,
// offO[v]) for every slot e = (u -> v); cu[cbase[u] + c] = the units of u's chunks before chunk c
// (a chunk's paths in units of 256); rec[u] = (offS[u], deg(u), T(u) = all two-hop paths of u,
// class) and rec2[u] = (cbase[u], units of u).  T(u) is below m (the paths of u are distinct
// edges), so 32 bits hold every prefix.  The class: split when T(u) reaches VS_MIN_PATHS or
// the bound on distinct two-hop neighbours, min(T(u), ns), passes what the large LDS table
// holds, else small or large by that bound, none without paths.
__global__ __launch_bounds__(256) void k_peelv_tables(const uint32_t* __restrict__ offS, const uint32_t* __restrict__ edgS,
                                                      const uint32_t* __restrict__ offO, int64_t ns,
                                                      const uint32_t* __restrict__ cbase, uint2* __restrict__ sv,
                                                      uint32_t* __restrict__ cu, uint4* __restrict__ rec, uint2* __restrict__ rec2) {
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  for (int64_t u = (int64_t)blockIdx.x * 4 + wid; u < ns; u += (int64_t)gridDim.x * 4) {
    const uint32_t e0 = offS[u], e1 = offS[u + 1], cb = cbase[u];
    uint32_t carry = 0, units = 0, c = 0;
    for (uint32_t p0 = e0; p0 < e1; p0 += 256, c++) {
      uint32_t d[4], vo[4];
#pragma unroll
      for (uint32_t i = 0; i < 4; i++) {
        const uint32_t p = p0 + 4 * lane + i;
        d[i] = 0;
        vo[i] = 0;
        if (p < e1) {
          const uint32_t v = edgS[p];
          vo[i] = offO[v];
          d[i] = offO[v + 1] - vo[i];
        }
      }
      const uint32_t s = d[0] + d[1] + d[2] + d[3];
      const uint32_t incl = pbf_wave_incl_scan(s);
      uint32_t x = carry + incl - s;
#pragma unroll
      for (uint32_t i = 0; i < 4; i++) {
        const uint32_t p = p0 + 4 * lane + i;
        if (p < e1) sv[p] = make_uint2(x, vo[i]);
        x += d[i];
      }
      const uint32_t chunk = __shfl(incl, 63, 64);
      if (lane == 0) cu[cb + c] = units;
      units += (chunk + VS_PCH - 1u) / VS_PCH;
      carry += chunk;
    }
    if (lane == 0) {
      const uint32_t T = carry;
      const uint32_t bnd = (int64_t)T < ns ? T : (uint32_t)ns;
      const uint32_t cls = T == 0 ? FC_NONE
                         : (bnd > V_LARGE_MAX || T >= VS_MIN_PATHS) ? FC_SPLIT
                         : bnd <= V_SMALL_MAX ? FC_SMALL : FC_LARGE;
      rec[u] = make_uint4(e0, e1 - e0, T, cls);
      rec2[u] = make_uint2(cb, units);
    }
  }
}

// Chunks of 256 centres per S vertex.
__global__ void k_peelv_nchunks(const uint32_t* __restrict__ offS, int64_t ns, uint32_t* __restrict__ nch) {
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < ns; u += (int64_t)gridDim.x * blockDim.x)
    nch[u] = (offS[u + 1] - offS[u] + 255u) / 256u;
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

// Items x of the lanes with add set are appended to the touched list unless already flagged:
// every lane's flag exchanges are in flight together, then the wavefront takes one slot range.
// Every lane of the wavefront calls it together.
template <uint32_t N>
__device__ __forceinline__ void touch_n(const uint32_t (&x)[N], const bool (&add)[N], uint32_t* tflag, uint32_t* touched,
                                        uint32_t* ntouched) {
  uint32_t old[N];
#pragma unroll
  for (uint32_t i = 0; i < N; i++) old[i] = add[i] ? atomicExch(&tflag[x[i]], 1u) : 1u;
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t i = 0; i < N; i++) cnt += old[i] == 0u ? 1u : 0u;
  const uint32_t incl = pbf_wave_incl_scan(cnt);
  const uint32_t tot = (uint32_t)__shfl((int)incl, 63, 64);
  if (tot == 0u) return;
  uint32_t base = 0;
  if (__lane_id() == 0) base = atomicAdd(ntouched, tot);
  base = (uint32_t)__shfl((int)base, 0, 64) + incl - cnt;
#pragma unroll
  for (uint32_t i = 0; i < N; i++)
    if (old[i] == 0u) touched[base++] = x[i];
}

// The number of entries of the sorted list arr[0, n) that are at most d (binary lifting from the
// largest power of two top <= n; every probe of a level in flight together for the PER
// searches), and whether d is among them (the last accepted probe equals d).
template <uint32_t PER, typename Arr>
__device__ __forceinline__ void count_le(const Arr& arr, uint32_t n, uint32_t top, const uint32_t (&d)[PER],
                                         uint32_t (&pos)[PER], bool (&hit)[PER]) {
  uint32_t last[PER];
#pragma unroll
  for (uint32_t i = 0; i < PER; i++) { pos[i] = 0; last[i] = 0xFFFFFFFFu; }
  for (uint32_t step = top; step; step >>= 1) {
    uint32_t v[PER];
#pragma unroll
    for (uint32_t i = 0; i < PER; i++) {
      const uint32_t q = pos[i] + step - 1u;
      v[i] = arr[q < n ? q : n - 1u];
    }
#pragma unroll
    for (uint32_t i = 0; i < PER; i++) {
      if (pos[i] + step <= n && v[i] <= d[i]) { pos[i] += step; last[i] = v[i]; }
    }
  }
#pragma unroll
  for (uint32_t i = 0; i < PER; i++) hit[i] = pos[i] > 0u && last[i] == d[i];
}

struct GlobalList {
  const uint32_t* p;
  __device__ __forceinline__ uint32_t operator[](uint32_t q) const { return p[q]; }
};
struct LdsList {
  const uint32_t* p;
  __device__ __forceinline__ uint32_t operator[](uint32_t q) const { return p[q]; }
};

// Walk of a peeled edge e = (v, u).  Every butterfly through e is {v, v2} x {u, u2}; it is found
// either from v (u2 in N(v), v2 in N(u2), test v2 in N(u)) or, symmetrically, from u (v2 in
// N(u), u2 in N(v2), test u2 in N(v)), whichever end has the smaller two-hop work.  Below, a is
// the starting end, b the other one: c runs over N(a) \ {b}, d over N(c) \ {a}, and d must lie in
// N(b) (staged in LDS when it fits).  The edge (a, c) loses the butterflies of its segment at
// once; (c, d) and (b, d) lose one each.  A thread takes EW_PER paths at a time, and each stage
// (the path's entry and edge id, the search of N(b), the two edges' states, the touched flags)
// has all of its loads in flight together.
constexpr uint32_t EW_PER = 8;
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
    if (bdeg == 0) continue;  // no butterfly (uniform; e itself lies in N(b), so this never happens)
    const bool cached = bdeg <= (uint32_t)UCAP;
    const uint32_t top = 1u << (31 - __clz(bdeg));  // the largest power of two <= bdeg
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
      uint8_t sc = 0;
      if (j < adeg) {
        const uint32_t q = ao + j;  // edge (a, c)
        const uint32_t c = nbrX[q];
        fc = idX[q];
        sc = a.state[fc];
        if (c != vb && (sc == 1 || (sc == 2 && fc > e))) {
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
      for (uint32_t w0 = 0; w0 < T; w0 += TEAM * EW_PER) {
        // wavefront-uniform early exit (every lane runs the wavefront operations below)
        if ((uint32_t)__builtin_amdgcn_readfirstlane((int)*found) >= target) break;
        uint32_t wq[EW_PER], sg[EW_PER], d[EW_PER], fd[EW_PER];
        bool ok[EW_PER];
#pragma unroll
        for (uint32_t i = 0; i < EW_PER; i++) {
          const uint32_t w = w0 + i * TEAM + tid;
          ok[i] = w < T;
          wq[i] = w < T ? w : T - 1u;
          sg[i] = 0;
        }
        // segment of each path: the number of P entries at most w (fixed depth)
#pragma unroll
        for (uint32_t step = TEAM / 2; step; step >>= 1) {
#pragma unroll
          for (uint32_t i = 0; i < EW_PER; i++) sg[i] = sh.P[sg[i] + step - 1u] <= wq[i] ? sg[i] + step : sg[i];
        }
#pragma unroll
        for (uint32_t i = 0; i < EW_PER; i++) {
          const uint32_t s = sg[i];
          const uint32_t p2 = sh.vs[s] + wq[i] - (s ? sh.P[s - 1u] : 0u);
          d[i] = nbrY[p2];
          fd[i] = idY[p2];
        }
        // edge (b, d): d in N(b)?
        uint32_t pos[EW_PER];
        bool hit[EW_PER];
#pragma unroll
        for (uint32_t i = 0; i < EW_PER; i++) ok[i] = ok[i] & (d[i] != va);
        if (cached) {
          bool any = false;
#pragma unroll
          for (uint32_t i = 0; i < EW_PER; i++) {
            const uint32_t h = (d[i] * 0x9E3779B1u) >> (32 - BM_LOG);
            ok[i] = ok[i] & (((nb_bits[h >> 5] >> (h & 31)) & 1u) != 0u);
            any |= ok[i];
          }
          if (__ballot(any) == 0ull) continue;
          count_le<EW_PER>(LdsList{nb_v}, bdeg, top, d, pos, hit);
        } else {
          count_le<EW_PER>(GlobalList{nbrY + bo}, bdeg, top, d, pos, hit);
        }
        uint32_t fb[EW_PER];
#pragma unroll
        for (uint32_t i = 0; i < EW_PER; i++) {
          ok[i] = ok[i] & hit[i];
          const uint32_t q = pos[i] ? pos[i] - 1u : 0u;
          fb[i] = cached ? nb_f[q] : idY[bo + q];
        }
        // edge (c, d) and edge (b, d) still there?
        uint32_t sd[EW_PER], sb[EW_PER];
#pragma unroll
        for (uint32_t i = 0; i < EW_PER; i++) {
          sd[i] = a.state[fd[i]];
          sb[i] = a.state[fb[i]];
        }
#pragma unroll
        for (uint32_t i = 0; i < EW_PER; i++) { pbf_ready(sd[i]); pbf_ready(sb[i]); }
        uint32_t tx[2 * EW_PER];
        bool tadd[2 * EW_PER];
#pragma unroll
        for (uint32_t i = 0; i < EW_PER; i++) {
          tx[2 * i] = fd[i];
          tx[2 * i + 1] = fb[i];
          tadd[2 * i] = false;
          tadd[2 * i + 1] = false;
          if (!(ok[i] && sd[i] != 0 && sb[i] != 0)) continue;
          atomicAdd((uint32_t*)found, 1u);  // a butterfly through e
          // charged by a smaller peeled edge?
          if (!(sd[i] == 1 || fd[i] > e) || !(sb[i] == 1 || fb[i] > e)) continue;
          atomicAdd(&sh.acc[sg[i]], 1ull);
          if (sd[i] == 1) { atomicAdd(&a.B[fd[i]], ~0ull); tadd[2 * i] = true; }
          if (sb[i] == 1) { atomicAdd(&a.B[fb[i]], ~0ull); tadd[2 * i + 1] = true; }
        }
        touch_n<2 * EW_PER>(tx, tadd, a.tflag, a.touched, a.ntouched);
      }
      __syncthreads();
      {
        const unsigned long long cnt = j < adeg ? sh.acc[tid] : 0ull;
        const bool add = cnt && sc == 1;
        if (add) atomicAdd(&a.B[fc], (unsigned long long)(-(long long)cnt));
        const uint32_t tx1[1] = {fc};
        const bool ta1[1] = {add};
        touch_n<1>(tx1, ta1, a.tflag, a.touched, a.ntouched);
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
  VPeel a{};
  a.offS = use_v ? g.offU : g.offV;
  a.edgS = use_v ? g.edgU : g.edgV;
  a.offO = use_v ? g.offV : g.offU;
  a.edgO = use_v ? g.edgV : g.edgU;
  a.ns = ns;
  a.B = pb.B;
  a.D = pb.D;
  a.state = pb.state;
  a.k = pb.c.k;
  a.live = pb.c.live;
  a.ctl = pb.c.len;  // [0] frontier entries, [1] split rows, [2] error flag (zeroed by k_ctl_init)
  hipLaunchKernelGGL(k_init_peel, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<const unsigned long long*>(counts), pb.B, pb.D, pb.state, pb.tflag, ns);
  hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, ctx->stream, pb.c);
  // split walks: up to vrows rows per round, each a dense row of ns counters (at most 2 GB of
  // rows); the static path prefix of every S slot and the record of every S item
  const uint32_t vrows = (uint32_t)std::max<int64_t>(1, std::min<int64_t>(VS_MAX_ROWS, (int64_t(1) << 31) / (4 * ns)));
  uint32_t nb, bs_log = 10;
  while (((ns - 1) >> bs_log) + 1 > 256) bs_log++;
  nb = (uint32_t)(((ns - 1) >> bs_log) + 1);
  {
    const size_t row_words = (size_t)vrows * (size_t)ns;
    // chunks of 256 centres: at most m / 256 + ns of them
    const size_t nchunk_max = (size_t)g.m / 256 + (size_t)ns + 1;
    const size_t words = row_words + 4 * (size_t)ns + 2 * (size_t)ns + 2 * (size_t)g.m + 2 * (size_t)ns + nchunk_max +
                         FR_CAP + 256 * (2 + 1 + 1 + NARG) + 64;
    void* sp;
    PBF_CHECK(pbf_scratch(ctx, S_PEEL7, 4 * words, &sp));
    uint32_t* w = static_cast<uint32_t*>(sp);
    a.rows = w;
    w += row_words;
    w = reinterpret_cast<uint32_t*>((reinterpret_cast<uintptr_t>(w) + 15) & ~uintptr_t(15));
    a.rec = reinterpret_cast<const uint4*>(w);
    w += 4 * (size_t)ns;
    a.rec2 = reinterpret_cast<const uint2*>(w);
    w += 2 * (size_t)ns;
    a.sv = reinterpret_cast<const uint2*>(w);
    w += 2 * (size_t)g.m;
    a.vmin = reinterpret_cast<unsigned long long*>(w);
    w = reinterpret_cast<uint32_t*>(a.vmin + 256);
    a.vcnt = w; w += 256;
    a.vflag = w; w += 256;
    a.varg = w; w += 256 * NARG;
    a.fr = w; w += FR_CAP;
    uint32_t* nch = w; w += ns;   // chunks per item, then their exclusive prefix (the item's first chunk)
    uint32_t* cbase = w; w += ns;
    a.cu = w;
    // rows start at zero; every prepare pass leaves the ones it credited so
    PBF_CHECK(pbf_fill_u32(ctx, a.rows, 0u, (int64_t)row_words));
    PBF_CHECK(pbf_fill_u32(ctx, a.vflag, 1u, nb));
    hipLaunchKernelGGL(k_peelv_nchunks, dim3(grid_for(ns, 256, 16384)), dim3(256), 0, ctx->stream, a.offS, ns, nch);
    PBF_CHECK(pbf_scan_excl_u32(ctx, nch, cbase, ns, nullptr));
    hipLaunchKernelGGL(k_peelv_tables, dim3(grid_for(ns, 4, 16384)), dim3(256), 0, ctx->stream, a.offS, a.edgS, a.offO, ns,
                       cbase, const_cast<uint2*>(a.sv), const_cast<uint32_t*>(a.cu), const_cast<uint4*>(a.rec),
                       const_cast<uint2*>(a.rec2));
  }
  a.nb = nb;
  a.bs_log = bs_log;
  a.vrows = vrows;
  PBF_HIP(hipGetLastError());
  const unsigned round_grid = (unsigned)ctx->num_cus * V_ROUND_PER_CU;
  auto round = [&]() -> int {
    hipLaunchKernelGGL(k_vpeel_prepare, dim3(nb), dim3(256), 0, ctx->stream, a);
    hipLaunchKernelGGL(k_vpeel_round, dim3(round_grid), dim3(256), 0, ctx->stream, a);
    PBF_HIP(hipGetLastError());
    return PBF_OK;
  };
  int64_t rounds = 0;
  PBF_CHECK(run_rounds(ctx, pb, ns, round, &rounds));
  ctx->timing.rounds = rounds;
  // a walk table that filled up (the classes rule it out) fails the peel rather than its result
  PBF_HIP(hipMemcpyAsync(ctx->pinned + 34, pb.c.len + 2, 4, hipMemcpyDeviceToHost, ctx->stream));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  if (reinterpret_cast<const uint32_t*>(ctx->pinned + 34)[0] != 0u) return PBF_ERR_CAPACITY;
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
    const unsigned sel_grid = grid_for(entries, 256 * SEL_PER, 8192);
    hipLaunchKernelGGL(k_round_begin, dim3(1), dim3(64), 0, ctx->stream, pb.c);
    hipLaunchKernelGGL(k_select_touched, dim3(ctx->num_cus * 4), dim3(256), 0, ctx->stream, pb.D, pb.B, pb.state, pb.touched,
                       pb.work, small_max, small_list, large_list, pb.c);
    hipLaunchKernelGGL(k_min_live, dim3(grid_for(entries, 1024, 2048)), dim3(256), 0, ctx->stream, pb.D, pb.state, cur.idO,
                       entries, pb.c);
    hipLaunchKernelGGL(k_select, dim3(sel_grid), dim3(256), 0, ctx->stream, pb.D, pb.B, pb.state, cur.idO, entries, pb.work,
                       small_max, small_list, large_list, pb.c);
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
