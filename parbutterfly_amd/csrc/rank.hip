// Ranking and relabelling on the GPU.
//
//  * cmpWedgeCounts (apps/hyper/butterfly_utils.h:1061-1086): two reductions of C(d,2).
//  * ADEG key (getApproxDegRanks, :275-284): key = floor(log2 d) + 1 (0 for d = 0), i.e. the
//    bit length of d.  ACOCORE key (getApproxCoCoreRanks, :448-503): the log-degree level at
//    which the vertex is removed by the decreasing-bucket peel (k_acore_* below).
//  * getRanks (:249-265) sorts ids by key, descending.  Here that is one stable counting-sort
//    pass over <= 33 key values (per-block histograms, device scan, stable scatter with
//    wavefront match masks), so ties are broken by ascending id.
//  * rankGraph / rankGraphEdges (:579-700): entries (rank << 1) | flag, lists ascending.  With a
//    stable key sort, a neighbour list in original (ascending id) order is already ascending in
//    rank inside each key class, so each ranked list is a stable counting sort of the original
//    list by key -- no comparison sort is needed.  The same pass yields in(r), the number of
//    lower-ranked neighbours (rankWedgeF, :216-237): every neighbour in a higher key class,
//    plus, for a U vertex, the neighbours of its own class (V ids precede U ids).
#include "pbf_device.h"

namespace {
constexpr int KEYS = 33;  // key values 0..32
constexpr uint32_t RL_CHUNK = 1024;  // lists longer than this are relabelled in chunks
constexpr uint32_t RL_SMALL = 16;    // lists up to this long are relabelled by one thread each
constexpr uint32_t RL_REG = 4;       // chunks of 64 entries of a list that the relabel keeps in registers

// Adds a 256-thread workgroup's values into *acc with one atomic: atomics of every wavefront
// on one address serialise, thousands of them cost more than the sum they carry.
__device__ __forceinline__ void block_add(uint64_t s, unsigned long long* acc) {
  __shared__ uint64_t part[4];
  s = pbf_wave_sum(s);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint64_t t = part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(acc, (unsigned long long)t);
  }
}

__global__ __launch_bounds__(256) void k_side_partial(const uint32_t* __restrict__ off, int64_t n, unsigned long long* acc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t s = 0;
  for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    s += (uint64_t)pbf_side_term(off[i + 1] - off[i]);
  }
  block_add(s, acc);
}

__device__ __forceinline__ uint32_t vertex_degree(const DevGraph& g, int64_t x) {
  return x < g.nv ? g.offV[x + 1] - g.offV[x] : g.offU[x - g.nv + 1] - g.offU[x - g.nv];
}

// SIDE: every vertex of the count side S before every vertex of O (ties by id), i.e. the
// wedges are exactly those with both endpoints on S and the centre on O -- the side-based
// enumeration the reference runs when no ranking is requested (butterfly_count.h:1639-1640).
__global__ void k_side_key(DevGraph g, int use_v, const uint8_t* __restrict__ sflag, uint8_t* __restrict__ key) {
  int64_t n = g.nv + g.nu;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    const bool in_s = sflag ? sflag[x] != 0 : (use_v ? x >= g.nv : x < g.nv);
    key[x] = in_s ? 2 : 1;
  }
}

__global__ void k_adeg_key(DevGraph g, uint8_t* __restrict__ key) {
  int64_t n = g.nv + g.nu;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x)
    key[x] = (uint8_t)pbf_bitlen(vertex_degree(g, x));
}

// Per-block key histogram; hist is laid out [KEYS][nblocks] in DESCENDING key order so that an
// exclusive scan over it gives each (key, block) its first output rank.
__global__ __launch_bounds__(1024) void k_key_hist(const uint8_t* __restrict__ key, int64_t n, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[KEYS];
  if (threadIdx.x < KEYS) h[threadIdx.x] = 0;
  __syncthreads();
  int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) atomicAdd(&h[key[x]], 1u);
  __syncthreads();
  if (threadIdx.x < KEYS) hist[(size_t)(KEYS - 1 - threadIdx.x) * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(1024) void k_key_scatter(const uint8_t* __restrict__ key, int64_t n,
                                                      const uint32_t* __restrict__ hist_scanned,
                                                      uint32_t* __restrict__ order, uint32_t* __restrict__ rank_of) {
  __shared__ uint32_t wc[16][KEYS];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int t = threadIdx.x; t < 16 * KEYS; t += blockDim.x) (&wc[0][0])[t] = 0;
  __syncthreads();
  int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool act = x < n;
  uint32_t k = act ? key[x] : 0;
  uint64_t active = __ballot(act);
  uint64_t mm = pbf_match_key6(k, active);
  uint32_t rank_in_wave = __popcll(mm & pbf_lanemask_lt());
  if (act && rank_in_wave == 0) wc[wid][k] = __popcll(mm);
  __syncthreads();
  if (act) {
    uint32_t before = 0;
    for (int w = 0; w < wid; w++) before += wc[w][k];
    uint32_t r = hist_scanned[(size_t)(KEYS - 1 - k) * gridDim.x + blockIdx.x] + before + rank_in_wave;
    order[r] = (uint32_t)x;
    rank_of[x] = r;
  }
  (void)lane;
}

__global__ void k_rank_key(const uint32_t* __restrict__ rank_of, const uint8_t* __restrict__ key, int64_t n,
                           uint2* __restrict__ rk) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x)
    rk[x] = make_uint2(rank_of[x], key[x]);
}

__global__ void k_ranked_degree(DevGraph g, const uint32_t* __restrict__ order, uint32_t* __restrict__ rdeg) {
  int64_t n = g.nv + g.nu;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    rdeg[r] = vertex_degree(g, order[r]);
}

// One wavefront per original vertex x: stable counting sort of x's neighbour list by key
// (descending), written at roff[rank_of[x]].  flag = 1 iff the neighbour lies on S.
// Flag of x's entries: 1 iff the neighbour lies on S, i.e. iff x itself lies on O.
__device__ __forceinline__ uint32_t entry_flag(int use_v, const uint8_t* __restrict__ sflag, int64_t x, bool xv) {
  if (sflag) return sflag[x] ? 0u : 1u;
  return use_v ? (xv ? 1u : 0u) : (xv ? 0u : 1u);
}

__global__ __launch_bounds__(256) void k_relabel(DevGraph g, int use_v, const uint8_t* __restrict__ sflag,
                                                 const uint2* __restrict__ rk, const uint32_t* __restrict__ roff,
                                                 const uint32_t* __restrict__ lo, uint32_t* __restrict__ redg,
                                                 uint32_t* __restrict__ rpos, uint32_t* __restrict__ inpref,
                                                 uint32_t* __restrict__ rown) {
  __shared__ uint32_t cnt[4][KEYS + 1];
  __shared__ uint32_t run[4][KEYS + 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t n = g.nv + g.nu;
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  for (int64_t x = (int64_t)blockIdx.x * 4 + wid; x < n; x += nwaves) {
    const bool xv = x < g.nv;
    const uint32_t obase = xv ? g.offV[x] : g.offU[x - g.nv];
    const uint32_t deg = xv ? g.offV[x + 1] - obase : g.offU[x - g.nv + 1] - obase;
    if (deg <= RL_SMALL || deg > RL_CHUNK) continue;  // short lists: k_relabel_small; long: k_rl_*
    const uint32_t* adj = xv ? g.edgV + obase : g.edgU + obase;
    const int64_t nbr_shift = xv ? g.nv : 0;  // combined id of neighbour = adj[j] + shift
    const uint2 rkx = rk[x];
    // the first RL_REG chunks of 64 entries (the whole list, for most lists) are loaded once, at
    // indices clamped to the last entry, then their neighbours' rank and key (one 8-byte gather
    // each), every load of a stage in flight together, and both passes use them from registers;
    // the chunks of a longer list past those are loaded in each pass
    const uint32_t nt = (uint32_t)__builtin_amdgcn_readfirstlane((int)((deg + 63u) / 64u));
    uint32_t nb[RL_REG];
    uint2 ry[RL_REG];
#pragma unroll
    for (uint32_t t = 0; t < RL_REG; t++) {
      const uint32_t j = t * 64u + lane;
      nb[t] = adj[j < deg ? j : deg - 1u];
    }
    // x's rank arrives first: its list offset is loaded with the gathers
    const uint32_t r = rkx.x;
    const uint32_t rbase = roff[r];
    const uint32_t lox = lo ? lo[r] : 0u;  // graph-local entries (a union ranked graph by graph)
#pragma unroll
    for (uint32_t t = 0; t < RL_REG; t++) ry[t] = rk[nb[t] + nbr_shift];
    const uint32_t flag = entry_flag(use_v, sflag, x, xv);
    if (lane < KEYS + 1) { cnt[wid][lane] = 0; run[wid][lane] = 0; }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // pass A: key histogram of the neighbours
#pragma unroll
    for (uint32_t t = 0; t < RL_REG; t++)
      if (t * 64u + lane < deg) atomicAdd(&cnt[wid][ry[t].y], 1u);
    for (uint32_t j = RL_REG * 64u + lane; j < deg; j += 64) atomicAdd(&cnt[wid][rk[adj[j] + nbr_shift].y], 1u);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // exclusive prefix in descending key order: start[k] = sum_{k' > k} cnt[k']
    uint32_t c = lane < KEYS ? cnt[wid][KEYS - 1 - lane] : 0;
    uint32_t incl = pbf_wave_incl_scan(c);
    if (lane < KEYS) run[wid][KEYS - 1 - lane] = incl - c;  // run[k] = start of class k
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
      const uint32_t kself = rkx.y;
      // in(r): neighbours of strictly higher key, plus same-key neighbours when x is in U
      inpref[r] = run[wid][kself] + (xv ? 0u : cnt[wid][kself]);
    }
    // pass B: stable scatter, 64 neighbours at a time
    auto scatter = [&](uint32_t t, uint2 v) {
      const uint32_t j = t * 64u + lane;
      const bool act = j < deg;
      const uint32_t kj = act ? v.y : 0u;
      const uint64_t active = __ballot(act);
      const uint64_t mm = pbf_match_key6(kj, active);
      const uint32_t rkn = __popcll(mm & pbf_lanemask_lt());
      if (act) {
        const uint32_t slot = rbase + run[wid][kj] + rkn;
        redg[slot] = ((v.x - lox) << 1) | flag;
        rown[slot] = r;
        if (rpos) rpos[slot] = obase + j;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (act && rkn == 0) run[wid][kj] += __popcll(mm);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    };
#pragma unroll
    for (uint32_t t = 0; t < RL_REG; t++)
      if (t < nt) scatter(t, ry[t]);  // uniform
    for (uint32_t t = RL_REG; t < nt; t++) {
      const uint32_t j = t * 64u + lane;
      scatter(t, rk[adj[j < deg ? j : deg - 1u] + nbr_shift]);
    }
  }
}

// Short lists (at most RL_SMALL entries), one thread each: the same stable sort by key, with
// the position of entry j counted directly (entries of a higher key, then earlier entries of
// the same key), so a wavefront serves 64 vertices instead of one.
__global__ __launch_bounds__(256) void k_relabel_small(DevGraph g, int use_v, const uint8_t* __restrict__ sflag,
                                                       const uint2* __restrict__ rk,
                                                       const uint32_t* __restrict__ roff, const uint32_t* __restrict__ lo,
                                                       uint32_t* __restrict__ redg, uint32_t* __restrict__ rpos,
                                                       uint32_t* __restrict__ inpref, uint32_t* __restrict__ rown) {
  const int64_t n = g.nv + g.nu;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    const bool xv = x < g.nv;
    const uint32_t obase = xv ? g.offV[x] : g.offU[x - g.nv];
    const uint32_t deg = (xv ? g.offV[x + 1] : g.offU[x - g.nv + 1]) - obase;
    // the largest short list of the wavefront (before the long lists leave)
    const uint32_t dmax = (uint32_t)__builtin_amdgcn_readfirstlane((int)pbf_wave_max(deg <= RL_SMALL ? deg : 0u));
    if (deg > RL_SMALL) continue;
    const uint32_t* adj = xv ? g.edgV + obase : g.edgU + obase;
    const int64_t sh = xv ? g.nv : 0;
    const uint2 rkx = rk[x];
    // each neighbour's rank and key (one 8-byte gather): the list's entries first, at indices
    // clamped to the last one, then the gathers (of x's own record for an empty list), so that
    // each stage's loads are in flight together; entries past deg are never used
    // in batches of 4 entries up to the largest list of the wavefront; an empty list reads a
    // valid word (offV[0]) and gathers x's own record
    const uint32_t* src = deg ? adj : g.offV;
    uint32_t nb[RL_SMALL], yr[RL_SMALL], ky[RL_SMALL];
#pragma unroll
    for (uint32_t b = 0; b < RL_SMALL; b += 4) {
      if (b == 0 || b < dmax) {
#pragma unroll
        for (uint32_t j = b; j < b + 4; j++) {
          const uint32_t v = src[j < deg ? j : (deg ? deg - 1u : 0u)];
          nb[j] = deg ? v + (uint32_t)sh : (uint32_t)x;
        }
      }
    }
    // x's rank arrives first: its list offset is loaded with the gathers
    const uint32_t r = rkx.x;
    const uint32_t rbase = roff[r];
    const uint32_t lox = lo ? lo[r] : 0u;
#pragma unroll
    for (uint32_t b = 0; b < RL_SMALL; b += 4) {
      if (b == 0 || b < dmax) {
#pragma unroll
        for (uint32_t j = b; j < b + 4; j++) {
          const uint2 ry = rk[nb[j]];
          yr[j] = ry.x;
          ky[j] = ry.y;
        }
      }
    }
    const uint32_t flag = entry_flag(use_v, sflag, x, xv);
    if (deg == 0) { inpref[r] = 0u; continue; }
    const uint32_t kself = rkx.y;
    // in(r): neighbours of strictly higher key, plus same-key neighbours when x is in U
    uint32_t in = 0;
#pragma unroll
    for (uint32_t j = 0; j < RL_SMALL; j++)
      in += (j < deg && (ky[j] > kself || (!xv && ky[j] == kself))) ? 1u : 0u;
    inpref[r] = in;
#pragma unroll
    for (uint32_t j = 0; j < RL_SMALL; j++) {
      if (j < deg) {
        uint32_t pos = 0;
#pragma unroll
        for (uint32_t k = 0; k < RL_SMALL; k++)
          pos += (k < deg && (ky[k] > ky[j] || (k < j && ky[k] == ky[j]))) ? 1u : 0u;
        const uint32_t slot = rbase + pos;
        redg[slot] = ((yr[j] - lox) << 1) | flag;
        rown[slot] = r;
        if (rpos) rpos[slot] = obase + j;
      }
    }
  }
}

// Long lists (more than RL_CHUNK entries) are relabelled in chunks so that a hub's list is
// spread over many wavefronts: (1) each chunk's key histogram, (2) per list, each chunk's
// starting offset inside every key class plus the class starts, (3) a stable scatter of each
// chunk.  Chunk c of long vertex b is item chunk_base[b] + c.

__device__ __forceinline__ uint32_t find_item_owner(const uint32_t* __restrict__ chunk_base, uint32_t nlong, uint32_t item) {
  uint32_t lo = 0, hi = nlong;  // last b with chunk_base[b] <= item
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (chunk_base[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}

struct LongLists {
  const uint32_t* list;        // original ids of long vertices
  const uint32_t* chunk_base;  // exclusive scan of chunk counts
  const uint32_t* nlong;       // device count of long vertices
  const uint32_t* nitems;      // device total of chunks
  uint32_t* hist;              // [item][KEYS] counts, then offsets
};

__device__ __forceinline__ void vertex_list(const DevGraph& g, uint32_t x, uint32_t* obase, uint32_t* deg,
                                            const uint32_t** adj, int64_t* shift) {
  const bool xv = x < g.nv;
  *obase = xv ? g.offV[x] : g.offU[x - g.nv];
  *deg = xv ? g.offV[x + 1] - *obase : g.offU[x - g.nv + 1] - *obase;
  *adj = xv ? g.edgV + *obase : g.edgU + *obase;
  *shift = xv ? g.nv : 0;
}

__global__ __launch_bounds__(256) void k_rl_hist(DevGraph g, const uint2* __restrict__ rk, LongLists L) {
  __shared__ uint32_t h[4][KEYS + 1];
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint32_t nl = *L.nlong, tot = *L.nitems;
  for (uint32_t item = blockIdx.x * 4 + wid; item < tot; item += gridDim.x * 4) {
    const uint32_t b = find_item_owner(L.chunk_base, nl, item);
    const uint32_t x = L.list[b], c = item - L.chunk_base[b];
    uint32_t obase, deg;
    const uint32_t* adj;
    int64_t sh;
    vertex_list(g, x, &obase, &deg, &adj, &sh);
    if (lane < KEYS + 1) h[wid][lane] = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t j1 = min(deg, (c + 1) * RL_CHUNK);
    for (uint32_t j = c * RL_CHUNK + lane; j < j1; j += 64) atomicAdd(&h[wid][rk[adj[j] + sh].y], 1u);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane < KEYS) L.hist[(size_t)item * KEYS + lane] = h[wid][lane];
    __builtin_amdgcn_wave_barrier();
  }
}

// One wavefront per long vertex; lane k owns key k: running offsets over the chunks, then the
// class starts (descending key order) are added, and in(r) is written.
__global__ __launch_bounds__(256) void k_rl_offsets(DevGraph g, const uint8_t* __restrict__ key,
                                                    const uint32_t* __restrict__ rank_of, LongLists L,
                                                    uint32_t* __restrict__ inpref) {
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint32_t nl = *L.nlong;
  for (uint32_t b = blockIdx.x * 4 + wid; b < nl; b += gridDim.x * 4) {
    const uint32_t x = L.list[b];
    const uint32_t i0 = L.chunk_base[b];
    const uint32_t i1 = (b + 1 < nl) ? L.chunk_base[b + 1] : *L.nitems;
    uint32_t run = 0;
    if (lane < KEYS)
      for (uint32_t it = i0; it < i1; it++) {
        const uint32_t c = L.hist[(size_t)it * KEYS + lane];
        L.hist[(size_t)it * KEYS + lane] = run;
        run += c;
      }
    // class start of key k = sum of totals of keys > k
    const uint32_t rv = lane < KEYS ? run : 0u;
    const uint32_t srcl = lane < KEYS ? (KEYS - 1 - lane) : lane;  // reversed order for the scan
    const uint32_t rrev = __shfl(rv, (int)srcl, 64);
    const uint32_t inc = pbf_wave_incl_scan(lane < KEYS ? rrev : 0u);
    const uint32_t start_rev = inc - (lane < KEYS ? rrev : 0u);  // start of key KEYS-1-lane
    const uint32_t start = __shfl(start_rev, (int)srcl, 64);    // start of key lane
    if (lane < KEYS)
      for (uint32_t it = i0; it < i1; it++) L.hist[(size_t)it * KEYS + lane] += start;
    const uint32_t kx = key[x];
    const uint32_t sx = __shfl(start, (int)kx, 64), cx = __shfl(rv, (int)kx, 64);
    if (lane == 0) inpref[rank_of[x]] = sx + (x < g.nv ? 0u : cx);
  }
}

__global__ __launch_bounds__(256) void k_rl_scatter(DevGraph g, int use_v, const uint8_t* __restrict__ sflag,
                                                    const uint2* __restrict__ rk, const uint32_t* __restrict__ roff,
                                                    const uint32_t* __restrict__ lo, LongLists L, uint32_t* __restrict__ redg,
                                                    uint32_t* __restrict__ rpos, uint32_t* __restrict__ rown) {
  __shared__ uint32_t run[4][KEYS + 1];
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint32_t nl = *L.nlong, tot = *L.nitems;
  for (uint32_t item = blockIdx.x * 4 + wid; item < tot; item += gridDim.x * 4) {
    const uint32_t b = find_item_owner(L.chunk_base, nl, item);
    const uint32_t x = L.list[b], c = item - L.chunk_base[b];
    uint32_t obase, deg;
    const uint32_t* adj;
    int64_t sh;
    vertex_list(g, x, &obase, &deg, &adj, &sh);
    const bool xv = x < g.nv;
    const uint32_t flag = entry_flag(use_v, sflag, x, xv);
    const uint32_t rx = rk[x].x;
    const uint32_t lox = lo ? lo[rx] : 0u;
    const uint32_t rbase = roff[rx];
    if (lane < KEYS) run[wid][lane] = L.hist[(size_t)item * KEYS + lane];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t j1 = min(deg, (c + 1) * RL_CHUNK);
    for (uint32_t j0 = c * RL_CHUNK; j0 < j1; j0 += 64) {
      const uint32_t j = j0 + lane;
      const bool act = j < j1;
      uint2 ry = make_uint2(0u, 0u);
      if (act) ry = rk[adj[j] + sh];
      const uint32_t ky = ry.y;
      const uint64_t mm = pbf_match_key6(ky, __ballot(act));
      const uint32_t rkn = __popcll(mm & pbf_lanemask_lt());
      if (act) {
        const uint32_t slot = rbase + run[wid][ky] + rkn;
        redg[slot] = ((ry.x - lox) << 1) | flag;
        rown[slot] = rx;
        if (rpos) rpos[slot] = obase + j;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (act && rkn == 0) run[wid][ky] += __popcll(mm);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// Lists vertices whose lists exceed RL_CHUNK entries, with their chunk counts.
__global__ void k_rl_find_long(DevGraph g, uint32_t* __restrict__ list, uint32_t* __restrict__ chunks,
                               uint32_t* __restrict__ nlong) {
  const int64_t n = g.nv + g.nu;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < n; base += stride) {
    const int64_t x = base + (threadIdx.x & 63);
    uint32_t d = 0;
    if (x < n) d = vertex_degree(g, x);
    const bool lg = d > RL_CHUNK;
    const uint64_t m = __ballot(lg);
    uint32_t b = 0;
    if ((threadIdx.x & 63) == 0 && m) b = atomicAdd(nlong, (uint32_t)__popcll(m));
    b = __shfl(b, 0, 64);
    if (lg) {
      const uint32_t slot = b + __popcll(m & pbf_lanemask_lt());
      list[slot] = (uint32_t)x;
      chunks[slot] = (d + RL_CHUNK - 1) / RL_CHUNK;
    }
  }
}

// W = sum_r in(r) * out(r) + C(in(r), 2) (rankWedgeF, butterfly_utils.h:216-237).
__global__ __launch_bounds__(256) void k_rank_wedges(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ inpref,
                                                     int64_t n, unsigned long long* acc) {
  uint64_t s = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    uint64_t d = roff[r + 1] - roff[r], in = inpref[r], out = d - in;
    s += in * out + in * (in - 1) / 2;
  }
  block_add(s, acc);
}

// edgeToIdxs (butterfly_utils.h:2034-2062): one thread per S-CSR position, binary search.
__global__ void k_edge_to_idxs(const uint32_t* __restrict__ offS, const uint32_t* __restrict__ edgS, int64_t ns,
                               const uint32_t* __restrict__ offO, const uint32_t* __restrict__ edgO, int64_t m,
                               uint32_t* __restrict__ eti) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m; p += (int64_t)gridDim.x * blockDim.x) {
    // owner s of position p: binary search in offS
    int64_t lo = 0, hi = ns;  // find last s with offS[s] <= p
    while (hi - lo > 1) {
      int64_t mid = (lo + hi) >> 1;
      if (offS[mid] <= (uint32_t)p) lo = mid; else hi = mid;
    }
    uint32_t s = (uint32_t)lo, o = edgS[p];
    uint32_t a = offO[o], b = offO[o + 1];
    while (a < b) {
      uint32_t mid = (a + b) >> 1;
      if (edgO[mid] < s) a = mid + 1; else b = mid;
    }
    eti[p] = a;
  }
}

// ---------------- ACOCORE: decreasing log-degree buckets ----------------
// level[x] = bitlen(remaining degree); per round, every unfinished vertex at the maximum level
// k is removed, and each unfinished neighbour below k loses one remaining degree per removed
// neighbour (getApproxCoCoreRanks, butterfly_utils.h:466-496).  lvl_cnt[l] counts unfinished
// vertices per level; state[0] holds the current k (or 0xFFFFFFFF when all are finished).
// Level counters are aggregated per workgroup in LDS (one global atomic per level and
// workgroup): a global atomic per vertex on a few dozen addresses serialises.
__global__ __launch_bounds__(256) void k_acore_init(DevGraph g, uint8_t* __restrict__ lvl, uint32_t* __restrict__ rem,
                                                    uint8_t* __restrict__ done, uint32_t* __restrict__ lvl_cnt) {
  __shared__ uint32_t h[KEYS];
  if (threadIdx.x < KEYS) h[threadIdx.x] = 0;
  __syncthreads();
  int64_t n = g.nv + g.nu;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    uint32_t d = vertex_degree(g, x);
    rem[x] = d;
    uint32_t l = pbf_bitlen(d);
    lvl[x] = (uint8_t)l;
    done[x] = 0;
    atomicAdd(&h[l], 1u);
  }
  __syncthreads();
  if (threadIdx.x < KEYS && h[threadIdx.x]) atomicAdd(&lvl_cnt[threadIdx.x], h[threadIdx.x]);
}

__global__ void k_acore_pick(uint32_t* __restrict__ lvl_cnt, uint32_t* __restrict__ state) {
  if (threadIdx.x == 0) {
    uint32_t k = 0xFFFFFFFFu;
    for (int l = KEYS - 1; l >= 0; l--)
      if (lvl_cnt[l]) { k = (uint32_t)l; break; }
    state[0] = k;
  }
}

// Remove the active vertices (level == k): one wavefront per vertex walks its neighbours and
// counts, per unfinished neighbour below k, the removed neighbours (no return value is needed,
// so the atomics do not wait; a shared append counter would serialise the whole round).
__global__ __launch_bounds__(256) void k_acore_remove(DevGraph g, const uint32_t* __restrict__ state, uint8_t* __restrict__ lvl,
                                                      uint8_t* __restrict__ done, uint32_t* __restrict__ removed,
                                                      uint32_t* __restrict__ lvl_cnt) {
  const uint32_t k = state[0];
  if (k == 0xFFFFFFFFu) return;
  __shared__ uint32_t nrem;
  if (threadIdx.x == 0) nrem = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t n = g.nv + g.nu;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  __shared__ uint32_t s_inc[4][64], s_ob[4][64];
  for (int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + wid) * 64; base < n; base += nwaves * 64) {
    // each wave scans 64 ids for actives, then walks all their lists together: entry e of the
    // wave's lists belongs to the lane whose inclusive degree prefix first exceeds e
    const int64_t x = base + lane;
    const int64_t xc = x < n ? x : n - 1;
    const uint8_t dn = done[xc], lv = lvl[xc];  // both loads in flight
    const bool a = (x < n) & (dn == 0) & (lv == k);
    const uint64_t am = __ballot(a);
    if (a) { done[x] = 1; }
    if (am && lane == 0) atomicAdd(&nrem, (uint32_t)__popcll(am));
    if (!am) continue;
    const bool xv = xc < g.nv;
    const uint64_t vm = __ballot(xv);
    uint32_t ob = 0, d = 0;
    if (a) {
      ob = xv ? g.offV[xc] : g.offU[xc - g.nv];
      d = (xv ? g.offV[xc + 1] : g.offU[xc - g.nv + 1]) - ob;
    }
    const uint32_t inc = pbf_wave_incl_scan(d);
    s_inc[wid][lane] = inc;
    s_ob[wid][lane] = ob;
    const uint32_t E = (uint32_t)__shfl((int)inc, 63);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    constexpr uint32_t PER = 4;
    for (uint32_t e0 = 0; e0 < E; e0 += 64u * PER) {
      uint32_t ec[PER], own[PER], y[PER];
#pragma unroll
      for (uint32_t i = 0; i < PER; i++) {
        const uint32_t e = e0 + i * 64u + lane;
        ec[i] = e < E ? e : E - 1u;
        own[i] = 0;
      }
      // owner: the number of lanes whose inclusive prefix is at most e (fixed depth, every
      // path's LDS read of a level in flight together)
#pragma unroll
      for (uint32_t step = 32; step; step >>= 1) {
#pragma unroll
        for (uint32_t i = 0; i < PER; i++) own[i] = s_inc[wid][own[i] + step - 1u] <= ec[i] ? own[i] + step : own[i];
      }
#pragma unroll
      for (uint32_t i = 0; i < PER; i++) {
        const uint32_t o = own[i];
        const uint32_t first = o ? s_inc[wid][o - 1u] : 0u;
        const bool ov = (vm >> o) & 1ull;
        const uint32_t* edg = ov ? g.edgV : g.edgU;
        y[i] = edg[s_ob[wid][o] + (ec[i] - first)] + (ov ? (uint32_t)g.nv : 0u);
      }
      uint32_t ly[PER];
#pragma unroll
      for (uint32_t i = 0; i < PER; i++) ly[i] = lvl[y[i]];
      // every level has arrived before the first atomic is issued (a wait placed after an
      // atomic would wait for the atomic as well)
#pragma unroll
      for (uint32_t i = 0; i < PER; i++) pbf_ready(ly[i]);
#pragma unroll
      for (uint32_t i = 0; i < PER; i++)
        // unfinished and below k (finished / active vertices have level >= k)
        if ((e0 + i * 64u + lane < E) & (ly[i] < k)) atomicAdd(&removed[y[i]], 1u);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();  // the LDS rows are rewritten by the next step
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();
  if (threadIdx.x == 0 && nrem) atomicSub(&lvl_cnt[k], nrem);
}

// Every vertex that lost neighbours this round: remaining degree and level (a scan of the
// removal counters).
__global__ __launch_bounds__(256) void k_acore_update(const uint32_t* __restrict__ state, int64_t n, uint8_t* __restrict__ lvl,
                                                      uint32_t* __restrict__ rem, uint32_t* __restrict__ removed,
                                                      uint32_t* __restrict__ lvl_cnt) {
  if (state[0] == 0xFFFFFFFFu) return;
  __shared__ int h[KEYS];
  if (threadIdx.x < KEYS) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t y = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; y < n; y += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t rv = removed[y];
    if (rv == 0) continue;
    const uint32_t r = rem[y] - rv;
    removed[y] = 0;
    rem[y] = r;
    uint32_t ol = lvl[y], nl = pbf_bitlen(r);
    if (nl != ol) {
      lvl[y] = (uint8_t)nl;
      atomicSub(&h[ol], 1);
      atomicAdd(&h[nl], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < KEYS && h[threadIdx.x]) atomicAdd(&lvl_cnt[threadIdx.x], (uint32_t)h[threadIdx.x]);
}

__global__ void k_fill_u32b(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) p[x] = v;
}

inline unsigned grid_for(int64_t n, int threads, int max_blocks = 65535 * 4) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)b;
}
}  // namespace

int pbf_dev_side(pbf_context* ctx, const DevGraph& g, int* use_v, int64_t* wedges) {
  void* mp;
  PBF_CHECK(pbf_scratch(ctx, S_MISC, 64 * sizeof(uint64_t), &mp));
  unsigned long long* acc = static_cast<unsigned long long*>(mp);
  PBF_CHECK(pbf_fill_u64(ctx, reinterpret_cast<uint64_t*>(acc), 0, 2));
  if (g.nv) hipLaunchKernelGGL(k_side_partial, dim3(grid_for(g.nv, 256, ctx->num_cus * 4)), dim3(256), 0, ctx->stream, g.offV, g.nv, acc);
  if (g.nu) hipLaunchKernelGGL(k_side_partial, dim3(grid_for(g.nu, 256, ctx->num_cus * 4)), dim3(256), 0, ctx->stream, g.offU, g.nu, acc + 1);
  PBF_HIP(hipGetLastError());
  PBF_HIP(hipMemcpyAsync(ctx->pinned, acc, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  int64_t wv = (int64_t)ctx->pinned[0], wu = (int64_t)ctx->pinned[1];
  *use_v = wv <= wu;
  *wedges = wv <= wu ? wv : wu;
  return PBF_OK;
}

// Fills R->order / R->rank_of (and the key array in S_KEY) for ADEG or ACOCORE.
int pbf_dev_rank(pbf_context* ctx, const DevGraph& g, int rank_type, RankedGraph* R, int use_v, const uint8_t* sflag,
                 bool sort) {
  const int64_t n = g.nv + g.nu;
  R->n = n;
  R->m2 = 2 * g.m;
  void *kp, *op, *rp, *hp;
  PBF_CHECK(pbf_scratch(ctx, S_KEY, (size_t)(n ? n : 1), &kp));
  PBF_CHECK(pbf_scratch(ctx, S_ORDER, 4 * (size_t)(n ? n : 1), &op));
  PBF_CHECK(pbf_scratch(ctx, S_RANKOF, 4 * (size_t)(n ? n : 1), &rp));
  uint8_t* key = static_cast<uint8_t*>(kp);
  R->order = static_cast<uint32_t*>(op);
  R->rank_of = static_cast<uint32_t*>(rp);
  if (n == 0) return PBF_OK;
  if (rank_type == PBF_RANK_ADEG) {
    hipLaunchKernelGGL(k_adeg_key, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, g, key);
  } else if (rank_type == PBF_RANK_SIDE) {
    hipLaunchKernelGGL(k_side_key, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, g, use_v, sflag, key);
  } else if (rank_type == PBF_RANK_ACOCORE) {
    void *a0, *a1, *a2;
    PBF_CHECK(pbf_scratch(ctx, S_ACORE0, 4 * (size_t)n, &a0));          // rem
    PBF_CHECK(pbf_scratch(ctx, S_ACORE1, 4 * (size_t)n, &a1));          // removed neighbours this round
    PBF_CHECK(pbf_scratch(ctx, S_ACORE2, (size_t)n, &a2));              // finished flags
    uint32_t* rem = static_cast<uint32_t*>(a0);
    uint32_t* removed = static_cast<uint32_t*>(a1);
    uint8_t* done = static_cast<uint8_t*>(a2);
    void* mp;
    PBF_CHECK(pbf_scratch(ctx, S_MISC, 64 * sizeof(uint64_t), &mp));
    uint32_t* lvl_cnt = static_cast<uint32_t*>(mp) + 64;  // 33 level counters (u32 words 64..96)
    uint32_t* state = static_cast<uint32_t*>(mp) + 112;   // k
    PBF_CHECK(pbf_fill_u32(ctx, removed, 0u, n));
    PBF_CHECK(pbf_fill_u32(ctx, lvl_cnt, 0u, 48));
    hipLaunchKernelGGL(k_acore_init, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, g, key, rem, done, lvl_cnt);
    // at most KEYS non-empty levels, and k strictly decreases every round
    for (int round = 0; round < KEYS; round++) {
      hipLaunchKernelGGL(k_acore_pick, dim3(1), dim3(64), 0, ctx->stream, lvl_cnt, state);
      hipLaunchKernelGGL(k_acore_remove, dim3(grid_for((n + 63) / 64, 4, 8192)), dim3(256), 0, ctx->stream, g, state, key,
                         done, removed, lvl_cnt);
      hipLaunchKernelGGL(k_acore_update, dim3(grid_for(n, 256, 4096)), dim3(256), 0, ctx->stream, state, n, key, rem,
                         removed, lvl_cnt);
    }
  } else {
    return PBF_ERR_UNSUPPORTED;
  }
  if (!sort) return PBF_OK;
  const int64_t nb = (n + 1023) / 1024;
  PBF_CHECK(pbf_scratch(ctx, S_HIST, 4 * (size_t)(nb * KEYS), &hp));
  uint32_t* hist = static_cast<uint32_t*>(hp);
  hipLaunchKernelGGL(k_key_hist, dim3((unsigned)nb), dim3(1024), 0, ctx->stream, key, n, hist);
  PBF_CHECK(pbf_scan_excl_u32(ctx, hist, hist, nb * KEYS, nullptr));
  hipLaunchKernelGGL(k_key_scatter, dim3((unsigned)nb), dim3(1024), 0, ctx->stream, key, n, hist, R->order, R->rank_of);
  PBF_HIP(hipGetLastError());
  return PBF_OK;
}

int pbf_dev_relabel(pbf_context* ctx, const DevGraph& g, int use_v, bool want_pos, RankedGraph* R, const uint8_t* sflag,
                    bool sync) {
  const int64_t n = g.nv + g.nu, m2 = 2 * g.m;
  void *ro, *re, *rpp, *ip, *mp;
  PBF_CHECK(pbf_scratch(ctx, S_ROFF, 4 * (size_t)(n + 1), &ro));
  PBF_CHECK(pbf_scratch(ctx, S_REDG, 4 * (size_t)(m2 ? m2 : 1), &re));
  PBF_CHECK(pbf_scratch(ctx, S_INPREF, 4 * (size_t)(n ? n : 1), &ip));
  PBF_CHECK(pbf_scratch(ctx, S_MISC, 64 * sizeof(uint64_t), &mp));
  void* rwp;
  PBF_CHECK(pbf_scratch(ctx, S_ROWN, 4 * (size_t)(m2 ? m2 : 1), &rwp));
  R->rown = static_cast<uint32_t*>(rwp);
  R->roff = static_cast<uint32_t*>(ro);
  R->redg = static_cast<uint32_t*>(re);
  R->inpref = static_cast<uint32_t*>(ip);
  R->rpos = nullptr;
  if (want_pos) {
    PBF_CHECK(pbf_scratch(ctx, S_RPOS, 4 * (size_t)(m2 ? m2 : 1), &rpp));
    R->rpos = static_cast<uint32_t*>(rpp);
  }
  if (n == 0) { R->wedges = 0; return PBF_OK; }
  const uint8_t* key = static_cast<const uint8_t*>(ctx->buf[S_KEY]);
  // (rank, key) per vertex in one 8-byte record: a neighbour's two are one gather (one line)
  void* rkp;
  PBF_CHECK(pbf_scratch(ctx, S_RKEY, 8 * (size_t)n, &rkp));
  const uint2* rk = static_cast<const uint2*>(rkp);
  hipLaunchKernelGGL(k_rank_key, dim3(grid_for(n, 256, 16384)), dim3(256), 0, ctx->stream, R->rank_of, key, n,
                     const_cast<uint2*>(rk));
  hipLaunchKernelGGL(k_ranked_degree, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, g, R->order, R->roff);
  PBF_CHECK(pbf_scan_excl_u32(ctx, R->roff, R->roff, n + 1, nullptr));
  hipLaunchKernelGGL(k_relabel_small, dim3(grid_for(n, 256, 16384)), dim3(256), 0, ctx->stream, g, use_v, sflag, rk,
                     R->roff, R->lo, R->redg, R->rpos, R->inpref, R->rown);
  hipLaunchKernelGGL(k_relabel, dim3(grid_for(n, 4, 16384)), dim3(256), 0, ctx->stream, g, use_v, sflag, rk, R->roff,
                     R->lo, R->redg, R->rpos, R->inpref, R->rown);
  // long lists, chunk-parallel
  {
    const int64_t max_items = 2 * (2 * g.m / RL_CHUNK) + 64;
    void *lp, *hp2;
    PBF_CHECK(pbf_scratch(ctx, S_RLONG, 4 * (size_t)(3 * n + 64), &lp));
    PBF_CHECK(pbf_scratch(ctx, S_RHIST, 4 * (size_t)max_items * KEYS, &hp2));
    uint32_t* llist = static_cast<uint32_t*>(lp);
    uint32_t* lchunks = llist + n;
    uint32_t* lbase = llist + 2 * n;
    uint32_t* lctl = llist + 3 * n;  // [0] nlong, [1] nitems
    hipLaunchKernelGGL(k_fill_u32b, dim3(grid_for(n + 64, 256, 4096)), dim3(256), 0, ctx->stream, lchunks, 0u, n);
    hipLaunchKernelGGL(k_fill_u32b, dim3(1), dim3(64), 0, ctx->stream, lctl, 0u, (int64_t)2);
    hipLaunchKernelGGL(k_rl_find_long, dim3(grid_for(n, 256, 4096)), dim3(256), 0, ctx->stream, g, llist, lchunks, lctl);
    PBF_CHECK(pbf_scan_excl_u32(ctx, lchunks, lbase, n, lctl + 1));
    LongLists L{llist, lbase, lctl, lctl + 1, static_cast<uint32_t*>(hp2)};
    hipLaunchKernelGGL(k_rl_hist, dim3(ctx->num_cus * 8), dim3(256), 0, ctx->stream, g, rk, L);
    hipLaunchKernelGGL(k_rl_offsets, dim3(grid_for(n, 4, 2048)), dim3(256), 0, ctx->stream, g, key, R->rank_of, L, R->inpref);
    hipLaunchKernelGGL(k_rl_scatter, dim3(ctx->num_cus * 8), dim3(256), 0, ctx->stream, g, use_v, sflag, rk, R->roff, R->lo,
                       L, R->redg, R->rpos, R->rown);
  }
  unsigned long long* acc = static_cast<unsigned long long*>(mp) + 8;
  PBF_CHECK(pbf_fill_u64(ctx, reinterpret_cast<uint64_t*>(acc), 0, 1));
  hipLaunchKernelGGL(k_rank_wedges, dim3(grid_for(n, 256, ctx->num_cus * 4)), dim3(256), 0, ctx->stream, R->roff, R->inpref, n, acc);
  PBF_HIP(hipGetLastError());
  R->wedges_dev = acc;
  R->wedges = -1;
  if (!sync) return PBF_OK;
  PBF_HIP(hipMemcpyAsync(ctx->pinned + 8, acc, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  R->wedges = (int64_t)ctx->pinned[8];
  return PBF_OK;
}

int pbf_dev_edge_to_idxs(pbf_context* ctx, const DevGraph& g, int use_v, uint32_t* eti) {
  if (g.m == 0) return PBF_OK;
  const uint32_t* offS = use_v ? g.offU : g.offV;
  const uint32_t* edgS = use_v ? g.edgU : g.edgV;
  const uint32_t* offO = use_v ? g.offV : g.offU;
  const uint32_t* edgO = use_v ? g.edgV : g.edgU;
  int64_t ns = use_v ? g.nu : g.nv;
  hipLaunchKernelGGL(k_edge_to_idxs, dim3(grid_for(g.m, 256, 16384)), dim3(256), 0, ctx->stream, offS, edgS, ns, offO, edgO,
                     g.m, eti);
  PBF_HIP(hipGetLastError());
  return PBF_OK;
}
