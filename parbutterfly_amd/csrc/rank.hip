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

__global__ void k_side_partial(const uint32_t* __restrict__ off, int64_t n, unsigned long long* acc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t s = 0;
  for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t d = off[i + 1] - off[i];
    s += d * (d - 1) / 2;
  }
  s = pbf_wave_sum(s);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(acc, (unsigned long long)s);
}

__device__ __forceinline__ uint32_t vertex_degree(const DevGraph& g, int64_t x) {
  return x < g.nv ? g.offV[x + 1] - g.offV[x] : g.offU[x - g.nv + 1] - g.offU[x - g.nv];
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

// Lanes of the wavefront holding the same 6-bit key (inactive lanes excluded via `active`).
__device__ __forceinline__ uint64_t match_key6(uint32_t k, uint64_t active) {
  uint64_t m = active;
#pragma unroll
  for (int b = 0; b < 6; b++) {
    uint64_t bal = __ballot((k >> b) & 1u);
    m &= ((k >> b) & 1u) ? bal : ~bal;
  }
  return m;
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
  uint64_t mm = match_key6(k, active);
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

__global__ void k_ranked_degree(DevGraph g, const uint32_t* __restrict__ order, uint32_t* __restrict__ rdeg) {
  int64_t n = g.nv + g.nu;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    rdeg[r] = vertex_degree(g, order[r]);
}

// One wavefront per original vertex x: stable counting sort of x's neighbour list by key
// (descending), written at roff[rank_of[x]].  flag = 1 iff the neighbour lies on S.
__global__ __launch_bounds__(256) void k_relabel(DevGraph g, int use_v, const uint8_t* __restrict__ key,
                                                 const uint32_t* __restrict__ rank_of, const uint32_t* __restrict__ roff,
                                                 uint32_t* __restrict__ redg, uint32_t* __restrict__ rpos,
                                                 uint32_t* __restrict__ inpref) {
  __shared__ uint32_t cnt[4][KEYS + 1];
  __shared__ uint32_t run[4][KEYS + 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t n = g.nv + g.nu;
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  for (int64_t x = (int64_t)blockIdx.x * 4 + wid; x < n; x += nwaves) {
    const bool xv = x < g.nv;
    const uint32_t obase = xv ? g.offV[x] : g.offU[x - g.nv];
    const uint32_t deg = xv ? g.offV[x + 1] - obase : g.offU[x - g.nv + 1] - obase;
    const uint32_t* adj = xv ? g.edgV + obase : g.edgU + obase;
    const int64_t nbr_shift = xv ? g.nv : 0;  // combined id of neighbour = adj[j] + shift
    const uint32_t flag = use_v ? (xv ? 1u : 0u) : (xv ? 0u : 1u);
    const uint32_t r = rank_of[x];
    if (lane < KEYS + 1) { cnt[wid][lane] = 0; run[wid][lane] = 0; }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // pass A: key histogram of the neighbours
    for (uint32_t j = lane; j < deg; j += 64) atomicAdd(&cnt[wid][key[adj[j] + nbr_shift]], 1u);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // exclusive prefix in descending key order: start[k] = sum_{k' > k} cnt[k']
    uint32_t c = lane < KEYS ? cnt[wid][KEYS - 1 - lane] : 0;
    uint32_t incl = pbf_wave_incl_scan(c);
    if (lane < KEYS) run[wid][KEYS - 1 - lane] = incl - c;  // run[k] = start of class k
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
      const uint32_t kself = key[x];
      // in(r): neighbours of strictly higher key, plus same-key neighbours when x is in U
      inpref[r] = run[wid][kself] + (xv ? 0u : cnt[wid][kself]);
    }
    const uint32_t rbase = roff[r];
    // pass B: stable scatter, 64 neighbours at a time
    for (uint32_t j0 = 0; j0 < deg; j0 += 64) {
      const uint32_t j = j0 + lane;
      const bool act = j < deg;
      uint32_t y = 0, ky = 0;
      if (act) { y = (uint32_t)(adj[j] + nbr_shift); ky = key[y]; }
      const uint64_t active = __ballot(act);
      const uint64_t mm = match_key6(ky, active);
      const uint32_t rk = __popcll(mm & pbf_lanemask_lt());
      uint32_t slot = 0;
      if (act) {
        slot = rbase + run[wid][ky] + rk;
        redg[slot] = (rank_of[y] << 1) | flag;
        if (rpos) rpos[slot] = obase + j;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (act && rk == 0) run[wid][ky] += __popcll(mm);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// W = sum_r in(r) * out(r) + C(in(r), 2) (rankWedgeF, butterfly_utils.h:216-237).
__global__ void k_rank_wedges(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ inpref, int64_t n,
                              unsigned long long* acc) {
  uint64_t s = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    uint64_t d = roff[r + 1] - roff[r], in = inpref[r], out = d - in;
    s += in * out + in * (in - 1) / 2;
  }
  s = pbf_wave_sum(s);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(acc, (unsigned long long)s);
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
__global__ void k_acore_init(DevGraph g, uint8_t* __restrict__ lvl, uint32_t* __restrict__ rem,
                             uint8_t* __restrict__ done, uint32_t* __restrict__ lvl_cnt) {
  int64_t n = g.nv + g.nu;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    uint32_t d = vertex_degree(g, x);
    rem[x] = d;
    uint32_t l = pbf_bitlen(d);
    lvl[x] = (uint8_t)l;
    done[x] = 0;
    atomicAdd(&lvl_cnt[l], 1u);
  }
}

__global__ void k_acore_pick(uint32_t* __restrict__ lvl_cnt, uint32_t* __restrict__ state) {
  if (threadIdx.x == 0) {
    uint32_t k = 0xFFFFFFFFu;
    for (int l = KEYS - 1; l >= 0; l--)
      if (lvl_cnt[l]) { k = (uint32_t)l; break; }
    state[0] = k;
  }
}

// Remove the active vertices (level == k): one wavefront per vertex walks its neighbours.
__global__ __launch_bounds__(256) void k_acore_remove(DevGraph g, const uint32_t* __restrict__ state, uint8_t* __restrict__ lvl,
                                                      uint8_t* __restrict__ done, uint32_t* __restrict__ removed,
                                                      uint32_t* __restrict__ lvl_cnt, uint32_t* __restrict__ touched,
                                                      uint32_t* __restrict__ ntouched) {
  const uint32_t k = state[0];
  if (k == 0xFFFFFFFFu) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t n = g.nv + g.nu;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + wid) * 64; base < n; base += nwaves * 64) {
    // each wave scans 64 ids for actives, then walks their lists cooperatively
    int64_t x = base + lane;
    bool a = x < n && !done[x] && lvl[x] == k;
    uint64_t am = __ballot(a);
    if (a) { done[x] = 1; }
    if (am && lane == 0) atomicSub(&lvl_cnt[k], (uint32_t)__popcll(am));
    while (am) {
      int src = __ffsll((long long)am) - 1;
      am &= am - 1;
      int64_t xa = base + src;
      const bool xv = xa < g.nv;
      const uint32_t ob = xv ? g.offV[xa] : g.offU[xa - g.nv];
      const uint32_t d = xv ? g.offV[xa + 1] - ob : g.offU[xa - g.nv + 1] - ob;
      const uint32_t* adj = xv ? g.edgV + ob : g.edgU + ob;
      const int64_t sh = xv ? g.nv : 0;
      for (uint32_t j = lane; j < d; j += 64) {
        uint32_t y = (uint32_t)(adj[j] + sh);
        if (lvl[y] < k) {  // unfinished and below k (finished / active vertices have level >= k)
          if (atomicAdd(&removed[y], 1u) == 0) touched[atomicAdd(ntouched, 1u)] = y;
        }
      }
    }
  }
}

__global__ void k_acore_update(const uint32_t* __restrict__ touched, const uint32_t* __restrict__ ntouched,
                               uint8_t* __restrict__ lvl, uint32_t* __restrict__ rem, uint32_t* __restrict__ removed,
                               uint32_t* __restrict__ lvl_cnt) {
  uint32_t nt = *ntouched;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    uint32_t y = touched[t];
    uint32_t r = rem[y] - removed[y];
    removed[y] = 0;
    rem[y] = r;
    uint32_t ol = lvl[y], nl = pbf_bitlen(r);
    if (nl != ol) {
      lvl[y] = (uint8_t)nl;
      atomicSub(&lvl_cnt[ol], 1u);
      atomicAdd(&lvl_cnt[nl], 1u);
    }
  }
}

__global__ void k_reset_u32(uint32_t* p) { *p = 0; }

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
  PBF_HIP(hipMemsetAsync(acc, 0, 2 * sizeof(uint64_t), ctx->stream));
  if (g.nv) hipLaunchKernelGGL(k_side_partial, dim3(grid_for(g.nv, 256, 4096)), dim3(256), 0, ctx->stream, g.offV, g.nv, acc);
  if (g.nu) hipLaunchKernelGGL(k_side_partial, dim3(grid_for(g.nu, 256, 4096)), dim3(256), 0, ctx->stream, g.offU, g.nu, acc + 1);
  PBF_HIP(hipGetLastError());
  PBF_HIP(hipMemcpyAsync(ctx->pinned, acc, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  int64_t wv = (int64_t)ctx->pinned[0], wu = (int64_t)ctx->pinned[1];
  *use_v = wv <= wu;
  *wedges = wv <= wu ? wv : wu;
  return PBF_OK;
}

// Fills R->order / R->rank_of (and the key array in S_KEY) for ADEG or ACOCORE.
int pbf_dev_rank(pbf_context* ctx, const DevGraph& g, int rank_type, RankedGraph* R) {
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
  } else if (rank_type == PBF_RANK_ACOCORE) {
    void *a0, *a1, *a2;
    PBF_CHECK(pbf_scratch(ctx, S_ACORE0, 4 * (size_t)n, &a0));          // rem
    PBF_CHECK(pbf_scratch(ctx, S_ACORE1, 4 * (size_t)n + 4096, &a1));   // removed | touched? (split below)
    PBF_CHECK(pbf_scratch(ctx, S_ACORE2, 4 * (size_t)n + (size_t)n + 1024, &a2));
    uint32_t* rem = static_cast<uint32_t*>(a0);
    uint32_t* removed = static_cast<uint32_t*>(a1);
    uint32_t* touched = reinterpret_cast<uint32_t*>(a2);
    uint8_t* done = reinterpret_cast<uint8_t*>(touched + n);
    void* mp;
    PBF_CHECK(pbf_scratch(ctx, S_MISC, 64 * sizeof(uint64_t), &mp));
    uint32_t* lvl_cnt = static_cast<uint32_t*>(mp) + 64;  // 33 level counters (u32 words 64..96)
    uint32_t* state = static_cast<uint32_t*>(mp) + 112;   // k, ntouched
    PBF_HIP(hipMemsetAsync(removed, 0, 4 * (size_t)n, ctx->stream));
    PBF_HIP(hipMemsetAsync(lvl_cnt, 0, 4 * 48, ctx->stream));
    hipLaunchKernelGGL(k_acore_init, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, g, key, rem, done, lvl_cnt);
    // at most KEYS non-empty levels, and k strictly decreases every round
    for (int round = 0; round < KEYS; round++) {
      hipLaunchKernelGGL(k_acore_pick, dim3(1), dim3(64), 0, ctx->stream, lvl_cnt, state);
      hipLaunchKernelGGL(k_reset_u32, dim3(1), dim3(1), 0, ctx->stream, state + 1);
      hipLaunchKernelGGL(k_acore_remove, dim3(grid_for((n + 63) / 64, 4, 8192)), dim3(256), 0, ctx->stream, g, state, key,
                         done, removed, lvl_cnt, touched, state + 1);
      hipLaunchKernelGGL(k_acore_update, dim3(grid_for(n, 256, 2048)), dim3(256), 0, ctx->stream, touched, state + 1, key,
                         rem, removed, lvl_cnt);
    }
  } else {
    return PBF_ERR_UNSUPPORTED;
  }
  const int64_t nb = (n + 1023) / 1024;
  PBF_CHECK(pbf_scratch(ctx, S_HIST, 4 * (size_t)(nb * KEYS), &hp));
  uint32_t* hist = static_cast<uint32_t*>(hp);
  hipLaunchKernelGGL(k_key_hist, dim3((unsigned)nb), dim3(1024), 0, ctx->stream, key, n, hist);
  PBF_CHECK(pbf_scan_excl_u32(ctx, hist, hist, nb * KEYS, nullptr));
  hipLaunchKernelGGL(k_key_scatter, dim3((unsigned)nb), dim3(1024), 0, ctx->stream, key, n, hist, R->order, R->rank_of);
  PBF_HIP(hipGetLastError());
  return PBF_OK;
}

int pbf_dev_relabel(pbf_context* ctx, const DevGraph& g, int use_v, bool want_pos, RankedGraph* R) {
  const int64_t n = g.nv + g.nu, m2 = 2 * g.m;
  void *ro, *re, *rpp, *ip, *mp;
  PBF_CHECK(pbf_scratch(ctx, S_ROFF, 4 * (size_t)(n + 1), &ro));
  PBF_CHECK(pbf_scratch(ctx, S_REDG, 4 * (size_t)(m2 ? m2 : 1), &re));
  PBF_CHECK(pbf_scratch(ctx, S_INPREF, 4 * (size_t)(n ? n : 1), &ip));
  PBF_CHECK(pbf_scratch(ctx, S_MISC, 64 * sizeof(uint64_t), &mp));
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
  hipLaunchKernelGGL(k_ranked_degree, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, g, R->order, R->roff);
  PBF_CHECK(pbf_scan_excl_u32(ctx, R->roff, R->roff, n + 1, nullptr));
  hipLaunchKernelGGL(k_relabel, dim3(grid_for(n, 4, 16384)), dim3(256), 0, ctx->stream, g, use_v, key, R->rank_of, R->roff,
                     R->redg, R->rpos, R->inpref);
  unsigned long long* acc = static_cast<unsigned long long*>(mp) + 8;
  PBF_HIP(hipMemsetAsync(acc, 0, sizeof(uint64_t), ctx->stream));
  hipLaunchKernelGGL(k_rank_wedges, dim3(grid_for(n, 256, 4096)), dim3(256), 0, ctx->stream, R->roff, R->inpref, n, acc);
  PBF_HIP(hipGetLastError());
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
