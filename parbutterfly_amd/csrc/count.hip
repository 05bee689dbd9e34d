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

// Binning: one wavefront per endpoint i computes UB_i = sum_{v in N(i)} in(v) (>= wedges of
// i) and bound_i = min(UB_i, i) (>= distinct u2, since u2 < i), then appends i to the small
// (bound <= small_max) or large list.  Endpoints without wedges are skipped.
__global__ __launch_bounds__(256) void k_bin(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ redg,
                                             const uint32_t* __restrict__ inpref, int64_t n, uint32_t small_max,
                                             uint32_t* __restrict__ bound, uint32_t* __restrict__ small_list,
                                             uint32_t* __restrict__ large_list, uint32_t* __restrict__ lens) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  for (int64_t i = (int64_t)blockIdx.x * 4 + wid; i < n; i += nwaves) {
    const uint32_t io = roff[i], deg = roff[i + 1] - io;
    uint64_t ub = 0;
    for (uint32_t j = lane; j < deg; j += 64) ub += inpref[redg[io + j] >> 1];
    ub = pbf_wave_sum(ub);
    if (lane == 0 && ub) {
      const uint32_t b = (uint32_t)(ub < (uint64_t)i ? ub : (uint64_t)i);
      bound[i] = b;
      if (b == 0) continue;
      if (b <= small_max) small_list[atomicAdd(&lens[0], 1u)] = (uint32_t)i;
      else large_list[atomicAdd(&lens[1], 1u)] = (uint32_t)i;
    }
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

// CountRank's conversion (butterfly_count.h:1600-1610): out[x] = B[rank(x)] for x in S.
__global__ void k_convert_vert(const unsigned long long* __restrict__ B, const uint32_t* __restrict__ rank_of, int64_t base,
                               int64_t ns, unsigned long long* __restrict__ out) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < ns; x += (int64_t)gridDim.x * blockDim.x)
    out[x] = B[rank_of[base + x]];
}

inline unsigned grid_for(int64_t n, int threads, int max_blocks) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)b;
}

constexpr int SMALL_TEAM = 64, SMALL_CAP = 1024;
constexpr int LARGE_TEAM = 256, LARGE_CAP = 8192;
constexpr int HBM_TEAM = 256, HBM_BLOCKS = 32;

template <int MODE>
int run_count(pbf_context* ctx, const RankedGraph& R, CountArgs a, uint32_t* lists, uint32_t* lens) {
  const int64_t n = R.n;
  const uint32_t small_max = (SMALL_CAP / 4) * 3 - SMALL_TEAM;
  // lists: [small n][large n][ovf n]; lens: [small, large, ovf, q_small, q_large, q_ovf]
  PBF_HIP(hipMemsetAsync(lens, 0, 8 * sizeof(uint32_t), ctx->stream));
  hipLaunchKernelGGL(k_bin, dim3(grid_for(n, 4, 16384)), dim3(256), 0, ctx->stream, R.roff, R.redg, R.inpref, n, small_max,
                     const_cast<uint32_t*>(a.bound), lists, lists + n, lens);
  PBF_HIP(hipGetLastError());
  // small endpoints: one wavefront each
  a.list = lists; a.list_len = lens + 0; a.queue = lens + 3;
  a.ovf_list = lists + 2 * n; a.ovf_len = lens + 2;
  hipLaunchKernelGGL((k_count<MODE, SMALL_TEAM, SMALL_CAP, true, T_LDS>), dim3(ctx->num_cus * 12), dim3(SMALL_TEAM), 0,
                     ctx->stream, a);
  // heavy endpoints: one 256-thread workgroup each
  a.list = lists + n; a.list_len = lens + 1; a.queue = lens + 4;
  hipLaunchKernelGGL((k_count<MODE, LARGE_TEAM, LARGE_CAP, false, T_LDS>), dim3(ctx->num_cus * 2), dim3(LARGE_TEAM), 0,
                     ctx->stream, a);
  PBF_HIP(hipGetLastError());
  // overflowed endpoints (distinct u2 beyond the LDS table): dense HBM histograms
  PBF_HIP(hipMemcpyAsync(ctx->pinned + 16, lens + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  const uint32_t novf = (uint32_t)(ctx->pinned[16] & 0xFFFFFFFFu);
  ctx->timing.hub_vertices = novf;
  if (novf) {
    void* pp;
    const int blocks = (int)(novf < (uint32_t)HBM_BLOCKS ? novf : (uint32_t)HBM_BLOCKS);
    PBF_CHECK(pbf_scratch(ctx, S_POOL, sizeof(uint32_t) * 2 * (size_t)n * blocks, &pp));
    PBF_HIP(hipMemsetAsync(pp, 0, sizeof(uint32_t) * 2 * (size_t)n * blocks, ctx->stream));
    a.pool = static_cast<uint32_t*>(pp);
    a.list = lists + 2 * n; a.list_len = lens + 2; a.queue = lens + 5;
    a.ovf_list = lists; a.ovf_len = lens + 6;  // never written by the HBM kernel
    hipLaunchKernelGGL((k_count<MODE, HBM_TEAM, 1, true, T_HBM>), dim3(blocks), dim3(HBM_TEAM), 0, ctx->stream, a);
    PBF_HIP(hipGetLastError());
  }
  return PBF_OK;
}
}  // namespace

int pbf_dev_count(pbf_context* ctx, const DevGraph& g, int use_v, const RankedGraph& R, int mode, uint64_t* out_dev,
                  const uint32_t* eti) {
  const int64_t n = R.n, m2 = R.m2;
  void *bp, *lp, *cp, *mp, *op = nullptr;
  PBF_CHECK(pbf_scratch(ctx, S_BOUND, 4 * (size_t)(n ? n : 1), &bp));
  PBF_CHECK(pbf_scratch(ctx, S_BINS, 4 * (size_t)(3 * n + 64), &lp));
  PBF_CHECK(pbf_scratch(ctx, S_MISC, 64 * sizeof(uint64_t), &mp));
  uint32_t* lists = static_cast<uint32_t*>(lp);
  uint32_t* lens = lists + 3 * n;
  CountArgs a{};
  a.roff = R.roff; a.redg = R.redg; a.inpref = R.inpref; a.bound = static_cast<uint32_t*>(bp);
  a.n = n;
  a.total = static_cast<unsigned long long*>(mp) + 12;
  if (mode == MODE_VERT) {
    PBF_CHECK(pbf_scratch(ctx, S_COUNTS, 8 * (size_t)(n ? n : 1), &cp));
    a.B = static_cast<unsigned long long*>(cp);
    PBF_HIP(hipMemsetAsync(a.B, 0, 8 * (size_t)n, ctx->stream));
    PBF_CHECK(run_count<MODE_VERT>(ctx, R, a, lists, lens));
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
    PBF_HIP(hipMemsetAsync(out_dev, 0, 8 * (size_t)g.m, ctx->stream));
    PBF_CHECK(run_count<MODE_EDGE>(ctx, R, a, lists, lens));
  } else {
    PBF_HIP(hipMemsetAsync(a.total, 0, 8, ctx->stream));
    PBF_CHECK(run_count<MODE_TOTAL>(ctx, R, a, lists, lens));
    PBF_HIP(hipMemcpyAsync(out_dev, a.total, 8, hipMemcpyDeviceToDevice, ctx->stream));
  }
  PBF_HIP(hipGetLastError());
  return PBF_OK;
}
