// Batched counting of many independent bipartite graphs (no counterpart in the reference; the
// per-graph results equal those of Count / CountRank, butterfly_count.h:1561-1665, graph by
// graph).
//
// PBF_BATCH_UNION: the graphs are laid out as one disjoint union -- V ids of all graphs first,
// then all U ids, each graph keeping its own order -- with a per-vertex flag for the count side
// S of its graph (cmpWedgeCounts per graph, butterfly_utils.h:1061-1086).  One pass of the
// ranking, relabel and counting kernels then serves every graph: wedges never cross graphs,
// the global ranking restricted to a graph is that graph's own ranking (ties by ascending id,
// V before U), and the persistent work-queue kernels balance the uneven graphs among
// themselves.  Per-graph outputs are gathered at the end.  With PBF_BATCH_CAPTURE the pass is
// captured in a HIP graph (no host synchronisation, no allocation, no memset inside) and
// replayed by identical calls.
// PBF_BATCH_PLAIN: the same kernels, launched graph by graph from a host loop (the baseline).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pbf_device.h"

namespace {
struct BGraph {
  const uint32_t* offV;
  const uint32_t* edgV;
  const uint32_t* offU;
  const uint32_t* edgU;
  uint32_t nv, nu, m, pad;
  uint64_t vbase, ubase, ebase;  // union: V ids vbase.., U ids (within the U side) ubase.., edges ebase..
};

// One workgroup per graph: union offsets and edges (neighbour ids shifted to the union).
__global__ __launch_bounds__(256) void k_union_build(const BGraph* __restrict__ gs, int64_t count, uint32_t* __restrict__ offV,
                                                     uint32_t* __restrict__ edgV, uint32_t* __restrict__ offU,
                                                     uint32_t* __restrict__ edgU, int64_t NV, int64_t NU, int64_t M) {
  for (int64_t g = blockIdx.x; g < count; g += gridDim.x) {
    const BGraph b = gs[g];
    const uint32_t eb = (uint32_t)b.ebase, ub = (uint32_t)b.ubase, vb = (uint32_t)b.vbase;
    for (uint32_t v = threadIdx.x; v < b.nv; v += blockDim.x) offV[vb + v] = eb + b.offV[v];
    for (uint32_t u = threadIdx.x; u < b.nu; u += blockDim.x) offU[ub + u] = eb + b.offU[u];
    for (uint32_t e = threadIdx.x; e < b.m; e += blockDim.x) {
      edgV[eb + e] = ub + b.edgV[e];
      edgU[eb + e] = vb + b.edgU[e];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    offV[NV] = (uint32_t)M;
    offU[NU] = (uint32_t)M;
  }
}

__device__ __forceinline__ unsigned long long block_sum(unsigned long long x, unsigned long long* sm) {
  x = pbf_wave_sum(x);
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sm[w] = x;
  __syncthreads();
  unsigned long long t = 0;
  for (uint32_t k = 0; k < (blockDim.x >> 6); k++) t += sm[k];
  __syncthreads();
  return t;
}

// One workgroup per graph: wV, wU (cmpWedgeCounts), the side choice, the per-vertex S flags of
// the union and |S|.
__global__ __launch_bounds__(256) void k_batch_side(const BGraph* __restrict__ gs, int64_t count, int64_t NV,
                                                    uint8_t* __restrict__ sflag, unsigned long long* __restrict__ sn,
                                                    uint8_t* __restrict__ usev, unsigned long long* __restrict__ wvu) {
  __shared__ unsigned long long sm[4];
  for (int64_t g = blockIdx.x; g < count; g += gridDim.x) {
    const BGraph b = gs[g];
    unsigned long long wv = 0, wu = 0;
    for (uint32_t v = threadIdx.x; v < b.nv; v += blockDim.x) wv += (unsigned long long)pbf_side_term(b.offV[v + 1] - b.offV[v]);
    for (uint32_t u = threadIdx.x; u < b.nu; u += blockDim.x) wu += (unsigned long long)pbf_side_term(b.offU[u + 1] - b.offU[u]);
    wv = block_sum(wv, sm);
    wu = block_sum(wu, sm);
    const bool use_v = wv <= wu;  // ties count on U
    for (uint32_t v = threadIdx.x; v < b.nv; v += blockDim.x) sflag[b.vbase + v] = use_v ? 0 : 1;
    for (uint32_t u = threadIdx.x; u < b.nu; u += blockDim.x) sflag[NV + b.ubase + u] = use_v ? 1 : 0;
    if (threadIdx.x == 0) {
      sn[g] = use_v ? b.nu : b.nv;
      usev[g] = use_v ? 1 : 0;
      if (wvu) { wvu[2 * g] = wv; wvu[2 * g + 1] = wu; }
    }
  }
}

// One workgroup per graph: the graph's ranking in its own block of ranks.  A stable counting
// sort of the graph's vertices by key (descending; ties in id order, V before U), placed at
// ranks [base, base + nv + nu), base = the vertices of the graphs before it.  Restricted to
// the graph this is the order it gets when ranked alone, and lo[r] = base lets the relabel
// store graph-local ranks, so the counting kernels size and key their tables per graph.
constexpr uint32_t RKEYS = 33;
__global__ __launch_bounds__(256) void k_local_rank(const BGraph* __restrict__ gs, int64_t count, int64_t NV,
                                                    const uint8_t* __restrict__ key, uint32_t* __restrict__ order,
                                                    uint32_t* __restrict__ rank_of, uint32_t* __restrict__ lo) {
  __shared__ uint32_t start[RKEYS];
  __shared__ uint32_t wc[4][RKEYS];
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int64_t g = blockIdx.x; g < count; g += gridDim.x) {
    const BGraph b = gs[g];
    const uint32_t nv = b.nv, n = b.nv + b.nu;
    const uint32_t base = (uint32_t)(b.vbase + b.ubase);
    const uint32_t vb = (uint32_t)b.vbase, ub = (uint32_t)(NV + (int64_t)b.ubase);
    if (threadIdx.x < RKEYS) start[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) atomicAdd(&start[key[t < nv ? vb + t : ub + (t - nv)]], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t acc = 0;
      for (int k = (int)RKEYS - 1; k >= 0; k--) { const uint32_t c = start[k]; start[k] = acc; acc += c; }
    }
    __syncthreads();
    for (uint32_t t0 = 0; t0 < n; t0 += blockDim.x) {
      const uint32_t t = t0 + threadIdx.x;
      const bool act = t < n;
      const uint32_t x = act ? (t < nv ? vb + t : ub + (t - nv)) : 0u;
      const uint32_t k = act ? key[x] : 0u;
      const uint64_t mm = pbf_match_key6(k, __ballot(act));
      const uint32_t rk = __popcll(mm & pbf_lanemask_lt());
      if (lane < RKEYS) wc[wid][lane] = 0;
      __syncthreads();
      if (act && rk == 0) wc[wid][k] = __popcll(mm);
      __syncthreads();
      if (act) {
        uint32_t before = start[k];
        for (uint32_t w = 0; w < wid; w++) before += wc[w][k];
        const uint32_t r = base + before + rk;
        order[r] = x;
        rank_of[x] = r;
        lo[r] = base;
      }
      __syncthreads();
      if (threadIdx.x < RKEYS)
        start[threadIdx.x] += wc[0][threadIdx.x] + wc[1][threadIdx.x] + wc[2][threadIdx.x] + wc[3][threadIdx.x];
      __syncthreads();
    }
  }
}

// One workgroup per graph: CountRank's conversion (out[x] = B[rank(x)] for x in S) into the
// graph's output segment, and the graph's total = sum over S / 2.
__global__ __launch_bounds__(256) void k_batch_out(const BGraph* __restrict__ gs, int64_t count, int64_t NV,
                                                   const uint8_t* __restrict__ usev, const unsigned long long* __restrict__ voff,
                                                   const unsigned long long* __restrict__ B, const uint32_t* __restrict__ rank_of,
                                                   unsigned long long* __restrict__ out, unsigned long long* __restrict__ totals) {
  __shared__ unsigned long long sm[4];
  for (int64_t g = blockIdx.x; g < count; g += gridDim.x) {
    const BGraph b = gs[g];
    const bool use_v = usev[g] != 0;
    const uint32_t ns = use_v ? b.nu : b.nv;
    const uint64_t base = use_v ? (uint64_t)NV + b.ubase : b.vbase;
    unsigned long long* o = out + voff[g];
    unsigned long long s = 0;
    for (uint32_t x = threadIdx.x; x < ns; x += blockDim.x) {
      const unsigned long long c = B[rank_of[base + x]];
      o[x] = c;
      s += c;
    }
    s = block_sum(s, sm);
    if (threadIdx.x == 0) totals[g] = s / 2;
  }
}

// Plain mode: total of one graph from its per-vertex counts.
__global__ __launch_bounds__(256) void k_total_of(const unsigned long long* __restrict__ v, int64_t n,
                                                  unsigned long long* __restrict__ total) {
  __shared__ unsigned long long sm[4];
  unsigned long long s = 0;
  for (int64_t x = threadIdx.x; x < n; x += blockDim.x) s += v[x];
  s = block_sum(s, sm);
  if (threadIdx.x == 0) *total = s / 2;
}

// ---- selector features ----
// in(x) = neighbours of x ranked below x (thread per vertex over the union).
__global__ __launch_bounds__(256) void k_inpref_union(DevGraph u, const uint32_t* __restrict__ rank_of,
                                                      uint32_t* __restrict__ in) {
  const int64_t n = u.nv + u.nu;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    const bool xv = x < u.nv;
    const uint32_t* off = xv ? u.offV : u.offU;
    const uint32_t* edg = xv ? u.edgV : u.edgU;
    const int64_t lx = xv ? x : x - u.nv, shift = xv ? u.nv : 0;
    const uint32_t r = rank_of[x];
    uint32_t c = 0;
    for (uint32_t p = off[lx]; p < off[lx + 1]; p++) c += rank_of[edg[p] + shift] < r ? 1u : 0u;
    in[x] = c;
  }
}

// One workgroup per graph: rankWedgeF summed over the graph's vertices (W under a ranking).
__global__ __launch_bounds__(256) void k_graph_wedges(const BGraph* __restrict__ gs, int64_t count, int64_t NV,
                                                      const uint32_t* __restrict__ in, unsigned long long* __restrict__ W) {
  __shared__ unsigned long long sm[4];
  for (int64_t g = blockIdx.x; g < count; g += gridDim.x) {
    const BGraph b = gs[g];
    unsigned long long s = 0;
    for (uint32_t v = threadIdx.x; v < b.nv + b.nu; v += blockDim.x) {
      const bool isv = v < b.nv;
      const uint32_t lv = isv ? v : v - b.nv;
      const unsigned long long d = isv ? b.offV[lv + 1] - b.offV[lv] : b.offU[lv + 1] - b.offU[lv];
      const unsigned long long i = in[isv ? b.vbase + lv : NV + b.ubase + lv];
      s += i * (d - i) + (i ? i * (i - 1) / 2 : 0);
    }
    s = block_sum(s, sm);
    if (threadIdx.x == 0) W[g] = s;
  }
}

// One workgroup per graph: the feature row (see pbf_batch_features).
__global__ __launch_bounds__(256) void k_features(const BGraph* __restrict__ gs, int64_t count,
                                                  const unsigned long long* __restrict__ wvu, const uint8_t* __restrict__ usev,
                                                  const unsigned long long* __restrict__ Wad, const unsigned long long* __restrict__ Waco,
                                                  float* __restrict__ F) {
  __shared__ unsigned long long sm[4];
  __shared__ uint32_t mx[2];
  for (int64_t g = blockIdx.x; g < count; g += gridDim.x) {
    const BGraph b = gs[g];
    if (threadIdx.x < 2) mx[threadIdx.x] = 0;
    __syncthreads();
    unsigned long long s1 = 0, s2 = 0, t1 = 0, t2 = 0;
    for (uint32_t v = threadIdx.x; v < b.nv; v += blockDim.x) {
      const unsigned long long d = b.offV[v + 1] - b.offV[v];
      s1 += d; s2 += d * d;
      atomicMax(&mx[0], (uint32_t)d);
    }
    for (uint32_t u = threadIdx.x; u < b.nu; u += blockDim.x) {
      const unsigned long long d = b.offU[u + 1] - b.offU[u];
      t1 += d; t2 += d * d;
      atomicMax(&mx[1], (uint32_t)d);
    }
    s1 = block_sum(s1, sm); s2 = block_sum(s2, sm); t1 = block_sum(t1, sm); t2 = block_sum(t2, sm);
    if (threadIdx.x == 0) {
      float* f = F + g * PBF_MLP_IN;
      const double nv = b.nv ? b.nv : 1, nu = b.nu ? b.nu : 1;
      const unsigned long long wv = wvu[2 * g], wu = wvu[2 * g + 1];
      f[0] = (float)log2(b.nv + 1.0);
      f[1] = (float)log2(b.nu + 1.0);
      f[2] = (float)log2(b.m + 1.0);
      f[3] = (float)log2(s1 / nv + 1.0);
      f[4] = (float)log2(t1 / nu + 1.0);
      f[5] = (float)log2(s2 / nv + 1.0);
      f[6] = (float)log2(t2 / nu + 1.0);
      f[7] = (float)log2(mx[0] + 1.0);
      f[8] = (float)log2(mx[1] + 1.0);
      f[9] = (float)log2((double)wv + 1.0);
      f[10] = (float)log2((double)wu + 1.0);
      f[11] = (float)log2((double)(wv < wu ? wv : wu) + 1.0);  // SIDE: wedges centred on O
      f[12] = (float)log2((double)Wad[g] + 1.0);
      f[13] = (float)log2((double)Waco[g] + 1.0);
      f[14] = usev[g] ? 1.f : 0.f;
      f[15] = (float)log2((double)b.m / (nv * nu) + 1e-9);
    }
    __syncthreads();
  }
}

struct BatchState {
  uint64_t sig = 0;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
};

uint64_t mix(uint64_t h, uint64_t x) {
  h ^= x + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h;
}

float elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) { (void)hipGetLastError(); return 0; }
  return ms;
}

// Everything after the table upload: device work only, no synchronisation (capturable).
int union_pass(pbf_context* ctx, const BGraph* tab, int64_t count, int64_t NV, int64_t NU, int64_t M, int64_t nmax,
               int rank_type, uint32_t* ubuf, uint8_t* sflag, unsigned long long* sn, unsigned long long* voff, uint8_t* usev,
               unsigned long long* vout, unsigned long long* totals) {
  DevGraph u;
  u.nv = NV; u.nu = NU; u.m = M;
  uint32_t* offV = ubuf;
  uint32_t* edgV = offV + NV + 1;
  uint32_t* offU = edgV + M;
  uint32_t* edgU = offU + NU + 1;
  u.offV = offV; u.edgV = edgV; u.offU = offU; u.edgU = edgU;
  const unsigned gb = (unsigned)std::min<int64_t>(count, (int64_t)ctx->num_cus * 16);
  hipLaunchKernelGGL(k_union_build, dim3(gb), dim3(256), 0, ctx->stream, tab, count, offV, edgV, offU, edgU, NV, NU, M);
  hipLaunchKernelGGL(k_batch_side, dim3(gb), dim3(256), 0, ctx->stream, tab, count, NV, sflag, sn, usev,
                     (unsigned long long*)nullptr);
  PBF_HIP(hipGetLastError());
  RankedGraph R{};
  PBF_CHECK(pbf_dev_rank(ctx, u, rank_type, &R, 0, sflag, false));  // keys only
  void* lp;
  PBF_CHECK(pbf_scratch(ctx, S_RLO, 4 * (size_t)(NV + NU), &lp));
  R.lo = static_cast<uint32_t*>(lp);
  hipLaunchKernelGGL(k_local_rank, dim3(gb), dim3(256), 0, ctx->stream, tab, count, NV,
                     static_cast<const uint8_t*>(ctx->buf[S_KEY]), R.order, R.rank_of, R.lo);
  PBF_CHECK(pbf_dev_relabel(ctx, u, 0, false, &R, sflag, false));
  CountOpts opts;
  opts.sync = false;
  opts.rank_space = true;
  opts.key_range = nmax;
  unsigned long long* B = nullptr;
  opts.B_out = &B;
  PBF_CHECK(pbf_dev_count(ctx, u, 0, R, MODE_VERT, nullptr, nullptr, opts));
  PBF_CHECK(pbf_scan_excl_u64(ctx, reinterpret_cast<const uint64_t*>(sn), reinterpret_cast<uint64_t*>(voff), count,
                              reinterpret_cast<uint64_t*>(voff + count)));
  hipLaunchKernelGGL(k_batch_out, dim3(gb), dim3(256), 0, ctx->stream, tab, count, NV, usev, voff, B, R.rank_of, vout, totals);
  PBF_HIP(hipGetLastError());
  return PBF_OK;
}
}  // namespace

extern "C" int pbf_count_batch(pbf_context* ctx, int64_t count, const pbf_bipartite_csr* graphs, int mem, int rank_type,
                               int flags, uint64_t* vert_out, int64_t vert_cap, uint64_t* vert_off, uint64_t* totals,
                               int out_mem) {
  if (!ctx || count < 0 || (count > 0 && (!graphs || !vert_out || !vert_off || !totals))) return PBF_ERR_ARG;
  if (rank_type != PBF_RANK_ADEG && rank_type != PBF_RANK_ACOCORE && rank_type != PBF_RANK_SIDE) return PBF_ERR_UNSUPPORTED;
  if (mem != PBF_MEM_DEVICE && mem != PBF_MEM_HOST) return PBF_ERR_ARG;
  if (out_mem != PBF_MEM_DEVICE && out_mem != PBF_MEM_HOST) return PBF_ERR_ARG;
  const int kind = flags & 15;
  if (kind != PBF_BATCH_UNION && kind != PBF_BATCH_PLAIN) return PBF_ERR_ARG;
  const bool capture = (flags & PBF_BATCH_CAPTURE) != 0;
  if (capture && (kind != PBF_BATCH_UNION || mem != PBF_MEM_DEVICE || out_mem != PBF_MEM_DEVICE)) return PBF_ERR_ARG;
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  if (prev != ctx->device) (void)hipSetDevice(ctx->device);
  struct Restore {
    int p, d;
    ~Restore() { if (p >= 0 && p != d) (void)hipSetDevice(p); }
  } restore{prev, ctx->device};

  // sizes and union layout (host side)
  std::vector<BGraph> tab((size_t)count);
  int64_t NV = 0, NU = 0, M = 0, cap_need = 0, nmax = 0;
  for (int64_t g = 0; g < count; g++) {
    const pbf_bipartite_csr& d = graphs[g];
    if (d.nv < 0 || d.nu < 0 || d.m < 0 || !d.offsetsV || !d.offsetsU || (d.m && (!d.edgesV || !d.edgesU)))
      return PBF_ERR_ARG;
    if (mem == PBF_MEM_HOST) PBF_CHECK(pbf_validate(&d));
    tab[g] = BGraph{d.offsetsV, d.edgesV, d.offsetsU, d.edgesU, (uint32_t)d.nv, (uint32_t)d.nu, (uint32_t)d.m, 0,
                    (uint64_t)NV, (uint64_t)NU, (uint64_t)M};
    NV += d.nv; NU += d.nu; M += d.m;
    cap_need += std::max(d.nv, d.nu);
    nmax = std::max(nmax, d.nv + d.nu);
  }
  if (NV + NU >= (int64_t(1) << 31) || 2 * M >= (int64_t(1) << 31)) return PBF_ERR_CAPACITY;
  if (vert_cap < cap_need) return PBF_ERR_ARG;
  memset(&ctx->timing, 0, sizeof(ctx->timing));
  hipEvent_t* ev = ctx->ev;
  if (count == 0) return PBF_OK;

  // output staging for host outputs
  unsigned long long *vout = reinterpret_cast<unsigned long long*>(vert_out), *voff = nullptr,
                     *tot = reinterpret_cast<unsigned long long*>(totals);
  void* op = nullptr;
  if (out_mem == PBF_MEM_HOST) {
    PBF_CHECK(pbf_scratch(ctx, S_OUT, 8 * (size_t)(cap_need + 2 * count + 1), &op));
    vout = static_cast<unsigned long long*>(op);
    tot = vout + cap_need;
  }

  if (kind == PBF_BATCH_PLAIN) {
    std::vector<uint64_t> off((size_t)count + 1, 0);
    PBF_HIP(hipEventRecord(ev[0], ctx->stream));
    for (int64_t g = 0; g < count; g++) {
      const pbf_bipartite_csr& d = graphs[g];
      DevGraph dg;
      dg.nv = d.nv; dg.nu = d.nu; dg.m = d.m;
      if (mem == PBF_MEM_DEVICE) {
        dg.offV = d.offsetsV; dg.edgV = d.edgesV; dg.offU = d.offsetsU; dg.edgU = d.edgesU;
      } else {
        const size_t words = size_t(d.nv + 1) + size_t(d.m) + size_t(d.nu + 1) + size_t(d.m);
        void* p;
        PBF_CHECK(pbf_scratch(ctx, S_GRAPH, 4 * words, &p));
        uint32_t* w = static_cast<uint32_t*>(p);
        PBF_CHECK(pbf_h2d(ctx, w, d.offsetsV, 4 * size_t(d.nv + 1)));
        PBF_CHECK(pbf_h2d(ctx, w + d.nv + 1, d.edgesV, 4 * size_t(d.m)));
        PBF_CHECK(pbf_h2d(ctx, w + d.nv + 1 + d.m, d.offsetsU, 4 * size_t(d.nu + 1)));
        PBF_CHECK(pbf_h2d(ctx, w + d.nv + 1 + d.m + d.nu + 1, d.edgesU, 4 * size_t(d.m)));
        dg.offV = w; dg.edgV = w + d.nv + 1; dg.offU = w + d.nv + 1 + d.m; dg.edgU = w + d.nv + 1 + d.m + d.nu + 1;
      }
      int use_v;
      int64_t wsum;
      PBF_CHECK(pbf_dev_side(ctx, dg, &use_v, &wsum));
      const int64_t ns = use_v ? d.nu : d.nv;
      off[g + 1] = off[g] + ns;
      RankedGraph R{};
      PBF_CHECK(pbf_dev_rank(ctx, dg, rank_type, &R, use_v));
      PBF_CHECK(pbf_dev_relabel(ctx, dg, use_v, false, &R));
      PBF_CHECK(pbf_dev_count(ctx, dg, use_v, R, MODE_VERT, reinterpret_cast<uint64_t*>(vout + off[g]), nullptr));
      hipLaunchKernelGGL(k_total_of, dim3(1), dim3(256), 0, ctx->stream, vout + off[g], ns, tot + g);
      PBF_HIP(hipGetLastError());
    }
    PBF_HIP(hipEventRecord(ev[4], ctx->stream));
    if (out_mem == PBF_MEM_HOST) {
      PBF_CHECK(pbf_d2h(ctx, vert_out, vout, 8 * (size_t)off[count]));
      PBF_CHECK(pbf_d2h(ctx, totals, tot, 8 * (size_t)count));
      PBF_HIP(hipStreamSynchronize(ctx->stream));
      memcpy(vert_off, off.data(), 8 * (size_t)(count + 1));
    } else {
      PBF_CHECK(pbf_h2d(ctx, vert_off, off.data(), 8 * (size_t)(count + 1)));
      PBF_HIP(hipStreamSynchronize(ctx->stream));
    }
    ctx->timing.total_ms = elapsed_ms(ev[0], ev[4]);
    ctx->timing.count_ms = ctx->timing.total_ms;
    return PBF_OK;
  }

  // union: staging of host graphs, the table, the union CSR, flags and per-graph scalars
  if (mem == PBF_MEM_HOST) {
    size_t words = 0;
    for (int64_t g = 0; g < count; g++) words += size_t(tab[g].nv + 1) + size_t(tab[g].nu + 1) + 2 * size_t(tab[g].m);
    void* sp;
    PBF_CHECK(pbf_scratch(ctx, S_BSTAGE, 4 * words, &sp));
    uint32_t* w = static_cast<uint32_t*>(sp);
    for (int64_t g = 0; g < count; g++) {
      const pbf_bipartite_csr& d = graphs[g];
      BGraph& b = tab[g];
      PBF_CHECK(pbf_h2d(ctx, w, d.offsetsV, 4 * size_t(d.nv + 1)));
      b.offV = w; w += d.nv + 1;
      PBF_CHECK(pbf_h2d(ctx, w, d.edgesV, 4 * size_t(d.m)));
      b.edgV = w; w += d.m;
      PBF_CHECK(pbf_h2d(ctx, w, d.offsetsU, 4 * size_t(d.nu + 1)));
      b.offU = w; w += d.nu + 1;
      PBF_CHECK(pbf_h2d(ctx, w, d.edgesU, 4 * size_t(d.m)));
      b.edgU = w; w += d.m;
    }
  }
  void *tp, *up, *fp, *ap;
  PBF_CHECK(pbf_scratch(ctx, S_BTAB, sizeof(BGraph) * (size_t)count, &tp));
  PBF_CHECK(pbf_scratch(ctx, S_BUNION, 4 * (size_t)(NV + 1 + NU + 1 + 2 * M), &up));
  PBF_CHECK(pbf_scratch(ctx, S_BFLAG, (size_t)(NV + NU + 1), &fp));
  PBF_CHECK(pbf_scratch(ctx, S_BAUX, 8 * (size_t)(3 * count + 2), &ap));
  unsigned long long* sn = static_cast<unsigned long long*>(ap);
  unsigned long long* offd = sn + count;  // count + 1 entries
  uint8_t* usev = reinterpret_cast<uint8_t*>(offd + count + 1);
  if (out_mem == PBF_MEM_DEVICE) offd = reinterpret_cast<unsigned long long*>(vert_off);
  BGraph* dtab = static_cast<BGraph*>(tp);

  uint64_t sig = 0;
  if (capture) {
    sig = mix(sig, (uint64_t)count);
    for (const BGraph& b : tab) {
      sig = mix(sig, reinterpret_cast<uint64_t>(b.offV)); sig = mix(sig, reinterpret_cast<uint64_t>(b.edgV));
      sig = mix(sig, reinterpret_cast<uint64_t>(b.offU)); sig = mix(sig, reinterpret_cast<uint64_t>(b.edgU));
      sig = mix(sig, ((uint64_t)b.nv << 32) | b.nu); sig = mix(sig, b.m);
    }
    sig = mix(sig, reinterpret_cast<uint64_t>(vout)); sig = mix(sig, reinterpret_cast<uint64_t>(offd));
    sig = mix(sig, reinterpret_cast<uint64_t>(tot)); sig = mix(sig, (uint64_t)rank_type);
    sig = mix(sig, ((uint64_t)ctx->tune.light_max << 40) ^ ((uint64_t)ctx->tune.light_work << 8) ^ ctx->tune.blocks_per_cu);
    sig = mix(sig, reinterpret_cast<uint64_t>(up)); sig = mix(sig, reinterpret_cast<uint64_t>(ctx->buf[S_BINS]));
    BatchState* st = static_cast<BatchState*>(ctx->batch);
    if (st && st->exec && st->sig == sig) {
      PBF_HIP(hipEventRecord(ev[0], ctx->stream));
      PBF_HIP(hipGraphLaunch(st->exec, ctx->stream));
      PBF_HIP(hipEventRecord(ev[4], ctx->stream));
      PBF_HIP(hipStreamSynchronize(ctx->stream));
      ctx->timing.total_ms = elapsed_ms(ev[0], ev[4]);
      ctx->timing.count_ms = ctx->timing.total_ms;
      ctx->timing.wedges = -1;
      return PBF_OK;
    }
  }
  PBF_CHECK(pbf_h2d(ctx, dtab, tab.data(), sizeof(BGraph) * (size_t)count));
  PBF_HIP(hipStreamSynchronize(ctx->stream));  // the table's host copy is released below
  PBF_HIP(hipEventRecord(ev[0], ctx->stream));
  PBF_CHECK(union_pass(ctx, dtab, count, NV, NU, M, nmax, rank_type, static_cast<uint32_t*>(up), static_cast<uint8_t*>(fp), sn,
                       offd, usev, vout, tot));
  PBF_HIP(hipEventRecord(ev[4], ctx->stream));
  if (out_mem == PBF_MEM_HOST) {
    std::vector<uint64_t> off((size_t)count + 1);
    PBF_CHECK(pbf_d2h(ctx, off.data(), offd, 8 * (size_t)(count + 1)));
    PBF_HIP(hipStreamSynchronize(ctx->stream));
    PBF_CHECK(pbf_d2h(ctx, vert_out, vout, 8 * (size_t)off[count]));
    PBF_CHECK(pbf_d2h(ctx, totals, tot, 8 * (size_t)count));
    PBF_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(vert_off, off.data(), 8 * (size_t)(count + 1));
  } else {
    PBF_HIP(hipStreamSynchronize(ctx->stream));
  }
  ctx->timing.total_ms = elapsed_ms(ev[0], ev[4]);
  ctx->timing.count_ms = ctx->timing.total_ms;
  ctx->timing.wedges = -1;
  if (capture) {
    // the eager pass sized every scratch buffer: capture the same pass for replays
    BatchState* st = static_cast<BatchState*>(ctx->batch);
    if (!st) { st = new BatchState(); ctx->batch = st; }
    if (st->exec) { (void)hipGraphExecDestroy(st->exec); st->exec = nullptr; }
    if (st->graph) { (void)hipGraphDestroy(st->graph); st->graph = nullptr; }
    PBF_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    const int rc = union_pass(ctx, dtab, count, NV, NU, M, nmax, rank_type, static_cast<uint32_t*>(up), static_cast<uint8_t*>(fp),
                              sn, offd, usev, vout, tot);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != PBF_OK) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (ec != hipSuccess) return PBF_ERR_HIP;
    st->graph = graph;
    PBF_HIP(hipGraphInstantiate(&st->exec, graph, nullptr, nullptr, 0));
    st->sig = sig;
  }
  return PBF_OK;
}

extern "C" void pbf_batch_release(pbf_context* ctx) {
  if (!ctx || !ctx->batch) return;
  BatchState* st = static_cast<BatchState*>(ctx->batch);
  if (st->exec) (void)hipGraphExecDestroy(st->exec);
  if (st->graph) (void)hipGraphDestroy(st->graph);
  delete st;
  ctx->batch = nullptr;
}

// Autotuner: times the union pass under each candidate setting of the counting kernels' knobs
// and keeps the fastest in the context.
extern "C" int pbf_batch_autotune(pbf_context* ctx, int64_t count, const pbf_bipartite_csr* graphs, int mem, int rank_type,
                                  uint64_t* vert_out, int64_t vert_cap, uint64_t* vert_off, uint64_t* totals, int out_mem,
                                  int* best, float* best_ms) {
  if (!ctx) return PBF_ERR_ARG;
  static const CountTune cands[] = {
      {767, 1u << 14, 4}, {767, 1u << 14, 2}, {1535, 1u << 14, 4}, {767, 1u << 13, 4},
      {383, 1u << 12, 4}, {511, 1u << 13, 4}, {767, 1u << 13, 2}, {767, 1u << 15, 4},
  };
  const int nc = (int)(sizeof(cands) / sizeof(cands[0]));
  int bi = 0;
  float bms = 1e30f;
  const CountTune saved = ctx->tune;
  for (int c = 0; c < nc; c++) {
    ctx->tune = cands[c];
    float ms = 0;
    for (int rep = 0; rep < 2; rep++) {  // first run warms the scratch
      const int rc = pbf_count_batch(ctx, count, graphs, mem, rank_type, PBF_BATCH_UNION, vert_out, vert_cap, vert_off,
                                     totals, out_mem);
      if (rc != PBF_OK) { ctx->tune = saved; return rc; }
      ms = ctx->timing.total_ms;
    }
    if (ms < bms) { bms = ms; bi = c; }
  }
  ctx->tune = cands[bi];
  if (best) *best = bi;
  if (best_ms) *best_ms = bms;
  return PBF_OK;
}

extern "C" int pbf_variant_tuning(int variant, int* rank_type, int* light_max, int* light_work, int* blocks_per_cu) {
  static const int ranks[3] = {PBF_RANK_SIDE, PBF_RANK_ADEG, PBF_RANK_ACOCORE};
  static const int knobs[3][3] = {{767, 1 << 14, 4}, {1535, 1 << 14, 4}, {383, 1 << 12, 2}};
  if (variant < 0 || variant >= 9) return PBF_ERR_ARG;
  if (rank_type) *rank_type = ranks[variant / 3];
  if (light_max) *light_max = knobs[variant % 3][0];
  if (light_work) *light_work = knobs[variant % 3][1];
  if (blocks_per_cu) *blocks_per_cu = knobs[variant % 3][2];
  return PBF_OK;
}

extern "C" int pbf_batch_features(pbf_context* ctx, int64_t count, const pbf_bipartite_csr* graphs, int mem, float* features) {
  if (!ctx || count < 0 || (count > 0 && (!graphs || !features))) return PBF_ERR_ARG;
  if (mem != PBF_MEM_DEVICE && mem != PBF_MEM_HOST) return PBF_ERR_ARG;
  if (count == 0) return PBF_OK;
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  if (prev != ctx->device) (void)hipSetDevice(ctx->device);
  struct Restore {
    int p, d;
    ~Restore() { if (p >= 0 && p != d) (void)hipSetDevice(p); }
  } restore{prev, ctx->device};
  std::vector<BGraph> tab((size_t)count);
  int64_t NV = 0, NU = 0, M = 0;
  size_t stage_words = 0;
  for (int64_t g = 0; g < count; g++) {
    const pbf_bipartite_csr& d = graphs[g];
    if (d.nv < 0 || d.nu < 0 || d.m < 0 || !d.offsetsV || !d.offsetsU || (d.m && (!d.edgesV || !d.edgesU)))
      return PBF_ERR_ARG;
    tab[g] = BGraph{d.offsetsV, d.edgesV, d.offsetsU, d.edgesU, (uint32_t)d.nv, (uint32_t)d.nu, (uint32_t)d.m, 0,
                    (uint64_t)NV, (uint64_t)NU, (uint64_t)M};
    NV += d.nv; NU += d.nu; M += d.m;
    stage_words += size_t(d.nv + 1) + size_t(d.nu + 1) + 2 * size_t(d.m);
  }
  if (NV + NU >= (int64_t(1) << 31) || 2 * M >= (int64_t(1) << 31)) return PBF_ERR_CAPACITY;
  if (mem == PBF_MEM_HOST) {
    void* sp;
    PBF_CHECK(pbf_scratch(ctx, S_BSTAGE, 4 * stage_words, &sp));
    uint32_t* w = static_cast<uint32_t*>(sp);
    for (int64_t g = 0; g < count; g++) {
      const pbf_bipartite_csr& d = graphs[g];
      BGraph& bg = tab[g];
      PBF_CHECK(pbf_h2d(ctx, w, d.offsetsV, 4 * size_t(d.nv + 1)));
      bg.offV = w; w += d.nv + 1;
      PBF_CHECK(pbf_h2d(ctx, w, d.edgesV, 4 * size_t(d.m)));
      bg.edgV = w; w += d.m;
      PBF_CHECK(pbf_h2d(ctx, w, d.offsetsU, 4 * size_t(d.nu + 1)));
      bg.offU = w; w += d.nu + 1;
      PBF_CHECK(pbf_h2d(ctx, w, d.edgesU, 4 * size_t(d.m)));
      bg.edgU = w; w += d.m;
    }
  }
  void *tp, *up, *fp, *ap, *xp;
  PBF_CHECK(pbf_scratch(ctx, S_BTAB, sizeof(BGraph) * (size_t)count, &tp));
  PBF_CHECK(pbf_scratch(ctx, S_BUNION, 4 * (size_t)(NV + 1 + NU + 1 + 2 * M), &up));
  PBF_CHECK(pbf_scratch(ctx, S_BFLAG, (size_t)(NV + NU + 1), &fp));
  PBF_CHECK(pbf_scratch(ctx, S_BAUX, 8 * (size_t)(3 * count + 2), &ap));
  PBF_CHECK(pbf_scratch(ctx, S_BFEAT, 8 * (size_t)(4 * count) + 4 * (size_t)count * PBF_MLP_IN + 4 * (size_t)(NV + NU + 1), &xp));
  BGraph* dtab = static_cast<BGraph*>(tp);
  unsigned long long* sn = static_cast<unsigned long long*>(ap);
  uint8_t* usev = reinterpret_cast<uint8_t*>(sn + 2 * count + 1);
  unsigned long long* wvu = static_cast<unsigned long long*>(xp);
  unsigned long long* Wad = wvu + 2 * count;
  unsigned long long* Waco = Wad + count;
  float* F = reinterpret_cast<float*>(Waco + count);
  uint32_t* inb = reinterpret_cast<uint32_t*>(F + (size_t)count * PBF_MLP_IN);
  PBF_CHECK(pbf_h2d(ctx, dtab, tab.data(), sizeof(BGraph) * (size_t)count));
  DevGraph u;
  u.nv = NV; u.nu = NU; u.m = M;
  uint32_t* offV = static_cast<uint32_t*>(up);
  u.offV = offV; u.edgV = offV + NV + 1; u.offU = u.edgV + M; u.edgU = u.offU + NU + 1;
  const unsigned gb = (unsigned)std::min<int64_t>(count, (int64_t)ctx->num_cus * 16);
  hipLaunchKernelGGL(k_union_build, dim3(gb), dim3(256), 0, ctx->stream, dtab, count, offV, const_cast<uint32_t*>(u.edgV),
                     const_cast<uint32_t*>(u.offU), const_cast<uint32_t*>(u.edgU), NV, NU, M);
  hipLaunchKernelGGL(k_batch_side, dim3(gb), dim3(256), 0, ctx->stream, dtab, count, NV, static_cast<uint8_t*>(fp), sn, usev,
                     wvu);
  const int64_t n = NV + NU;
  const unsigned nb = (unsigned)std::min<int64_t>((n + 255) / 256 + 1, 16384);
  const int ranks[2] = {PBF_RANK_ADEG, PBF_RANK_ACOCORE};
  unsigned long long* Wout[2] = {Wad, Waco};
  for (int k = 0; k < 2; k++) {
    RankedGraph R{};
    PBF_CHECK(pbf_dev_rank(ctx, u, ranks[k], &R, 0, static_cast<uint8_t*>(fp)));
    hipLaunchKernelGGL(k_inpref_union, dim3(nb), dim3(256), 0, ctx->stream, u, R.rank_of, inb);
    hipLaunchKernelGGL(k_graph_wedges, dim3(gb), dim3(256), 0, ctx->stream, dtab, count, NV, inb, Wout[k]);
  }
  hipLaunchKernelGGL(k_features, dim3(gb), dim3(256), 0, ctx->stream, dtab, count, wvu, usev, Wad, Waco, F);
  PBF_HIP(hipGetLastError());
  PBF_CHECK(pbf_d2h(ctx, features, F, 4 * (size_t)count * PBF_MLP_IN));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  return PBF_OK;
}

extern "C" int pbf_set_tuning(pbf_context* ctx, int light_max, int light_work, int blocks_per_cu) {
  if (!ctx || light_max < 0 || light_max > 1535 || light_work < 0 || blocks_per_cu < 1 || blocks_per_cu > 8) return PBF_ERR_ARG;
  ctx->tune.light_max = (uint32_t)light_max;
  ctx->tune.light_work = (uint32_t)light_work;
  ctx->tune.blocks_per_cu = (uint32_t)blocks_per_cu;
  return PBF_OK;
}

// ---------------- learned selection (no counterpart in the reference) ----------------
namespace {
constexpr int FIXED_VARIANT = 3;  // ADEG with the default knobs: the method without the selector

// Where one graph's outputs go: its part's packed segment (src) to the caller's layout (dst).
struct Place {
  uint64_t src, dst;      // first count in the parts' packed array / in the caller's array
  uint32_t len, src_tot;  // counts of the graph (n_S) / its total in the parts' totals
  uint32_t dst_tot, pad;
};

// One workgroup per graph: copies its counts and total into the caller's layout.
__global__ __launch_bounds__(256) void k_place(const Place* __restrict__ pl, int64_t count,
                                               const unsigned long long* __restrict__ pv,
                                               const unsigned long long* __restrict__ pt,
                                               unsigned long long* __restrict__ vout, unsigned long long* __restrict__ tot) {
  for (int64_t g = blockIdx.x; g < count; g += gridDim.x) {
    const Place p = pl[g];
    for (uint32_t x = threadIdx.x; x < p.len; x += blockDim.x) vout[p.dst + x] = pv[p.src + x];
    if (threadIdx.x == 0) tot[p.dst_tot] = pt[p.src_tot];
  }
}
}  // namespace

extern "C" int pbf_select(pbf_context* ctx, int64_t count, const pbf_bipartite_csr* graphs, int mem,
                          const pbf_selector_model* model, int* variants) {
  if (!ctx || count < 0 || (count > 0 && (!graphs || !variants))) return PBF_ERR_ARG;
  if (count == 0) return PBF_OK;
  if (!model) {
    for (int64_t g = 0; g < count; g++) variants[g] = FIXED_VARIANT;
    return PBF_OK;
  }
  std::vector<float> F((size_t)count * PBF_MLP_IN), Y((size_t)count * PBF_MLP_OUT);
  PBF_CHECK(pbf_batch_features(ctx, count, graphs, mem, F.data()));
  for (int64_t g = 0; g < count; g++)
    for (int f = 0; f < PBF_MLP_IN; f++) {
      float& x = F[(size_t)g * PBF_MLP_IN + f];
      x = (x - model->fmu[f]) / model->fsd[f];
    }
  PBF_CHECK(pbf_mlp_predict(ctx, model->params, F.data(), count, Y.data()));
  for (int64_t g = 0; g < count; g++) {
    // predicted log2 runtimes (plain float multiply and add, no fused multiply-add, so that
    // a caller's float32 y * tsd + tmu gives the same values); a prediction that is not finite
    // is rejected and the graph keeps the fixed variant
#pragma clang fp contract(off)
    int best = FIXED_VARIANT;
    float bt = 0;
    bool ok = true;
    for (int v = 0; v < PBF_MLP_OUT; v++) {
      const float t = Y[(size_t)g * PBF_MLP_OUT + v] * model->tsd[v] + model->tmu[v];
      if (!std::isfinite(t)) { ok = false; break; }
      if (v == 0 || t < bt) { bt = t; best = v; }
    }
    variants[g] = ok ? best : FIXED_VARIANT;
  }
  return PBF_OK;
}

extern "C" int pbf_count_batch_learned(pbf_context* ctx, int64_t count, const pbf_bipartite_csr* graphs, int mem,
                                       const pbf_selector_model* model, uint64_t* vert_out, int64_t vert_cap,
                                       uint64_t* vert_off, uint64_t* totals, int out_mem, int* variants) {
  if (!ctx || count < 0 || (count > 0 && (!graphs || !vert_out || !vert_off || !totals))) return PBF_ERR_ARG;
  if (mem != PBF_MEM_DEVICE && mem != PBF_MEM_HOST) return PBF_ERR_ARG;
  if (out_mem != PBF_MEM_DEVICE && out_mem != PBF_MEM_HOST) return PBF_ERR_ARG;
  int64_t cap_need = 0;
  for (int64_t g = 0; g < count; g++) {
    const pbf_bipartite_csr& d = graphs[g];
    if (d.nv < 0 || d.nu < 0 || d.m < 0) return PBF_ERR_ARG;
    cap_need += std::max(d.nv, d.nu);
  }
  if (vert_cap < cap_need) return PBF_ERR_ARG;
  if (count == 0) return PBF_OK;
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  if (prev != ctx->device) (void)hipSetDevice(ctx->device);
  struct Restore {
    int p, d;
    ~Restore() { if (p >= 0 && p != d) (void)hipSetDevice(p); }
  } restore{prev, ctx->device};
  // events 6 and 7 bracket the whole call (the union passes inside use events 0 and 4)
  hipEvent_t t0 = ctx->ev[6], t1 = ctx->ev[7];
  PBF_HIP(hipEventRecord(t0, ctx->stream));
  std::vector<int> var((size_t)count);
  PBF_CHECK(pbf_select(ctx, count, graphs, mem, model, var.data()));
  if (variants) std::copy(var.begin(), var.end(), variants);
  // parts: the graphs of each chosen ranking, each run as one union pass with the part's most
  // chosen knob setting
  std::vector<std::vector<int64_t>> parts(3);
  for (int64_t g = 0; g < count; g++) parts[(size_t)var[(size_t)g] / 3].push_back(g);
  int nparts = 0;
  for (auto& p : parts) nparts += p.empty() ? 0 : 1;
  const CountTune saved = ctx->tune;
  auto use_knobs = [&](const std::vector<int64_t>& idx) {
    int votes[3] = {0, 0, 0};
    for (int64_t g : idx) votes[var[(size_t)g] % 3]++;
    const int k = (int)(std::max_element(votes, votes + 3) - votes);
    int rt, lm, lw, bpc;
    (void)pbf_variant_tuning(k, &rt, &lm, &lw, &bpc);
    ctx->tune.light_max = (uint32_t)lm;
    ctx->tune.light_work = (uint32_t)lw;
    ctx->tune.blocks_per_cu = (uint32_t)bpc;
  };
  struct Untune {
    pbf_context* c;
    CountTune t;
    ~Untune() { c->tune = t; }
  } untune{ctx, saved};
  static const int rank_of_part[3] = {PBF_RANK_SIDE, PBF_RANK_ADEG, PBF_RANK_ACOCORE};
  if (nparts == 1) {
    for (int r = 0; r < 3; r++)
      if (!parts[r].empty()) {
        use_knobs(parts[r]);
        PBF_CHECK(pbf_count_batch(ctx, count, graphs, mem, rank_of_part[r], PBF_BATCH_UNION, vert_out, vert_cap, vert_off,
                                  totals, out_mem));
      }
    PBF_HIP(hipEventRecord(t1, ctx->stream));
    PBF_HIP(hipEventSynchronize(t1));
    ctx->timing.total_ms = elapsed_ms(t0, t1);
    ctx->timing.count_ms = ctx->timing.total_ms;
    return PBF_OK;
  }
  // several parts: each part's outputs packed into S_LEARN0, then placed in the caller's layout
  void *lp, *pp;
  const size_t part_words = (size_t)cap_need + 2 * (size_t)count + 3;
  PBF_CHECK(pbf_scratch(ctx, S_LEARN0, 8 * part_words, &lp));
  unsigned long long* pv = static_cast<unsigned long long*>(lp);
  unsigned long long* pt = pv + cap_need;
  unsigned long long* po = pt + count;  // per part: its count + 1 offsets, one part after another
  std::vector<Place> place((size_t)count);
  std::vector<uint64_t> ns((size_t)count);
  std::vector<pbf_bipartite_csr> sub;
  int64_t vbase = 0, tbase = 0;
  for (int r = 0; r < 3; r++) {
    const auto& idx = parts[r];
    if (idx.empty()) continue;
    sub.clear();
    int64_t cap = 0;
    for (int64_t g : idx) { sub.push_back(graphs[g]); cap += std::max(graphs[g].nv, graphs[g].nu); }
    const int64_t c = (int64_t)idx.size();
    use_knobs(idx);
    PBF_CHECK(pbf_count_batch(ctx, c, sub.data(), mem, rank_of_part[r], PBF_BATCH_UNION,
                              reinterpret_cast<uint64_t*>(pv + vbase), cap, reinterpret_cast<uint64_t*>(po + tbase + r),
                              reinterpret_cast<uint64_t*>(pt + tbase), PBF_MEM_DEVICE));
    std::vector<uint64_t> off((size_t)c + 1);
    PBF_CHECK(pbf_d2h(ctx, off.data(), po + tbase + r, 8 * (size_t)(c + 1)));
    for (int64_t j = 0; j < c; j++) {
      const int64_t g = idx[(size_t)j];
      ns[(size_t)g] = off[(size_t)j + 1] - off[(size_t)j];
      place[(size_t)g].src = (uint64_t)vbase + off[(size_t)j];
      place[(size_t)g].len = (uint32_t)ns[(size_t)g];
      place[(size_t)g].src_tot = (uint32_t)(tbase + j);
      place[(size_t)g].dst_tot = (uint32_t)g;
    }
    vbase += cap;
    tbase += c;
  }
  std::vector<uint64_t> goff((size_t)count + 1, 0);
  for (int64_t g = 0; g < count; g++) {
    place[(size_t)g].dst = goff[(size_t)g];
    goff[(size_t)g + 1] = goff[(size_t)g] + ns[(size_t)g];
  }
  PBF_CHECK(pbf_scratch(ctx, S_LEARN1, sizeof(Place) * (size_t)count, &pp));
  PBF_CHECK(pbf_h2d(ctx, pp, place.data(), sizeof(Place) * (size_t)count));
  unsigned long long *vdst = reinterpret_cast<unsigned long long*>(vert_out), *tdst = reinterpret_cast<unsigned long long*>(totals);
  if (out_mem == PBF_MEM_HOST) {
    void* fp;
    PBF_CHECK(pbf_scratch(ctx, S_LEARN2, 8 * ((size_t)goff[(size_t)count] + (size_t)count), &fp));
    vdst = static_cast<unsigned long long*>(fp);
    tdst = vdst + goff[(size_t)count];
  }
  const unsigned gb = (unsigned)std::min<int64_t>(count, (int64_t)ctx->num_cus * 16);
  hipLaunchKernelGGL(k_place, dim3(gb), dim3(256), 0, ctx->stream, static_cast<const Place*>(pp), count, pv, pt, vdst, tdst);
  PBF_HIP(hipGetLastError());
  PBF_HIP(hipEventRecord(t1, ctx->stream));
  if (out_mem == PBF_MEM_HOST) {
    PBF_CHECK(pbf_d2h(ctx, vert_out, vdst, 8 * (size_t)goff[(size_t)count]));
    PBF_CHECK(pbf_d2h(ctx, totals, tdst, 8 * (size_t)count));
    memcpy(vert_off, goff.data(), 8 * (size_t)(count + 1));
  } else {
    PBF_CHECK(pbf_h2d(ctx, vert_off, goff.data(), 8 * (size_t)(count + 1)));
  }
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  ctx->timing.total_ms = elapsed_ms(t0, t1);
  ctx->timing.count_ms = ctx->timing.total_ms;
  ctx->timing.wedges = -1;
  return PBF_OK;
}
