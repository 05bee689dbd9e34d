// Sparsification for approximate total counts: eSparseBipartite / clrSparseBipartite followed
// by delZeroDeg (apps/hyper/butterfly_utils.h:744-996), then CountTotal on the sparsified graph
// with the driver's scaling (apps/hyper/BipartiteButterflies.C:213-214).
//
// Each side of the graph is handled as a stream of list entries (slots):
//   k_tile_owner  owner vertex of the first slot of every 2048-slot tile (binary search);
//   k_sp_mark     keep flag of every slot (the tile's offsets staged in LDS for the owner
//                 search; the colour is recomputed from the ids, no gathers);
//   scan          exclusive scan of the flags in place: pos[p] = new position of slot p,
//                 pos[m] = edges kept;
//   k_sp_vert     kept-vertex flags (a vertex stays when pos moves across its list), scanned
//                 in place into the new ids;
//   k_sp_off      new offsets of the kept vertices;
//   k_sp_edges    kept entries moved to pos[p], the neighbour renamed by the other side's ids.
// Every pass streams its arrays once (the edge pass also gathers the other side's new ids).
#include <cmath>

#include "owner.h"

namespace {

constexpr int SP_THREADS = 256, SP_PER = 8, SP_TILE = SP_THREADS * SP_PER;
constexpr int SP_LDS = SP_TILE + 1;

// hashInt(ulong) (ligra/utils.h:389-397).  The edges a seed keeps are defined by this mix, so
// its shifts and constants are the reference's.
__device__ __forceinline__ uint64_t ref_hash64(uint64_t a) {
  a = (a + 0x7ed55d166bef7a1dull) + (a << 12);
  a = (a ^ 0xc761c23c510fa2ddull) ^ (a >> 9);
  a = (a + 0x165667b183a9c0e1ull) + (a << 59);
  a = (a + 0xd3a2646cab3487e3ull) ^ (a << 49);
  a = (a + 0xfd7046c5ef9ab54cull) + (a << 3);
  a = (a ^ 0xb55a4f090dd4a67bull) ^ (a >> 32);
  return a;
}

struct Rule {
  int type;       // PBF_SPARSE_EDGE or PBF_SPARSE_COLOR
  uint32_t ncol;  // max(1, denom) as the reference's uintE
  uint64_t seed;
  uint64_t salt;  // (seed << 6) + (seed >> 2)
  uint64_t inv;   // floor((2^64 - 1) / ncol)
};

// a % r.ncol without a 64-bit division: q = mulhi(a, inv) satisfies a/d - 3 < q <= a/d, so
// a - q d lies in [0, 3d) and two conditional subtractions finish it.
__device__ __forceinline__ uint64_t mod_ncol(const Rule& r, uint64_t a) {
  const uint64_t q = __umul64hi(a, r.inv);
  uint64_t m = a - q * r.ncol;
  if (m >= r.ncol) m -= r.ncol;
  if (m >= r.ncol) m -= r.ncol;
  return m;
}

// Keep edge (v in V, u in U)?
//  EDGE  (butterfly_utils.h:802-810): concat = (v + 0x9e3779b9 + salt) ^ (u + 0x9e3779b9 + salt),
//        where u + 0x9e3779b9 wraps in 32 bits (unsigned int arithmetic) and the rest is 64-bit;
//        the edge stays when hashInt(seed + concat) % ncol == 0.  The U side's copy of the
//        colour (:813-825, a binary search for the same edge) is the same function of (v, u).
//  COLOR (:911-914, :921-937): colour(x) = hashInt(seed + x) % ncol on each side (ids within
//        the side); the edge stays when both ends have the same colour.
__device__ __forceinline__ bool keep_edge(const Rule& r, uint32_t v, uint32_t u) {
  if (r.type == PBF_SPARSE_EDGE) {
    const uint64_t a = (uint64_t)v + 0x9e3779b9ull + r.salt;
    const uint64_t b = (uint64_t)(uint32_t)(u + 0x9e3779b9u) + r.salt;
    return mod_ncol(r, ref_hash64(r.seed + (a ^ b))) == 0;
  }
  return mod_ncol(r, ref_hash64(r.seed + v)) == mod_ncol(r, ref_hash64(r.seed + u));
}

// flag[p] = 1 when slot p's edge is kept (p < m); flag[m] = 0.  One tile of SP_TILE slots per
// block; the owners of the tile lie in [xb[t], xb[t+1]].
template <bool VSIDE>
__global__ __launch_bounds__(SP_THREADS) void k_sp_mark(const uint32_t* __restrict__ off, const uint32_t* __restrict__ edg,
                                                        int64_t m, Rule r, const uint32_t* __restrict__ xb,
                                                        uint32_t* __restrict__ flag) {
  __shared__ uint32_t s_off[SP_LDS];
  const int64_t t = blockIdx.x;
  TileOwners own;
  own.stage(off, xb, t, s_off, SP_LDS);
  const int64_t base = t * SP_TILE;
#pragma unroll
  for (int k = 0; k < SP_PER; k++) {
    const int64_t p = base + (int64_t)k * SP_THREADS + threadIdx.x;
    if (p < m) {
      const uint32_t x = own.owner(off, s_off, (uint32_t)p);
      const uint32_t y = edg[p];
      flag[p] = (VSIDE ? keep_edge(r, x, y) : keep_edge(r, y, x)) ? 1u : 0u;
    }
  }
  if (t == 0 && threadIdx.x == 0) flag[m] = 0;
}

// keep[x] = 1 when vertex x keeps an edge (x < n); keep[n] = 0.  pos = scanned slot flags.
__global__ void k_sp_vert(const uint32_t* __restrict__ off, int64_t n, const uint32_t* __restrict__ pos,
                          uint32_t* __restrict__ keep) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x <= n; x += (int64_t)gridDim.x * blockDim.x)
    keep[x] = (x < n && pos[off[x + 1]] > pos[off[x]]) ? 1u : 0u;
}

// New offsets: off2[id[x]] = pos[off[x]] for every kept x, and off2[n'] = m' (x = n).
__global__ void k_sp_off(const uint32_t* __restrict__ off, int64_t n, const uint32_t* __restrict__ pos,
                         const uint32_t* __restrict__ id, uint32_t* __restrict__ off2) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x <= n; x += (int64_t)gridDim.x * blockDim.x)
    if (x == n || id[x + 1] > id[x]) off2[id[x]] = pos[off[x]];
}

// Kept entries to their new positions, neighbours renamed (delZeroDeg, :766-771).
__global__ void k_sp_edges(const uint32_t* __restrict__ edg, int64_t m, const uint32_t* __restrict__ pos,
                           const uint32_t* __restrict__ id_other, uint32_t* __restrict__ edg2) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m; p += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t q = pos[p];
    if (pos[p + 1] > q) edg2[q] = id_other[edg[p]];
  }
}

inline unsigned sp_grid(int64_t n, int max_blocks) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)b;
}

}  // namespace

int pbf_dev_sparsify(pbf_context* ctx, const DevGraph& g, int sparse_type, int64_t denom, int64_t seed, DevGraph* out,
                     uint32_t* offV2, uint32_t* edgV2, uint32_t* offU2, uint32_t* edgU2) {
  if (sparse_type != PBF_SPARSE_EDGE && sparse_type != PBF_SPARSE_COLOR) return PBF_ERR_ARG;
  if (denom < 1 || denom >= (int64_t(1) << 32) || seed < 0 || seed >= (int64_t(1) << 56)) return PBF_ERR_ARG;
  Rule r;
  r.type = sparse_type;
  r.ncol = (uint32_t)denom;
  r.seed = (uint64_t)seed;
  r.salt = ((uint64_t)seed << 6) + ((uint64_t)seed >> 2);
  r.inv = ~0ull / r.ncol;
  const int64_t m = g.m, ntiles = (m + SP_TILE - 1) / SP_TILE;
  const int64_t words = 2 * (m + 1) + (g.nv + 1) + (g.nu + 1) + 2 * (ntiles + 1);
  void* sp;
  PBF_CHECK(pbf_scratch(ctx, S_SPARSE0, 4 * (size_t)words, &sp));
  uint32_t* posV = static_cast<uint32_t*>(sp);
  uint32_t* posU = posV + (m + 1);
  uint32_t* idV = posU + (m + 1);
  uint32_t* idU = idV + (g.nv + 1);
  uint32_t* xbV = idU + (g.nu + 1);
  uint32_t* xbU = xbV + (ntiles + 1);
  const int64_t n_side[2] = {g.nv, g.nu};
  const uint32_t* offs[2] = {g.offV, g.offU};
  const uint32_t* edgs[2] = {g.edgV, g.edgU};
  uint32_t* pos[2] = {posV, posU};
  uint32_t* ids[2] = {idV, idU};
  uint32_t* xbs[2] = {xbV, xbU};
  for (int s = 0; s < 2; s++) {
    if (m > 0) {
      hipLaunchKernelGGL(k_tile_owner<SP_TILE>, dim3(sp_grid(ntiles + 1, 1024)), dim3(256), 0, ctx->stream, offs[s], n_side[s], m,
                         ntiles, xbs[s]);
      if (s == 0)
        hipLaunchKernelGGL(k_sp_mark<true>, dim3((unsigned)ntiles), dim3(SP_THREADS), 0, ctx->stream, offs[s], edgs[s], m, r,
                           xbs[s], pos[s]);
      else
        hipLaunchKernelGGL(k_sp_mark<false>, dim3((unsigned)ntiles), dim3(SP_THREADS), 0, ctx->stream, offs[s], edgs[s], m,
                           r, xbs[s], pos[s]);
    } else {
      PBF_CHECK(pbf_fill_u32(ctx, pos[s], 0u, 1));
    }
    PBF_CHECK(pbf_scan_excl_u32(ctx, pos[s], pos[s], m + 1, nullptr));
    hipLaunchKernelGGL(k_sp_vert, dim3(sp_grid(n_side[s] + 1, 8192)), dim3(256), 0, ctx->stream, offs[s], n_side[s], pos[s],
                       ids[s]);
    PBF_CHECK(pbf_scan_excl_u32(ctx, ids[s], ids[s], n_side[s] + 1, nullptr));
  }
  PBF_HIP(hipGetLastError());
  // sizes of the sparsified graph: m' on each side (they agree on a valid transpose), nv', nu'
  uint32_t* sz = reinterpret_cast<uint32_t*>(ctx->pinned + 48);
  PBF_HIP(hipMemcpyAsync(sz + 0, posV + m, 4, hipMemcpyDeviceToHost, ctx->stream));
  PBF_HIP(hipMemcpyAsync(sz + 1, posU + m, 4, hipMemcpyDeviceToHost, ctx->stream));
  PBF_HIP(hipMemcpyAsync(sz + 2, idV + g.nv, 4, hipMemcpyDeviceToHost, ctx->stream));
  PBF_HIP(hipMemcpyAsync(sz + 3, idU + g.nu, 4, hipMemcpyDeviceToHost, ctx->stream));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  if (sz[0] != sz[1]) return PBF_ERR_FORMAT;
  out->m = sz[0];
  out->nv = sz[2];
  out->nu = sz[3];
  uint32_t* off2[2] = {offV2, offU2};
  uint32_t* edg2[2] = {edgV2, edgU2};
  for (int s = 0; s < 2; s++) {
    hipLaunchKernelGGL(k_sp_off, dim3(sp_grid(n_side[s] + 1, 8192)), dim3(256), 0, ctx->stream, offs[s], n_side[s], pos[s],
                       ids[s], off2[s]);
    if (m > 0)
      hipLaunchKernelGGL(k_sp_edges, dim3(sp_grid(m, 16384)), dim3(256), 0, ctx->stream, edgs[s], m, pos[s], ids[1 - s],
                         edg2[s]);
  }
  PBF_HIP(hipGetLastError());
  out->offV = offV2;
  out->edgV = edgV2;
  out->offU = offU2;
  out->edgU = edgU2;
  return PBF_OK;
}

// The driver's scaling (BipartiteButterflies.C:213-214): the long count times pow(denom, k),
// a double, converted back to long.
int pbf_sparse_scale(uint64_t total, int sparse_type, int64_t denom, int64_t* approx) {
  if (sparse_type == PBF_SPARSE_NONE) {
    if (total >= (uint64_t(1) << 63)) return PBF_ERR_DOMAIN;
    *approx = (int64_t)total;
    return PBF_OK;
  }
  const double s = (double)(int64_t)total * std::pow((double)denom, sparse_type == PBF_SPARSE_COLOR ? 3.0 : 4.0);
  if (!(s < 9223372036854775808.0)) return PBF_ERR_DOMAIN;
  *approx = (int64_t)s;
  return PBF_OK;
}
