// Owner lookup for CSR entries: which vertex's list holds entry p.  A launch that walks the
// entries of one side in tiles of TILE consecutive entries first finds the owner of each tile's
// first entry (k_tile_owner, one binary search per tile); a block then stages the offsets of the
// tile's owners in LDS and searches there.  Searches stay in bounds for any offsets array (they
// only narrow an index range), so they are safe to run before the offsets are validated.
#pragma once
#include "pbf_device.h"

// Last x in [lo, hi] with off[x] <= p (lo when there is none).
__device__ __forceinline__ uint32_t pbf_owner_in(const uint32_t* off, uint32_t lo, uint32_t hi, uint32_t p) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo + 1) >> 1);
    if (off[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// xb[t] = owner of entry t * TILE for every tile t < ntiles; xb[ntiles] = owner of entry m - 1.
// n >= 1 and m >= 1.
template <int TILE>
__global__ void k_tile_owner(const uint32_t* __restrict__ off, int64_t n, int64_t m, int64_t ntiles,
                             uint32_t* __restrict__ xb) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t <= ntiles; t += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t p = (uint32_t)(t < ntiles ? t * TILE : m - 1);
    xb[t] = pbf_owner_in(off, 0, (uint32_t)(n - 1), p);
  }
}

// Per-block owner search over one tile: the offsets of owners x0..x1 are staged in s_off when
// they fit (span <= cap), else the search runs on the offsets in HBM.  Call stage() once per
// block (it synchronises the block), then owner(p) for entries p of the tile.
struct TileOwners {
  uint32_t x0, x1, span;
  bool staged;
  __device__ void stage(const uint32_t* off, const uint32_t* xb, int64_t t, uint32_t* s_off, uint32_t cap) {
    x0 = xb[t];
    x1 = xb[t + 1];
    if (x1 < x0) x1 = x0;  // only with offsets that are not monotone
    span = x1 - x0 + 1;
    staged = span <= cap;
    if (staged)
      for (uint32_t k = threadIdx.x; k < span; k += blockDim.x) s_off[k] = off[x0 + k];
    __syncthreads();
  }
  __device__ uint32_t owner(const uint32_t* off, const uint32_t* s_off, uint32_t p) const {
    return staged ? x0 + pbf_owner_in(s_off, 0, span - 1, p) : pbf_owner_in(off, x0, x1, p);
  }
};
