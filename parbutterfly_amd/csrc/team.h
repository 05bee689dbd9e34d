// Team-level device helpers shared by the counting and peeling kernels: a "team" is one
// workgroup (a single wavefront or several) that owns one vertex at a time.  Provides team
// scans / sums, the segment search of the flattened two-hop enumeration, and the LDS
// open-addressing hash table (u32 key -> u32 count) with bounded probing.
#pragma once
#include "pbf_device.h"

namespace pbf {
constexpr uint32_t EMPTY = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t hash_slot(uint32_t key, uint32_t shift) {
  return (key * 0x9E3779B1u) >> shift;
}

template <int TEAM>
struct TeamShared {
  uint32_t P[TEAM];    // inclusive scan of segment lengths
  uint32_t vs[TEAM];   // first redg index of segment
  uint32_t vv[TEAM];   // centre vertex v of segment
  unsigned long long acc[TEAM];
  uint32_t scan[TEAM / 64 + 2];
  unsigned long long red[TEAM / 64 + 1];
  uint32_t item, nused, ovf, pad;
};

template <int TEAM>
__device__ __forceinline__ void team_sync() {
  __syncthreads();  // a team is a whole workgroup (TEAM == blockDim.x)
}

// Inclusive team scan; *total receives the team total.
template <int TEAM>
__device__ __forceinline__ uint32_t team_incl_scan(uint32_t x, TeamShared<TEAM>& sh, uint32_t* total) {
  uint32_t incl = pbf_wave_incl_scan(x);
  if constexpr (TEAM == 64) {
    *total = __shfl(incl, 63, 64);
    return incl;
  } else {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 63) sh.scan[wid] = incl;
    __syncthreads();
    uint32_t add = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < TEAM / 64; w++) {
      uint32_t s = sh.scan[w];
      if (w < wid) add += s;
      tot += s;
    }
    __syncthreads();
    *total = tot;
    return incl + add;
  }
}

template <int TEAM>
__device__ __forceinline__ unsigned long long team_sum(unsigned long long x, TeamShared<TEAM>& sh) {
  x = pbf_wave_sum(x);
  if constexpr (TEAM == 64) {
    return x;
  } else {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh.red[wid] = x;
    __syncthreads();
    unsigned long long t = 0;
#pragma unroll
    for (int w = 0; w < TEAM / 64; w++) t += sh.red[w];
    __syncthreads();
    return t;
  }
}

// First index s in [0, TEAM) with P[s] > w (P non-decreasing, P[TEAM-1] > w).
template <int TEAM>
__device__ __forceinline__ uint32_t seg_of(const uint32_t* P, uint32_t w) {
  uint32_t lo = 0, hi = TEAM - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (P[mid] > w) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// Number of entries of N(v)[0, len) whose rank is below i (entries are (rank << 1) | flag).
__device__ __forceinline__ uint32_t rank_lower_bound(const uint32_t* __restrict__ a, uint32_t len, uint32_t i) {
  const uint32_t key = i << 1;
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Table operations -------------------------------------------------------------------------
struct LdsTable {
  uint32_t* keys;
  uint32_t* cnts;
  uint32_t* used;   // may be null
  uint32_t shift, mask, maxfill;
};

template <int TEAM>
__device__ __forceinline__ void lds_insert(const LdsTable& t, TeamShared<TEAM>& sh, uint32_t key) {
  uint32_t h = hash_slot(key, t.shift);
  // at most mask+1 probes: a full table (never reached while occupancy < maxfill + TEAM)
  // raises the overflow flag instead of spinning
  for (uint32_t probe = 0; probe <= t.mask; probe++) {
    uint32_t old = atomicCAS(&t.keys[h], EMPTY, key);
    if (old == EMPTY) {
      atomicAdd(&t.cnts[h], 1u);
      uint32_t pos = atomicAdd(&sh.nused, 1u);
      if (t.used && pos <= t.mask) t.used[pos] = h;
      if (pos >= t.maxfill) sh.ovf = 1;
      return;
    }
    if (old == key) {
      atomicAdd(&t.cnts[h], 1u);
      return;
    }
    h = (h + 1) & t.mask;
  }
  sh.ovf = 1;
}

// Count of a key inserted in pass 1 (bounded probe; 0 if absent).
__device__ __forceinline__ uint32_t lds_lookup(const LdsTable& t, uint32_t key) {
  uint32_t h = hash_slot(key, t.shift);
  for (uint32_t probe = 0; probe <= t.mask; probe++) {
    const uint32_t k = t.keys[h];
    if (k == key) return t.cnts[h];
    if (k == EMPTY) return 0;
    h = (h + 1) & t.mask;
  }
  return 0;
}

}  // namespace pbf
