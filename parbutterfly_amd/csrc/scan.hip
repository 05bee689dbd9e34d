// Device-wide exclusive prefix sums (u32 / u64): per-block totals, a single-block scan of the
// totals, then a per-block scan with the block's carry-in.  Used for the ranked offsets, the
// key histograms of the ranking sort and the bin offsets.
#include "pbf_device.h"

namespace {
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// Block-wide exclusive scan of one value per thread; returns the block total in *total.
template <typename T>
__device__ T block_excl_scan(T x, T* sm /* >= 16 entries */, T* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  T incl = pbf_wave_incl_scan(x);
  if (lane == 63) sm[wid] = incl;
  __syncthreads();
  if (wid == 0) {
    T w = lane < (int)(blockDim.x >> 6) ? sm[lane] : T(0);
    T wi = pbf_wave_incl_scan(w);
    if (lane < (int)(blockDim.x >> 6)) sm[lane] = wi - w;
    if (lane == (int)(blockDim.x >> 6) - 1) sm[16] = wi;
  }
  __syncthreads();
  T res = incl - x + sm[wid];
  *total = sm[16];
  __syncthreads();
  return res;
}

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_tile_sums(const T* __restrict__ in, T* __restrict__ sums, int64_t n) {
  __shared__ T sm[17];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  T s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    int64_t idx = base + (int64_t)k * SCAN_THREADS + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  T tot;
  block_excl_scan(s, sm, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Single block: exclusive scan of nb tile sums in place, total to *total_out (may be null).
template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_sums(T* sums, int64_t nb, T* total_out) {
  __shared__ T sm[17];
  T carry = 0;
  for (int64_t base = 0; base < nb; base += SCAN_THREADS) {
    int64_t idx = base + threadIdx.x;
    T x = idx < nb ? sums[idx] : T(0);
    T tot;
    T ex = block_excl_scan(x, sm, &tot);
    if (idx < nb) sums[idx] = ex + carry;
    carry += tot;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const T* in, T* out, const T* __restrict__ sums, int64_t n) {
  __shared__ T sm[17];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  T v[SCAN_ITEMS];
  T s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    int64_t idx = base + k;
    v[k] = idx < n ? in[idx] : T(0);
    s += v[k];
  }
  T tot;
  T ex = block_excl_scan(s, sm, &tot) + sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    int64_t idx = base + k;
    if (idx < n) out[idx] = ex;
    ex += v[k];
  }
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, int64_t n) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) p[x] = v;
}

template <typename T>
int fill_impl(pbf_context* ctx, T* p, T v, int64_t n) {
  if (n <= 0) return PBF_OK;
  int64_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)b), dim3(256), 0, ctx->stream, p, v, n);
  PBF_HIP(hipGetLastError());
  return PBF_OK;
}

template <typename T>
int scan_impl(pbf_context* ctx, const T* in, T* out, int64_t n, T* total_dev) {
  if (n <= 0) {
    if (total_dev) PBF_CHECK(fill_impl<T>(ctx, total_dev, T(0), 1));
    return PBF_OK;
  }
  int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  void* sp;
  PBF_CHECK(pbf_scratch(ctx, S_SCAN, sizeof(T) * (size_t)nb, &sp));
  T* sums = static_cast<T*>(sp);
  // The tile sums are taken per tile of SCAN_TILE contiguous items; k_scan_tiles reads its
  // tile as SCAN_ITEMS consecutive items per thread, which covers the same tile.
  hipLaunchKernelGGL(k_tile_sums<T>, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, ctx->stream, in, sums, n);
  hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(SCAN_THREADS), 0, ctx->stream, sums, nb, total_dev);
  hipLaunchKernelGGL(k_scan_tiles<T>, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, ctx->stream, in, out, sums, n);
  PBF_HIP(hipGetLastError());
  return PBF_OK;
}
}  // namespace

int pbf_scan_excl_u32(pbf_context* ctx, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev) {
  return scan_impl<uint32_t>(ctx, in, out, n, total_dev);
}
int pbf_scan_excl_u64(pbf_context* ctx, const uint64_t* in, uint64_t* out, int64_t n, uint64_t* total_dev) {
  return scan_impl<uint64_t>(ctx, in, out, n, total_dev);
}
int pbf_fill_u32(pbf_context* ctx, uint32_t* p, uint32_t v, int64_t n) { return fill_impl<uint32_t>(ctx, p, v, n); }
int pbf_fill_u64(pbf_context* ctx, uint64_t* p, uint64_t v, int64_t n) { return fill_impl<uint64_t>(ctx, p, v, n); }
