// C ABI of libparbutterfly_amd: context management and the count / rank / peel entry points.
// Each entry point mirrors a reference function (cited in include/parbutterfly.h) and returns
// a pbf_status code instead of printing and exiting.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "pbf_device.h"

// Scratch of pbf_count_vert_noalloc: one array of PBF_NOALLOC_SCRATCH_BYTES per device, held
// by every device the library runs on until the process ends.  In host code its name is only a
// stand-in, never a device address: the address is looked up on each device (noalloc_address,
// called by pbf_create with the device current) and only that pointer is used.
static __device__ __align__(256) unsigned char scratch[PBF_NOALLOC_SCRATCH_BYTES];

namespace {
int alloc_dev(void** p, size_t bytes) {
  if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return PBF_ERR_OOM;
  }
  return PBF_OK;
}

// Validates the descriptor sizes (the graph contents are validated by pbf_validate on host
// graphs; device graphs are trusted to satisfy the same contract).
int check_desc(const pbf_bipartite_csr* g) {
  if (!g || g->nv < 0 || g->nu < 0 || g->m < 0) return PBF_ERR_ARG;
  if (!g->offsetsV || !g->offsetsU || (g->m > 0 && (!g->edgesV || !g->edgesU))) return PBF_ERR_ARG;
  if (g->nv + g->nu >= (int64_t(1) << 31) || 2 * g->m >= (int64_t(1) << 31)) return PBF_ERR_CAPACITY;
  return PBF_OK;
}

constexpr size_t STAGE_BYTES = size_t(8) << 20;

int stage_init(pbf_context* ctx) {
  for (int k = 0; k < 2; k++) {
    if (!ctx->stage[k]) PBF_HIP(hipHostMalloc(&ctx->stage[k], STAGE_BYTES, hipHostMallocDefault));
    if (!ctx->stage_ev[k]) PBF_HIP(hipEventCreateWithFlags(&ctx->stage_ev[k], hipEventDisableTiming));
  }
  return PBF_OK;
}

int stage_wait(pbf_context* ctx, int k) {
  if (ctx->stage_busy[k]) PBF_HIP(hipEventSynchronize(ctx->stage_ev[k]));
  ctx->stage_busy[k] = false;
  return PBF_OK;
}
}  // namespace

int pbf_h2d(pbf_context* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return PBF_OK;
  PBF_CHECK(stage_init(ctx));
  int k = 0;
  for (size_t off = 0; off < bytes; off += STAGE_BYTES, k ^= 1) {
    const size_t n = std::min(STAGE_BYTES, bytes - off);
    PBF_CHECK(stage_wait(ctx, k));  // this buffer's previous DMA is done
    memcpy(ctx->stage[k], static_cast<const char*>(src) + off, n);
    PBF_HIP(hipMemcpyAsync(static_cast<char*>(dst) + off, ctx->stage[k], n, hipMemcpyHostToDevice, ctx->stream));
    PBF_HIP(hipEventRecord(ctx->stage_ev[k], ctx->stream));
    ctx->stage_busy[k] = true;
  }
  return PBF_OK;
}

int pbf_d2h(pbf_context* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) {
    PBF_HIP(hipStreamSynchronize(ctx->stream));
    return PBF_OK;
  }
  PBF_CHECK(stage_init(ctx));
  for (int k = 0; k < 2; k++) PBF_CHECK(stage_wait(ctx, k));
  const size_t nch = (bytes + STAGE_BYTES - 1) / STAGE_BYTES;
  auto issue = [&](size_t c) -> int {
    const int k = (int)(c & 1);
    const size_t off = c * STAGE_BYTES, n = std::min(STAGE_BYTES, bytes - off);
    PBF_HIP(hipMemcpyAsync(ctx->stage[k], static_cast<const char*>(src) + off, n, hipMemcpyDeviceToHost, ctx->stream));
    PBF_HIP(hipEventRecord(ctx->stage_ev[k], ctx->stream));
    ctx->stage_busy[k] = true;
    return PBF_OK;
  };
  PBF_CHECK(issue(0));
  for (size_t c = 0; c < nch; c++) {
    if (c + 1 < nch) PBF_CHECK(issue(c + 1));  // its buffer was emptied in the previous step
    const int k = (int)(c & 1);
    PBF_CHECK(stage_wait(ctx, k));
    const size_t off = c * STAGE_BYTES, n = std::min(STAGE_BYTES, bytes - off);
    memcpy(static_cast<char*>(dst) + off, ctx->stage[k], n);
  }
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  return PBF_OK;
}

namespace {
// Resolves the graph to device pointers, uploading a host graph into the context.
int resolve_graph(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, DevGraph* dg) {
  PBF_CHECK(check_desc(g));
  dg->nv = g->nv; dg->nu = g->nu; dg->m = g->m;
  if (mem == PBF_MEM_DEVICE) {
    dg->offV = g->offsetsV; dg->edgV = g->edgesV; dg->offU = g->offsetsU; dg->edgU = g->edgesU;
    return PBF_OK;
  }
  if (mem != PBF_MEM_HOST) return PBF_ERR_ARG;
  PBF_CHECK(pbf_validate(g));
  size_t words = size_t(g->nv + 1) + size_t(g->m) + size_t(g->nu + 1) + size_t(g->m);
  void* p;
  PBF_CHECK(pbf_scratch(ctx, S_GRAPH, 4 * words, &p));
  uint32_t* d = static_cast<uint32_t*>(p);
  PBF_CHECK(pbf_h2d(ctx, d, g->offsetsV, 4 * size_t(g->nv + 1)));
  PBF_CHECK(pbf_h2d(ctx, d + g->nv + 1, g->edgesV, 4 * size_t(g->m)));
  PBF_CHECK(pbf_h2d(ctx, d + g->nv + 1 + g->m, g->offsetsU, 4 * size_t(g->nu + 1)));
  PBF_CHECK(pbf_h2d(ctx, d + g->nv + 1 + g->m + g->nu + 1, g->edgesU, 4 * size_t(g->m)));
  dg->offV = d;
  dg->edgV = d + g->nv + 1;
  dg->offU = d + g->nv + 1 + g->m;
  dg->edgU = d + g->nv + 1 + g->m + g->nu + 1;
  return PBF_OK;
}

// Output buffer: the caller's device pointer, or staging in the context for host outputs.
int out_buffer(pbf_context* ctx, int mem, void* user, size_t bytes, void** dev) {
  if (mem == PBF_MEM_DEVICE) { *dev = user; return PBF_OK; }
  return pbf_scratch(ctx, S_OUT, bytes, dev);
}

int finish_output(pbf_context* ctx, int mem, void* user, const void* dev, size_t bytes) {
  if (mem == PBF_MEM_HOST) return pbf_d2h(ctx, user, dev, bytes);
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  return PBF_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) { (void)hipGetLastError(); return 0; }
  return ms;
}

bool rank_supported(int rank_type) {
  return rank_type == PBF_RANK_ADEG || rank_type == PBF_RANK_ACOCORE || rank_type == PBF_RANK_SIDE;
}
bool count_supported(int count_type) { return count_type >= PBF_COUNT_SORT && count_type <= PBF_COUNT_GPU; }

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// Shared driver for the three count entry points.
int count_common(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v, int count_type, int rank_type,
                 int mode, void* out_user) {
  if (!ctx || !out_user) return PBF_ERR_ARG;
  if (!count_supported(count_type)) return PBF_ERR_ARG;
  if (!rank_supported(rank_type)) return PBF_ERR_UNSUPPORTED;
  DeviceGuard guard(ctx->device);
  DevGraph dg;
  PBF_CHECK(resolve_graph(ctx, g, mem, &dg));
  const int64_t ns = use_v ? g->nu : g->nv;
  size_t out_bytes = mode == MODE_VERT ? 8 * (size_t)ns : mode == MODE_EDGE ? 8 * (size_t)g->m : 8;
  void* out_dev;
  if (mode == MODE_TOTAL) PBF_CHECK(pbf_scratch(ctx, S_OUT, 8, &out_dev));
  else PBF_CHECK(out_buffer(ctx, mem, out_user, out_bytes, &out_dev));
  memset(&ctx->timing, 0, sizeof(ctx->timing));
  hipEvent_t* ev = ctx->ev;
  PBF_HIP(hipEventRecord(ev[0], ctx->stream));
  uint32_t* eti = nullptr;
  if (mode == MODE_EDGE) {
    void* ep;
    PBF_CHECK(pbf_scratch(ctx, S_ETI, 4 * (size_t)(g->m ? g->m : 1), &ep));
    eti = static_cast<uint32_t*>(ep);
    PBF_CHECK(pbf_dev_edge_to_idxs(ctx, dg, use_v, eti));
  }
  PBF_HIP(hipEventRecord(ev[1], ctx->stream));
  RankedGraph R{};
  PBF_CHECK(pbf_dev_rank(ctx, dg, rank_type, &R, use_v));
  PBF_HIP(hipEventRecord(ev[2], ctx->stream));
  PBF_CHECK(pbf_dev_relabel(ctx, dg, use_v, mode == MODE_EDGE, &R));
  PBF_HIP(hipEventRecord(ev[3], ctx->stream));
  PBF_CHECK(pbf_dev_count(ctx, dg, use_v, R, mode, static_cast<uint64_t*>(out_dev), eti));
  PBF_HIP(hipEventRecord(ev[4], ctx->stream));
  if (mode == MODE_TOTAL) {
    if (mem == PBF_MEM_DEVICE) PBF_HIP(hipMemcpyAsync(out_user, out_dev, 8, hipMemcpyDeviceToDevice, ctx->stream));
    else PBF_CHECK(pbf_d2h(ctx, out_user, out_dev, 8));
    PBF_HIP(hipStreamSynchronize(ctx->stream));
  } else {
    PBF_CHECK(finish_output(ctx, mem, out_user, out_dev, out_bytes));
  }
  ctx->timing.convert_ms = elapsed(ev[0], ev[1]);
  ctx->timing.rank_ms = elapsed(ev[1], ev[2]);
  ctx->timing.relabel_ms = elapsed(ev[2], ev[3]);
  ctx->timing.count_ms = elapsed(ev[3], ev[4]);
  ctx->timing.total_ms = elapsed(ev[0], ev[4]);
  ctx->timing.wedges = R.wedges;
  return PBF_OK;
}
}  // namespace

int pbf_scratch(pbf_context* ctx, int slot, size_t bytes, void** out) {
  if (slot < 0 || slot >= S_NUM_SLOTS) return PBF_ERR_ARG;
  if (ctx->fixed.base) {
    // fixed region: a slot keeps its piece while it is large enough, else gets a new piece
    // past the last one (the old piece is not reused within the call)
    if (ctx->cap[slot] < bytes || !ctx->buf[slot]) {
      const size_t off = (ctx->fixed.used + 255) & ~size_t(255);
      if (off > ctx->fixed.cap || bytes > ctx->fixed.cap - off) return PBF_ERR_CAPACITY;
      ctx->buf[slot] = ctx->fixed.base + off;
      ctx->cap[slot] = bytes;
      ctx->fixed.used = off + bytes;
    }
    *out = ctx->buf[slot];
    return PBF_OK;
  }
  if (ctx->cap[slot] < bytes || !ctx->buf[slot]) {
    if (ctx->buf[slot]) {
      // the stream may still use the old buffer
      PBF_HIP(hipStreamSynchronize(ctx->stream));
      (void)hipFree(ctx->buf[slot]);
      ctx->buf[slot] = nullptr;
      ctx->cap[slot] = 0;
    }
    size_t want = bytes + bytes / 8 + 256;
    PBF_CHECK(alloc_dev(&ctx->buf[slot], want));
    ctx->cap[slot] = want;
  }
  *out = ctx->buf[slot];
  return PBF_OK;
}

size_t pbf_scratch_room(const pbf_context* ctx) {
  if (!ctx->fixed.base) return SIZE_MAX;
  const size_t off = (ctx->fixed.used + 255) & ~size_t(255);
  return off < ctx->fixed.cap ? ctx->fixed.cap - off : 0;
}

namespace {
constexpr int NOALLOC_DEVICES = 64;
std::mutex noalloc_mu;
bool noalloc_done[NOALLOC_DEVICES] = {};
void* noalloc_ptr[NOALLOC_DEVICES] = {};

// Address of the fixed scratch array on `device`, which must be current: looked up once per
// device (each device holds its own copy of the array at its own address).  Null on failure.
void* noalloc_address(int device) {
  if (device < 0 || device >= NOALLOC_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (!noalloc_done[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
    }
    noalloc_ptr[device] = p;
    noalloc_done[device] = true;
  }
  return noalloc_ptr[device];
}
}  // namespace

extern "C" int pbf_create(pbf_context** out, int device) {
  if (!out) return PBF_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { (void)hipGetLastError(); return PBF_ERR_HIP; }
  if (device < 0 || device >= ndev) return PBF_ERR_ARG;
  DeviceGuard guard(device);
  pbf_context* ctx = new (std::nothrow) pbf_context();
  if (!ctx) return PBF_ERR_OOM;
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    ctx->num_cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return PBF_ERR_HIP; }
  ctx->stream = ctx->own_stream;
  for (auto& e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) { pbf_destroy(ctx); return PBF_ERR_HIP; }
  if (hipHostMalloc(reinterpret_cast<void**>(&ctx->pinned), 4096, hipHostMallocDefault) != hipSuccess) {
    pbf_destroy(ctx);
    return PBF_ERR_OOM;
  }
  // the no-allocation entry point's array on this device (a failed lookup only disables that
  // entry point, which then returns PBF_ERR_HIP)
  ctx->noalloc = static_cast<unsigned char*>(noalloc_address(device));
  *out = ctx;
  return PBF_OK;
}

extern "C" int pbf_trim(pbf_context* ctx) {
  if (!ctx) return PBF_ERR_ARG;
  DeviceGuard guard(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  pbf_batch_release(ctx);  // a captured batch refers to the scratch buffers
  for (int s = 0; s < S_NUM_SLOTS; s++) {
    if (ctx->buf[s]) (void)hipFree(ctx->buf[s]);
    ctx->buf[s] = nullptr;
    ctx->cap[s] = 0;
  }
  ctx->hot_clear = false;
  ctx->pool_clean = 0;
  return PBF_OK;
}

extern "C" int pbf_set_limits(pbf_context* ctx, int64_t max_wedges, int64_t max_array_size) {
  if (!ctx || max_wedges < 0 || max_array_size < 0) return PBF_ERR_ARG;
  ctx->max_wedges = max_wedges;
  ctx->max_array = max_array_size;
  return PBF_OK;
}

extern "C" int pbf_destroy(pbf_context* ctx) {
  if (!ctx) return PBF_ERR_ARG;
  DeviceGuard guard(ctx->device);
  pbf_trim(ctx);
  for (auto& e : ctx->ev) if (e) (void)hipEventDestroy(e);
  for (int k = 0; k < 2; k++) {
    if (ctx->stage_ev[k]) (void)hipEventDestroy(ctx->stage_ev[k]);
    if (ctx->stage[k]) (void)hipHostFree(ctx->stage[k]);
  }
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  if (ctx->stream && ctx->stream != ctx->own_stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return PBF_OK;
}

extern "C" void* pbf_stream(pbf_context* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

extern "C" int pbf_set_stream(pbf_context* ctx, void* stream) {
  if (!ctx) return PBF_ERR_ARG;
  DeviceGuard guard(ctx->device);
  hipStream_t next = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
  if (next == ctx->stream) return PBF_OK;
  // the previous stream's work (and the staging copies recorded on it) completes first
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < 2; k++) ctx->stage_busy[k] = false;
  ctx->stream = next;
  return PBF_OK;
}

extern "C" int pbf_synchronize(pbf_context* ctx) {
  if (!ctx) return PBF_ERR_ARG;
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  return PBF_OK;
}

extern "C" int pbf_last_timing(pbf_context* ctx, pbf_timing* out) {
  if (!ctx || !out) return PBF_ERR_ARG;
  *out = ctx->timing;
  return PBF_OK;
}

extern "C" int pbf_device_alloc(pbf_context* ctx, size_t bytes, void** ptr) {
  if (!ctx || !ptr) return PBF_ERR_ARG;
  DeviceGuard guard(ctx->device);
  return alloc_dev(ptr, bytes);
}

extern "C" int pbf_device_free(pbf_context* ctx, void* ptr) {
  if (!ctx) return PBF_ERR_ARG;
  if (!ptr) return PBF_OK;
  DeviceGuard guard(ctx->device);
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  PBF_HIP(hipFree(ptr));
  return PBF_OK;
}

extern "C" int pbf_copy_to_device(pbf_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return PBF_ERR_ARG;
  DeviceGuard guard(ctx->device);
  if (bytes) PBF_CHECK(pbf_h2d(ctx, dst, src, bytes));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  return PBF_OK;
}

namespace {
// Streaming copy: 16-byte vectors, grid-stride; the tail (bytes % 16) is copied bytewise.
__global__ __launch_bounds__(256) void k_copy(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n16,
                                              uint8_t* __restrict__ dtail, const uint8_t* __restrict__ stail, size_t ntail) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16; k += stride) dst[k] = src[k];
  if (blockIdx.x == 0 && threadIdx.x < ntail) dtail[threadIdx.x] = stail[threadIdx.x];
}
}  // namespace

extern "C" int pbf_copy_device(pbf_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return PBF_ERR_ARG;
  if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u) return PBF_ERR_ARG;
  DeviceGuard guard(ctx->device);
  if (bytes) {
    const size_t n16 = bytes / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((n16 + 255) / 256 + 1, (size_t)ctx->num_cus * 16);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, ctx->stream, static_cast<uint4*>(dst),
                       static_cast<const uint4*>(src), n16, static_cast<uint8_t*>(dst) + n16 * 16,
                       static_cast<const uint8_t*>(src) + n16 * 16, bytes - n16 * 16);
    PBF_HIP(hipGetLastError());
  }
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  return PBF_OK;
}

extern "C" int pbf_copy_to_host(pbf_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return PBF_ERR_ARG;
  DeviceGuard guard(ctx->device);
  if (bytes) PBF_CHECK(pbf_d2h(ctx, dst, src, bytes));
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  return PBF_OK;
}

extern "C" int pbf_side(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int* use_v, int64_t* num_wedges) {
  if (!ctx || !use_v || !num_wedges) return PBF_ERR_ARG;
  DeviceGuard guard(ctx->device);
  DevGraph dg;
  PBF_CHECK(resolve_graph(ctx, g, mem, &dg));
  return pbf_dev_side(ctx, dg, use_v, num_wedges);
}

extern "C" int pbf_rank(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int rank_type, uint32_t* order,
                        uint32_t* rank_of, int64_t* wedges) {
  if (!ctx) return PBF_ERR_ARG;
  if (!rank_supported(rank_type)) return PBF_ERR_UNSUPPORTED;
  DeviceGuard guard(ctx->device);
  DevGraph dg;
  PBF_CHECK(resolve_graph(ctx, g, mem, &dg));
  int use_v = 1;
  if (rank_type == PBF_RANK_SIDE) {
    int64_t w;
    PBF_CHECK(pbf_dev_side(ctx, dg, &use_v, &w));
  }
  RankedGraph R{};
  PBF_CHECK(pbf_dev_rank(ctx, dg, rank_type, &R, use_v));
  if (wedges) {
    PBF_CHECK(pbf_dev_relabel(ctx, dg, use_v, false, &R));
    *wedges = R.wedges;
  }
  const size_t nb = 4 * (size_t)R.n;
  if (mem == PBF_MEM_DEVICE) {
    if (order && nb) PBF_HIP(hipMemcpyAsync(order, R.order, nb, hipMemcpyDeviceToDevice, ctx->stream));
    if (rank_of && nb) PBF_HIP(hipMemcpyAsync(rank_of, R.rank_of, nb, hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    if (order) PBF_CHECK(pbf_d2h(ctx, order, R.order, nb));
    if (rank_of) PBF_CHECK(pbf_d2h(ctx, rank_of, R.rank_of, nb));
  }
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  return PBF_OK;
}

extern "C" int pbf_count_vert(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v, int count_type,
                              int rank_type, uint64_t* out) {
  return count_common(ctx, g, mem, use_v, count_type, rank_type, MODE_VERT, out);
}

extern "C" int pbf_count_vert_noalloc(pbf_context* ctx, const pbf_bipartite_csr* g, int use_v, int count_type,
                                      int rank_type, uint64_t* out) {
  if (!ctx || !g || !out) return PBF_ERR_ARG;
  if (!ctx->noalloc) return PBF_ERR_HIP;
  // The arena hands out pieces of the device's fixed array for this call only: the context's
  // own buffers are set aside and restored afterwards, untouched.  The array keeps what the
  // last call left, so the buffers that must start at zero (hub accumulators, HBM-fallback
  // counters) are cleared again.
  struct Session {
    pbf_context* c;
    void* buf[S_NUM_SLOTS];
    size_t cap[S_NUM_SLOTS];
    bool hot_clear;
    size_t pool_clean;
    explicit Session(pbf_context* ctx) : c(ctx), hot_clear(ctx->hot_clear), pool_clean(ctx->pool_clean) {
      memcpy(buf, c->buf, sizeof(buf));
      memcpy(cap, c->cap, sizeof(cap));
      memset(c->buf, 0, sizeof(c->buf));
      memset(c->cap, 0, sizeof(c->cap));
      c->hot_clear = false;
      c->pool_clean = 0;
      c->fixed.base = c->noalloc;
      c->fixed.cap = PBF_NOALLOC_SCRATCH_BYTES;
      c->fixed.used = 0;
    }
    ~Session() {
      c->fixed.base = nullptr;
      c->fixed.cap = c->fixed.used = 0;
      memcpy(c->buf, buf, sizeof(buf));
      memcpy(c->cap, cap, sizeof(cap));
      c->hot_clear = hot_clear;
      c->pool_clean = pool_clean;
    }
  } session(ctx);
  return count_common(ctx, g, PBF_MEM_DEVICE, use_v, count_type, rank_type, MODE_VERT, out);
}

extern "C" int pbf_count_edge(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v, int count_type,
                              int rank_type, uint64_t* out) {
  return count_common(ctx, g, mem, use_v, count_type, rank_type, MODE_EDGE, out);
}

extern "C" int pbf_count_total(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v, int count_type,
                               int rank_type, uint64_t* total) {
  return count_common(ctx, g, mem, use_v, count_type, rank_type, MODE_TOTAL, total);
}

extern "C" int pbf_edge_to_idxs(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v, uint32_t* eti) {
  if (!ctx || !eti) return PBF_ERR_ARG;
  DeviceGuard guard(ctx->device);
  DevGraph dg;
  PBF_CHECK(resolve_graph(ctx, g, mem, &dg));
  void* dev;
  PBF_CHECK(out_buffer(ctx, mem, eti, 4 * (size_t)g->m, &dev));
  PBF_CHECK(pbf_dev_edge_to_idxs(ctx, dg, use_v, static_cast<uint32_t*>(dev)));
  return finish_output(ctx, mem, eti, dev, 4 * (size_t)g->m);
}

namespace {
// Shared driver of the two peel entry points: counts in, peel numbers out (n items).
int peel_common(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v, int peel_type,
                const uint64_t* counts, uint64_t* out, bool edges) {
  if (!ctx || !counts || !out) return PBF_ERR_ARG;
  if (peel_type < PBF_PEEL_SORT || peel_type > PBF_PEEL_GPU) return PBF_ERR_ARG;
  DeviceGuard guard(ctx->device);
  DevGraph dg;
  PBF_CHECK(resolve_graph(ctx, g, mem, &dg));
  const int64_t n = edges ? g->m : (use_v ? g->nu : g->nv);
  const size_t bytes = 8 * (size_t)n;
  const uint64_t* cdev = counts;
  void* odev;
  if (mem == PBF_MEM_HOST) {
    void* cp;
    PBF_CHECK(pbf_scratch(ctx, S_COUNTS, bytes ? bytes : 8, &cp));
    if (bytes) PBF_CHECK(pbf_h2d(ctx, cp, counts, bytes));
    cdev = static_cast<const uint64_t*>(cp);
  }
  PBF_CHECK(out_buffer(ctx, mem, out, bytes ? bytes : 8, &odev));
  memset(&ctx->timing, 0, sizeof(ctx->timing));
  PBF_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
  if (edges) PBF_CHECK(pbf_dev_peel_edge(ctx, dg, use_v, cdev, static_cast<uint64_t*>(odev)));
  else PBF_CHECK(pbf_dev_peel_vert(ctx, dg, use_v, cdev, static_cast<uint64_t*>(odev)));
  PBF_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
  PBF_CHECK(finish_output(ctx, mem, out, odev, bytes));
  ctx->timing.peel_ms = elapsed(ctx->ev[0], ctx->ev[1]);
  ctx->timing.total_ms = ctx->timing.peel_ms;
  return PBF_OK;
}
}  // namespace

extern "C" int pbf_peel_vert(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v, int peel_type,
                             const uint64_t* counts, uint64_t* tip) {
  return peel_common(ctx, g, mem, use_v, peel_type, counts, tip, false);
}

extern "C" int pbf_peel_edge(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v, int peel_type,
                             const uint64_t* counts, uint64_t* wing) {
  return peel_common(ctx, g, mem, use_v, peel_type, counts, wing, true);
}

extern "C" int pbf_sparsify(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int sparse_type, int64_t denom,
                            int64_t seed, int64_t* nv_out, int64_t* nu_out, int64_t* m_out, uint32_t* offsetsV,
                            uint32_t* edgesV, uint32_t* offsetsU, uint32_t* edgesU) {
  if (!ctx || !nv_out || !nu_out || !m_out || !offsetsV || !offsetsU) return PBF_ERR_ARG;
  DeviceGuard guard(ctx->device);
  DevGraph dg, sg;
  PBF_CHECK(resolve_graph(ctx, g, mem, &dg));
  if (dg.m > 0 && (!edgesV || !edgesU)) return PBF_ERR_ARG;
  uint32_t *oV = offsetsV, *eV = edgesV, *oU = offsetsU, *eU = edgesU;
  if (mem == PBF_MEM_HOST) {
    void* p;
    PBF_CHECK(pbf_scratch(ctx, S_SPARSE1, 4 * (size_t)(dg.nv + 1 + dg.m + dg.nu + 1 + dg.m), &p));
    oV = static_cast<uint32_t*>(p);
    eV = oV + dg.nv + 1;
    oU = eV + dg.m;
    eU = oU + dg.nu + 1;
  }
  memset(&ctx->timing, 0, sizeof(ctx->timing));
  PBF_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
  PBF_CHECK(pbf_dev_sparsify(ctx, dg, sparse_type, denom, seed, &sg, oV, eV, oU, eU));
  PBF_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
  if (mem == PBF_MEM_HOST) {
    PBF_CHECK(pbf_d2h(ctx, offsetsV, oV, 4 * (size_t)(sg.nv + 1)));
    PBF_CHECK(pbf_d2h(ctx, edgesV, eV, 4 * (size_t)sg.m));
    PBF_CHECK(pbf_d2h(ctx, offsetsU, oU, 4 * (size_t)(sg.nu + 1)));
    PBF_CHECK(pbf_d2h(ctx, edgesU, eU, 4 * (size_t)sg.m));
  }
  PBF_HIP(hipStreamSynchronize(ctx->stream));
  ctx->timing.sparsify_ms = elapsed(ctx->ev[0], ctx->ev[1]);
  ctx->timing.total_ms = ctx->timing.sparsify_ms;
  *nv_out = sg.nv;
  *nu_out = sg.nu;
  *m_out = sg.m;
  return PBF_OK;
}

extern "C" int pbf_count_total_sparse(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int sparse_type,
                                      int64_t denom, int64_t seed, int count_type, int rank_type, int64_t* approx,
                                      uint64_t* sparse_total) {
  if (!ctx || !approx) return PBF_ERR_ARG;
  if (sparse_type < PBF_SPARSE_NONE || sparse_type > PBF_SPARSE_EDGE) return PBF_ERR_ARG;
  if (!count_supported(count_type)) return PBF_ERR_ARG;
  if (!rank_supported(rank_type)) return PBF_ERR_UNSUPPORTED;
  DeviceGuard guard(ctx->device);
  DevGraph dg, sg;
  PBF_CHECK(resolve_graph(ctx, g, mem, &dg));
  memset(&ctx->timing, 0, sizeof(ctx->timing));
  hipEvent_t* ev = ctx->ev;
  PBF_HIP(hipEventRecord(ev[0], ctx->stream));
  if (sparse_type == PBF_SPARSE_NONE) {
    sg = dg;
  } else {
    void* p;
    PBF_CHECK(pbf_scratch(ctx, S_SPARSE1, 4 * (size_t)(dg.nv + 1 + dg.m + dg.nu + 1 + dg.m), &p));
    uint32_t* oV = static_cast<uint32_t*>(p);
    PBF_CHECK(pbf_dev_sparsify(ctx, dg, sparse_type, denom, seed, &sg, oV, oV + dg.nv + 1, oV + dg.nv + 1 + dg.m,
                               oV + dg.nv + 1 + dg.m + dg.nu + 1));
  }
  PBF_HIP(hipEventRecord(ev[1], ctx->stream));
  uint64_t total = 0;
  int64_t wedges = 0;
  RankedGraph R{};
  if (sg.m > 0) {
    int use_v = 0;
    PBF_CHECK(pbf_dev_side(ctx, sg, &use_v, &wedges));
    PBF_HIP(hipEventRecord(ev[2], ctx->stream));
    PBF_CHECK(pbf_dev_rank(ctx, sg, rank_type, &R, use_v));
    PBF_HIP(hipEventRecord(ev[3], ctx->stream));
    PBF_CHECK(pbf_dev_relabel(ctx, sg, use_v, false, &R));
    PBF_HIP(hipEventRecord(ev[4], ctx->stream));
    void* out_dev;
    PBF_CHECK(pbf_scratch(ctx, S_OUT, 8, &out_dev));
    PBF_CHECK(pbf_dev_count(ctx, sg, use_v, R, MODE_TOTAL, static_cast<uint64_t*>(out_dev), nullptr));
    PBF_HIP(hipEventRecord(ev[5], ctx->stream));
    PBF_HIP(hipMemcpyAsync(ctx->pinned + 50, out_dev, 8, hipMemcpyDeviceToHost, ctx->stream));
    PBF_HIP(hipStreamSynchronize(ctx->stream));
    total = ctx->pinned[50];
    ctx->timing.rank_ms = elapsed(ev[2], ev[3]);
    ctx->timing.relabel_ms = elapsed(ev[3], ev[4]);
    ctx->timing.count_ms = elapsed(ev[4], ev[5]);
    ctx->timing.total_ms = elapsed(ev[0], ev[5]);
    ctx->timing.wedges = R.wedges;
  } else {
    PBF_HIP(hipStreamSynchronize(ctx->stream));
    ctx->timing.total_ms = elapsed(ev[0], ev[1]);
  }
  ctx->timing.sparsify_ms = sparse_type == PBF_SPARSE_NONE ? 0.0 : elapsed(ev[0], ev[1]);
  if (sparse_total) *sparse_total = total;
  return pbf_sparse_scale(total, sparse_type, denom, approx);
}
