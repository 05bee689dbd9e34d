// Reading a graph file straight into HBM: readBipartite (apps/hyper/butterfly_utils.h:1008-1058)
// with the tokenising and number parsing on the GPU, and the library's binary CSR layout copied
// as is; both then validated on the GPU with the checks (and the error codes, first error in
// the order the host reader meets them) of pbf_read_bipartite / pbf_validate.
//
// Text path: the host maps the file and reads the five header tokens; the whole file is copied
// to HBM through pinned staging; then
//   k_tok_count   per 4096-byte tile, the number of token starts (a non-space byte after a
//                 space; the whitespace set of ligra/IO.h:75-84);
//   scan          tile bases = index of each tile's first token;
//   k_tok_parse   every token start parses its decimal and stores it at its index in the
//                 flat layout offsetsV | edgesV | offsetsU | edgesU (neighbour ranges checked).
// Validation (text and binary): k_val_off, k_val_start + k_val_entries per side, and
// k_val_transpose (each V entry (v, u) is found in u's list by binary search).  Every check
// posts (phase, position, code) to one 64-bit atomic minimum, so the reported error is the one
// a serial reader meets first; every access is bounds-checked, so the checks run together
// even on malformed input.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstring>
#include <string>

#include "owner.h"

namespace {

constexpr int TOK_THREADS = 256, TOK_BYTES = 16, TOK_TILE = TOK_THREADS * TOK_BYTES;  // 4096
constexpr int VAL_TILE = 2048, VAL_LDS = VAL_TILE + 1;

__device__ __forceinline__ bool is_space(uint32_t c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == 0; }

__device__ __forceinline__ unsigned long long err_key(uint64_t phase, uint64_t pos, uint64_t code) {
  return (unsigned long long)((phase << 56) | (pos << 4) | code);
}

// Non-space bytes of this thread's 16 (bytes before `body` count as spaces) and the token
// starts among them as a bit mask.
__device__ __forceinline__ uint32_t token_starts(const uint8_t* buf, int64_t base, int64_t body) {
  const uint4 w = *reinterpret_cast<const uint4*>(buf + base);
  const uint32_t word[4] = {w.x, w.y, w.z, w.w};
  bool prev_space = base == 0 || base - 1 < body || is_space(buf[base - 1]);
  uint32_t starts = 0;
#pragma unroll
  for (int j = 0; j < TOK_BYTES; j++) {
    const uint32_t c = (word[j >> 2] >> (8 * (j & 3))) & 0xFFu;
    const bool sp = base + j < body || is_space(c);
    if (!sp && prev_space) starts |= 1u << j;
    prev_space = sp;
  }
  return starts;
}

__global__ __launch_bounds__(TOK_THREADS) void k_tok_count(const uint8_t* __restrict__ buf, int64_t body,
                                                           uint32_t* __restrict__ tile_cnt) {
  __shared__ uint32_t s_w[TOK_THREADS / 64];
  const int64_t base = (int64_t)blockIdx.x * TOK_TILE + (int64_t)threadIdx.x * TOK_BYTES;
  const uint32_t c = __popc(token_starts(buf, base, body));
  const uint32_t ws = pbf_wave_sum(c);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = ws;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int k = 0; k < TOK_THREADS / 64; k++) t += s_w[k];
    tile_cnt[blockIdx.x] = t;
  }
}

struct ParseArgs {
  const uint8_t* buf;
  int64_t nbytes, body;
  const uint32_t* tile_base;
  int64_t nv, nu, m;
  uint32_t* flat;  // offsetsV[nv+1] | edgesV[m] | offsetsU[nu+1] | edgesU[m]
  unsigned long long* key;
};

__global__ __launch_bounds__(TOK_THREADS) void k_tok_parse(ParseArgs a) {
  __shared__ uint32_t s_w[TOK_THREADS / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * TOK_TILE + (int64_t)threadIdx.x * TOK_BYTES;
  uint32_t starts = token_starts(a.buf, base, a.body);
  const uint32_t c = __popc(starts);
  // block-exclusive scan of the counts
  const uint32_t incl = pbf_wave_incl_scan(c);
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) s_w[wid] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (int k = 0; k < wid; k++) before += s_w[k];
  int64_t k = (int64_t)a.tile_base[blockIdx.x] + before + incl - c;
  const int64_t expected = a.nv + a.m + a.nu + a.m;
  while (starts) {
    const int j = __ffs(starts) - 1;
    starts &= starts - 1;
    // the token runs to the next space (the buffer ends in spaces)
    uint64_t v = 0;
    bool digits = true;
    for (int64_t q = base + j; q < a.nbytes; q++) {
      const uint32_t ch = a.buf[q];
      if (is_space(ch)) break;
      if (ch < '0' || ch > '9') digits = false;
      v = v * 10 + (uint64_t)(ch - '0');
    }
    if (k >= expected) {
      atomicMin(a.key, err_key(0, (uint64_t)expected, PBF_ERR_FORMAT));  // trailing tokens
    } else if (!digits) {
      atomicMin(a.key, err_key(0, (uint64_t)k, PBF_ERR_FORMAT));
    } else if (k < a.nv) {
      a.flat[k] = (uint32_t)v;
    } else if (k < a.nv + a.m) {
      if (v >= (uint64_t)a.nu) atomicMin(a.key, err_key(0, (uint64_t)k, PBF_ERR_RANGE));
      else a.flat[k + 1] = (uint32_t)v;
    } else if (k < a.nv + a.m + a.nu) {
      a.flat[k + 1] = (uint32_t)v;
    } else {
      if (v >= (uint64_t)a.nv) atomicMin(a.key, err_key(0, (uint64_t)k, PBF_ERR_RANGE));
      else a.flat[k + 2] = (uint32_t)v;
    }
    k++;
  }
}

// off[n] = m (the reader's sentinel) for both sides.
__global__ void k_sentinels(uint32_t* offV, int64_t nv, uint32_t* offU, int64_t nu, int64_t m) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    offV[nv] = (uint32_t)m;
    offU[nu] = (uint32_t)m;
  }
}

// validate_side's offset checks: off[0] == 0 and off[n] == m (phase p0), no decrease (p0 + 1).
__global__ void k_val_off(const uint32_t* __restrict__ off, int64_t n, int64_t m, uint64_t p0,
                          unsigned long long* __restrict__ key) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    if (x == 0 && (off[0] != 0 || (int64_t)off[n] != m)) atomicMin(key, err_key(p0, 0, PBF_ERR_FORMAT));
    if (off[x + 1] < off[x]) atomicMin(key, err_key(p0 + 1, (uint64_t)x, PBF_ERR_UNSORTED));
  }
  if (n == 0 && blockIdx.x == 0 && threadIdx.x == 0 && (off[0] != 0 || m != 0))
    atomicMin(key, err_key(p0, 0, PBF_ERR_FORMAT));
}

// start[p] = 1 where a non-empty list begins (start[] cleared to 0 beforehand).
__global__ void k_val_start(const uint32_t* __restrict__ off, int64_t n, int64_t m, uint8_t* __restrict__ start) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t b = off[x], e = off[x + 1];
    if (b < e && (int64_t)e <= m) start[b] = 1;
  }
}

// validate_side's entry checks (phase p): neighbour below `other` (RANGE), lists strictly
// ascending (UNSORTED); at one position RANGE is met first.
__global__ void k_val_entries(const uint32_t* __restrict__ edg, int64_t m, int64_t other, const uint8_t* __restrict__ start,
                              uint64_t p, unsigned long long* __restrict__ key) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t e = edg[q];
    if ((int64_t)e >= other) atomicMin(key, err_key(p, (uint64_t)q, PBF_ERR_RANGE));
    else if (q > 0 && !start[q] && e <= edg[q - 1]) atomicMin(key, err_key(p, (uint64_t)q, PBF_ERR_UNSORTED));
  }
}

// The U lists are the transpose of the V lists (phase 7): with both sides strictly ascending
// and m entries each, it suffices that every V entry (v, u) is found in u's list.
__global__ __launch_bounds__(256) void k_val_transpose(DevGraph g, const uint32_t* __restrict__ xb,
                                                       unsigned long long* __restrict__ key) {
  __shared__ uint32_t s_off[VAL_LDS];
  const int64_t t = blockIdx.x;
  TileOwners own;
  own.stage(g.offV, xb, t, s_off, VAL_LDS);
  const int64_t base = t * VAL_TILE;
  for (int k = 0; k < VAL_TILE / 256; k++) {
    const int64_t p = base + (int64_t)k * 256 + threadIdx.x;
    if (p >= g.m) break;
    const uint32_t v = own.owner(g.offV, s_off, (uint32_t)p);
    const uint32_t u = g.edgV[p];
    bool found = false;
    if ((int64_t)u < g.nu) {
      uint32_t lo = g.offU[u], hi = g.offU[u + 1];
      if (lo <= hi && (int64_t)hi <= g.m) {
        while (lo < hi) {
          const uint32_t mid = lo + ((hi - lo) >> 1);
          if (g.edgU[mid] < v) lo = mid + 1;
          else hi = mid;
        }
        found = lo < g.offU[u + 1] && g.edgU[lo] == v;
      }
    }
    if (!found) atomicMin(key, err_key(7, (uint64_t)p, PBF_ERR_FORMAT));
  }
}

inline unsigned grid256(int64_t n, int max_blocks) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)b;
}

struct MappedFile {
  const char* p = nullptr;
  size_t n = 0;
  int fd = -1;
  bool open(const char* path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (fstat(fd, &st) != 0) return false;
    n = (size_t)st.st_size;
    if (n == 0) { p = ""; return true; }
    void* q = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    if (q == MAP_FAILED) { p = nullptr; return false; }
    p = static_cast<const char*>(q);
    return true;
  }
  ~MappedFile() {
    if (p && n) munmap(const_cast<char*>(p), n);
    if (fd >= 0) close(fd);
  }
};

bool host_space(char c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == 0; }

// Next whitespace-separated word of [*pos, n); false at the end.
bool next_word(const char* s, size_t n, size_t* pos, size_t* ws, size_t* we) {
  size_t i = *pos;
  while (i < n && host_space(s[i])) i++;
  if (i >= n) return false;
  *ws = i;
  while (i < n && !host_space(s[i])) i++;
  *we = i;
  *pos = i;
  return true;
}

bool parse_u64(const char* s, size_t a, size_t b, uint64_t* out) {
  uint64_t v = 0;
  for (size_t i = a; i < b; i++) {
    if (s[i] < '0' || s[i] > '9') return false;
    v = v * 10 + uint64_t(s[i] - '0');
  }
  *out = v;
  return true;
}

int decode_key(unsigned long long k) { return k == ~0ull ? PBF_OK : (int)(k & 0xF); }

// The validation checks of both sides and the transpose, all posted to *key (not read here).
int validate_dev(pbf_context* ctx, const DevGraph& g, unsigned long long* key) {
  const int64_t sides_n[2] = {g.nv, g.nu}, sides_o[2] = {g.nu, g.nv};
  const uint32_t* offs[2] = {g.offV, g.offU};
  const uint32_t* edgs[2] = {g.edgV, g.edgU};
  const int64_t ntiles = (g.m + VAL_TILE - 1) / VAL_TILE;
  void* sp;
  PBF_CHECK(pbf_scratch(ctx, S_TEXT1, (size_t)(g.m + 16) + 4 * (size_t)(ntiles + 1) + 16, &sp));
  uint8_t* start = static_cast<uint8_t*>(sp);
  uint32_t* xb = reinterpret_cast<uint32_t*>(start + ((g.m + 16 + 3) & ~int64_t(3)));
  for (int s = 0; s < 2; s++) {
    const uint64_t p0 = 1 + 3 * (uint64_t)s;
    hipLaunchKernelGGL(k_val_off, dim3(grid256(sides_n[s], 8192)), dim3(256), 0, ctx->stream, offs[s], sides_n[s], g.m, p0, key);
    if (g.m > 0) {
      PBF_CHECK(pbf_fill_u32(ctx, reinterpret_cast<uint32_t*>(start), 0u, (g.m + 4) / 4));
      hipLaunchKernelGGL(k_val_start, dim3(grid256(sides_n[s], 8192)), dim3(256), 0, ctx->stream, offs[s], sides_n[s], g.m,
                         start);
      hipLaunchKernelGGL(k_val_entries, dim3(grid256(g.m, 16384)), dim3(256), 0, ctx->stream, edgs[s], g.m, sides_o[s], start,
                         p0 + 2, key);
    }
  }
  if (g.m > 0 && g.nv > 0) {
    hipLaunchKernelGGL(k_tile_owner<VAL_TILE>, dim3(grid256(ntiles + 1, 1024)), dim3(256), 0, ctx->stream, g.offV, g.nv, g.m,
                       ntiles, xb);
    hipLaunchKernelGGL(k_val_transpose, dim3((unsigned)ntiles), dim3(256), 0, ctx->stream, g, xb, key);
  }
  PBF_HIP(hipGetLastError());
  return PBF_OK;
}

}  // namespace

extern "C" int pbf_read_bipartite_device(pbf_context* ctx, const char* path, pbf_bipartite_csr* out, void** storage) {
  if (!ctx || !path || !out || !storage) return PBF_ERR_ARG;
  memset(out, 0, sizeof(*out));
  *storage = nullptr;
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  if (prev != ctx->device) (void)hipSetDevice(ctx->device);
  struct Restore {
    int dev;
    ~Restore() { if (dev >= 0) (void)hipSetDevice(dev); }
  } restore{prev != ctx->device ? prev : -1};

  MappedFile f;
  if (!f.open(path)) return PBF_ERR_IO;
  const bool binary = f.n >= 8 && memcmp(f.p, "PBFCSR1", 8) == 0;
  int64_t nv, nu, m;
  size_t body = 0;
  if (binary) {
    if (f.n < 32) return PBF_ERR_FORMAT;
    int64_t hdr[3];
    memcpy(hdr, f.p + 8, 24);
    nv = hdr[0]; nu = hdr[1]; m = hdr[2];
    if (nv < 0 || nu < 0 || m < 0) return PBF_ERR_FORMAT;
    if (nv >= (int64_t(1) << 31) || nu >= (int64_t(1) << 31) || 2 * m >= (int64_t(1) << 31)) return PBF_ERR_CAPACITY;
    if (f.n != 32 + 4 * (size_t(nv + 1) + size_t(m) + size_t(nu + 1) + size_t(m))) return PBF_ERR_FORMAT;
  } else {
    size_t pos = 0, ws, we;
    if (!next_word(f.p, f.n, &pos, &ws, &we) || std::string(f.p + ws, we - ws) != "AdjacencyHypergraph")
      return PBF_ERR_FORMAT;
    uint64_t h[4];
    for (int k = 0; k < 4; k++)
      if (!next_word(f.p, f.n, &pos, &ws, &we) || !parse_u64(f.p, ws, we, &h[k])) return PBF_ERR_FORMAT;
    if (h[1] != h[3]) return PBF_ERR_FORMAT;
    if (h[0] >= (uint64_t(1) << 31) || h[2] >= (uint64_t(1) << 31) || 2 * h[1] >= (uint64_t(1) << 31))
      return PBF_ERR_CAPACITY;
    nv = (int64_t)h[0]; m = (int64_t)h[1]; nu = (int64_t)h[2];
    body = pos;
  }
  const size_t words = size_t(nv + 1) + size_t(m) + size_t(nu + 1) + size_t(m);
  void* blk = nullptr;
  if (hipMalloc(&blk, 4 * words) != hipSuccess) { (void)hipGetLastError(); return PBF_ERR_OOM; }
  uint32_t* flat = static_cast<uint32_t*>(blk);
  DevGraph g;
  g.nv = nv; g.nu = nu; g.m = m;
  uint32_t* oV = flat;
  uint32_t* eV = oV + nv + 1;
  uint32_t* oU = eV + m;
  uint32_t* eU = oU + nu + 1;
  g.offV = oV; g.edgV = eV; g.offU = oU; g.edgU = eU;
  auto fail = [&](int st) { (void)hipStreamSynchronize(ctx->stream); (void)hipFree(blk); return st; };

  void* kp;
  if (pbf_scratch(ctx, S_TEXT2, 64, &kp) != PBF_OK) return fail(PBF_ERR_OOM);
  unsigned long long* key = static_cast<unsigned long long*>(kp);
  if (pbf_fill_u64(ctx, reinterpret_cast<uint64_t*>(key), ~0ull, 1) != PBF_OK) return fail(PBF_ERR_HIP);
  int64_t total_tokens = 0;
  if (binary) {
    if (pbf_h2d(ctx, flat, f.p + 32, 4 * words) != PBF_OK) return fail(PBF_ERR_HIP);
  } else {
    // the whole file, padded with zero bytes (spaces) to whole tiles
    const int64_t ntiles = (int64_t)((f.n + TOK_TILE - 1) / TOK_TILE);
    const size_t padded = (size_t)(ntiles ? ntiles : 1) * TOK_TILE;
    void* tp;
    if (pbf_scratch(ctx, S_TEXT0, padded + 4 * (size_t)(ntiles + 1) + 16, &tp) != PBF_OK) return fail(PBF_ERR_OOM);
    uint8_t* buf = static_cast<uint8_t*>(tp);
    uint32_t* tile = reinterpret_cast<uint32_t*>(buf + padded);
    if (pbf_fill_u32(ctx, reinterpret_cast<uint32_t*>(buf + padded - TOK_TILE), 0u, TOK_TILE / 4) != PBF_OK)
      return fail(PBF_ERR_HIP);
    if (pbf_h2d(ctx, buf, f.p, f.n) != PBF_OK) return fail(PBF_ERR_HIP);
    if (ntiles > 0) {
      hipLaunchKernelGGL(k_tok_count, dim3((unsigned)ntiles), dim3(TOK_THREADS), 0, ctx->stream, buf, (int64_t)body, tile);
      if (pbf_fill_u32(ctx, tile + ntiles, 0u, 1) != PBF_OK) return fail(PBF_ERR_HIP);
      if (pbf_scan_excl_u32(ctx, tile, tile, ntiles + 1, nullptr) != PBF_OK) return fail(PBF_ERR_HIP);
      ParseArgs a{buf, (int64_t)f.n, (int64_t)body, tile, nv, nu, m, flat, key};
      hipLaunchKernelGGL(k_tok_parse, dim3((unsigned)ntiles), dim3(TOK_THREADS), 0, ctx->stream, a);
    } else {
      if (pbf_fill_u32(ctx, tile, 0u, 1) != PBF_OK) return fail(PBF_ERR_HIP);
    }
    hipLaunchKernelGGL(k_sentinels, dim3(1), dim3(64), 0, ctx->stream, oV, nv, oU, nu, m);
    if (hipGetLastError() != hipSuccess) return fail(PBF_ERR_HIP);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(ctx->pinned + 52);
    if (hipMemcpyAsync(cnt, tile + ntiles, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) return fail(PBF_ERR_HIP);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(PBF_ERR_HIP);
    total_tokens = *cnt;
  }
  if (validate_dev(ctx, g, key) != PBF_OK) return fail(PBF_ERR_HIP);
  if (hipMemcpyAsync(ctx->pinned + 53, key, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) return fail(PBF_ERR_HIP);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(PBF_ERR_HIP);
  unsigned long long k = ctx->pinned[53];
  const int64_t expected = nv + m + nu + m;
  if (!binary && total_tokens < expected) {
    // a missing token: the serial reader fails at index total_tokens
    const unsigned long long miss = ((unsigned long long)total_tokens << 4) | PBF_ERR_FORMAT;
    if (miss < k) k = miss;
  }
  const int st = decode_key(k);
  if (st != PBF_OK) return fail(st);
  out->nv = nv; out->nu = nu; out->m = m;
  out->offsetsV = oV; out->edgesV = eV; out->offsetsU = oU; out->edgesU = eU;
  *storage = blk;
  return PBF_OK;
}
