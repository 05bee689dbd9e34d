// Host side of libparbutterfly_amd: graph storage, AdjacencyHypergraph IO, validation and
// the seeded synthetic generators.  Pure C++ (OpenMP for the generators' sort), no GPU.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <parallel/algorithm>
#include <string>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <omp.h>

#include "parbutterfly.h"
#include "pbf_internal.h"

extern "C" const char* pbf_status_str(int s) {
  switch (s) {
    case PBF_OK: return "ok";
    case PBF_ERR_ARG: return "invalid argument";
    case PBF_ERR_FORMAT: return "bad input format";
    case PBF_ERR_RANGE: return "neighbour id out of range";
    case PBF_ERR_UNSORTED: return "adjacency list not strictly ascending";
    case PBF_ERR_OOM: return "out of memory";
    case PBF_ERR_CAPACITY: return "size beyond supported capacity";
    case PBF_ERR_DOMAIN: return "input outside the parity domain";
    case PBF_ERR_HIP: return "HIP runtime error";
    case PBF_ERR_UNSUPPORTED: return "unsupported option combination";
    case PBF_ERR_IO: return "file IO error";
    default: return "unknown status";
  }
}

extern "C" int pbf_abi_version(void) { return PBF_ABI_VERSION; }

// ------------------------------------------------------------------------------------------
// Graph storage: one malloc'd block holding offsetsV, edgesV, offsetsU, edgesU.
// ------------------------------------------------------------------------------------------
extern "C" int pbf_graph_alloc(int64_t nv, int64_t nu, int64_t m, pbf_graph* out) {
  if (!out || nv < 0 || nu < 0 || m < 0) return PBF_ERR_ARG;
  if (nv >= (int64_t(1) << 31) || nu >= (int64_t(1) << 31) || 2 * m >= (int64_t(1) << 31))
    return PBF_ERR_CAPACITY;
  size_t words = size_t(nv + 1) + size_t(m) + size_t(nu + 1) + size_t(m);
  uint32_t* blk = static_cast<uint32_t*>(calloc(words ? words : 1, sizeof(uint32_t)));
  if (!blk) return PBF_ERR_OOM;
  out->storage = blk;
  out->csr.nv = nv; out->csr.nu = nu; out->csr.m = m;
  out->csr.offsetsV = blk;
  out->csr.edgesV = blk + (nv + 1);
  out->csr.offsetsU = blk + (nv + 1) + m;
  out->csr.edgesU = blk + (nv + 1) + m + (nu + 1);
  return PBF_OK;
}

extern "C" void pbf_graph_free(pbf_graph* g) {
  if (!g) return;
  free(g->storage);
  g->storage = nullptr;
  memset(&g->csr, 0, sizeof(g->csr));
}

static inline uint32_t* mut(const uint32_t* p) { return const_cast<uint32_t*>(p); }

// Builds both CSRs from a list of (v, u) keys already sorted ascending and duplicate free.
static int csr_from_sorted_keys(int64_t nv, int64_t nu, const std::vector<uint64_t>& keys, pbf_graph* out) {
  int64_t m = (int64_t)keys.size();
  int st = pbf_graph_alloc(nv, nu, m, out);
  if (st) return st;
  uint32_t* oV = mut(out->csr.offsetsV);
  uint32_t* eV = mut(out->csr.edgesV);
  uint32_t* oU = mut(out->csr.offsetsU);
  uint32_t* eU = mut(out->csr.edgesU);
  for (int64_t i = 0; i < m; i++) {
    uint32_t v = uint32_t(keys[i] >> 32), u = uint32_t(keys[i]);
    oV[v + 1]++;
    oU[u + 1]++;
    eV[i] = u;
  }
  for (int64_t v = 0; v < nv; v++) oV[v + 1] += oV[v];
  for (int64_t u = 0; u < nu; u++) oU[u + 1] += oU[u];
  std::vector<uint32_t> cur(oU, oU + nu);
  // keys are ordered by (v, u), so each U list receives its V neighbours in ascending order
  for (int64_t i = 0; i < m; i++) {
    uint32_t v = uint32_t(keys[i] >> 32), u = uint32_t(keys[i]);
    eU[cur[u]++] = v;
  }
  return PBF_OK;
}

extern "C" int pbf_graph_from_edges(int64_t nv, int64_t nu, int64_t ne, const uint32_t* v, const uint32_t* u,
                                    pbf_graph* out) {
  if (!out || (ne > 0 && (!v || !u)) || nv < 0 || nu < 0 || ne < 0) return PBF_ERR_ARG;
  std::vector<uint64_t> keys((size_t)ne);
  for (int64_t i = 0; i < ne; i++) {
    if (v[i] >= (uint64_t)nv || u[i] >= (uint64_t)nu) return PBF_ERR_RANGE;
    keys[i] = (uint64_t(v[i]) << 32) | u[i];
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  return csr_from_sorted_keys(nv, nu, keys, out);
}

// ------------------------------------------------------------------------------------------
// Validation: the reference checks only the header, mv == mu and id ranges
// (butterfly_utils.h:1013-1051); the algorithms additionally rely on sorted, duplicate-free
// lists (every "else break" in the wedge loops) and on the two CSRs describing the same edges.
// ------------------------------------------------------------------------------------------
// Lowest index x in [0, n) with bad(x), or n; checked in parallel, so the error reported is the
// one a serial scan would meet first.
template <class F>
static int64_t first_bad(int64_t n, F bad) {
  int64_t first = n;
#pragma omp parallel for schedule(static, 4096) reduction(min : first)
  for (int64_t x = 0; x < n; x++)
    if (x < first && bad(x)) first = x;
  return first;
}

static int validate_side(int64_t n, int64_t other, int64_t m, const uint32_t* off, const uint32_t* edg) {
  if (off[0] != 0 || (int64_t)off[n] != m) return PBF_ERR_FORMAT;
  // offsets first, so that the list scan below reads only inside [0, m)
  if (first_bad(n, [&](int64_t x) { return off[x + 1] < off[x]; }) < n) return PBF_ERR_UNSORTED;
  int64_t x = first_bad(n, [&](int64_t x) {
    for (uint32_t p = off[x]; p < off[x + 1]; p++)
      if ((int64_t)edg[p] >= other || (p > off[x] && edg[p] <= edg[p - 1])) return true;
    return false;
  });
  if (x == n) return PBF_OK;
  for (uint32_t p = off[x]; p < off[x + 1]; p++) {
    if ((int64_t)edg[p] >= other) return PBF_ERR_RANGE;
    if (p > off[x] && edg[p] <= edg[p - 1]) return PBF_ERR_UNSORTED;
  }
  return PBF_OK;
}

extern "C" int pbf_validate(const pbf_bipartite_csr* g) {
  if (!g || g->nv < 0 || g->nu < 0 || g->m < 0) return PBF_ERR_ARG;
  if (g->nv >= (int64_t(1) << 31) || g->nu >= (int64_t(1) << 31) || 2 * g->m >= (int64_t(1) << 31))
    return PBF_ERR_CAPACITY;
  if (!g->offsetsV || !g->offsetsU || (g->m > 0 && (!g->edgesV || !g->edgesU))) return PBF_ERR_ARG;
  int st = validate_side(g->nv, g->nu, g->m, g->offsetsV, g->edgesV);
  if (st) return st;
  st = validate_side(g->nu, g->nv, g->m, g->offsetsU, g->edgesU);
  if (st) return st;
  // The U lists must be the transpose of the V lists: walking the V lists in order visits each
  // U list's entries in ascending V order. Each thread owns a range of U ids (cut at edge
  // boundaries) and walks all V lists, following only the entries in its range.
  const uint32_t *offU = g->offsetsU, *edgU = g->edgesU, *offV = g->offsetsV, *edgV = g->edgesV;
  int ok = 1;
#pragma omp parallel reduction(& : ok)
  {
    int nt = omp_get_num_threads(), t = omp_get_thread_num();
    int64_t m = g->m;
    // thread t owns the U ids whose lists start in edge range [m t / nt, m (t + 1) / nt)
    auto cut = [&](int k) -> uint32_t {
      if (k == 0) return 0;
      if (k == nt) return (uint32_t)g->nu;
      return (uint32_t)(std::lower_bound(offU, offU + g->nu, (uint32_t)(m * k / nt)) - offU);
    };
    uint32_t lo = cut(t), hi = cut(t + 1);
    std::vector<uint32_t> cur(offU + lo, offU + hi);
    for (int64_t v = 0; v < g->nv && ok; v++)
      for (uint32_t p = offV[v]; p < offV[v + 1]; p++) {
        uint32_t u = edgV[p];
        if (u < lo || u >= hi) continue;
        uint32_t& c = cur[u - lo];
        if (c >= offU[u + 1] || edgU[c] != (uint32_t)v) { ok = 0; break; }
        c++;
      }
  }
  return ok ? PBF_OK : PBF_ERR_FORMAT;
}

extern "C" int pbf_side_host(const pbf_bipartite_csr* g, int* use_v, int64_t* num_wedges) {
  if (!g || !use_v || !num_wedges) return PBF_ERR_ARG;
  int64_t wv = 0, wu = 0;
  for (int64_t v = 0; v < g->nv; v++) wv += pbf_side_term(g->offsetsV[v + 1] - g->offsetsV[v]);
  for (int64_t u = 0; u < g->nu; u++) wu += pbf_side_term(g->offsetsU[u + 1] - g->offsetsU[u]);
  *use_v = wv <= wu;
  *num_wedges = wv <= wu ? wv : wu;
  return PBF_OK;
}

// ------------------------------------------------------------------------------------------
// AdjacencyHypergraph text IO (layout of utils/graphIO.h:779-800, read as butterfly_utils.h:
// 1008-1058): "AdjacencyHypergraph", nv, mv, nu, mu, offsetsV[nv], edgesV[mv], offsetsU[nu],
// edgesU[mu], whitespace separated.  PBFCSR1 binary: magic, int64 nv, nu, m, then the arrays.
// ------------------------------------------------------------------------------------------
namespace {
struct Mapped {
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
    if (q == MAP_FAILED) return false;
    p = static_cast<const char*>(q);
    return true;
  }
  ~Mapped() {
    if (p && n) munmap(const_cast<char*>(p), n);
    if (fd >= 0) close(fd);
  }
};

// The whitespace of the reference's tokeniser (isSpace, ligra/IO.h:75-84).
inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == 0; }

struct Tok {
  const char* p;
  const char* e;
  bool next_word(const char** ws, const char** we) {
    while (p < e && is_space(*p)) p++;
    if (p >= e) return false;
    *ws = p;
    while (p < e && !is_space(*p)) p++;
    *we = p;
    return true;
  }
  // Parses a non-negative decimal; returns false on a non-numeric token or end of input.
  bool next_u64(uint64_t* out) {
    const char *ws, *we;
    if (!next_word(&ws, &we)) return false;
    uint64_t v = 0;
    for (const char* q = ws; q < we; q++) {
      if (*q < '0' || *q > '9') return false;
      v = v * 10 + uint64_t(*q - '0');
    }
    *out = v;
    return true;
  }
};
}  // namespace

static int read_binary(const Mapped& f, pbf_graph* out) {
  if (f.n < 32) return PBF_ERR_FORMAT;
  int64_t hdr[3];
  memcpy(hdr, f.p + 8, 24);
  int64_t nv = hdr[0], nu = hdr[1], m = hdr[2];
  if (nv < 0 || nu < 0 || m < 0) return PBF_ERR_FORMAT;
  size_t need = 32 + 4 * (size_t(nv + 1) + size_t(m) + size_t(nu + 1) + size_t(m));
  if (f.n != need) return PBF_ERR_FORMAT;
  int st = pbf_graph_alloc(nv, nu, m, out);
  if (st) return st;
  memcpy(out->storage, f.p + 32, need - 32);
  st = pbf_validate(&out->csr);
  if (st) pbf_graph_free(out);
  return st;
}

extern "C" int pbf_read_bipartite(const char* path, pbf_graph* out) {
  if (!path || !out) return PBF_ERR_ARG;
  memset(out, 0, sizeof(*out));
  Mapped f;
  if (!f.open(path)) return PBF_ERR_IO;
  if (f.n >= 8 && memcmp(f.p, "PBFCSR1", 8) == 0) return read_binary(f, out);
  Tok t{f.p, f.p + f.n};
  const char *ws, *we;
  if (!t.next_word(&ws, &we) || std::string(ws, we) != "AdjacencyHypergraph") return PBF_ERR_FORMAT;
  uint64_t nv, mv, nu, mu;
  if (!t.next_u64(&nv) || !t.next_u64(&mv) || !t.next_u64(&nu) || !t.next_u64(&mu)) return PBF_ERR_FORMAT;
  if (mv != mu) return PBF_ERR_FORMAT;
  int st = pbf_graph_alloc((int64_t)nv, (int64_t)nu, (int64_t)mv, out);
  if (st) return st;
  uint32_t* oV = mut(out->csr.offsetsV);
  uint32_t* eV = mut(out->csr.edgesV);
  uint32_t* oU = mut(out->csr.offsetsU);
  uint32_t* eU = mut(out->csr.edgesU);
  uint64_t x;
  st = PBF_OK;
  for (uint64_t i = 0; i < nv && !st; i++) { if (!t.next_u64(&x)) st = PBF_ERR_FORMAT; else oV[i] = (uint32_t)x; }
  oV[nv] = (uint32_t)mv;
  for (uint64_t i = 0; i < mv && !st; i++) {
    if (!t.next_u64(&x)) st = PBF_ERR_FORMAT;
    else if (x >= nu) st = PBF_ERR_RANGE;
    else eV[i] = (uint32_t)x;
  }
  for (uint64_t i = 0; i < nu && !st; i++) { if (!t.next_u64(&x)) st = PBF_ERR_FORMAT; else oU[i] = (uint32_t)x; }
  oU[nu] = (uint32_t)mu;
  for (uint64_t i = 0; i < mu && !st; i++) {
    if (!t.next_u64(&x)) st = PBF_ERR_FORMAT;
    else if (x >= nv) st = PBF_ERR_RANGE;
    else eU[i] = (uint32_t)x;
  }
  // trailing tokens mean the header's lengths do not match the body (butterfly_utils.h:1024)
  if (!st && t.next_word(&ws, &we)) st = PBF_ERR_FORMAT;
  if (!st) st = pbf_validate(&out->csr);
  if (st) pbf_graph_free(out);
  return st;
}

// KONECT bipartite edge lists (readKONECT, utils/graphIO.h:1032-1098, turned into a graph as
// utils/KONECTtoHyperAdj.C does through hypergraphFromHyperedges, :378-440).  Leading lines
// that start with '%' are skipped; the number of tokens on the first remaining line is the
// number of columns (a token starts at a non-space after a space of isSpace, and C isspace
// ends one, as there); the rest is tokenised on isSpace and cut into rows of that many tokens
// (a partial last row is dropped); each row's first column is a V id and its second a U id,
// read as atol reads them (further columns, such as weights or times, are ignored).  Each
// side's ids are renumbered in increasing order over the ids that occur; the lists are
// sorted and duplicate edges dropped.  Ids must be below 2^32 and not negative.
extern "C" int pbf_read_konect(const char* path, pbf_graph* out) {
  if (!path || !out) return PBF_ERR_ARG;
  memset(out, 0, sizeof(*out));
  Mapped f;
  if (!f.open(path)) return PBF_ERR_IO;
  const char* p = f.p;
  const size_t n = f.n;
  size_t k = 0;
  while (k < n && p[k] == '%') {
    while (k < n && p[k] != '\n') k++;
    if (k < n) k++;
  }
  int64_t cols = 0;
  bool prev_space = true;
  for (size_t t = k; t < n && p[t] != '\n'; t++) {
    const char c = p[t];
    if (!is_space(c) && prev_space) { prev_space = false; cols++; }
    else if (c == ' ' || c == '\t' || c == '\v' || c == '\f' || c == '\r') prev_space = true;
  }
  if (cols < 2) return PBF_ERR_FORMAT;
  std::vector<uint32_t> vv, uu;
  Tok t{p + k, p + n};
  const char *ws, *we;
  int64_t col = 0;
  uint32_t row[2] = {0, 0};
  // atol of a token: an optional '+' and the leading digits (0 when there are none)
  auto value = [](const char* a, const char* b, uint32_t* x) -> int {
    if (a < b && *a == '-') {
      for (const char* q = a + 1; q < b && *q >= '0' && *q <= '9'; q++)
        if (*q != '0') return PBF_ERR_RANGE;
      *x = 0;
      return PBF_OK;
    }
    if (a < b && *a == '+') a++;
    uint64_t v = 0;
    for (; a < b && *a >= '0' && *a <= '9'; a++) {
      v = v * 10 + uint64_t(*a - '0');
      if (v >> 32) return PBF_ERR_RANGE;
    }
    *x = (uint32_t)v;
    return PBF_OK;
  };
  while (t.next_word(&ws, &we)) {
    if (col < 2) {
      const int st = value(ws, we, &row[col]);
      if (st) return st;
    }
    if (++col == cols) {
      vv.push_back(row[0]);
      uu.push_back(row[1]);
      col = 0;
    }
  }
  const int64_t m = (int64_t)vv.size();
  if (2 * m >= (int64_t(1) << 31)) return PBF_ERR_CAPACITY;
  // dense ids per side, in increasing order of the original ids
  auto compact = [](std::vector<uint32_t>& ids) -> int64_t {
    std::vector<uint32_t> seen(ids);
    __gnu_parallel::sort(seen.begin(), seen.end());
    seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)ids.size(); i++)
      ids[i] = (uint32_t)(std::lower_bound(seen.begin(), seen.end(), ids[i]) - seen.begin());
    return (int64_t)seen.size();
  };
  const int64_t nv = compact(vv), nu = compact(uu);
  return pbf_graph_from_edges(nv, nu, m, vv.data(), uu.data(), out);
}

extern "C" int pbf_write_bipartite(const char* path, const pbf_bipartite_csr* g) {
  if (!path || !g) return PBF_ERR_ARG;
  FILE* f = fopen(path, "w");
  if (!f) return PBF_ERR_IO;
  std::string buf;
  buf.reserve(1 << 20);
  auto flush = [&]() { fwrite(buf.data(), 1, buf.size(), f); buf.clear(); };
  auto put = [&](uint64_t x) {
    char tmp[24];
    int n = snprintf(tmp, sizeof(tmp), "%llu\n", (unsigned long long)x);
    buf.append(tmp, n);
    if (buf.size() > (1 << 20)) flush();
  };
  buf += "AdjacencyHypergraph\n";
  put(g->nv); put(g->m); put(g->nu); put(g->m);
  for (int64_t i = 0; i < g->nv; i++) put(g->offsetsV[i]);
  for (int64_t i = 0; i < g->m; i++) put(g->edgesV[i]);
  for (int64_t i = 0; i < g->nu; i++) put(g->offsetsU[i]);
  for (int64_t i = 0; i < g->m; i++) put(g->edgesU[i]);
  flush();
  int bad = ferror(f);
  fclose(f);
  return bad ? PBF_ERR_IO : PBF_OK;
}

extern "C" int pbf_write_binary(const char* path, const pbf_bipartite_csr* g) {
  if (!path || !g) return PBF_ERR_ARG;
  FILE* f = fopen(path, "wb");
  if (!f) return PBF_ERR_IO;
  int64_t hdr[3] = {g->nv, g->nu, g->m};
  fwrite("PBFCSR1", 1, 8, f);
  fwrite(hdr, 8, 3, f);
  fwrite(g->offsetsV, 4, g->nv + 1, f);
  fwrite(g->edgesV, 4, g->m, f);
  fwrite(g->offsetsU, 4, g->nu + 1, f);
  fwrite(g->edgesU, 4, g->m, f);
  int bad = ferror(f);
  fclose(f);
  return bad ? PBF_ERR_IO : PBF_OK;
}

// ------------------------------------------------------------------------------------------
// Seeded generators.  Draw k of stream s under seed is rng(seed, s, k) (a SplitMix64-style
// hash of a counter), so generation is order independent and reproducible.  The graph is the
// first m DISTINCT (v,u) pairs in draw order: draws are made in growing prefixes, each pair
// keeps the index of its first occurrence, and the m earliest first occurrences are kept.
// ------------------------------------------------------------------------------------------
namespace {
struct DrawKey {
  uint64_t key;
  uint64_t k;
};

template <class DrawF>
int first_m_distinct(int64_t nv, int64_t nu, int64_t m, double overdraw, DrawF draw, pbf_graph* out) {
  if (m < 0 || nv <= 0 || nu <= 0) return m == 0 ? csr_from_sorted_keys(nv, nu, {}, out) : PBF_ERR_ARG;
  if ((double)m > 0.75 * (double)nv * (double)nu) return PBF_ERR_CAPACITY;
  int64_t K = int64_t(double(m) * overdraw) + 1024;
  std::vector<DrawKey> d;
  for (int attempt = 0; attempt < 64; attempt++) {
    d.resize((size_t)K);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < K; k++) d[k] = DrawKey{draw((uint64_t)k), (uint64_t)k};
    __gnu_parallel::sort(d.begin(), d.end(), [](const DrawKey& a, const DrawKey& b) {
      return a.key < b.key || (a.key == b.key && a.k < b.k);
    });
    size_t w = 0;
    for (size_t i = 0; i < d.size(); i++)
      if (i == 0 || d[i].key != d[i - 1].key) d[w++] = d[i];
    d.resize(w);
    if ((int64_t)w >= m) {
      std::nth_element(d.begin(), d.begin() + m, d.end(),
                       [](const DrawKey& a, const DrawKey& b) { return a.k < b.k; });
      d.resize((size_t)m);
      std::vector<uint64_t> keys((size_t)m);
      for (int64_t i = 0; i < m; i++) keys[i] = d[i].key;
      __gnu_parallel::sort(keys.begin(), keys.end());
      return csr_from_sorted_keys(nv, nu, keys, out);
    }
    double ratio = double(m) / double(std::max<size_t>(w, 1));
    K = int64_t(double(K) * std::min(ratio * 1.1, 4.0)) + 1024;
  }
  return PBF_ERR_CAPACITY;
}

// Inverse-CDF sampler over cumulative weights.
struct CumSampler {
  std::vector<double> cw;
  double total = 0;
  explicit CumSampler(int64_t n, double gamma) : cw((size_t)n) {
    double ex = -1.0 / (gamma - 1.0), s = 0;
    for (int64_t i = 0; i < n; i++) { s += std::pow(double(i + 1), ex); cw[i] = s; }
    total = s;
  }
  uint32_t sample(uint64_t r) const {
    double x = pbf_u01(r) * total;
    size_t i = size_t(std::upper_bound(cw.begin(), cw.end(), x) - cw.begin());
    if (i >= cw.size()) i = cw.size() - 1;
    return (uint32_t)i;
  }
};
}  // namespace

extern "C" int pbf_gen_uniform(int64_t nv, int64_t nu, int64_t m, uint64_t seed, pbf_graph* out) {
  if (!out) return PBF_ERR_ARG;
  memset(out, 0, sizeof(*out));
  auto draw = [=](uint64_t k) -> uint64_t {
    uint64_t v = pbf_bounded(pbf_rng(seed, 0, k), (uint64_t)nv);
    uint64_t u = pbf_bounded(pbf_rng(seed, 1, k), (uint64_t)nu);
    return (v << 32) | u;
  };
  return first_m_distinct(nv, nu, m, 1.05, draw, out);
}

extern "C" int pbf_gen_chung_lu(int64_t nv, int64_t nu, int64_t m, double gamma_v, double gamma_u, uint64_t seed,
                                pbf_graph* out) {
  if (!out || !(gamma_v > 1.0) || !(gamma_u > 1.0)) return PBF_ERR_ARG;
  memset(out, 0, sizeof(*out));
  if (nv <= 0 || nu <= 0) return first_m_distinct(nv, nu, m, 1.0, [](uint64_t) { return uint64_t(0); }, out);
  CumSampler sv(nv, gamma_v), su(nu, gamma_u);
  auto draw = [&](uint64_t k) -> uint64_t {
    uint64_t v = sv.sample(pbf_rng(seed, 2, k));
    uint64_t u = su.sample(pbf_rng(seed, 3, k));
    return (v << 32) | u;
  };
  return first_m_distinct(nv, nu, m, 1.3, draw, out);
}

extern "C" int pbf_gen_rmat(int log_nv, int log_nu, int64_t m, double a, double b, double c, double d, uint64_t seed,
                            pbf_graph* out) {
  if (!out || log_nv < 0 || log_nu < 0 || log_nv > 30 || log_nu > 30) return PBF_ERR_ARG;
  memset(out, 0, sizeof(*out));
  double s = a + b + c + d;
  if (!(s > 0)) return PBF_ERR_ARG;
  a /= s; b /= s; c /= s; d /= s;
  int levels = std::max(log_nv, log_nu);
  auto draw = [=](uint64_t k) -> uint64_t {
    uint64_t v = 0, u = 0;
    for (int l = 0; l < levels; l++) {
      double x = pbf_u01(pbf_rng(seed, 4 + (uint64_t)l, k));
      bool vb = l < log_nv, ub = l < log_nu;
      int bv = 0, bu = 0;
      if (vb && ub) {
        if (x < a) { bv = 0; bu = 0; }
        else if (x < a + b) { bv = 0; bu = 1; }
        else if (x < a + b + c) { bv = 1; bu = 0; }
        else { bv = 1; bu = 1; }
      } else if (vb) {
        bv = x >= a + b;
      } else if (ub) {
        bu = x >= a + c;
      }
      if (vb) v = (v << 1) | (uint64_t)bv;
      if (ub) u = (u << 1) | (uint64_t)bu;
    }
    return (v << 32) | u;
  };
  return first_m_distinct(int64_t(1) << log_nv, int64_t(1) << log_nu, m, 1.1, draw, out);
}

extern "C" int pbf_batch_graph_params(uint64_t seed, int64_t* nv, int64_t* nu, int64_t* m, double* gamma) {
  if (!nv || !nu || !m || !gamma) return PBF_ERR_ARG;
  const double lo = std::log(64.0), hi = std::log(65536.0);
  const double dlo = std::log(2.0), dhi = std::log(32.0);
  *nv = (int64_t)std::llround(std::exp(lo + (hi - lo) * pbf_u01(pbf_rng(seed, 100, 0))));
  *nu = (int64_t)std::llround(std::exp(lo + (hi - lo) * pbf_u01(pbf_rng(seed, 100, 1))));
  double avg = std::exp(dlo + (dhi - dlo) * pbf_u01(pbf_rng(seed, 100, 2)));
  int64_t mm = (int64_t)std::llround(avg * double(*nv + *nu) / 2.0);
  int64_t cap = (*nv) * (*nu) / 2;
  *m = mm < cap ? mm : cap;
  *gamma = 2.0 + pbf_u01(pbf_rng(seed, 100, 3));
  return PBF_OK;
}

extern "C" int pbf_gen_batch_graph(uint64_t seed, pbf_graph* out) {
  int64_t nv, nu, m;
  double gamma;
  int st = pbf_batch_graph_params(seed, &nv, &nu, &m, &gamma);
  if (st) return st;
  return pbf_gen_chung_lu(nv, nu, m, gamma, gamma, seed, out);
}
