// bipartite_butterflies: the command-line driver, with the flags and output lines of the
// reference's apps/hyper/BipartiteButterflies.C (Compute :75-277, parallel_main :283-355),
// running every count / peel on the GPU through libparbutterfly_amd's C ABI.
//
//   bipartite_butterflies [-countType T | -t N] [-rankType R | -w N] [-peelType P | -tp N]
//                         [-per VERT|EDGE|TOTAL | -e N | -total] [-r rounds] [-m max_wedges]
//                         [-a max_array_size] [-nopeel] [-sparseType EDGE|COLOR | -s N] [-d denom]
//                         [-device D] [-hostRead] <inFile>
//
// Count types are the reference's (ASORT ... SERIAL, numeric aliases of :96-107) plus GPU;
// they are aggregation strategies with identical results, and all of them run the GPU kernels
// here.  Rank types ADEG, ACOCORE and SIDE are supported on the GPU; DEG / COCORE return
// PBF_ERR_UNSUPPORTED.  The graph is re-read every round, as the reference does (:331-341),
// straight into HBM (pbf_read_bipartite_device; -hostRead reads it into host memory instead);
// with -sparseType it is sparsified (seed 0) on the GPU after each read (:332-337), the count
// becomes TOTAL and the printed total is scaled by denom^3 (COLOR) or denom^4 (EDGE)
// (:193, :213-214).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "parbutterfly.h"

namespace {
struct Args {
  int argc;
  char** argv;
  bool has(const char* opt) const {
    for (int i = 1; i < argc - 1; i++) if (!strcmp(argv[i], opt)) return true;
    return false;
  }
  std::string str(const char* opt, const std::string& def) const {
    for (int i = 1; i < argc - 1; i++) if (!strcmp(argv[i], opt)) return argv[i + 1];
    return def;
  }
  long num(const char* opt, long def) const {
    for (int i = 1; i < argc - 1; i++) if (!strcmp(argv[i], opt)) return atol(argv[i + 1]);
    return def;
  }
};

const long NONE = -0x7fffffffffffffffL;

// CountTypeToStr (BipartiteButterflies.C:41-55), plus the GPU mode.
const char* count_label(int t) {
  switch (t) {
    case PBF_COUNT_ASORT: return "Sort";
    case PBF_COUNT_SORT: return "SortCE";
    case PBF_COUNT_AHASH: return "Hash";
    case PBF_COUNT_HASH: return "HashCE";
    case PBF_COUNT_AHIST: return "Hist";
    case PBF_COUNT_HIST: return "HistCE";
    case PBF_COUNT_BATCHS: return "Par";
    case PBF_COUNT_SERIAL: return "Serial";
    case PBF_COUNT_BATCHWA: return "WedgePar";
    case PBF_COUNT_GPU: return "GPU";
  }
  return "";
}
// The timing line of a count or a peel: every type runs the GPU kernels, so the line says so
// and names the requested CPU type beside it (a script keyed on "Par :" does not take GPU times
// for the reference's).
void print_time(const char* gpu_label, const char* requested, int is_gpu, double sec) {
  if (is_gpu) printf("%s : %.3g\n", gpu_label, sec);
  else printf("%s (requested %s) : %.3g\n", gpu_label, requested, sec);
}

const char* peel_label(int t) {
  switch (t) {
    case PBF_PEEL_SORT: return "SortCE Peel";
    case PBF_PEEL_HASH: return "HashCE Peel";
    case PBF_PEEL_HIST: return "HistCE Peel";
    case PBF_PEEL_BATCHS: return "Par Peel";
    case PBF_PEEL_BATCHWA: return "WedgePar Peel";
    case PBF_PEEL_GPU: return "GPU Peel";
  }
  return "";
}

// SparseType from -s / -sparseType (:177-192).
int sparse_type(const Args& P) {
  const long s = P.num("-s", NONE);
  if (s == NONE) {
    const std::string t = P.str("-sparseType", "");
    return t == "EDGE" ? PBF_SPARSE_EDGE : t == "COLOR" ? PBF_SPARSE_COLOR : PBF_SPARSE_NONE;
  }
  return s == 1 ? PBF_SPARSE_COLOR : s == 2 ? PBF_SPARSE_EDGE : PBF_SPARSE_NONE;
}

// num_butterflies *= pow(denom, 3 or 4) on a long (:213-214).
int scale_total(uint64_t total, int sparse, long denom, int64_t* out) {
  if (sparse == PBF_SPARSE_NONE) { *out = (int64_t)total; return 0; }
  const double s = (double)(int64_t)total * std::pow((double)denom, sparse == PBF_SPARSE_COLOR ? 3.0 : 4.0);
  if (!(s < 9223372036854775808.0)) return PBF_ERR_DOMAIN;
  *out = (int64_t)s;
  return 0;
}

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

int fail(const char* what, int st) {
  fprintf(stderr, "%s: %s (%d)\n", what, pbf_status_str(st), st);
  return st ? st : 1;
}

// A count / peel output in the graph's memory space (mem), readable on the host after fetch().
struct OutBuf {
  pbf_context* ctx;
  int mem;
  size_t n;
  std::vector<uint64_t> host;
  void* dev = nullptr;
  OutBuf(pbf_context* c, int m, size_t count) : ctx(c), mem(m), n(count), host(count ? count : 1) {}
  int init() { return mem == PBF_MEM_DEVICE ? pbf_device_alloc(ctx, 8 * (n ? n : 1), &dev) : PBF_OK; }
  uint64_t* ptr() { return mem == PBF_MEM_DEVICE ? static_cast<uint64_t*>(dev) : host.data(); }
  int fetch() { return mem == PBF_MEM_DEVICE ? pbf_copy_to_host(ctx, host.data(), dev, 8 * (n ? n : 1)) : PBF_OK; }
  ~OutBuf() { if (dev) pbf_device_free(ctx, dev); }
};

int compute(pbf_context* ctx, const pbf_bipartite_csr& g, int mem, const Args& P) {
  // count type (:79-109)
  long tl = P.num("-t", NONE);
  std::string ts = P.str("-countType", "");
  int ty = PBF_COUNT_GPU;
  if (tl == NONE) {
    if (ts == "ASORT") ty = PBF_COUNT_ASORT; else if (ts == "SORT") ty = PBF_COUNT_SORT;
    else if (ts == "AHASH") ty = PBF_COUNT_AHASH; else if (ts == "HASH") ty = PBF_COUNT_HASH;
    else if (ts == "AHIST") ty = PBF_COUNT_AHIST; else if (ts == "HIST") ty = PBF_COUNT_HIST;
    else if (ts == "BATCHS") ty = PBF_COUNT_BATCHS; else if (ts == "BATCHWA") ty = PBF_COUNT_BATCHWA;
    else if (ts == "SERIAL") ty = PBF_COUNT_SERIAL; else ty = PBF_COUNT_GPU;
  } else {
    switch (tl) {
      case 0: ty = PBF_COUNT_ASORT; break; case 1: ty = PBF_COUNT_SORT; break;
      case 2: ty = PBF_COUNT_AHASH; break; case 3: ty = PBF_COUNT_HASH; break;
      case 4: ty = PBF_COUNT_AHIST; break; case 6: ty = PBF_COUNT_HIST; break;
      case 5: case 7: case 11: ty = PBF_COUNT_BATCHS; break;
      case 9: case 12: ty = PBF_COUNT_SERIAL; break;
      case 8: ty = PBF_COUNT_BATCHWA; break;
      default: ty = PBF_COUNT_GPU; break;
    }
  }
  // rank type (:111-129)
  long wl = P.num("-w", NONE);
  std::string ws = P.str("-rankType", "");
  int tw = PBF_RANK_ADEG;
  if (wl == NONE) {
    if (ws == "SIDE") tw = PBF_RANK_SIDE; else if (ws == "COCORE") tw = PBF_RANK_COCORE;
    else if (ws == "ACOCORE") tw = PBF_RANK_ACOCORE; else if (ws == "DEG") tw = PBF_RANK_DEG;
    else tw = PBF_RANK_ADEG;
  } else if (wl >= 0 && wl <= 4) {
    tw = (int)wl;
  }
  // peel type (:131-153)
  long pl = P.num("-tp", NONE);
  std::string ps = P.str("-peelType", "");
  int tp = PBF_PEEL_GPU;
  bool nopeel = false;
  if (pl == NONE) {
    if (ps == "SORT") tp = PBF_PEEL_SORT; else if (ps == "HASH") tp = PBF_PEEL_HASH;
    else if (ps == "HIST") tp = PBF_PEEL_HIST; else if (ps == "BATCHS") tp = PBF_PEEL_BATCHS;
    else if (ps == "BATCHWA") tp = PBF_PEEL_BATCHWA; else if (ps == "GPU") tp = PBF_PEEL_GPU;
    else nopeel = true;
  } else {
    nopeel = P.has("-nopeel");
    switch (pl) {
      case 0: tp = PBF_PEEL_HASH; break; case 1: tp = PBF_PEEL_SORT; break; case 2: tp = PBF_PEEL_HIST; break;
      case 3: case 4: tp = PBF_PEEL_BATCHS; break; case 5: tp = PBF_PEEL_BATCHWA; break;
      case 6: tp = PBF_PEEL_GPU; break;
      default: nopeel = true; break;
    }
  }
  // per type (:155-168)
  std::string pers = P.str("-per", "");
  long te = P.num("-e", NONE);
  int per = PBF_PER_VERT;
  if (te == NONE) {
    if (pers == "EDGE") per = PBF_PER_EDGE; else if (pers == "TOTAL") per = PBF_PER_TOTAL;
  } else {
    per = P.has("-total") ? PBF_PER_TOTAL : (te == 0 ? PBF_PER_VERT : PBF_PER_EDGE);
  }
  const int sparse = sparse_type(P);
  const long denom = P.num("-d", 25);
  if (sparse != PBF_SPARSE_NONE) per = PBF_PER_TOTAL;
  if (per == PBF_PER_TOTAL) nopeel = true;

  double t0 = now();
  int use_v = 0;
  int64_t nw = 0;
  int st = pbf_side(ctx, &g, mem, &use_v, &nw);
  if (st) return fail("pbf_side", st);
  printf("preprocess (wedge counts) : %.3g\n", now() - t0);

  if (per == PBF_PER_TOTAL) {
    OutBuf t(ctx, mem, 1);
    if ((st = t.init())) return fail("pbf_device_alloc", st);
    t0 = now();
    st = pbf_count_total(ctx, &g, mem, use_v, ty, tw, t.ptr());
    if (st) return fail("pbf_count_total", st);
    print_time("GPU", count_label(ty), ty == PBF_COUNT_GPU, now() - t0);
    if ((st = t.fetch())) return fail("pbf_copy_to_host", st);
    const uint64_t total = t.host[0];
    int64_t scaled = 0;
    st = scale_total(total, sparse, denom, &scaled);
    if (st) return fail("scaling", st);
    printf("number of butterflies: %lld\n", (long long)scaled);
    return 0;
  }
  if (per == PBF_PER_VERT) {
    const int64_t ns = use_v ? g.nu : g.nv;
    OutBuf b(ctx, mem, (size_t)ns);
    if ((st = b.init())) return fail("pbf_device_alloc", st);
    t0 = now();
    st = pbf_count_vert(ctx, &g, mem, use_v, ty, tw, b.ptr());
    if (st) return fail("pbf_count_vert", st);
    print_time("GPU", count_label(ty), ty == PBF_COUNT_GPU, now() - t0);
    if ((st = b.fetch())) return fail("pbf_copy_to_host", st);
    unsigned long long sum = 0;
    for (int64_t i = 0; i < ns; i++) sum += b.host[i];
    printf("number of butterflies: %llu\n", sum / 2);
    if (!nopeel) {
      OutBuf tip(ctx, mem, (size_t)ns);
      if ((st = tip.init())) return fail("pbf_device_alloc", st);
      t0 = now();
      st = pbf_peel_vert(ctx, &g, mem, use_v, tp, b.ptr(), tip.ptr());
      if (st) return fail("pbf_peel_vert", st);
      print_time("GPU Peel", peel_label(tp), tp == PBF_PEEL_GPU, now() - t0);
    }
    return 0;
  }
  OutBuf e(ctx, mem, (size_t)g.m);
  if ((st = e.init())) return fail("pbf_device_alloc", st);
  t0 = now();
  st = pbf_count_edge(ctx, &g, mem, use_v, ty, tw, e.ptr());
  if (st) return fail("pbf_count_edge", st);
  print_time("GPU", count_label(ty), ty == PBF_COUNT_GPU, now() - t0);
  if ((st = e.fetch())) return fail("pbf_copy_to_host", st);
  unsigned long long sum = 0;
  for (int64_t i = 0; i < g.m; i++) sum += e.host[i];
  printf("number of edge butterflies: %llu\n", sum / 4);
  if (!nopeel) {
    OutBuf wing(ctx, mem, (size_t)g.m);
    if ((st = wing.init())) return fail("pbf_device_alloc", st);
    t0 = now();
    st = pbf_peel_edge(ctx, &g, mem, use_v, tp, e.ptr(), wing.ptr());
    if (st) return fail("pbf_peel_edge", st);
    print_time("GPU Peel", peel_label(tp), tp == PBF_PEEL_GPU, now() - t0);
  }
  return 0;
}
}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s [options] <inFile>\n", argv[0]);
    return 2;
  }
  Args P{argc, argv};
  const char* file = argv[argc - 1];
  long rounds = P.num("-r", 3);
  pbf_context* ctx = nullptr;
  int st = pbf_create(&ctx, (int)P.num("-device", 0));
  if (st) return fail("pbf_create", st);
  // -m / -a (:174-175) bound the counting passes only when given: the reference's defaults
  // exist to bound its CPU scratch, which the GPU path does not have
  if (P.has("-m") || P.has("-a")) {
    st = pbf_set_limits(ctx, P.num("-m", 0), P.num("-a", 0));
    if (st) { pbf_destroy(ctx); return fail("pbf_set_limits", st); }
  }
  int rc = 0;
  const bool host_read = P.has("-hostRead");
  const int mem = host_read ? PBF_MEM_HOST : PBF_MEM_DEVICE;
  const int sparse = sparse_type(P);
  for (long r = 0; r < rounds + 1 && rc == 0; r++) {
    // the graph: in host memory (pbf_read_bipartite) or read straight into HBM
    pbf_graph hg{};
    pbf_bipartite_csr g{};
    void* blk = nullptr;
    st = host_read ? pbf_read_bipartite(file, &hg) : pbf_read_bipartite_device(ctx, file, &g, &blk);
    if (st) { rc = fail("readBipartite", st); break; }
    if (host_read) g = hg.csr;
    if (sparse != PBF_SPARSE_NONE) {
      // sparsified with seed 0 after each read (:332-337), into arrays of the input's sizes
      const size_t words = size_t(g.nv + 1) + size_t(g.m) + size_t(g.nu + 1) + size_t(g.m);
      pbf_graph hs{};
      void* bs = nullptr;
      st = host_read ? pbf_graph_alloc(g.nv, g.nu, g.m, &hs) : pbf_device_alloc(ctx, 4 * words, &bs);
      pbf_bipartite_csr s{};
      if (!st) {
        uint32_t* w = host_read ? const_cast<uint32_t*>(hs.csr.offsetsV) : static_cast<uint32_t*>(bs);
        uint32_t *oV = w, *eV = oV + g.nv + 1, *oU = eV + g.m, *eU = oU + g.nu + 1;
        st = pbf_sparsify(ctx, &g, mem, sparse, P.num("-d", 25), 0, &s.nv, &s.nu, &s.m, oV, eV, oU, eU);
        s.offsetsV = oV; s.edgesV = eV; s.offsetsU = oU; s.edgesU = eU;  // prefixes of those arrays
      }
      if (host_read) pbf_graph_free(&hg); else pbf_device_free(ctx, blk);
      hg = hs;
      blk = bs;
      g = s;
      if (st) {
        if (host_read) pbf_graph_free(&hg); else pbf_device_free(ctx, blk);
        rc = fail("sparsify", st);
        break;
      }
    }
    rc = compute(ctx, g, mem, P);
    if (host_read) pbf_graph_free(&hg); else pbf_device_free(ctx, blk);
  }
  pbf_destroy(ctx);
  return rc;
}
