// Test infrastructure (tests/test_host_asan.py): CPU stand-in for the device entry points the
// CLI calls, for an AddressSanitizer run of the CLI's host-memory path (-hostRead): each stub
// validates and reads exactly the host bytes the library's host path reads (resolve_graph:
// pbf_validate + the four arrays), and writes exactly the bytes it writes (the output
// arrays), with results from the CPU oracle.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "parbutterfly.h"
extern "C" {
typedef struct { int64_t nv, nu, m; const uint32_t *offV, *edgV, *offU, *edgU; } orc_graph;
int orc_count_vert(const orc_graph* g, int use_v, int rank_type, uint64_t* out);
int orc_count_edge(const orc_graph* g, int use_v, int rank_type, uint64_t* out);
int orc_count_total(const orc_graph* g, int use_v, int rank_type, uint64_t* total);
void orc_peel_vert(const orc_graph* g, int use_v, const uint64_t* counts, uint64_t* tip);
void orc_peel_edge(const orc_graph* g, int use_v, const uint64_t* counts, uint64_t* wing);
int orc_sparsify(const orc_graph* g, int type, int64_t denom, int64_t seed, int64_t* nv_out, int64_t* nu_out,
                 int64_t* m_out, uint32_t* offV, uint32_t* edgV, uint32_t* offU, uint32_t* edgU);
}
struct pbf_context { std::vector<uint32_t> graph; };
static int resolve(pbf_context* c, const pbf_bipartite_csr* g, int mem, orc_graph* o) {
  if (mem != PBF_MEM_HOST) return PBF_ERR_UNSUPPORTED;
  int st = pbf_validate(g);
  if (st) return st;
  // the upload: every word of the four arrays, copied out of the caller's memory
  c->graph.assign(size_t(g->nv + 1) + size_t(g->m) * 2 + size_t(g->nu + 1), 0);
  uint32_t* d = c->graph.data();
  memcpy(d, g->offsetsV, 4 * size_t(g->nv + 1));
  memcpy(d + g->nv + 1, g->edgesV, 4 * size_t(g->m));
  memcpy(d + g->nv + 1 + g->m, g->offsetsU, 4 * size_t(g->nu + 1));
  memcpy(d + g->nv + 1 + g->m + g->nu + 1, g->edgesU, 4 * size_t(g->m));
  *o = orc_graph{g->nv, g->nu, g->m, d, d + g->nv + 1, d + g->nv + 1 + g->m, d + g->nv + 1 + g->m + g->nu + 1};
  return PBF_OK;
}
static int rk(int rank_type) { return rank_type == PBF_RANK_ACOCORE ? 2 : 4; }
extern "C" {
int pbf_create(pbf_context** out, int) { *out = new pbf_context(); return PBF_OK; }
int pbf_destroy(pbf_context* c) { delete c; return PBF_OK; }
int pbf_set_limits(pbf_context*, int64_t, int64_t) { return PBF_OK; }
int pbf_side(pbf_context* c, const pbf_bipartite_csr* g, int mem, int* use_v, int64_t* w) {
  orc_graph o;
  int st = resolve(c, g, mem, &o);
  return st ? st : pbf_side_host(g, use_v, w);
}
int pbf_count_vert(pbf_context* c, const pbf_bipartite_csr* g, int mem, int use_v, int, int rank_type, uint64_t* out) {
  orc_graph o;
  int st = resolve(c, g, mem, &o);
  if (st) return st;
  std::vector<uint64_t> r(size_t(use_v ? g->nu : g->nv) + 1);
  orc_count_vert(&o, use_v, rk(rank_type), r.data());
  memcpy(out, r.data(), 8 * size_t(use_v ? g->nu : g->nv));
  return PBF_OK;
}
int pbf_count_edge(pbf_context* c, const pbf_bipartite_csr* g, int mem, int use_v, int, int rank_type, uint64_t* out) {
  orc_graph o;
  int st = resolve(c, g, mem, &o);
  if (st) return st;
  std::vector<uint64_t> r(size_t(g->m) + 1);
  orc_count_edge(&o, use_v, rk(rank_type), r.data());
  memcpy(out, r.data(), 8 * size_t(g->m));
  return PBF_OK;
}
int pbf_count_total(pbf_context* c, const pbf_bipartite_csr* g, int mem, int use_v, int, int rank_type, uint64_t* total) {
  orc_graph o;
  int st = resolve(c, g, mem, &o);
  if (st) return st;
  uint64_t t = 0;
  if (g->m) orc_count_total(&o, use_v, rk(rank_type), &t);
  memcpy(total, &t, 8);
  return PBF_OK;
}
static int peel(pbf_context* c, const pbf_bipartite_csr* g, int mem, int use_v, const uint64_t* counts, uint64_t* out, bool e) {
  orc_graph o;
  int st = resolve(c, g, mem, &o);
  if (st) return st;
  const size_t n = e ? size_t(g->m) : size_t(use_v ? g->nu : g->nv);
  std::vector<uint64_t> in(counts, counts + n), r(n + 1);  // the counts' upload
  if (e) orc_peel_edge(&o, use_v, in.data(), r.data()); else orc_peel_vert(&o, use_v, in.data(), r.data());
  memcpy(out, r.data(), 8 * n);
  return PBF_OK;
}
int pbf_peel_vert(pbf_context* c, const pbf_bipartite_csr* g, int mem, int use_v, int, const uint64_t* cnt, uint64_t* tip) {
  return peel(c, g, mem, use_v, cnt, tip, false);
}
int pbf_peel_edge(pbf_context* c, const pbf_bipartite_csr* g, int mem, int use_v, int, const uint64_t* cnt, uint64_t* wing) {
  return peel(c, g, mem, use_v, cnt, wing, true);
}
int pbf_sparsify(pbf_context* c, const pbf_bipartite_csr* g, int mem, int type, int64_t denom, int64_t seed, int64_t* nv,
                 int64_t* nu, int64_t* m, uint32_t* oV, uint32_t* eV, uint32_t* oU, uint32_t* eU) {
  orc_graph o;
  int st = resolve(c, g, mem, &o);
  if (st) return st;
  // the library writes the sparsified arrays' own lengths into the caller's arrays
  std::vector<uint32_t> a(g->nv + 1), b(g->m + 1), x(g->nu + 1), y(g->m + 1);
  st = orc_sparsify(&o, type, denom, seed, nv, nu, m, a.data(), b.data(), x.data(), y.data());
  if (st) return PBF_ERR_ARG;
  memcpy(oV, a.data(), 4 * size_t(*nv + 1)); memcpy(eV, b.data(), 4 * size_t(*m));
  memcpy(oU, x.data(), 4 * size_t(*nu + 1)); memcpy(eU, y.data(), 4 * size_t(*m));
  return PBF_OK;
}
int pbf_device_alloc(pbf_context*, size_t, void**) { return PBF_ERR_UNSUPPORTED; }
int pbf_device_free(pbf_context*, void*) { return PBF_OK; }
int pbf_copy_to_host(pbf_context*, void*, const void*, size_t) { return PBF_ERR_UNSUPPORTED; }
int pbf_read_bipartite_device(pbf_context*, const char*, pbf_bipartite_csr*, void**) { return PBF_ERR_UNSUPPORTED; }
}
