/*
 * CPU oracle for butterfly counting and peeling -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of ParButterfly's SERIAL ranked path and of its bucketed
 * peeling semantics.  It exists to check the HIP implementation; nothing in the product
 * (the parbutterfly_amd package, its shared library or its CLI) may link, load or call it.
 * Only tests/, the smoke check and the CPU-baseline leg of bench.py use it, as a checker.
 *
 * Every function names the reference code it restates (paths relative to the reference
 * repository root).  Pinning (how each function is checked independently of its own text):
 *   orc_side            - closed form on K_{a,b}; golden vectors from the compiled reference.
 *   orc_rank_adeg/acocore, orc_rank_wedges
 *                       - W under the reference's own order (dumped by oracle/_ref) must equal
 *                         the reference's num_ccwedges; the key must be non-increasing along
 *                         the reference's order.
 *   orc_count_*         - brute force (orc_brute_*), closed forms, golden vectors from the
 *                         compiled reference's SERIAL and BATCHS modes.
 *   orc_peel_vert/edge  - textbook one-at-a-time peeling (orc_brute_peel_*), K_{a,b} closed
 *                         forms, golden vectors from the compiled reference's PBATCHS mode.
 *   orc_sparsify        - the graphs the compiled reference's eSparseBipartite /
 *                         clrSparseBipartite (+ delZeroDeg) write for 40 (graph, sparsifier,
 *                         denom, seed) cases, array for array; one colour keeps every edge;
 *                         an independent numpy restatement of the colour rule.
 *
 * Graph convention (as the reference's bipartiteCSR, butterfly_utils.h:184-197): V has nv
 * vertices, U has nu; offsetsV[nv+1]/edgesV[m] list U-neighbours of V vertices, offsetsU[nu+1]/
 * edgesU[m] list V-neighbours of U vertices; adjacency lists sorted ascending, duplicate free.
 * In the combined numbering used for ranking, V vertex v is id v and U vertex u is id nv+u.
 * "use_v" (cmpWedgeCounts) true means counts live on U (S = U, O = V); false means S = V.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int64_t nv, nu, m;
  const uint32_t *offV, *edgV, *offU, *edgU;
} orc_graph;

static inline int64_t deg_of(const orc_graph* g, int64_t x) {
  return x < g->nv ? (int64_t)g->offV[x + 1] - g->offV[x]
                   : (int64_t)g->offU[x - g->nv + 1] - g->offU[x - g->nv];
}
static inline const uint32_t* adj_of(const orc_graph* g, int64_t x) {
  return x < g->nv ? g->edgV + g->offV[x] : g->edgU + g->offU[x - g->nv];
}
/* combined id of the j-th neighbour of x */
static inline int64_t nbr_id(const orc_graph* g, int64_t x, int64_t j) {
  return x < g->nv ? (int64_t)g->nv + adj_of(g, x)[j] : (int64_t)adj_of(g, x)[j];
}
/* floor(log2(d)) + 1 for d > 0, 0 for d == 0 (butterfly_utils.h:279-280) */
static inline uint32_t log_key(int64_t d) {
  uint32_t k = 0;
  while (d > 0) { k++; d >>= 1; }
  return k;
}

/* wedgeF (butterfly_utils.h:1061-1069): C(d,2) with the product d * (d - 1) taken in the
 * 32-bit degree type, so it wraps modulo 2^32 for d > 65536 before the halving. */
static inline int64_t side_term(uint32_t d) { return (int64_t)((uint32_t)(d * (d - 1u)) / 2u); }

/* cmpWedgeCounts (butterfly_utils.h:1080-1086): wV = sum_V wedgeF, wU = sum_U wedgeF (64-bit
 * sums of the wrapped terms); use_v = (wV <= wU); the returned wedge count is the smaller one. */
void orc_side(const orc_graph* g, int* use_v, int64_t* wedges) {
  int64_t wv = 0, wu = 0;
  for (int64_t v = 0; v < g->nv; v++) wv += side_term(g->offV[v + 1] - g->offV[v]);
  for (int64_t u = 0; u < g->nu; u++) wu += side_term(g->offU[u + 1] - g->offU[u]);
  *use_v = (wv <= wu);
  *wedges = (wv <= wu) ? wv : wu;
}

/* Stable counting sort of ids by key, descending key; ties by ascending id.  The reference
 * sorts with sampleSort (getRanks, butterfly_utils.h:249-265), whose tie order is unspecified;
 * counts and peel numbers do not depend on it. order[r] = id of rank r; rank_of[id] = r. */
static void order_by_key_desc(int64_t n, const uint32_t* key, uint32_t* order, uint32_t* rank_of) {
  uint32_t maxk = 0;
  for (int64_t i = 0; i < n; i++) if (key[i] > maxk) maxk = key[i];
  int64_t* start = (int64_t*)calloc((size_t)maxk + 2, sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) start[maxk - key[i] + 1]++;
  for (uint32_t b = 0; b <= maxk; b++) start[b + 1] += start[b];
  for (int64_t i = 0; i < n; i++) {
    int64_t r = start[maxk - key[i]]++;
    order[r] = (uint32_t)i;
    rank_of[i] = (uint32_t)r;
  }
  free(start);
}

/* ADEG: getApproxDegRanks (butterfly_utils.h:275-284) -> getRanks (:249-265). */
void orc_rank_adeg(const orc_graph* g, uint32_t* order, uint32_t* rank_of) {
  int64_t n = g->nv + g->nu;
  uint32_t* key = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
  for (int64_t x = 0; x < n; x++) key[x] = log_key(deg_of(g, x));
  order_by_key_desc(n, key, order, rank_of);
  free(key);
}

/* ACOCORE: getApproxCoCoreRanks (butterfly_utils.h:448-503).  Buckets of log-degree are
 * processed in decreasing order (make_buckets(n, D, decreasing), :464); each round takes every
 * unprocessed vertex whose D equals the current maximum k, removes it, and for each neighbour
 * y still below k subtracts the number of removed neighbours from y's remaining degree and sets
 * D[y] = log_key(remaining) (:470-489; the min with k there never binds since degrees only
 * fall).  The final order sorts by D descending (:499-502).  The D values are written to
 * core_out when it is non-null. */
void orc_rank_acocore(const orc_graph* g, uint32_t* order, uint32_t* rank_of, uint32_t* core_out) {
  int64_t n = g->nv + g->nu;
  size_t nn = (size_t)(n ? n : 1);
  uint32_t* D = (uint32_t*)malloc(sizeof(uint32_t) * nn);
  int64_t* Dact = (int64_t*)malloc(sizeof(int64_t) * nn);
  uint8_t* done = (uint8_t*)calloc(nn, 1);
  int64_t* removed = (int64_t*)calloc(nn, sizeof(int64_t));
  int64_t* active = (int64_t*)malloc(sizeof(int64_t) * nn);
  int64_t* touched = (int64_t*)malloc(sizeof(int64_t) * nn);
  for (int64_t x = 0; x < n; x++) { Dact[x] = deg_of(g, x); D[x] = log_key(Dact[x]); }
  int64_t finished = 0;
  while (finished < n) {
    uint32_t k = 0;
    int first = 1;
    for (int64_t x = 0; x < n; x++)
      if (!done[x] && (first || D[x] > k)) { k = D[x]; first = 0; }
    int64_t na = 0;
    for (int64_t x = 0; x < n; x++) if (!done[x] && D[x] == k) active[na++] = x;
    finished += na;
    int64_t nt = 0;
    for (int64_t a = 0; a < na; a++) {
      int64_t x = active[a];
      done[x] = 1;
      int64_t d = deg_of(g, x);
      for (int64_t j = 0; j < d; j++) {
        int64_t y = nbr_id(g, x, j);
        if (removed[y]++ == 0) touched[nt++] = y;
      }
    }
    for (int64_t t = 0; t < nt; t++) {
      int64_t y = touched[t];
      if (D[y] < k) {
        Dact[y] -= removed[y];
        uint32_t nl = log_key(Dact[y]);
        D[y] = nl < k ? nl : k;
      }
      removed[y] = 0;
    }
  }
  order_by_key_desc(n, D, order, rank_of);
  if (core_out) memcpy(core_out, D, sizeof(uint32_t) * (size_t)n);
  free(D); free(Dact); free(done); free(removed); free(active); free(touched);
}

/* rankWedgeF (butterfly_utils.h:216-237), summed over U and V as in Count
 * (butterfly_count.h:1654-1658): W = sum_x in(x)*out(x) + C(in(x),2), in(x) = neighbours of
 * lower rank.  Computed in 64 bits (the reference multiplies 32-bit intT values). */
int64_t orc_rank_wedges(const orc_graph* g, const uint32_t* rank_of) {
  int64_t n = g->nv + g->nu, W = 0;
  for (int64_t x = 0; x < n; x++) {
    int64_t d = deg_of(g, x), in = 0;
    for (int64_t j = 0; j < d; j++) if (rank_of[nbr_id(g, x, j)] < rank_of[x]) in++;
    int64_t out = d - in;
    W += in * out + in * (in - 1) / 2;
  }
  return W;
}

/* rankGraph / rankGraphEdges (butterfly_utils.h:579-623, 646-700): vertex r of the ranked
 * graph is order[r]; its list holds (rank_of[nbr] << 1) | flag, flag = 1 iff the neighbour is
 * on the count side S, sorted ascending (INVERSE).  pos[p] (optional) is the position of slot p
 * in the owner's original CSR (edgesV for V owners, edgesU for U owners). */
void orc_rank_graph(const orc_graph* g, int use_v, const uint32_t* order, const uint32_t* rank_of,
                    uint32_t* roff, uint32_t* redg, uint32_t* pos) {
  int64_t n = g->nv + g->nu;
  roff[0] = 0;
  for (int64_t r = 0; r < n; r++) roff[r + 1] = roff[r] + (uint32_t)deg_of(g, order[r]);
  for (int64_t r = 0; r < n; r++) {
    int64_t x = order[r], d = deg_of(g, x);
    uint32_t base = roff[r];
    uint32_t obase = x < g->nv ? g->offV[x] : g->offU[x - g->nv];
    int nbr_in_u = x < g->nv;          /* neighbours of a V vertex are U vertices */
    uint32_t flag = use_v ? (uint32_t)nbr_in_u : (uint32_t)!nbr_in_u;
    for (int64_t j = 0; j < d; j++) {
      uint32_t e = (rank_of[nbr_id(g, x, j)] << 1) | flag;
      /* insertion sort keeps the list ascending (lists are short in oracle-sized inputs) */
      int64_t t = j;
      while (t > 0 && redg[base + t - 1] > e) {
        redg[base + t] = redg[base + t - 1];
        if (pos) pos[base + t] = pos[base + t - 1];
        t--;
      }
      redg[base + t] = e;
      if (pos) pos[base + t] = obase + (uint32_t)j;
    }
  }
}

static int rank_any(const orc_graph* g, int rank_type, uint32_t* order, uint32_t* rank_of) {
  if (rank_type == 2) { orc_rank_acocore(g, order, rank_of, NULL); return 0; }  /* ACOCORE */
  if (rank_type == 4) { orc_rank_adeg(g, order, rank_of); return 0; }           /* ADEG */
  return -1;
}

typedef struct {
  int64_t n;
  uint32_t *order, *rank_of, *roff, *redg, *pos;
} ranked;

static int make_ranked(const orc_graph* g, int use_v, int rank_type, int want_pos, ranked* R) {
  int64_t n = g->nv + g->nu;
  size_t nn = (size_t)(n ? n : 1), mm = (size_t)(2 * g->m ? 2 * g->m : 1);
  R->n = n;
  R->order = (uint32_t*)malloc(4 * nn);
  R->rank_of = (uint32_t*)malloc(4 * nn);
  R->roff = (uint32_t*)malloc(4 * (nn + 1));
  R->redg = (uint32_t*)malloc(4 * mm);
  R->pos = want_pos ? (uint32_t*)malloc(4 * mm) : NULL;
  if (rank_any(g, rank_type, R->order, R->rank_of)) return -1;
  orc_rank_graph(g, use_v, R->order, R->rank_of, R->roff, R->redg, R->pos);
  return 0;
}
static void free_ranked(ranked* R) {
  free(R->order); free(R->rank_of); free(R->roff); free(R->redg); free(R->pos);
}

/* CountWorkEfficientSerial (butterfly_count.h:1402-1497) on the ranked graph, then the
 * rank -> original conversion of CountRank (:1600-1610).  out[x] for x in S (original id
 * within S).  rank_type: 2 = ACOCORE, 4 = ADEG (reference enum values). */
int orc_count_vert(const orc_graph* g, int use_v, int rank_type, uint64_t* out) {
  ranked R;
  if (make_ranked(g, use_v, rank_type, 0, &R)) return -1;
  int64_t n = R.n;
  size_t nn = (size_t)(n ? n : 1);
  uint64_t* B = (uint64_t*)calloc(nn, 8);
  uint32_t* wedges = (uint32_t*)calloc(nn, 4);
  uint32_t* used = (uint32_t*)malloc(4 * nn);
  for (int64_t i = 0; i < n; i++) {
    int64_t nused = 0;
    uint32_t io = R.roff[i], id = R.roff[i + 1] - io;
    for (uint32_t j = 0; j < id; j++) {
      uint32_t v = R.redg[io + j] >> 1, vo = R.roff[v], vd = R.roff[v + 1] - vo;
      for (uint32_t k = 0; k < vd; k++) {
        uint32_t e = R.redg[vo + k], u2 = e >> 1;
        if (!(u2 < i && u2 < v)) break;
        if (e & 1) { B[i] += wedges[u2]; B[u2] += wedges[u2]; }  /* running count sums C(c,2) */
        if (wedges[u2]++ == 0) used[nused++] = u2;
      }
    }
    for (uint32_t j = 0; j < id; j++) {
      uint32_t ev = R.redg[io + j], v = ev >> 1, vo = R.roff[v], vd = R.roff[v + 1] - vo;
      if (!(ev & 1)) continue;  /* centre credited only when it is on S */
      for (uint32_t k = 0; k < vd; k++) {
        uint32_t u2 = R.redg[vo + k] >> 1;
        if (!(u2 < i && u2 < v)) break;
        if (wedges[u2] > 1) B[v] += wedges[u2] - 1;
      }
    }
    for (int64_t t = 0; t < nused; t++) wedges[used[t]] = 0;
  }
  /* convert: out[x] = B[rank(x)] for x in S */
  int64_t ns = use_v ? g->nu : g->nv, base = use_v ? g->nv : 0;
  for (int64_t x = 0; x < ns; x++) out[x] = B[R.rank_of[base + x]];
  free(B); free(wedges); free(used); free_ranked(&R);
  return 0;
}

/* edgeToIdxs (butterfly_utils.h:2034-2062): for each position p of S's CSR, the position of
 * the same edge in O's CSR (binary search of the S vertex in the O vertex's list). */
void orc_edge_to_idxs(const orc_graph* g, int use_v, uint32_t* eti) {
  int64_t ns = use_v ? g->nu : g->nv;
  const uint32_t* offS = use_v ? g->offU : g->offV;
  const uint32_t* edgS = use_v ? g->edgU : g->edgV;
  const uint32_t* offO = use_v ? g->offV : g->offU;
  const uint32_t* edgO = use_v ? g->edgV : g->edgU;
  for (int64_t s = 0; s < ns; s++) {
    for (uint32_t p = offS[s]; p < offS[s + 1]; p++) {
      uint32_t o = edgS[p], lo = offO[o], hi = offO[o + 1];
      while (lo < hi) { uint32_t mid = lo + (hi - lo) / 2; if (edgO[mid] < (uint32_t)s) lo = mid + 1; else hi = mid; }
      eti[p] = lo;
    }
  }
}

/* CountEWorkEfficientSerial (butterfly_ecount.h:1438-1529) on the ranked graph (with the
 * original-position payload of rankGraphEdges), then CountERank's conversion (:1723-1741):
 * a slot whose flag is 1 belongs to an O vertex and its payload is already an O-CSR position;
 * otherwise the payload is an S-CSR position and eti maps it.  out has m entries indexed by
 * position in O's CSR (edgesV if use_v, else edgesU). */
int orc_count_edge(const orc_graph* g, int use_v, int rank_type, uint64_t* out) {
  ranked R;
  if (make_ranked(g, use_v, rank_type, 1, &R)) return -1;
  int64_t n = R.n, m2 = 2 * g->m;
  size_t nn = (size_t)(n ? n : 1);
  uint64_t* E = (uint64_t*)calloc((size_t)(m2 ? m2 : 1), 8);
  uint32_t* wedges = (uint32_t*)calloc(nn, 4);
  uint32_t* used = (uint32_t*)malloc(4 * nn);
  for (int64_t i = 0; i < n; i++) {
    int64_t nused = 0;
    uint32_t io = R.roff[i], id = R.roff[i + 1] - io;
    for (uint32_t j = 0; j < id; j++) {
      uint32_t v = R.redg[io + j] >> 1, vo = R.roff[v], vd = R.roff[v + 1] - vo;
      for (uint32_t k = 0; k < vd; k++) {
        uint32_t u2 = R.redg[vo + k] >> 1;
        if (!(u2 < i && u2 < v)) break;
        if (wedges[u2]++ == 0) used[nused++] = u2;
      }
    }
    for (uint32_t j = 0; j < id; j++) {
      uint32_t v = R.redg[io + j] >> 1, vo = R.roff[v], vd = R.roff[v + 1] - vo;
      for (uint32_t k = 0; k < vd; k++) {
        uint32_t u2 = R.redg[vo + k] >> 1;
        if (!(u2 < i && u2 < v)) break;
        if (wedges[u2] > 1) { E[vo + k] += wedges[u2] - 1; E[io + j] += wedges[u2] - 1; }
      }
    }
    for (int64_t t = 0; t < nused; t++) wedges[used[t]] = 0;
  }
  uint32_t* eti = (uint32_t*)malloc(4 * (size_t)(g->m ? g->m : 1));
  orc_edge_to_idxs(g, use_v, eti);
  memset(out, 0, 8 * (size_t)g->m);
  for (int64_t p = 0; p < m2; p++) {
    if (R.redg[p] & 1) out[R.pos[p]] += E[p];
    else out[eti[R.pos[p]]] += E[p];
  }
  free(eti); free(E); free(wedges); free(used); free_ranked(&R);
  return 0;
}

/* CountWorkEfficientSerialTotal (butterfly_count_total.h:833-906) via CountRankTotal
 * (:929-997): the running count nb += wedges[u2] over every rank-filtered wedge. */
int orc_count_total(const orc_graph* g, int use_v, int rank_type, uint64_t* total) {
  ranked R;
  if (make_ranked(g, use_v, rank_type, 0, &R)) return -1;
  int64_t n = R.n;
  size_t nn = (size_t)(n ? n : 1);
  uint32_t* wedges = (uint32_t*)calloc(nn, 4);
  uint32_t* used = (uint32_t*)malloc(4 * nn);
  uint64_t nb = 0;
  for (int64_t i = 0; i < n; i++) {
    int64_t nused = 0;
    uint32_t io = R.roff[i], id = R.roff[i + 1] - io;
    for (uint32_t j = 0; j < id; j++) {
      uint32_t v = R.redg[io + j] >> 1, vo = R.roff[v], vd = R.roff[v + 1] - vo;
      for (uint32_t k = 0; k < vd; k++) {
        uint32_t u2 = R.redg[vo + k] >> 1;
        if (!(u2 < i && u2 < v)) break;
        nb += wedges[u2];
        if (wedges[u2]++ == 0) used[nused++] = u2;
      }
    }
    for (int64_t t = 0; t < nused; t++) wedges[used[t]] = 0;
  }
  *total = nb;
  free(wedges); free(used); free_ranked(&R);
  return 0;
}

/* Tip numbers on S: Peel (butterfly_peel.h:896-934) with PeelOrigParallel (:647-711) and
 * updateBuckets (butterfly_putils.h:1027-1115) under the increasing buckets of ligra/bucket.h.
 * Round structure: k = min D over unpeeled; every unpeeled x with D[x] == k is peeled together;
 * for each peeled u and each u2 != u reached through a centre v, B[u2] -= C(c(u,u2),2) (running
 * count, :684-696); then each touched u2 with D[u2] > k gets D[u2] = max(B[u2], k).  The tip
 * number is the final D.  counts/tip are indexed by original id within S. */
void orc_peel_vert(const orc_graph* g, int use_v, const uint64_t* counts, uint64_t* tip) {
  int64_t ns = use_v ? g->nu : g->nv;
  const uint32_t* offS = use_v ? g->offU : g->offV;
  const uint32_t* edgS = use_v ? g->edgU : g->edgV;
  const uint32_t* offO = use_v ? g->offV : g->offU;
  const uint32_t* edgO = use_v ? g->edgV : g->edgU;
  size_t nn = (size_t)(ns ? ns : 1);
  int64_t* B = (int64_t*)malloc(8 * nn);
  uint64_t* D = tip;
  uint8_t* peeled = (uint8_t*)calloc(nn, 1);
  uint8_t* upd = (uint8_t*)calloc(nn, 1);
  uint32_t* cnt = (uint32_t*)calloc(nn, 4);
  int64_t* active = (int64_t*)malloc(8 * nn);
  int64_t* touched = (int64_t*)malloc(8 * nn);
  int64_t* used = (int64_t*)malloc(8 * nn);
  for (int64_t x = 0; x < ns; x++) { B[x] = (int64_t)counts[x]; D[x] = counts[x]; }
  int64_t finished = 0;
  while (finished < ns) {
    uint64_t k = 0;
    int first = 1;
    for (int64_t x = 0; x < ns; x++) if (!peeled[x] && (first || D[x] < k)) { k = D[x]; first = 0; }
    int64_t na = 0, nt = 0;
    for (int64_t x = 0; x < ns; x++) if (!peeled[x] && D[x] == k) active[na++] = x;
    finished += na;
    for (int64_t a = 0; a < na; a++) {
      int64_t u = active[a], nused = 0;
      for (uint32_t p = offS[u]; p < offS[u + 1]; p++) {
        uint32_t v = edgS[p];
        for (uint32_t q = offO[v]; q < offO[v + 1]; q++) {
          uint32_t u2 = edgO[q];
          if ((int64_t)u2 == u) continue;
          B[u2] -= cnt[u2];
          if (cnt[u2]++ == 0) {
            used[nused++] = u2;
            if (!upd[u2]) { upd[u2] = 1; touched[nt++] = u2; }
          }
        }
      }
      for (int64_t t = 0; t < nused; t++) cnt[used[t]] = 0;
    }
    for (int64_t a = 0; a < na; a++) peeled[active[a]] = 1;
    for (int64_t t = 0; t < nt; t++) {
      int64_t u2 = touched[t];
      upd[u2] = 0;
      if (D[u2] > k) {
        int64_t b = B[u2];
        D[u2] = (b > (int64_t)k) ? (uint64_t)b : k;
      }
    }
  }
  free(B); free(peeled); free(upd); free(cnt); free(active); free(touched); free(used);
}

/* Wing numbers: PeelE (butterfly_epeel.h:519-569) with PeelEOrigParallel (:339-443).  Edges are
 * indexed by position in O's CSR.  Round structure: k = min D over live edges; every live edge
 * with D == k is peeled together.  A butterfly whose four edges are live at the start of the
 * round and which contains at least one peeled edge is charged once, by its minimum-index peeled
 * edge (:383, :393, :399-400): its other three edges each lose one butterfly (:404-405, :425).
 * Then the peeled edges die (:432-435) and touched edges with D > k get D = max(B, k). */
void orc_peel_edge(const orc_graph* g, int use_v, const uint64_t* counts, uint64_t* wing) {
  int64_t m = g->m;
  int64_t ns = use_v ? g->nu : g->nv;
  int64_t no = use_v ? g->nv : g->nu;
  const uint32_t* offS = use_v ? g->offU : g->offV;
  const uint32_t* edgS = use_v ? g->edgU : g->edgV;
  const uint32_t* offO = use_v ? g->offV : g->offU;
  const uint32_t* edgO = use_v ? g->edgV : g->edgU;
  size_t mm = (size_t)(m ? m : 1);
  (void)ns; (void)no;
  uint32_t* eti = (uint32_t*)malloc(4 * mm);  /* S-CSR position -> O-CSR position */
  orc_edge_to_idxs(g, use_v, eti);
  uint32_t* owner_o = (uint32_t*)malloc(4 * mm); /* O vertex of each O-CSR position */
  for (int64_t v = 0; v < no; v++) for (uint32_t p = offO[v]; p < offO[v + 1]; p++) owner_o[p] = (uint32_t)v;
  int64_t* B = (int64_t*)malloc(8 * mm);
  uint64_t* D = wing;
  uint8_t* alive = (uint8_t*)malloc(mm);
  uint8_t* cur = (uint8_t*)calloc(mm, 1);
  uint8_t* upd = (uint8_t*)calloc(mm, 1);
  int64_t* active = (int64_t*)malloc(8 * mm);
  int64_t* touched = (int64_t*)malloc(8 * mm);
  for (int64_t e = 0; e < m; e++) { B[e] = (int64_t)counts[e]; D[e] = counts[e]; alive[e] = 1; }
  int64_t finished = 0;
  while (finished < m) {
    uint64_t k = 0;
    int first = 1;
    for (int64_t e = 0; e < m; e++) if (alive[e] && (first || D[e] < k)) { k = D[e]; first = 0; }
    int64_t na = 0, nt = 0;
    for (int64_t e = 0; e < m; e++) if (alive[e] && D[e] == k) { active[na++] = e; cur[e] = 1; }
    finished += na;
#define TOUCH(f) do { if (!upd[f]) { upd[f] = 1; touched[nt++] = (f); } } while (0)
    for (int64_t a = 0; a < na; a++) {
      int64_t e = active[a];               /* e = (v, u), v in O, u in S */
      uint32_t v = owner_o[e], u = edgO[e];
      for (uint32_t q = offO[v]; q < offO[v + 1]; q++) {   /* f1 = (v, u2) */
        uint32_t u2 = edgO[q];
        if (u2 == u || !alive[q]) continue;
        if (cur[q] && (int64_t)q < e) continue;
        for (uint32_t p2 = offS[u2]; p2 < offS[u2 + 1]; p2++) {   /* f2 = (v2, u2) */
          uint32_t v2 = edgS[p2], f2 = eti[p2];
          if (v2 == v || !alive[f2]) continue;
          if (cur[f2] && (int64_t)f2 < e) continue;
          /* f3 = (v2, u): search v2 in u's S list */
          uint32_t lo = offS[u], hi = offS[u + 1];
          while (lo < hi) { uint32_t mid = lo + (hi - lo) / 2; if (edgS[mid] < v2) lo = mid + 1; else hi = mid; }
          if (lo == offS[u + 1] || edgS[lo] != v2) continue;
          uint32_t f3 = eti[lo];
          if (!alive[f3]) continue;
          if (cur[f3] && (int64_t)f3 < e) continue;
          B[q]--; B[f2]--; B[f3]--;
          TOUCH(q); TOUCH(f2); TOUCH(f3);
        }
      }
    }
#undef TOUCH
    for (int64_t a = 0; a < na; a++) { alive[active[a]] = 0; cur[active[a]] = 0; }
    for (int64_t t = 0; t < nt; t++) {
      int64_t f = touched[t];
      upd[f] = 0;
      if (D[f] > k) {
        int64_t b = B[f];
        D[f] = (b > (int64_t)k) ? (uint64_t)b : k;
      }
    }
  }
  free(eti); free(owner_o); free(B); free(alive); free(cur); free(upd); free(active); free(touched);
}

/* ------------------------------------------------------------------------------------------
 * Brute-force checkers (mathematical definitions, independent of the reference's algorithm).
 * ------------------------------------------------------------------------------------------ */

static int64_t common_sorted(const uint32_t* a, int64_t na, const uint32_t* b, int64_t nb) {
  int64_t i = 0, j = 0, c = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) i++; else if (a[i] > b[j]) j++; else { c++; i++; j++; }
  }
  return c;
}

/* B[x] = sum over y != x in S of C(|N(x) & N(y)|, 2); total = sum over pairs.  O(n_S^2 d). */
uint64_t orc_brute_vert(const orc_graph* g, int use_v, uint64_t* out) {
  int64_t ns = use_v ? g->nu : g->nv;
  const uint32_t* offS = use_v ? g->offU : g->offV;
  const uint32_t* edgS = use_v ? g->edgU : g->edgV;
  uint64_t total = 0;
  if (out) memset(out, 0, 8 * (size_t)ns);
  for (int64_t x = 0; x < ns; x++)
    for (int64_t y = x + 1; y < ns; y++) {
      int64_t c = common_sorted(edgS + offS[x], offS[x + 1] - offS[x], edgS + offS[y], offS[y + 1] - offS[y]);
      uint64_t b = (uint64_t)(c * (c - 1) / 2);
      total += b;
      if (out) { out[x] += b; out[y] += b; }
    }
  return total;
}

/* E[e=(v,u)] = sum over u2 in N(v), u2 != u, of (|N(u) & N(u2)| - 1), indexed by O-CSR position. */
void orc_brute_edge(const orc_graph* g, int use_v, uint64_t* out) {
  int64_t no = use_v ? g->nv : g->nu;
  const uint32_t* offS = use_v ? g->offU : g->offV;
  const uint32_t* edgS = use_v ? g->edgU : g->edgV;
  const uint32_t* offO = use_v ? g->offV : g->offU;
  const uint32_t* edgO = use_v ? g->edgV : g->edgU;
  for (int64_t v = 0; v < no; v++)
    for (uint32_t p = offO[v]; p < offO[v + 1]; p++) {
      uint32_t u = edgO[p];
      uint64_t s = 0;
      for (uint32_t q = offO[v]; q < offO[v + 1]; q++) {
        uint32_t u2 = edgO[q];
        if (u2 == u) continue;
        int64_t c = common_sorted(edgS + offS[u], offS[u + 1] - offS[u], edgS + offS[u2], offS[u2 + 1] - offS[u2]);
        s += (uint64_t)(c - 1);
      }
      out[p] = s;
    }
}

/* Textbook tip decomposition, one vertex at a time: repeatedly remove an S vertex of minimum
 * current butterfly count b; its tip number is max(b, largest tip so far). */
void orc_brute_peel_vert(const orc_graph* g, int use_v, uint64_t* tip) {
  int64_t ns = use_v ? g->nu : g->nv;
  const uint32_t* offS = use_v ? g->offU : g->offV;
  const uint32_t* edgS = use_v ? g->edgU : g->edgV;
  size_t nn = (size_t)(ns ? ns : 1);
  int64_t* B = (int64_t*)calloc(nn, 8);
  uint8_t* gone = (uint8_t*)calloc(nn, 1);
  int64_t* C2 = (int64_t*)malloc(8 * nn * nn);
  for (int64_t x = 0; x < ns; x++)
    for (int64_t y = 0; y < ns; y++) {
      int64_t c = x == y ? 0 : common_sorted(edgS + offS[x], offS[x + 1] - offS[x], edgS + offS[y], offS[y + 1] - offS[y]);
      C2[x * ns + y] = c * (c - 1) / 2;
      B[x] += C2[x * ns + y];
    }
  int64_t level = 0;
  for (int64_t it = 0; it < ns; it++) {
    int64_t best = -1;
    for (int64_t x = 0; x < ns; x++) if (!gone[x] && (best < 0 || B[x] < B[best])) best = x;
    if (B[best] > level) level = B[best];
    tip[best] = (uint64_t)level;
    gone[best] = 1;
    for (int64_t y = 0; y < ns; y++) if (!gone[y]) B[y] -= C2[best * ns + y];
  }
  free(B); free(gone); free(C2);
}

/* Textbook wing decomposition, one edge at a time: repeatedly remove a live edge of minimum
 * current butterfly count; every butterfly through it loses its other three edges' credit. */
void orc_brute_peel_edge(const orc_graph* g, int use_v, uint64_t* wing) {
  int64_t m = g->m, no = use_v ? g->nv : g->nu;
  const uint32_t* offS = use_v ? g->offU : g->offV;
  const uint32_t* edgS = use_v ? g->edgU : g->edgV;
  const uint32_t* offO = use_v ? g->offV : g->offU;
  const uint32_t* edgO = use_v ? g->edgV : g->edgU;
  size_t mm = (size_t)(m ? m : 1);
  uint32_t* eti = (uint32_t*)malloc(4 * mm);
  orc_edge_to_idxs(g, use_v, eti);
  uint32_t* owner_o = (uint32_t*)malloc(4 * mm);
  for (int64_t v = 0; v < no; v++) for (uint32_t p = offO[v]; p < offO[v + 1]; p++) owner_o[p] = (uint32_t)v;
  uint8_t* alive = (uint8_t*)malloc(mm);
  int64_t* B = (int64_t*)calloc(mm, 8);
  for (int64_t e = 0; e < m; e++) alive[e] = 1;
  /* recount every edge from scratch over live edges each time it is needed */
  int64_t level = 0;
  for (int64_t it = 0; it < m; it++) {
    for (int64_t e = 0; e < m; e++) {
      if (!alive[e]) continue;
      uint32_t v = owner_o[e], u = edgO[e];
      int64_t s = 0;
      for (uint32_t q = offO[v]; q < offO[v + 1]; q++) {
        uint32_t u2 = edgO[q];
        if (u2 == u || !alive[q]) continue;
        for (uint32_t p2 = offS[u2]; p2 < offS[u2 + 1]; p2++) {
          uint32_t v2 = edgS[p2];
          if (v2 == v || !alive[eti[p2]]) continue;
          for (uint32_t p3 = offS[u]; p3 < offS[u + 1]; p3++)
            if (edgS[p3] == v2 && alive[eti[p3]]) s++;
        }
      }
      B[e] = s;
    }
    int64_t best = -1;
    for (int64_t e = 0; e < m; e++) if (alive[e] && (best < 0 || B[e] < B[best])) best = e;
    if (B[best] > level) level = B[best];
    wing[best] = (uint64_t)level;
    alive[best] = 0;
  }
  free(eti); free(owner_o); free(alive); free(B);
}

/* ---------------- sparsification (approximate TOTAL) ---------------- */

/* hashInt(ulong) (ligra/utils.h:389-397).  Which edges a seed keeps is defined by this mix, so
 * its steps and constants are the reference's. */
static uint64_t ref_hash64(uint64_t a) {
  a = (a + 0x7ed55d166bef7a1dull) + (a << 12);
  a = (a ^ 0xc761c23c510fa2ddull) ^ (a >> 9);
  a = (a + 0x165667b183a9c0e1ull) + (a << 59);
  a = (a + 0xd3a2646cab3487e3ull) ^ (a << 49);
  a = (a + 0xfd7046c5ef9ab54cull) + (a << 3);
  a = (a ^ 0xb55a4f090dd4a67bull) ^ (a >> 32);
  return a;
}
uint64_t orc_hash64(uint64_t a) { return ref_hash64(a); }

/* delZeroDeg (butterfly_utils.h:744-785) applied to the kept entries: vertices with a kept
 * entry are renumbered in id order on each side; lists keep their order. */
static void del_zero_deg(const orc_graph* g, const uint8_t* keepV, const uint8_t* keepU, int64_t* nv_out,
                         int64_t* nu_out, int64_t* m_out, uint32_t* offV, uint32_t* edgV, uint32_t* offU,
                         uint32_t* edgU) {
  uint32_t* idV = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(g->nv + 1));
  uint32_t* idU = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(g->nu + 1));
  int64_t nv2 = 0, nu2 = 0;
  for (int64_t v = 0; v < g->nv; v++) {
    int any = 0;
    for (uint32_t p = g->offV[v]; p < g->offV[v + 1]; p++) any |= keepV[p];
    idV[v] = any ? (uint32_t)nv2++ : 0xFFFFFFFFu;
  }
  for (int64_t u = 0; u < g->nu; u++) {
    int any = 0;
    for (uint32_t p = g->offU[u]; p < g->offU[u + 1]; p++) any |= keepU[p];
    idU[u] = any ? (uint32_t)nu2++ : 0xFFFFFFFFu;
  }
  int64_t mv = 0, mu = 0;
  for (int64_t v = 0; v < g->nv; v++) {
    if (idV[v] == 0xFFFFFFFFu) continue;
    offV[idV[v]] = (uint32_t)mv;
    for (uint32_t p = g->offV[v]; p < g->offV[v + 1]; p++)
      if (keepV[p]) edgV[mv++] = idU[g->edgV[p]];
  }
  offV[nv2] = (uint32_t)mv;
  for (int64_t u = 0; u < g->nu; u++) {
    if (idU[u] == 0xFFFFFFFFu) continue;
    offU[idU[u]] = (uint32_t)mu;
    for (uint32_t p = g->offU[u]; p < g->offU[u + 1]; p++)
      if (keepU[p]) edgU[mu++] = idV[g->edgU[p]];
  }
  offU[nu2] = (uint32_t)mu;
  *nv_out = nv2;
  *nu_out = nu2;
  *m_out = mv == mu ? mv : -1;
  free(idV);
  free(idU);
}

/* eSparseBipartite (butterfly_utils.h:796-898), type 2, and clrSparseBipartite (:909-996),
 * type 1, each followed by delZeroDeg.  numColors = max(1, (uint32) denom).
 *   ESPARSE: slot (i, u) of V's CSR gets colour hashInt(seed + concat) % numColors with
 *     concat = (i + 0x9e3779b9 + (seed << 6) + (seed >> 2))
 *              ^ ((uint32)(u + 0x9e3779b9) + (seed << 6) + (seed >> 2))   (:802-810);
 *     each U slot (i, v) copies the colour of the V slot holding i, found by binary search in
 *     N(v) (:813-825); entries of colour 0 stay.
 *   CLRSPARSE: colour(x) = hashInt(seed + x) % numColors per side (:911-914); an entry stays
 *     when its two ends have the same colour (:921-937).
 * Output arrays as large as the input's; returns 0, or -1 on a bad type. */
int orc_sparsify(const orc_graph* g, int type, int64_t denom, int64_t seed, int64_t* nv_out, int64_t* nu_out,
                 int64_t* m_out, uint32_t* offV, uint32_t* edgV, uint32_t* offU, uint32_t* edgU) {
  if (type != 1 && type != 2) return -1;
  const uint64_t ncol = (uint32_t)denom > 1u ? (uint32_t)denom : 1u;
  const uint64_t s = (uint64_t)seed, salt = (s << 6) + (s >> 2);
  size_t mm = (size_t)(g->m ? g->m : 1);
  uint8_t* keepV = (uint8_t*)malloc(mm);
  uint8_t* keepU = (uint8_t*)malloc(mm);
  if (type == 2) {
    uint32_t* colV = (uint32_t*)malloc(sizeof(uint32_t) * mm);
    for (int64_t i = 0; i < g->nv; i++)
      for (uint32_t p = g->offV[i]; p < g->offV[i + 1]; p++) {
        uint32_t u = g->edgV[p];
        uint64_t concat = (uint64_t)i + 0x9e3779b9ull + salt;
        concat ^= (uint64_t)(uint32_t)(u + 0x9e3779b9u) + salt;
        colV[p] = (uint32_t)(ref_hash64(s + concat) % ncol);
        keepV[p] = colV[p] == 0;
      }
    for (int64_t i = 0; i < g->nu; i++)
      for (uint32_t q = g->offU[i]; q < g->offU[i + 1]; q++) {
        uint32_t v = g->edgU[q];
        uint32_t lo = g->offV[v], hi = g->offV[v + 1];
        while (lo < hi) {  /* first position in N(v) holding a value >= i */
          uint32_t mid = lo + (hi - lo) / 2;
          if (g->edgV[mid] < (uint32_t)i) lo = mid + 1; else hi = mid;
        }
        keepU[q] = colV[lo] == 0;
      }
    free(colV);
  } else {
    for (int64_t v = 0; v < g->nv; v++) {
      uint64_t cv = ref_hash64(s + (uint64_t)v) % ncol;
      for (uint32_t p = g->offV[v]; p < g->offV[v + 1]; p++)
        keepV[p] = ref_hash64(s + g->edgV[p]) % ncol == cv;
    }
    for (int64_t u = 0; u < g->nu; u++) {
      uint64_t cu = ref_hash64(s + (uint64_t)u) % ncol;
      for (uint32_t p = g->offU[u]; p < g->offU[u + 1]; p++)
        keepU[p] = ref_hash64(s + g->edgU[p]) % ncol == cu;
    }
  }
  del_zero_deg(g, keepV, keepU, nv_out, nu_out, m_out, offV, edgV, offU, edgU);
  free(keepV);
  free(keepU);
  return 0;
}
