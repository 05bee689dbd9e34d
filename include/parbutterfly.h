/*
 * parbutterfly.h -- C ABI of the MI355X butterfly counting / peeling library
 * (libparbutterfly_amd.so).  Plain pointers and sizes only; no C++ or torch types.
 *
 * Every entry point returns a pbf_status code (0 = success) and never exits or prints,
 * replacing the reference's print-and-exit error convention (apps/hyper/butterfly_utils.h:
 * 1013-1016, 1039-1041, 1049-1051).
 *
 * Graph convention (apps/hyper/butterfly_utils.h:184-197, bipartiteCSR): V has nv vertices,
 * U has nu; offsetsV[nv+1] / edgesV[m] list the U-neighbours of each V vertex and
 * offsetsU[nu+1] / edgesU[m] the V-neighbours of each U vertex.  Every list must be strictly
 * ascending (sorted, duplicate free) -- the reference assumes this without checking; this
 * library checks it (pbf_validate) and returns PBF_ERR_UNSORTED.  Ids and offsets are 32-bit
 * as in the reference's default build (uintE / uintT, ligra/parallel.h:102-125); 2*m must be
 * below 2^31 (the reference's intT offsets, Appendix-A limit) else PBF_ERR_CAPACITY.
 *
 * Count side: use_v != 0 means counts live on U (S = U, O = V), else on V (S = V, O = U),
 * exactly the reference's cmpWedgeCounts flag (butterfly_utils.h:1080-1086).  Per-vertex
 * outputs have n_S entries indexed by original id within S; per-edge outputs have m entries
 * indexed by position in O's CSR (edgesV if use_v, else edgesU) -- the reference's output
 * contract (butterfly_count.h:1600-1610, butterfly_ecount.h:1723-1741).  Counts are uint64.
 *
 * Memory: `mem` says where BOTH the graph arrays and the output / count arrays live:
 * PBF_MEM_HOST (pageable host memory; the library copies) or PBF_MEM_DEVICE (HIP device
 * pointers on the context's device; nothing crosses PCIe).
 */
#ifndef PARBUTTERFLY_H
#define PARBUTTERFLY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBF_ABI_VERSION 2

typedef enum {
  PBF_OK = 0,
  PBF_ERR_ARG = 1,         /* null pointer, bad enum value, bad size */
  PBF_ERR_FORMAT = 2,      /* not an AdjacencyHypergraph file / length mismatch / mv != mu */
  PBF_ERR_RANGE = 3,       /* neighbour id out of range (butterfly_utils.h:1039-1041, 1049-1051) */
  PBF_ERR_UNSORTED = 4,    /* adjacency list not strictly ascending, or offsets decreasing */
  PBF_ERR_OOM = 5,         /* host or device allocation failed */
  PBF_ERR_CAPACITY = 6,    /* sizes beyond the 32-bit index domain, or a fixed-capacity limit */
  PBF_ERR_DOMAIN = 7,      /* input outside the documented parity domain */
  PBF_ERR_HIP = 8,         /* HIP runtime failure (no device, launch or copy error) */
  PBF_ERR_UNSUPPORTED = 9, /* combination of options not implemented by this library */
  PBF_ERR_IO = 10          /* file could not be opened / read / written */
} pbf_status;

/* Enums carry the reference's numeric values (apps/hyper/butterfly_utils.h:24-28); the GPU
 * modes are appended after the reference's last value. */
enum { PBF_COUNT_SORT = 0, PBF_COUNT_ASORT, PBF_COUNT_HASH, PBF_COUNT_AHASH, PBF_COUNT_HIST,
       PBF_COUNT_AHIST, PBF_COUNT_BATCHS, PBF_COUNT_BATCHWA, PBF_COUNT_SERIAL, PBF_COUNT_GPU = 9 };
enum { PBF_RANK_SIDE = 0, PBF_RANK_COCORE, PBF_RANK_ACOCORE, PBF_RANK_DEG, PBF_RANK_ADEG };
enum { PBF_PEEL_SORT = 0, PBF_PEEL_HASH, PBF_PEEL_HIST, PBF_PEEL_BATCHS, PBF_PEEL_BATCHWA,
       PBF_PEEL_GPU = 5 };
enum { PBF_PER_VERT = 0, PBF_PER_EDGE, PBF_PER_TOTAL };
enum { PBF_MEM_HOST = 0, PBF_MEM_DEVICE = 1 };

typedef struct {
  int64_t nv, nu, m;
  const uint32_t* offsetsV; /* nv+1 */
  const uint32_t* edgesV;   /* m    */
  const uint32_t* offsetsU; /* nu+1 */
  const uint32_t* edgesU;   /* m    */
} pbf_bipartite_csr;

/* Host graph owned by the library (reader / generators); release with pbf_graph_free. */
typedef struct {
  pbf_bipartite_csr csr;
  void* storage;
} pbf_graph;

/* Phase timings of the last count / peel call on a context, from HIP events (ms). */
typedef struct {
  double rank_ms;     /* ranking (ADEG / ACOCORE keys, sort, rank maps) */
  double relabel_ms;  /* ranked CSR build (rankGraph / rankGraphEdges) */
  double count_ms;    /* wedge enumeration + aggregation + count scatter */
  double convert_ms;  /* rank -> original id conversion */
  double peel_ms;     /* peeling rounds */
  double total_ms;    /* whole call, first kernel to last */
  int64_t wedges;     /* W = rank-filtered wedges processed (rankWedgeF) */
  int64_t rounds;     /* peeling rounds */
  int64_t hub_vertices;   /* endpoints that took the HBM-histogram path */
  double sparsify_ms;     /* sparsification (pbf_sparsify / pbf_count_total_sparse) */
  int64_t batches;        /* rank ranges the last count ran in (pbf_set_limits) */
} pbf_timing;

const char* pbf_status_str(int status);
int pbf_abi_version(void);

/* ---------------- host side: graphs, IO, generators (no GPU needed) ---------------- */

/* readBipartite (apps/hyper/butterfly_utils.h:1008-1058): AdjacencyHypergraph text, or the
 * library's binary "PBFCSR1" layout (detected by magic).  Validates like pbf_validate. */
int pbf_read_bipartite(const char* path, pbf_graph* out);
/* A KONECT bipartite edge list (readKONECT, utils/graphIO.h:1032-1098, as utils/
 * KONECTtoHyperAdj.C converts it): leading '%' lines skipped, the first two columns of every
 * row a V id and a U id (further columns ignored), each side's ids renumbered densely in
 * increasing order, lists sorted and duplicates dropped.  Ids are non-negative and below 2^32
 * (else PBF_ERR_RANGE); fewer than two columns is PBF_ERR_FORMAT. */
int pbf_read_konect(const char* path, pbf_graph* out);
/* Writes AdjacencyHypergraph text (utils/graphIO.h:779-800 layout) or PBFCSR1 binary. */
int pbf_write_bipartite(const char* path, const pbf_bipartite_csr* g);
int pbf_write_binary(const char* path, const pbf_bipartite_csr* g);
/* Range, order and size checks on a host graph. */
int pbf_validate(const pbf_bipartite_csr* g);
/* Allocate an empty host graph with the given sizes (offsets zeroed) / release any pbf_graph. */
int pbf_graph_alloc(int64_t nv, int64_t nu, int64_t m, pbf_graph* out);
void pbf_graph_free(pbf_graph* g);
/* Build both CSRs from an edge list (v[i] in [0,nv), u[i] in [0,nu)); duplicates removed. */
int pbf_graph_from_edges(int64_t nv, int64_t nu, int64_t ne, const uint32_t* v, const uint32_t* u,
                         pbf_graph* out);

/* Seeded generators.  Edges are drawn with a counter-based RNG; the graph is the first m
 * DISTINCT pairs in draw order, so the same arguments give the same graph on any machine. */
int pbf_gen_uniform(int64_t nv, int64_t nu, int64_t m, uint64_t seed, pbf_graph* out);
/* Chung-Lu: endpoint weights w_i = (i+1)^(-1/(gamma-1)) on each side, endpoints drawn
 * independently with probability proportional to weight. */
int pbf_gen_chung_lu(int64_t nv, int64_t nu, int64_t m, double gamma_v, double gamma_u,
                     uint64_t seed, pbf_graph* out);
/* Rectangular R-MAT over 2^log_nv x 2^log_nu with quadrant probabilities a, b, c, d. */
int pbf_gen_rmat(int log_nv, int log_nu, int64_t m, double a, double b, double c, double d,
                 uint64_t seed, pbf_graph* out);
/* Parameters of graph `index` of the mixed batch distribution: nv, nu ~ logU[64, 65536],
 * average degree ~ logU[2, 32], m = round(avg*(nv+nu)/2) capped at nv*nu/2, gamma ~ U[2,3]. */
int pbf_batch_graph_params(uint64_t seed, int64_t* nv, int64_t* nu, int64_t* m, double* gamma);
int pbf_gen_batch_graph(uint64_t seed, pbf_graph* out);

/* cmpWedgeCounts on a host graph (butterfly_utils.h:1080-1086). */
int pbf_side_host(const pbf_bipartite_csr* g, int* use_v, int64_t* num_wedges);

/* ---------------- GPU context ---------------- */

typedef struct pbf_context pbf_context;

/* Creates a context on HIP device `device` with its own default-priority stream.  Device
 * scratch is cached in the context and grows on demand; one context per host thread. */
int pbf_create(pbf_context** out, int device);
int pbf_destroy(pbf_context* ctx);
/* The HIP stream (hipStream_t, as an opaque pointer) the context queues its work on. */
void* pbf_stream(pbf_context* ctx);
/* Queues the context's work on the caller's HIP stream (a hipStream_t of the context's device,
 * as an opaque pointer) instead of the context's own; NULL returns to the context's own
 * stream.  Work queued on the previous stream is finished first.  Afterwards every entry point
 * queues its kernels and copies on that stream, and the synchronous ones wait for that stream
 * only, so a caller can order the library's work with its own.  The stream must outlive its
 * use by the context (set NULL before destroying it). */
int pbf_set_stream(pbf_context* ctx, void* stream);
int pbf_synchronize(pbf_context* ctx);
int pbf_last_timing(pbf_context* ctx, pbf_timing* out);
/* Releases cached device scratch. */
int pbf_trim(pbf_context* ctx);
/* Memory bounds of the counting passes, the reference's -m and -a (BipartiteButterflies.C:
 * 174-175; 0 = no bound, the default).  max_wedges > 0: the ranked endpoints are counted in
 * consecutive rank ranges of at most max_wedges wedges each, cut by the wedge prefix as
 * getNextWedgeIdx does (butterfly_utils.h:1700-1714, 1948-1959; butterfly_count.h:1578-1587);
 * an endpoint with more wedges on its own makes the count fail with PBF_ERR_CAPACITY (the
 * reference exits there).  max_array_size > 0: the dense HBM histograms of the endpoints no
 * LDS table holds use at most that many 32-bit counters, 2 (nv + nu) per concurrently counted
 * endpoint, and at least one such endpoint's worth.  Results do not depend on either. */
int pbf_set_limits(pbf_context* ctx, int64_t max_wedges, int64_t max_array_size);

/* Device buffers on the context's device (for PBF_MEM_DEVICE callers without their own HIP
 * allocator).  Copies are synchronous with respect to the host and ordered on the context's
 * stream.  No counterpart in the reference (it has no device memory). */
int pbf_device_alloc(pbf_context* ctx, size_t bytes, void** ptr);
int pbf_device_free(pbf_context* ctx, void* ptr);
int pbf_copy_to_device(pbf_context* ctx, void* dst, const void* src, size_t bytes);
int pbf_copy_to_host(pbf_context* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy by the library's own streaming kernel (16-byte vector loads and stores
 * over the whole device); synchronous; dst and src 16-byte aligned (else PBF_ERR_ARG).  It
 * moves exactly `bytes` in each direction, which makes it the reference for calibrating
 * hardware traffic counters. */
int pbf_copy_device(pbf_context* ctx, void* dst, const void* src, size_t bytes);

/* readBipartite (butterfly_utils.h:1008-1058) with its output in HBM: the file is copied to the
 * context's device as is and tokenised, parsed (text) and validated there, with the error codes
 * of pbf_read_bipartite (the first error in file order, then pbf_validate's order).  out's
 * arrays point into one device allocation returned in *storage (release it with
 * pbf_device_free).  Synchronous. */
int pbf_read_bipartite_device(pbf_context* ctx, const char* path, pbf_bipartite_csr* out, void** storage);

/* cmpWedgeCounts (butterfly_utils.h:1061-1086) on device or host graphs. */
int pbf_side(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int* use_v, int64_t* num_wedges);

/* getRanks / getApproxDegRanks / getApproxCoCoreRanks (butterfly_utils.h:249-284, 448-503)
 * plus rankWedgeF (:216-237).  order[r] = combined id of rank r (V ids 0..nv-1, U ids
 * nv..nv+nu-1); rank_of[id] = r; *wedges = W under this order.  Ties inside a key are broken
 * by ascending id (the reference's sampleSort leaves them unspecified).  Any of the output
 * pointers may be NULL.  rank_type: PBF_RANK_ADEG, PBF_RANK_ACOCORE, or PBF_RANK_SIDE (the
 * count side S first, then O, each by id: side-based wedges, butterfly_count.h:1639-1640). */
int pbf_rank(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int rank_type,
             uint32_t* order, uint32_t* rank_of, int64_t* wedges);

/* Count (butterfly_count.h:1634): per-vertex butterfly counts on S, n_S = use_v ? nu : nv. */
int pbf_count_vert(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v,
                   int count_type, int rank_type, uint64_t* out);
/* CountE (butterfly_ecount.h:1766): per-edge counts, m entries in O-CSR order. */
int pbf_count_edge(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v,
                   int count_type, int rank_type, uint64_t* out);
/* CountTotal (butterfly_count_total.h:1014): total count, written to host *total. */
int pbf_count_total(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v,
                    int count_type, int rank_type, uint64_t* total);
/* pbf_count_vert without allocating.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and results as pbf_count_vert with PBF_MEM_DEVICE
 * (graph and out are device pointers on the context's device; there is no mem argument); it
 * calls no hipMalloc / hipFree and no allocating library routine.  Its scratch is one fixed
 * array of PBF_NOALLOC_SCRATCH_BYTES per device inside the library, held by every device the
 * library runs on until the process ends, and its address is looked up by pbf_create.  A graph
 * whose scratch does not fit returns PBF_ERR_CAPACITY and leaves out unwritten (about 110 bytes
 * per vertex plus 32 per edge, plus up to 17 MB of fixed tables; every graph with nv + nu <=
 * 2^20 and m <= 2^21 fits).  All calls on one device share that device's array and it keeps
 * the last call's contents: calls that overlap on one device (other host threads, other
 * contexts on the device) race, and the caller must serialise them. */
#define PBF_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int pbf_count_vert_noalloc(pbf_context* ctx, const pbf_bipartite_csr* g, int use_v, int count_type,
                           int rank_type, uint64_t* out);
/* edgeToIdxs (butterfly_utils.h:2034-2062): eti[p] = O-CSR position of S-CSR position p. */
int pbf_edge_to_idxs(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v,
                     uint32_t* eti);
/* Peel (butterfly_peel.h:896): tip numbers on S from per-vertex counts (n_S entries). */
int pbf_peel_vert(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v,
                  int peel_type, const uint64_t* counts, uint64_t* tip);
/* PeelE (butterfly_epeel.h:519): wing numbers from per-edge counts (m entries, O-CSR order). */
int pbf_peel_edge(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int use_v,
                  int peel_type, const uint64_t* counts, uint64_t* wing);

/* ---------------- approximate totals by sparsification ----------------
 * SparseType (butterfly_utils.h:28) with the driver's numeric -s values
 * (BipartiteButterflies.C:186-191). */
enum { PBF_SPARSE_NONE = 0, PBF_SPARSE_COLOR = 1, PBF_SPARSE_EDGE = 2 };
/* eSparseBipartite (butterfly_utils.h:796-898) for PBF_SPARSE_EDGE: edge (v, u) stays when its
 * hashed colour among denom colours is 0.  clrSparseBipartite (:909-996) for PBF_SPARSE_COLOR:
 * every vertex gets a hashed colour and the edges whose ends share a colour stay.  Both then
 * drop the vertices left without edges and renumber each side in id order (delZeroDeg,
 * :744-785).  1 <= denom < 2^32, 0 <= seed < 2^56 (else PBF_ERR_ARG).  The output arrays are
 * the caller's, in `mem`, with room for the input's sizes (nv + 1, m, nu + 1, m); the sizes of
 * the sparsified graph go to host *nv_out, *nu_out, *m_out.  Synchronous. */
int pbf_sparsify(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int sparse_type, int64_t denom,
                 int64_t seed, int64_t* nv_out, int64_t* nu_out, int64_t* m_out, uint32_t* offsetsV,
                 uint32_t* edgesV, uint32_t* offsetsU, uint32_t* edgesU);
/* CountTotal on the sparsified graph (its side chosen by cmpWedgeCounts on that graph), scaled
 * as the driver scales it (BipartiteButterflies.C:213-214): *approx = total * denom^3 for
 * PBF_SPARSE_COLOR, total * denom^4 for PBF_SPARSE_EDGE, the product taken in double precision
 * and truncated to int64 as there (PBF_ERR_DOMAIN when it reaches 2^63).  *sparse_total (may
 * be NULL) = the exact count of the sparsified graph.  PBF_SPARSE_NONE counts the graph
 * itself.  Results are host values; synchronous. */
int pbf_count_total_sparse(pbf_context* ctx, const pbf_bipartite_csr* g, int mem, int sparse_type,
                           int64_t denom, int64_t seed, int count_type, int rank_type, int64_t* approx,
                           uint64_t* sparse_total);

/* ---------------- batches of graphs (no counterpart in the reference) ----------------
 * Per-vertex (VERT) and total butterfly counts of `count` independent graphs, each exactly as
 * pbf_count_vert / pbf_count_total would give them with the graph's own side
 * (cmpWedgeCounts).  graphs: host array of descriptors whose arrays are in `mem`.
 * flags: PBF_BATCH_UNION runs one pass of the kernels over the disjoint union of the graphs;
 * PBF_BATCH_PLAIN runs the per-graph pipeline graph by graph (the baseline);
 * PBF_BATCH_UNION | PBF_BATCH_CAPTURE (device graphs and outputs only) also captures the pass
 * in a HIP graph on the context's stream and replays it on later calls with the same graphs,
 * outputs, rank type and tuning (one captured batch per context).
 * Outputs (in out_mem): vert_out holds graph g's counts on its side S at
 * [vert_off[g], vert_off[g+1]) (vert_cap >= sum over graphs of max(nv, nu)); vert_off has
 * count + 1 entries; totals[g] = butterflies of graph g.  Synchronous. */
enum { PBF_BATCH_UNION = 0, PBF_BATCH_PLAIN = 1, PBF_BATCH_CAPTURE = 16 };
int pbf_count_batch(pbf_context* ctx, int64_t count, const pbf_bipartite_csr* graphs, int mem,
                    int rank_type, int flags, uint64_t* vert_out, int64_t vert_cap,
                    uint64_t* vert_off, uint64_t* totals, int out_mem);
/* Releases the context's captured batch graph (pbf_destroy does it too). */
void pbf_batch_release(pbf_context* ctx);
/* Autotuner of the counting kernels for a batch: times PBF_BATCH_UNION under each candidate
 * setting of (light-endpoint key bound, light-endpoint wedge bound, workgroups per CU), keeps
 * the fastest in the context and returns its index and time.  Arguments as pbf_count_batch. */
int pbf_batch_autotune(pbf_context* ctx, int64_t count, const pbf_bipartite_csr* graphs, int mem,
                       int rank_type, uint64_t* vert_out, int64_t vert_cap, uint64_t* vert_off,
                       uint64_t* totals, int out_mem, int* best, float* best_ms);
/* Sets the counting kernels' knobs directly (light_max <= 1535, 0 = the default of each count
 * mode: 767 per vertex / total, 1535 per edge; 1 <= blocks_per_cu <= 8). */
int pbf_set_tuning(pbf_context* ctx, int light_max, int light_work, int blocks_per_cu);

/* ---------------- learned variant selector (no counterpart in the reference) ----------------
 * Features of each graph of a batch (PBF_MLP_IN floats per graph, row-major): log2 of
 * nv + 1, nu + 1, m + 1; per side the log2 of the mean degree, of the mean squared degree and of
 * the maximum degree; log2 of wV + 1 and wU + 1 (cmpWedgeCounts); log2 of W + 1 under SIDE,
 * ADEG and ACOCORE (exact rank-filtered wedge counts, rankWedgeF butterfly_utils.h:216-237);
 * use_v; and log2 of the density m / (nv nu).  One pass over the disjoint union per ranking. */
#define PBF_MLP_IN 16
#define PBF_MLP_HIDDEN 64
#define PBF_MLP_OUT 9
#define PBF_MLP_PARAMS 6288
int pbf_batch_features(pbf_context* ctx, int64_t count, const pbf_bipartite_csr* graphs, int mem,
                       float* features);
/* The variants the selector scores: variant v = 3 * r + k runs ranking r (PBF_RANK_SIDE,
 * PBF_RANK_ADEG, PBF_RANK_ACOCORE) with counting knobs k (pbf_variant_tuning). */
int pbf_variant_tuning(int variant, int* rank_type, int* light_max, int* light_work, int* blocks_per_cu);
/* MLP 16 -> 64 -> 64 -> 9 (ReLU), FP32 parameters in the layout W1[16][64], b1[64], W2[64][64],
 * b2[64], W3[64][16] (columns 9..15 zero), b3[16] (entries 9..15 zero).  Products in BF16 on the
 * matrix cores with FP32 accumulation.  All arrays on the host; synchronous. */
int pbf_mlp_init(uint64_t seed, float* params);
/* Adam (lr, beta 0.9 / 0.999) on the mean squared error over n samples X[n][16] -> T[n][9],
 * minibatches of 128 in a seeded random order each epoch; params updated in place. */
int pbf_mlp_train(pbf_context* ctx, float* params, const float* X, const float* T, int64_t n, int epochs,
                  float lr, uint64_t seed, float* final_loss);
/* Gradient of that loss over n <= 128 samples (same layout as params) and the loss. */
int pbf_mlp_grad(pbf_context* ctx, const float* params, const float* X, const float* T, int64_t n,
                 float* grads, float* loss);
/* Predictions Y[n][9]. */
int pbf_mlp_predict(pbf_context* ctx, const float* params, const float* X, int64_t n, float* Y);

/* A trained selector: network parameters (pbf_mlp_* layout) and the standardisation of its
 * inputs (feature f enters as (x - fmu[f]) / fsd[f]) and outputs (variant v's predicted log2
 * runtime is y[v] * tsd[v] + tmu[v]). */
typedef struct {
  float params[PBF_MLP_PARAMS];
  float fmu[PBF_MLP_IN], fsd[PBF_MLP_IN];
  float tmu[PBF_MLP_OUT], tsd[PBF_MLP_OUT];
} pbf_selector_model;
/* The method without the selector: ADEG ranking with the default counting knobs. */
#define PBF_VARIANT_FIXED 3
/* Chooses a variant for every graph of a batch (host int variants[count]): features
 * (pbf_batch_features), one batched inference, and the variant of least predicted runtime.
 * model == NULL, or a prediction that is not finite, gives the fixed variant.  The choice only
 * orders the work: every variant counts exactly.  Arguments as pbf_batch_features. */
int pbf_select(pbf_context* ctx, int64_t count, const pbf_bipartite_csr* graphs, int mem,
               const pbf_selector_model* model, int* variants);
/* Batched + learned counting: pbf_select, then one union pass (PBF_BATCH_UNION) per chosen
 * ranking with the knob setting most chosen among its graphs; outputs in pbf_count_batch's
 * layout and order (graph g's counts at [vert_off[g], vert_off[g+1]), totals[g]), and the
 * chosen variants in variants[count] (host, may be NULL).  With model == NULL it is one fixed
 * ADEG union pass.  Synchronous. */
int pbf_count_batch_learned(pbf_context* ctx, int64_t count, const pbf_bipartite_csr* graphs, int mem,
                            const pbf_selector_model* model, uint64_t* vert_out, int64_t vert_cap,
                            uint64_t* vert_off, uint64_t* totals, int out_mem, int* variants);

#ifdef __cplusplus
}
#endif

#endif /* PARBUTTERFLY_H */
