// Test-infrastructure harness around the ParButterfly reference (not product code).
//
// This file is ours; it includes the reference's own headers from where they lie
// (-I <reference>/ligra -I <reference>/apps/hyper, see oracle/Makefile) and calls the
// reference's entry points unchanged:
//   Count      apps/hyper/butterfly_count.h:1634
//   CountE     apps/hyper/butterfly_ecount.h:1766
//   CountTotal apps/hyper/butterfly_count_total.h:1014
//   Peel       apps/hyper/butterfly_peel.h:896
//   PeelE      apps/hyper/butterfly_epeel.h:519
//   cmpWedgeCounts / edgeToIdxs / idxsToEdge / readBipartite  apps/hyper/butterfly_utils.h
// It mirrors what Compute() does in apps/hyper/BipartiteButterflies.C:75-277, but instead of
// printing only the totals it dumps the dense per-vertex / per-edge arrays and the tip / wing
// numbers (which Compute discards) so that they can serve as golden vectors and as the CPU
// baseline.  Built into oracle/_ref/ only; never shipped as part of the product.
//
// Usage:
//   ref_harness <op> <graph> <out.bin> [-countType T] [-rankType R] [-peelType P]
//               [-rounds N] [-a max_array_size] [-m max_wedges]
//   op    : side | rank | vert | edge | total | peelv | peele | sparse | vertlist
//           (rank: counts = order, peel = {W}; sparse: see below, options -sparseType EDGE|COLOR,
//            -d denom, -seed seed; vertlist: graph is a text file naming one graph per line,
//            each counted per vertex in turn, as the batched workload's per-graph loop:
//            counts = every graph's total, seconds = the whole loop, per round)
//   graph : AdjacencyHypergraph text file (readBipartite) or a PBFCSR1 binary file
// Output file layout (little endian):
//   int64 use_v, int64 side_wedges, int64 n_counts, int64 counts[n_counts],
//   int64 n_peel, int64 peel[n_peel]      (n_peel = 0 unless op is peelv/peele)
// A JSON line with the timed seconds of every round is printed on stdout.
#include <iostream>
#include <math.h>
#include <stdlib.h>
#include <string>
#include <assert.h>
#include <stdio.h>
#include <string.h>
#include <vector>
#include <algorithm>

#define HYPER 1
#define MCX16 1
#define INVERSE 1

#include "hypergraphIO.h"
#include "index_map.h"
#include "bucket.h"
#include "edgeMapReduce.h"
#include "parallel.h"
#include "gettime.h"
#include "utils.h"
#include "graph.h"
#include "parseCommandLine.h"
#include "vertex.h"
#include "sequence.h"
#include "sparseSet.h"
#include "sampleSort.h"
#ifndef OPENMP
#include "../../lib/histogram.h"
#endif
#include "butterfly_count.h"
#include "butterfly_ecount.h"
#include "butterfly_peel.h"
#include "butterfly_epeel.h"
#include "butterfly_count_total.h"

#include <chrono>

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// PBFCSR1 binary graph: "PBFCSR1\0", int64 nv, nu, m, u32 offsetsV[nv+1], u32 edgesV[m],
// u32 offsetsU[nu+1], u32 edgesU[m].
static bool load_binary(const char* fname, bipartiteCSR& G) {
  FILE* f = fopen(fname, "rb");
  if (!f) return false;
  char magic[8];
  if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "PBFCSR1", 8) != 0) { fclose(f); return false; }
  int64_t hdr[3];
  if (fread(hdr, 8, 3, f) != 3) { fclose(f); return false; }
  long nv = hdr[0], nu = hdr[1], m = hdr[2];
  uintT* oV = newA(uintT, nv + 1); uintT* oU = newA(uintT, nu + 1);
  uintE* eV = newA(uintE, m); uintE* eU = newA(uintE, m);
  bool ok = fread(oV, 4, nv + 1, f) == (size_t)(nv + 1) && fread(eV, 4, m, f) == (size_t)m &&
            fread(oU, 4, nu + 1, f) == (size_t)(nu + 1) && fread(eU, 4, m, f) == (size_t)m;
  fclose(f);
  if (!ok) return false;
  G = bipartiteCSR(oV, oU, eV, eU, nv, nu, m);
  return true;
}

static bipartiteCSR load_graph(char* fname) {
  bipartiteCSR G;
  if (load_binary(fname, G)) return G;
  return readBipartite(fname);  // AdjacencyHypergraph text (butterfly_utils.h:1008)
}

// Same string / numeric aliases as Compute() (BipartiteButterflies.C:79-150).
static CountType parse_count(const std::string& s) {
  if (s == "ASORT") return ASORT; if (s == "SORT") return SORT; if (s == "AHASH") return AHASH;
  if (s == "HASH") return HASH; if (s == "AHIST") return AHIST; if (s == "HIST") return HIST;
  if (s == "BATCHS") return BATCHS; if (s == "BATCHWA") return BATCHWA; return SERIAL;
}
static RankType parse_rank(const std::string& s) {
  if (s == "SIDE") return SIDE; if (s == "COCORE") return COCORE; if (s == "ACOCORE") return ACOCORE;
  if (s == "DEG") return DEG; return ADEG;
}
static PeelType parse_peel(const std::string& s) {
  if (s == "SORT") return PSORT; if (s == "HASH") return PHASH; if (s == "HIST") return PHIST;
  if (s == "BATCHWA") return PBATCHWA; return PBATCHS;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s side|vert|edge|total|peelv|peele graph out.bin [opts]\n", argv[0]);
    return 2;
  }
  std::string op = argv[1];
  char* gfile = argv[2];
  const char* ofile = argv[3];
  std::string ct = "BATCHS", rt = "ADEG", pt = "BATCHS";
  long rounds = 1;
  long max_array_size = 23090996160L;  // Compute's default -a (BipartiteButterflies.C:175)
  long max_wedges = 2577500000L;       // Compute's default -m (BipartiteButterflies.C:174)
  std::string sparse_type = "EDGE";   // op sparse: EDGE (eSparseBipartite) or COLOR
  long denom = 25;                     // Compute's default -d (BipartiteButterflies.C:176)
  long seed = 0;
  for (int i = 4; i + 1 < argc; i += 2) {
    std::string k = argv[i], v = argv[i + 1];
    if (k == "-countType") ct = v;
    else if (k == "-rankType") rt = v;
    else if (k == "-peelType") pt = v;
    else if (k == "-rounds") rounds = atol(v.c_str());
    else if (k == "-a") max_array_size = atol(v.c_str());
    else if (k == "-m") max_wedges = atol(v.c_str());
    else if (k == "-sparseType") sparse_type = v;
    else if (k == "-d") denom = atol(v.c_str());
    else if (k == "-seed") seed = atol(v.c_str());
  }
  CountType ty = parse_count(ct);
  RankType tw = parse_rank(rt);
  PeelType tp = parse_peel(pt);
  const size_t stride = 64 / sizeof(long);

  std::vector<double> secs, peel_secs;
  std::vector<long> counts, peel;
  long use_v_out = 0, side_w = 0;
  if (op == "vertlist") {
    // the graphs are read before the timed loop; each is counted as Compute does (side
    // choice, then Count with the run's count and rank types), its total = sum over S / 2
    std::vector<std::string> files;
    {
      FILE* lf = fopen(gfile, "r");
      if (!lf) { perror("open list"); return 1; }
      char line[4096];
      while (fgets(line, sizeof(line), lf)) {
        std::string f(line);
        while (!f.empty() && (f.back() == '\n' || f.back() == '\r')) f.pop_back();
        if (!f.empty()) files.push_back(f);
      }
      fclose(lf);
    }
    std::vector<bipartiteCSR> gs;
    for (auto& f : files) gs.push_back(load_graph(const_cast<char*>(f.c_str())));
    for (long r = 0; r < rounds; r++) {
      counts.clear();
      double t0 = now_s();
      for (auto& G : gs) {
        tuple<bool, long> uv = cmpWedgeCounts(G);
        bool use_v = get<0>(uv);
        long n_s = use_v ? G.nu : G.nv;
        long total = 0;
        if (G.numEdges) {
          long* b = Count(G, use_v, get<1>(uv), max_wedges, max_array_size, ty, tw);
          for (long i = 0; i < n_s; i++) total += b[stride * i];
          free(b);
        }
        counts.push_back(total / 2);
      }
      secs.push_back(now_s() - t0);
    }
    for (auto& G : gs) G.del();
    rounds = 0;
  }
  for (long r = 0; r < rounds; r++) {
    // The reference re-reads the graph every round (BipartiteButterflies.C:331-341),
    // because PeelE mutates it; the read is outside the timed region.
    bipartiteCSR G = load_graph(gfile);
    tuple<bool, long> uv = cmpWedgeCounts(G);
    bool use_v = get<0>(uv);
    use_v_out = use_v; side_w = get<1>(uv);
    long n_s = use_v ? G.nu : G.nv;
    counts.clear(); peel.clear();
    if (op == "side") {
      secs.push_back(0.0);
    } else if (op == "rank") {
      // The ranking step of Count (butterfly_count.h:1640-1662): order, then the wedge count
      // num_ccwedges under that order.  Dumped as counts = order[n], peel = {W}.
      double t0 = now_s();
      tuple<uintE*, uintE*, uintE*> rt3;
      if (tw == COCORE) rt3 = getCoCoreRanks(G);
      else if (tw == ACOCORE) rt3 = getApproxCoCoreRanks(G);
      else if (tw == DEG) rt3 = getDegRanks(G);
      else rt3 = getApproxDegRanks(G);
      long W = sequence::reduce<long>((long)0, G.nu, addF<long>(),
                                      rankWedgeF<long>(G.offsetsU, G.edgesU, get<2>(rt3), get<1>(rt3)));
      W += sequence::reduce<long>((long)0, G.nv, addF<long>(),
                                  rankWedgeF<long>(G.offsetsV, G.edgesV, get<1>(rt3), get<2>(rt3)));
      secs.push_back(now_s() - t0);
      counts.resize(G.nv + G.nu);
      for (long i = 0; i < G.nv + G.nu; i++) counts[i] = get<0>(rt3)[i];
      peel.push_back(W);
      free(get<0>(rt3)); free(get<1>(rt3)); free(get<2>(rt3));
    } else if (op == "total") {
      double t0 = now_s();
      long nb = CountTotal(G, use_v, side_w, max_wedges, max_array_size, ty, tw);
      secs.push_back(now_s() - t0);
      counts.push_back(nb);
    } else if (op == "vert" || op == "peelv") {
      double t0 = now_s();
      long* b = Count(G, use_v, side_w, max_wedges, max_array_size, ty, tw);
      secs.push_back(now_s() - t0);
      counts.resize(n_s);
      for (long i = 0; i < n_s; i++) counts[i] = b[stride * i];
      if (op == "peelv") {
        double t1 = now_s();
        auto D = Peel(G, use_v, b, max_wedges, tp, max_array_size);
        peel_secs.push_back(now_s() - t1);
        peel.resize(n_s);
        for (long i = 0; i < n_s; i++) peel[i] = D.s[i];

      }
      free(b);
    } else if (op == "edge" || op == "peele") {
      uintE* eti = edgeToIdxs(G, use_v);
      double t0 = now_s();
      long* b = CountE(eti, G, use_v, side_w, max_wedges, max_array_size, ty, tw);
      secs.push_back(now_s() - t0);
      counts.resize(G.numEdges);
      for (long i = 0; i < G.numEdges; i++) counts[i] = b[stride * i];
      if (op == "peele") {
        double t1 = now_s();
        uintE* ite = idxsToEdge(G, use_v);
        auto D = PeelE(eti, ite, G, use_v, b, max_wedges, tp);
        peel_secs.push_back(now_s() - t1);
        peel.resize(G.numEdges);
        for (long i = 0; i < G.numEdges; i++) peel[i] = D.s[i];

        free(ite);
      }
      free(eti);
      free(b);
    } else if (op == "sparse") {
      // eSparseBipartite / clrSparseBipartite (butterfly_utils.h:796-996, with delZeroDeg
      // :744-785), then CountTotal of the sparsified graph and the CLI's scaling
      // (BipartiteButterflies.C:213-214).  counts = {nv, nu, m, offsetsV[nv+1], edgesV[m],
      // offsetsU[nu+1], edgesU[m]} of the sparsified graph; peel = {total, scaled total}.
      double t0 = now_s();
      bipartiteCSR H = sparse_type == "COLOR" ? clrSparseBipartite(G, denom, seed) : eSparseBipartite(G, denom, seed);
      secs.push_back(now_s() - t0);
      counts.push_back(H.nv); counts.push_back(H.nu); counts.push_back(H.numEdges);
      for (long i = 0; i <= H.nv; i++) counts.push_back(H.offsetsV[i]);
      for (long i = 0; i < H.numEdges; i++) counts.push_back(H.edgesV[i]);
      for (long i = 0; i <= H.nu; i++) counts.push_back(H.offsetsU[i]);
      for (long i = 0; i < H.numEdges; i++) counts.push_back(H.edgesU[i]);
      tuple<bool, long> huv = cmpWedgeCounts(H);
      long nb = H.numEdges ? CountTotal(H, get<0>(huv), get<1>(huv), max_wedges, max_array_size, ty, tw) : 0;
      peel.push_back(nb);
      if (sparse_type == "COLOR") nb *= pow(denom, 3);
      else nb *= pow(denom, 4);
      peel.push_back(nb);
      H.del();
    } else {
      fprintf(stderr, "unknown op %s\n", op.c_str());
      return 2;
    }
    G.del();
  }

  FILE* f = fopen(ofile, "wb");
  if (!f) { perror("open out"); return 1; }
  int64_t hdr[3] = {use_v_out, side_w, (int64_t)counts.size()};
  fwrite(hdr, 8, 3, f);
  if (!counts.empty()) fwrite(counts.data(), 8, counts.size(), f);
  int64_t np = (int64_t)peel.size();
  fwrite(&np, 8, 1, f);
  if (np) fwrite(peel.data(), 8, peel.size(), f);
  fclose(f);

  printf("{\"op\": \"%s\", \"countType\": \"%s\", \"rankType\": \"%s\", \"peelType\": \"%s\", "
         "\"workers\": %d, \"seconds\": [", op.c_str(), ct.c_str(), rt.c_str(), pt.c_str(), getWorkers());
  for (size_t i = 0; i < secs.size(); i++) printf("%s%.6f", i ? ", " : "", secs[i]);
  printf("], \"peel_seconds\": [");
  for (size_t i = 0; i < peel_secs.size(); i++) printf("%s%.6f", i ? ", " : "", peel_secs[i]);
  printf("]}\n");
  return 0;
}
