// konect_to_hyperadj: converts a KONECT bipartite edge list into the AdjacencyHypergraph text
// that readBipartite reads, as the reference's utils/KONECTtoHyperAdj.C does (unweighted form:
// columns after the first two are ignored).  The binary CSR layout is written with -b.
//
//   konect_to_hyperadj [-b] <KONECT file> <output file>
#include <cstdio>
#include <cstring>

#include "parbutterfly.h"

int main(int argc, char** argv) {
  const bool binary = argc == 4 && !strcmp(argv[1], "-b");
  if (argc != 3 && !binary) {
    fprintf(stderr, "usage: %s [-b] <KONECT file> <output file>\n", argv[0]);
    return 2;
  }
  const char* in = argv[argc - 2];
  const char* out = argv[argc - 1];
  pbf_graph g{};
  int st = pbf_read_konect(in, &g);
  if (st) {
    fprintf(stderr, "readKONECT: %s (%d)\n", pbf_status_str(st), st);
    return st;
  }
  st = binary ? pbf_write_binary(out, &g.csr) : pbf_write_bipartite(out, &g.csr);
  pbf_graph_free(&g);
  if (st) {
    fprintf(stderr, "write: %s (%d)\n", pbf_status_str(st), st);
    return st;
  }
  return 0;
}
