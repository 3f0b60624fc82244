// la3_rmat — writes a seeded R-MAT graph as an LA3 binary edge file, the
// input format of Graph::load_binary (src/matrix/graph.hpp:38-117):
//   la3_rmat <scale> <edgefactor> <seed> <out> [-w] [-ho] [-a A -b B -c C]
//   -w    weighted: u32 triples (src, dst, weight), weight uniform in
//         [1, 128]; default u32 pairs (src, dst)
//   -ho   a 16-byte header first (u32 n, u32 n, u64 nedges), as mtx2bin -hio
//         writes it (src/tools/mtx2bin.cpp:80-82)
//   -a/-b/-c  quadrant probabilities (default 0.57 / 0.19 / 0.19)
// Vertex ids 1 .. 2^scale, self-loops and duplicate edges removed, records in
// the generator's order; the same records as la3_rmat_generate (include/la3.h)
// and la3_amd.rmat for the same arguments. Prints "n nedges" on stdout.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/la3.h"

int main(int argc, char** argv) {
  if (argc < 5) {
    fprintf(stderr, "Usage: %s <scale> <edgefactor> <seed> <out> [-w] [-ho] [-a A -b B -c C]\n", argv[0]);
    return 1;
  }
  const uint32_t scale = (uint32_t)atoi(argv[1]), ef = (uint32_t)atoi(argv[2]);
  const uint64_t seed = strtoull(argv[3], nullptr, 10);
  const char* path = argv[4];
  bool weighted = false, header = false;
  double a = 0.57, b = 0.19, c = 0.19;
  for (int i = 5; i < argc; i++) {
    const std::string f = argv[i];
    if (f == "-w") weighted = true;
    else if (f == "-ho") header = true;
    else if ((f == "-a" || f == "-b" || f == "-c") && i + 1 < argc) {
      const double v = atof(argv[++i]);
      (f == "-a" ? a : f == "-b" ? b : c) = v;
    } else {
      fprintf(stderr, "unknown option %s\n", argv[i]);
      return 1;
    }
  }
  if (scale == 0 || scale > 31 || ef == 0) {
    fprintf(stderr, "scale must be in [1, 31] and edgefactor positive\n");
    return 1;
  }
  uint64_t nrec = 0;
  la3_status st = la3_rmat_generate(scale, ef, a, b, c, seed, weighted, nullptr, &nrec);
  if (st) {
    fprintf(stderr, "la3_rmat_generate failed (%d)\n", (int)st);
    return 1;
  }
  const size_t width = weighted ? 3 : 2;
  std::vector<uint32_t> rec(nrec * width);
  st = la3_rmat_generate(scale, ef, a, b, c, seed, weighted, rec.data(), &nrec);
  if (st) {
    fprintf(stderr, "la3_rmat_generate failed (%d)\n", (int)st);
    return 1;
  }
  FILE* out = fopen(path, "wb");
  if (!out) {
    perror(path);
    return 1;
  }
  const uint32_t n = 1u << scale;
  bool ok = true;
  if (header) {
    const uint64_t m = nrec;
    ok = fwrite(&n, 4, 1, out) == 1 && fwrite(&n, 4, 1, out) == 1 && fwrite(&m, 8, 1, out) == 1;
  }
  ok = ok && fwrite(rec.data(), 4, rec.size(), out) == rec.size();
  ok = (fclose(out) == 0) && ok;
  if (!ok) {
    fprintf(stderr, "write failed: %s\n", path);
    return 1;
  }
  printf("%u %llu\n", n, (unsigned long long)nrec);
  return 0;
}
