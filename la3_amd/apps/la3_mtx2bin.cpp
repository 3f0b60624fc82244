// la3_mtx2bin: text edge list (Matrix Market body: "i j" or "i j w" lines)
// to the LA3 binary record format, with the flags of the reference's tool
// (src/tools/mtx2bin.cpp:14-30):
//   la3_mtx2bin <in> <out> [-hi|-hio] [-wi] [-woi|-wod]
//   -hi   the first non-comment line is a header "n m nnz"; -hio also writes
//         it as a 16-byte binary header (u32 n, u32 m, u64 nnz)
//   -wi   lines carry a weight (int or double)
//   -woi / -wod  write u32 / f64 weights; without -wi each edge gets
//         1 + rand() % 128 after srand(0) (mtx2bin.cpp:89,103), so on the
//         same C library the weights match the reference tool's.
// Lines starting with '%' or '#' before the data are skipped. The header is
// read from the first non-comment line (the reference seeks back into the
// last comment line there, which misreads files that have comments).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

int main(int argc, char* argv[]) {
  if (argc < 3) {
    fprintf(stderr, "Usage: %s <filepath_in> <filepath_out> [-hi[o]] [-wi] [-wo{i|d}]\n", argv[0]);
    return 1;
  }
  bool hin = false, hout = false, win = false, wout = false, wint = false;
  for (int a = 3; a < argc; a++) {
    const std::string f = argv[a];
    if (f == "-hi" || f == "-hio") hin = true;
    if (f == "-hio") hout = true;
    if (f == "-wi") win = true;
    if (f == "-woi" || f == "-wod") wout = true;
    if (f == "-woi") wint = true;
  }
  const bool wrand = !win && wout;
  FILE* in = fopen(argv[1], "r");
  if (!in) { perror(argv[1]); return 1; }
  FILE* out = fopen(argv[2], "wb");
  if (!out) { perror(argv[2]); return 1; }
  char line[4096];
  bool header_done = !hin;
  srand(0);
  uint64_t written = 0;
  while (fgets(line, sizeof(line), in)) {
    const char* p = line;
    while (*p == ' ' || *p == '\t') p++;
    if (*p == '%' || *p == '#' || *p == '\n' || *p == '\0') continue;
    if (!header_done) {
      unsigned long n = 0, m = 0, nnz = 0;
      if (sscanf(p, "%lu %lu %lu", &n, &m, &nnz) != 3) { fprintf(stderr, "bad header: %s", line); return 1; }
      printf("Header: %lu %lu %lu\n", n, m, nnz);
      if (hout) {
        const uint32_t n32 = (uint32_t)n, m32 = (uint32_t)m;
        const uint64_t z = nnz;
        fwrite(&n32, 4, 1, out);
        fwrite(&m32, 4, 1, out);
        fwrite(&z, 8, 1, out);
      }
      header_done = true;
      continue;
    }
    unsigned long i = 0, j = 0;
    double w = 1;
    const int got = win ? sscanf(p, "%lu %lu %lf", &i, &j, &w) : sscanf(p, "%lu %lu", &i, &j);
    if (got < 2 || (win && got < 3)) { fprintf(stderr, "bad line: %s", line); return 1; }
    const uint32_t ij[2] = {(uint32_t)i, (uint32_t)j};
    fwrite(ij, 4, 2, out);
    if (wout) {
      if (wrand) w = 1 + (rand() % 128);
      if (wint) {
        const uint32_t wi = (uint32_t)w;
        fwrite(&wi, 4, 1, out);
      } else {
        fwrite(&w, 8, 1, out);
      }
    }
    written++;
  }
  fclose(in);
  if (fclose(out) != 0) { perror(argv[2]); return 1; }
  printf("Records: %lu\n", (unsigned long)written);
  return 0;
}
