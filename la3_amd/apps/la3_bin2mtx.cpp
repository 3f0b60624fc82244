// la3_bin2mtx: LA3 binary records to a text edge list, with the flags of the
// reference's tool (src/tools/bin2mtx.cpp:14-28):
//   la3_bin2mtx <in> <out> [-hi|-hio] [-wi] [-wo|-wor]
//   -hi   the file starts with the 16-byte header (u32 n, u32 m, u64 nnz);
//         -hio also writes it as a first line "n m nnz"
//   -wi   records are triples (u32 weight)
//   -wo   write a weight per line (the record's, else 1); -wor writes
//         1 + rand() % 128 after srand(0) instead (bin2mtx.cpp:80).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>

int main(int argc, char* argv[]) {
  if (argc < 3) {
    fprintf(stderr, "Usage: %s <filepath_in> <filepath_out> [-hi[o]] [-wi] [-wo[r]]\n", argv[0]);
    return 1;
  }
  bool hin = false, hout = false, win = false, wout = false, wrand = false;
  for (int a = 3; a < argc; a++) {
    const std::string f = argv[a];
    if (f == "-hi" || f == "-hio") hin = true;
    if (f == "-hio") hout = true;
    if (f == "-wi") win = true;
    if (f == "-wo" || f == "-wor") wout = true;
    if (f == "-wor") wrand = true;
  }
  FILE* in = fopen(argv[1], "rb");
  if (!in) { perror(argv[1]); return 1; }
  FILE* out = fopen(argv[2], "w");
  if (!out) { perror(argv[2]); return 1; }
  if (hin) {
    uint32_t n = 0, m = 0;
    uint64_t nnz = 0;
    if (fread(&n, 4, 1, in) != 1 || fread(&m, 4, 1, in) != 1 || fread(&nnz, 8, 1, in) != 1) {
      fprintf(stderr, "short header\n");
      return 1;
    }
    printf("Header: %u %u %lu\n", n, m, (unsigned long)nnz);
    if (hout) fprintf(out, "%u %u %lu\n", n, m, (unsigned long)nnz);
  }
  srand(0);
  uint32_t rec[3];
  const size_t width = win ? 3 : 2;
  while (fread(rec, 4, width, in) == width) {
    uint32_t w = win ? rec[2] : 1;
    if (wout && wrand) w = 1 + (rand() % 128);
    if (wout) fprintf(out, "%u %u %u\n", rec[0], rec[1], w);
    else fprintf(out, "%u %u\n", rec[0], rec[1]);
  }
  fclose(in);
  if (fclose(out) != 0) { perror(argv[2]); return 1; }
  return 0;
}
