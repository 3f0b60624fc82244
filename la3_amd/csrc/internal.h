// Internal declarations shared by the host graph builder, the HIP kernels and
// the C-ABI implementation. Not installed; the public surface is include/la3.h.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace la3 {

constexpr uint32_t BFS_INF = 255;            // bfs.h:15 (u8 hops)
constexpr uint32_t SSSP_INF = 0x7fffffffu;   // sssp.h:14 (UINT32_MAX / 2)
constexpr uint32_t LONG_ROW = 64;            // rows longer than this get a whole wavefront
constexpr uint32_t CHUNK = 256;              // entries per traversal work item

// SimpleBucketHasher(n, nranks = 1) (src/matrix/hashers.h:39-67).
struct Hasher {
  bool bucket = true;
  int64_t nparts = 128, height = 0, max_range = 0;
  Hasher() = default;
  Hasher(int64_t max_domain, bool bucket_) : bucket(bucket_) {
    height = max_domain / nparts;
    max_range = height * nparts;
  }
  int64_t hash(int64_t v) const {
    if (!bucket || v >= max_range) return v;
    return v / nparts + (int64_t)((uint32_t)v % nparts) * height;
  }
  int64_t unhash(int64_t v) const {
    if (!bucket || v >= max_range) return v;
    return v / height + (v % height) * nparts;
  }
};

// Host-side result of ingest: every per-vertex array is in "slot" order
// [regular | sink | source | isolated], each class by increasing hashed id
// (src/utils/locator.h:84-120). Columns are the non-empty columns, numbered
// [regular | source]; rows are [regular | sink] = slots [0, nrows).
struct HostLayout {
  uint32_t n = 0, nslots = 0, nreg = 0, nsink = 0, nsrc = 0, niso = 0, ncols = 0, nrows = 0;
  bool directed = true, weighted = false;
  int mode = 0;
  Hasher h;
  uint64_t nrecords = 0, nnz = 0, nnz_reg = 0, nnz_sink = 0;
  std::vector<uint32_t> hid_of;     // slot -> hashed id
  std::vector<uint32_t> id_of;      // slot -> original id
  std::vector<uint32_t> slot_of_id; // original id -> slot
  // DCSC over non-empty columns: entries sorted by row slot, so that the
  // regular-row entries [col_ptr, col_split) precede the sink-row entries.
  std::vector<uint32_t> col_ptr, col_split, col_rows, col_w;
  // Row-major companion: per row, in-entries in the reference's message order
  // (source columns first, then regular columns, each ascending);
  // row_cols holds the SLOT of the sending vertex.
  std::vector<uint32_t> row_ptr, row_split, row_cols, row_w;
  std::vector<uint32_t> long_rows;  // rows with more than LONG_ROW in-entries
  // Rows with more than PSEG_ENTRIES in-entries, split into segments of that
  // many entries (batched PPR): row list (regular rows first, then sinks),
  // first segment of each row, and per segment its row and first entry.
  std::vector<uint32_t> plong, pseg_first, pseg_row, pseg_beg;
  uint32_t nplong_reg = 0, npseg_reg = 0;
  uint64_t pnnz_reg = 0;       // entries of the regular rows in plong
  uint64_t nnz_long_reg = 0;   // entries of the regular rows in long_rows
};

constexpr uint32_t PSEG_ENTRIES = 256;

// Builds the layout from LA3 binary records. Returns an empty string on
// success, else an error message (ids beyond n, nnz >= 2^32).
std::string build_layout(const uint32_t* rec, uint64_t nrec, uint32_t n, bool weighted, int mode,
                         bool bucket_hashing, HostLayout& L);

// R-MAT generator (host, multi-threaded, deterministic per seed).
std::vector<uint32_t> rmat_generate(uint32_t scale, uint32_t edgefactor, double a, double b,
                                    double c, uint64_t seed, bool weighted);

int host_threads();

}  // namespace la3
