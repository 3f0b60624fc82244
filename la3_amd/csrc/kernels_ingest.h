// Device ingest: LA3 binary records -> the graph layout of kernels.h, built
// on the GPU. Same result, byte for byte, as the host builder
// (host_build.cpp, which restates src/matrix/graph.hpp:151-191,
// processed_matrix2d.hpp:182-229, locator.h:84-120, csc_matrix2d.hpp:60-72
// and csc.h:94-125):
//   1. drop self-loops, orient (directed: row = destination; acyclic: row =
//      larger id, or smaller when reversed), hash, undirected both ways, in
//      record order (positions from a prefix sum of the kept records);
//   2. classes from row / column occupancy; slots by a prefix sum per class
//      in increasing hashed id;
//   3. entries sorted by (column, row slot) with a stable radix sort of the
//      record-order index, so the first record of a duplicate edge comes
//      first and is the one kept (first wins);
//   4. row-major companion sorted by (row slot, message-order key of the
//      column): source columns first, then regular columns;
//   5. long rows (> LONG_ROW entries) and the PPR segments of rows with more
//      than PSEG_ENTRIES entries.
// Included once, by capi.hip.
#pragma once

#include "kernels_dev.h"

namespace la3 {

struct IngestHash {
  uint32_t bucket;
  uint64_t nparts, height, max_range;
  __device__ uint32_t hash(uint32_t v) const {
    if (!bucket || v >= max_range) return v;
    return (uint32_t)(v / nparts + (uint64_t)(v % (uint32_t)nparts) * height);
  }
  __device__ uint32_t unhash(uint32_t v) const {
    if (!bucket || v >= max_range) return v;
    return (uint32_t)(v / height + (v % height) * nparts);
  }
};

#define GRID_STRIDE(i, n) \
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

// keep[i] = stored entries of record i (0, 1 or 2); *bad = 1 if an id > n.
__global__ void k_ing_keep(const uint32_t* rec, uint64_t nrec, uint32_t rw, uint32_t n, uint32_t per,
                           uint32_t* keep, uint32_t* bad) {
  GRID_STRIDE(i, nrec) {
    const uint32_t u = rec[i * rw], v = rec[i * rw + 1];
    if (u > n || v > n) atomicOr(bad, 1u);
    keep[i] = (u != v) ? per : 0;
  }
}

// mode flags: directed, reverse, acyclic (host_build.cpp step 2).
__global__ void k_ing_orient(const uint32_t* rec, uint64_t nrec, uint32_t rw, const uint32_t* pos, int directed,
                             int reverse, int acyclic, IngestHash h, uint32_t* R, uint32_t* C, uint32_t* W) {
  GRID_STRIDE(i, nrec) {
    uint32_t row = rec[i * rw], col = rec[i * rw + 1];
    if (row == col) continue;
    if (directed && !reverse) { const uint32_t t = row; row = col; col = t; }
    if (acyclic && ((!reverse && col > row) || (reverse && col < row))) { const uint32_t t = row; row = col; col = t; }
    row = h.hash(row);
    col = h.hash(col);
    const uint32_t o = pos[i];
    R[o] = row;
    C[o] = col;
    if (W) W[o] = rec[i * rw + 2];
    if (!directed) {
      R[o + 1] = col;
      C[o + 1] = row;
      if (W) W[o + 1] = rec[i * rw + 2];
    }
  }
}

__global__ void k_ing_occupy(const uint32_t* R, const uint32_t* C, uint64_t m, uint8_t* rowf, uint8_t* colf) {
  GRID_STRIDE(i, m) {
    rowf[R[i]] = 1;
    colf[C[i]] = 1;
  }
}

// cls[4][i]: indicator of class q (0 regular, 1 sink, 2 source, 3 isolated).
__global__ void k_ing_class(const uint8_t* rowf, const uint8_t* colf, uint32_t S, uint32_t* cls) {
  GRID_STRIDE(i, S) {
    const bool r = rowf[i], c = colf[i];
    const uint32_t q = r ? (c ? 0 : 1) : (c ? 2 : 3);
    for (uint32_t x = 0; x < 4; x++) cls[(uint64_t)x * S + i] = q == x;
  }
}

// slot of hashed id i = offset of its class + its rank in the class.
__global__ void k_ing_slots(const uint8_t* rowf, const uint8_t* colf, uint32_t S, const uint32_t* rank,
                            uint32_t off1, uint32_t off2, uint32_t off3, IngestHash h, uint32_t* slot,
                            uint32_t* hid_of, uint32_t* id_of, uint32_t* slot_of_id) {
  GRID_STRIDE(i, S) {
    const bool r = rowf[i], c = colf[i];
    const uint32_t q = r ? (c ? 0 : 1) : (c ? 2 : 3);
    const uint32_t base = q == 0 ? 0 : q == 1 ? off1 : q == 2 ? off2 : off3;
    const uint32_t p = base + rank[(uint64_t)q * S + i];
    slot[i] = p;
    hid_of[p] = (uint32_t)i;
    const uint32_t id = h.unhash((uint32_t)i);
    id_of[p] = id;
    slot_of_id[id] = p;
  }
}

// Sort keys of the entries: (column index << b) | row slot, value = record
// order index.
__global__ void k_ing_colkeys(const uint32_t* R, const uint32_t* C, uint64_t m, const uint32_t* slot, uint32_t nreg,
                              uint32_t nsink, uint32_t b, uint64_t* keys, uint32_t* idx) {
  GRID_STRIDE(i, m) {
    const uint32_t cs = slot[C[i]];
    const uint32_t col = cs < nreg ? cs : cs - nsink;
    keys[i] = ((uint64_t)col << b) | slot[R[i]];
    idx[i] = (uint32_t)i;
  }
}

// First occurrence of each key in the sorted order.
__global__ void k_ing_firsts(const uint64_t* keys, uint64_t m, uint32_t* first) {
  GRID_STRIDE(i, m) first[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_ing_compact(const uint64_t* keys, const uint32_t* idx, const uint32_t* first, const uint32_t* upos,
                              uint64_t m, uint32_t b, const uint32_t* W, uint32_t* col_rows, uint32_t* col_w,
                              uint32_t* col_of) {
  const uint64_t mask = (1ull << b) - 1;
  GRID_STRIDE(i, m) {
    if (!first[i]) continue;
    const uint32_t o = upos[i];
    col_rows[o] = (uint32_t)(keys[i] & mask);
    col_of[o] = (uint32_t)(keys[i] >> b);
    if (W) col_w[o] = W[idx[i]];
  }
}

// ptr[g] = first entry of group g for groups given by a sorted group id per
// entry (every group in [0, ngroups) is non-empty); ptr[ngroups] = nnz.
__global__ void k_ing_ptr(const uint32_t* group_of, uint64_t nnz, uint32_t ngroups, uint32_t* ptr) {
  GRID_STRIDE(j, nnz) {
    if (j == 0 || group_of[j] != group_of[j - 1]) ptr[group_of[j]] = (uint32_t)j;
    if (j == 0) ptr[ngroups] = (uint32_t)nnz;
  }
}

// split[g] = first entry of group g whose value >= thr (values ascending in
// a group), else ptr[g + 1].
__global__ void k_ing_split(const uint32_t* group_of, const uint32_t* val, const uint32_t* ptr, uint64_t nnz,
                            uint32_t ngroups, uint32_t thr, uint32_t* split) {
  GRID_STRIDE(g, ngroups) split[g] = ptr[g + 1];
}
__global__ void k_ing_split2(const uint32_t* group_of, const uint32_t* val, const uint32_t* ptr, uint64_t nnz,
                             uint32_t thr, uint32_t* split) {
  GRID_STRIDE(j, nnz) {
    const uint32_t g = group_of[j];
    if (val[j] >= thr && (j == ptr[g] || val[j - 1] < thr)) split[g] = (uint32_t)j;
  }
}

// Row-major sort keys: (row slot << b) | message-order key of the column.
__global__ void k_ing_rowkeys(const uint32_t* col_rows, const uint32_t* col_of, uint64_t nnz, uint32_t nreg,
                              uint32_t nsrc, uint32_t b, uint64_t* keys, uint32_t* idx) {
  GRID_STRIDE(j, nnz) {
    const uint32_t c = col_of[j];
    const uint32_t mk = c >= nreg ? c - nreg : c + nsrc;
    keys[j] = ((uint64_t)col_rows[j] << b) | mk;
    idx[j] = (uint32_t)j;
  }
}

// row_cols = sender slot, row_w, and the row / message-key parts of the key.
__global__ void k_ing_rowfill(const uint64_t* keys, const uint32_t* idx, uint64_t nnz, uint32_t b,
                              const uint32_t* col_of, const uint32_t* col_w, uint32_t nreg, uint32_t nsink,
                              uint32_t* row_cols, uint32_t* row_w, uint32_t* row_of, uint32_t* mkey) {
  const uint64_t mask = (1ull << b) - 1;
  GRID_STRIDE(x, nnz) {
    const uint32_t j = idx[x];
    const uint32_t c = col_of[j];
    row_cols[x] = c < nreg ? c : c + nsink;
    if (row_w) row_w[x] = col_w[j];
    row_of[x] = (uint32_t)(keys[x] >> b);
    mkey[x] = (uint32_t)(keys[x] & mask);
  }
}

// Per row r < nrows: flag and length of rows longer than `lim`.
__global__ void k_ing_longflag(const uint32_t* row_ptr, uint32_t nrows, uint32_t lim, uint32_t* flag) {
  GRID_STRIDE(r, nrows) flag[r] = row_ptr[r + 1] - row_ptr[r] > lim;
}

__global__ void k_ing_compact_ids(const uint32_t* flag, const uint32_t* pos, uint32_t n, uint32_t* out) {
  GRID_STRIDE(r, n) if (flag[r]) out[pos[r]] = (uint32_t)r;
}

// Segments per PPR long row (plong[i]) and their counts / entry sums.
__global__ void k_ing_segcount(const uint32_t* plong, uint32_t np, const uint32_t* row_ptr, uint32_t seg,
                               uint32_t* nseg) {
  GRID_STRIDE(i, np) {
    const uint32_t r = plong[i];
    nseg[i] = (row_ptr[r + 1] - row_ptr[r] + seg - 1) / seg;
  }
}

__global__ void k_ing_segfill(const uint32_t* plong, uint32_t np, const uint32_t* row_ptr, const uint32_t* first,
                              uint32_t seg, uint32_t* pseg_row, uint32_t* pseg_beg) {
  GRID_STRIDE(i, np) {
    const uint32_t r = plong[i];
    for (uint32_t s = first[i], k = 0; s < first[i + 1]; s++, k++) {
      pseg_row[s] = r;
      pseg_beg[s] = row_ptr[r] + k * seg;
    }
  }
}

// Sums over rows r < lim_row of the lengths of flagged rows (entries of long
// regular rows): one atomic per wavefront.
__global__ void k_ing_lensum(const uint32_t* rows, uint32_t n, const uint32_t* row_ptr, uint32_t lim_row,
                             unsigned long long* out) {
  unsigned long long s = 0;
  GRID_STRIDE(i, n) {
    const uint32_t r = rows[i];
    if (r < lim_row) s += row_ptr[r + 1] - row_ptr[r];
  }
  s = wave_sum(s);
  if (lane_id() == 0 && s) atomicAdd(out, s);
}

#undef GRID_STRIDE

// ------------------------------------------ message rows by out-degree ---
// A narrow PPR row sum (at most 32 queries, LPR f32 lanes per sender)
// gathers one message row per entry, and a 128-B line holds the rows of
// 32 / LPR senders. In slot order (hashed ids) the senders of a line are
// unrelated; ordered by decreasing out-degree, the senders of most entries
// share few lines, which stay in L2 (R-MAT scale 24, one query: the 263M
// gathers of an iteration take 2.4 ms against 4.1 ms in a random order).
// The order: (key ~outdeg, value slot) pairs in slot order, stably sorted
// by key; then rank_of[slot] = position.
__global__ void k_msg_keys(DG g, uint64_t* key, uint32_t* slot) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x) {
    uint32_t d = 0;
    if (slot_has_col(g, s)) { const uint32_t c = col_of_slot(g, s); d = g.col_ptr[c + 1] - g.col_ptr[c]; }
    key[s] = ~d;
    slot[s] = s;
  }
}

__global__ void k_msg_rank(const uint32_t* slot, uint32_t n, uint32_t* rank_of) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) rank_of[slot[i]] = i;
}

__global__ void k_msg_identity(uint32_t n, uint32_t* rank_of) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) rank_of[i] = i;
}

__global__ void k_msg_entries(const uint32_t* row_cols, uint64_t nnz, const uint32_t* rank_of, uint32_t* row_rank) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nnz; j += (uint64_t)gridDim.x * blockDim.x)
    row_rank[j] = rank_of[row_cols[j]];
}

}  // namespace la3
