// Device kernels for the vertex-program path on gfx950 (wave64, CDNA4).
//
// Data layout in HBM (one graph = one rank's tile set, src/matrix/matrix2d.hpp:6-8
// at np=1): every per-vertex array is indexed by SLOT, the reference's
// vertex-state order [regular | sink | source | isolated]
// (src/utils/locator.h:84-120). Non-empty columns c are numbered
// [regular | source], so slot(c) = c for regular and c + nsink for sources;
// rows are slots [0, nregular + nsink).
//   col_ptr[ncols+1], col_split[ncols], col_rows[nnz], col_w[nnz]
//       DCSC (src/utils/csc.h:61-73). Entries of a column are sorted by row
//       slot, so [col_ptr, col_split) are the regular-row entries (the
//       reference's `csc`) and [col_split, col_ptr+1) the sink-row entries
//       (`sink_csc`). Column streaming is coalesced: one wavefront reads 64
//       consecutive entries.
//   row_ptr[nrows+1], row_split[nrows], row_cols[nnz], row_w[nnz]
//       Row-major companion for pull steps. A row's entries are in the
//       reference's message order: source columns, then regular columns
//       (src/vprogram/vertex_program_execute.hpp:477-482); row_split marks
//       the first regular column. row_cols holds the sender's slot.
//   rank_of[nslots], row_rank[nnz]
//       Message rows of the narrow PPR kernels (at most 32 queries): slots
//       by decreasing out-degree (ties by slot), rank_of[slot] = message row,
//       row_rank[j] = rank_of[row_cols[j]]. The senders of most entries then
//       share few cache lines.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace la3 {

struct DG {
  uint32_t nslots, nreg, nsink, nsrc, ncols, nrows;
  uint64_t nnz;
  const uint32_t* col_ptr;
  const uint32_t* col_split;
  const uint32_t* col_rows;
  const uint32_t* col_w;
  const uint32_t* row_ptr;
  const uint32_t* row_split;
  const uint32_t* row_cols;
  const uint32_t* row_w;
  const uint32_t* rank_of;
  const uint32_t* row_rank;
  const uint32_t* id_of;
  const uint32_t* hid_of;
  const uint32_t* long_rows;
  uint32_t nlong;
  const uint32_t* slot_of_id;
  // batched-PPR row segments (see HostLayout)
  const uint32_t* plong;
  const uint32_t* pseg_first;
  const uint32_t* pseg_row;
  const uint32_t* pseg_beg;
  uint32_t nplong, nplong_reg, npseg, npseg_reg;
};

}  // namespace la3
