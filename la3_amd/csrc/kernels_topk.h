// Device top-k of result rows (VertexProgram::topk, vertex_program.hpp:607-680,
// per query row of a batched result). Values are mapped to unsigned keys
// whose order is the requested order (largest first: key order = value
// order; smallest first: complemented), then
//   1. radix select, 8 bits per pass from the top: per row, a 256-bin
//      histogram of the next digit of the keys that match the row's prefix
//      so far (LDS histogram per block, one block = one chunk of one row);
//      the host picks the digit holding the k-th key, until the k-th key T
//      is known exactly with the count of keys above it;
//   2. ordered compaction: keys above T all qualify; keys equal to T
//      qualify in increasing id order until k are taken (block-level
//      ballot scans keep the id order inside a chunk, per-chunk counts and
//      their prefix sums across chunks).
// Included once, by capi.hip.
#pragma once

#include "kernels_dev.h"

namespace la3 {

enum : uint32_t { TK_U8 = 0, TK_U32 = 1, TK_F32 = 2, TK_F64 = 3 };
constexpr uint32_t TK_CHUNK = 65536;  // ids per block (one row)

// Orderable key of element i of a row-major array; `largest` = keep the
// order, else complement it.
__device__ __forceinline__ uint64_t tk_key(const void* v, uint32_t dtype, uint64_t i, bool largest) {
  uint64_t k;
  if (dtype == TK_U8) {
    k = ((const uint8_t*)v)[i];
  } else if (dtype == TK_U32) {
    k = ((const uint32_t*)v)[i];
  } else if (dtype == TK_F32) {
    const uint32_t b = ((const uint32_t*)v)[i];
    k = (b & 0x80000000u) ? (uint32_t)~b : (b | 0x80000000u);
  } else {
    const uint64_t b = ((const uint64_t*)v)[i];
    k = (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
  }
  const uint32_t bits = dtype == TK_U8 ? 8 : dtype == TK_F64 ? 64 : 32;
  const uint64_t all = bits == 64 ? ~0ull : ((1ull << bits) - 1);
  return largest ? k : (~k & all);
}

// hist[row * 256 + digit] += keys of the row with (key & hi_mask) ==
// prefix[row], digit = (key >> shift) & 255. Grid: rows * chunks blocks.
__global__ void __launch_bounds__(BLOCK) k_topk_hist(const void* v, uint32_t dtype, int largest, uint64_t n,
                                                     uint32_t chunks, const uint64_t* prefix, uint64_t hi_mask,
                                                     uint32_t shift, uint32_t* hist) {
  __shared__ uint32_t h[256];
  const uint32_t row = blockIdx.x / chunks, ch = blockIdx.x % chunks;
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t beg = (uint64_t)ch * TK_CHUNK, end = beg + TK_CHUNK < n ? beg + TK_CHUNK : n;
  const uint64_t p = prefix[row];
  for (uint64_t i = beg + threadIdx.x; i < end; i += BLOCK) {
    const uint64_t k = tk_key(v, dtype, (uint64_t)row * n + i, largest != 0);
    if ((k & hi_mask) == p) atomicAdd(&h[(k >> shift) & 255], 1u);
  }
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[(size_t)row * 256 + threadIdx.x], h[threadIdx.x]);
}

// Per (row, chunk): number of keys equal to thr[row].
__global__ void __launch_bounds__(BLOCK) k_topk_count_eq(const void* v, uint32_t dtype, int largest, uint64_t n,
                                                         uint32_t chunks, const uint64_t* thr, uint32_t* ceq) {
  __shared__ uint32_t s;
  const uint32_t row = blockIdx.x / chunks, ch = blockIdx.x % chunks;
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  const uint64_t beg = (uint64_t)ch * TK_CHUNK, end = beg + TK_CHUNK < n ? beg + TK_CHUNK : n;
  const uint64_t t = thr[row];
  uint32_t c = 0;
  for (uint64_t i = beg + threadIdx.x; i < end; i += BLOCK)
    c += tk_key(v, dtype, (uint64_t)row * n + i, largest != 0) == t;
  c = wave_sum(c);
  if (lane_id() == 0 && c) atomicAdd(&s, c);
  __syncthreads();
  if (threadIdx.x == 0) ceq[blockIdx.x] = s;
}

// Writes the qualifying ids of each row into out_ids[row * k + j]: keys
// above thr[row] in slots [0, gtc[row]) (any order, taken from gt_next[row]),
// and keys equal to thr[row] whose rank among the row's equal keys in id
// order (eq_base of the chunk + rank inside it) is below need[row], in slot
// gtc[row] + rank.
__global__ void __launch_bounds__(BLOCK) k_topk_write(const void* v, uint32_t dtype, int largest, uint64_t n,
                                                      uint32_t chunks, const uint64_t* thr, const uint32_t* need,
                                                      const uint32_t* gtc, const uint32_t* eq_base,
                                                      uint32_t* gt_next, uint32_t k, uint32_t* out_ids) {
  __shared__ uint32_t wsum[BLOCK / WAVE];
  __shared__ uint32_t run;  // equal keys of this chunk before the current 256 ids
  const uint32_t row = blockIdx.x / chunks, ch = blockIdx.x % chunks;
  const uint64_t beg = (uint64_t)ch * TK_CHUNK, end = beg + TK_CHUNK < n ? beg + TK_CHUNK : n;
  const uint64_t t = thr[row];
  const uint32_t nd = need[row], base = eq_base[blockIdx.x];
  const uint32_t w = threadIdx.x / WAVE, lane = lane_id();
  if (threadIdx.x == 0) run = 0;
  __syncthreads();
  for (uint64_t i0 = beg; i0 < end; i0 += BLOCK) {  // block-uniform
    const uint64_t i = i0 + threadIdx.x;
    uint64_t key = 0;
    bool in = i < end;
    if (in) key = tk_key(v, dtype, (uint64_t)row * n + i, largest != 0);
    const bool gt = in && key > t, eq = in && key == t;
    if (gt) {
      const uint32_t slot = atomicAdd(&gt_next[row], 1u);
      if (slot < gtc[row]) out_ids[(size_t)row * k + slot] = (uint32_t)i;
    }
    const uint64_t b = __ballot(eq);
    const uint32_t before = __popcll(b & ((1ull << lane) - 1));
    if (lane == 0) wsum[w] = __popcll(b);
    __syncthreads();
    uint32_t off = run;
    for (uint32_t x = 0; x < w; x++) off += wsum[x];
    const uint32_t rank = base + off + before;
    if (eq && rank < nd) out_ids[(size_t)row * k + gtc[row] + rank] = (uint32_t)i;
    __syncthreads();
    if (threadIdx.x == 0) for (uint32_t x = 0; x < BLOCK / WAVE; x++) run += wsum[x];
    __syncthreads();
  }
}

// vals[row * k + j] = v[row * n + ids[row * k + j]] for j < cnt[row].
__global__ void k_topk_gather(const void* v, uint32_t esz, uint64_t n, uint32_t rows, uint32_t k,
                              const uint32_t* cnt, const uint32_t* ids, void* vals) {
  const uint64_t total = (uint64_t)rows * k;
  for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < total; x += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t row = (uint32_t)(x / k), j = (uint32_t)(x % k);
    if (j >= cnt[row]) continue;
    const uint64_t src = (uint64_t)row * n + ids[x];
    if (esz == 1) ((uint8_t*)vals)[x] = ((const uint8_t*)v)[src];
    else if (esz == 4) ((uint32_t*)vals)[x] = ((const uint32_t*)v)[src];
    else ((uint64_t*)vals)[x] = ((const uint64_t*)v)[src];
  }
}

// Per-row reductions of a device-resident result (the checksums the
// reference's drivers print): rows x n values of a type. Block b of row r
// reduces entries [b * RD_CHUNK, (b + 1) * RD_CHUNK) of the row in a fixed
// order (each thread over a fixed stride, then a fixed pairwise tree) into
// part[r * nb + b]; k_row_reduce_fin adds a row's partials in block order.
// Integer sums and counts are exact in 64 bits; floating-point values are
// summed in FP64, in the same order on every run (the last bits may differ
// from a sequential sum). op 0: sum; op 1: count of values != ne.
constexpr uint32_t RD_CHUNK = 1u << 16;

template <class T, class Acc>
__global__ void __launch_bounds__(BLOCK) k_row_reduce(const T* v, uint64_t n, uint32_t nb, uint32_t op, T ne,
                                                      Acc* part) {
  __shared__ Acc red[BLOCK];
  const uint32_t r = blockIdx.x / nb, b = blockIdx.x % nb;
  const uint64_t lo = (uint64_t)b * RD_CHUNK, hi = min(n, lo + RD_CHUNK);
  const T* row = v + (uint64_t)r * n;
  Acc acc = 0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += BLOCK) {
    const T x = row[i];
    acc += op == 0 ? (Acc)x : (Acc)(x != ne);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t h = BLOCK / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

template <class Acc>
__global__ void k_row_reduce_fin(const Acc* part, uint32_t rows, uint32_t nb, Acc* out) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
    Acc acc = 0;
    for (uint32_t b = 0; b < nb; b++) acc += part[(uint64_t)r * nb + b];
    out[r] = acc;
  }
}

// Rows of a result by original id reordered to the vertex-state slot order
// [regular | sink | source | isolated] (each class by hashed id):
// out[r * S + p] = in[r * S + id_of[p]], elements of eb bytes.
__global__ void k_rows_by_slot(const uint32_t* id_of, const uint8_t* in, uint8_t* out, uint64_t S, uint32_t rows,
                               uint32_t eb) {
  const uint64_t total = S * rows;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = i / S, p = i - r * S;
    const uint64_t from = r * S + id_of[p];
    if (eb == 1) out[i] = in[from];
    else if (eb == 2) ((uint16_t*)out)[i] = ((const uint16_t*)in)[from];
    else if (eb == 4) ((uint32_t*)out)[i] = ((const uint32_t*)in)[from];
    else ((uint64_t*)out)[i] = ((const uint64_t*)in)[from];
  }
}

}  // namespace la3
