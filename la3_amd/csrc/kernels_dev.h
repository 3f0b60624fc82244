// Device code. Included once, by capi.hip, so that launches and the
// per-launch event profiler live in the same translation unit.
//
// Conventions for every kernel here:
//  - blocks of 256 threads = 4 wavefronts of 64 lanes;
//  - loops whose body contains a wave-collective (ballot / shfl / emit) have a
//    wave-uniform trip count, so every lane reaches the collective;
//  - no floating-point atomics: sums are per-row sequential or fixed-shape
//    wavefront reductions, so results are run-to-run deterministic.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "internal.h"
#include "kernels.h"

namespace la3 {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (WAVE - 1); }
__device__ __forceinline__ uint32_t col_of_slot(const DG& g, uint32_t s) {
  return s < g.nreg ? s : s - g.nsink;
}
__device__ __forceinline__ uint32_t slot_of_col(const DG& g, uint32_t c) {
  return c < g.nreg ? c : c + g.nsink;
}
__device__ __forceinline__ bool slot_has_col(const DG& g, uint32_t s) {
  return s < g.nreg || (s >= g.nrows && s < g.nrows + g.nsrc);
}
// Position of column c in the reference's per-row message order: source
// columns first, then regular columns (vertex_program_execute.hpp:477-482).
__device__ __forceinline__ uint32_t msg_key(const DG& g, uint32_t c) {
  return c >= g.nreg ? c - g.nreg : c + g.nsrc;
}
__device__ __forceinline__ uint32_t col_of_key(const DG& g, uint32_t k) {
  return k < g.nsrc ? k + g.nreg : k - g.nsrc;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    uint32_t t = __shfl_up(v, d, WAVE);
    if (lane >= (uint32_t)d) v += t;
  }
  return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, WAVE);
  return v;
}

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Traversal frontier bookkeeping, one set per query.
// Counters are spread over NSLOT slots (block b adds into slot b % NSLOT) so
// that no single address takes an atomic from every wavefront; the host sums
// the slots. Work items go through a per-wavefront LDS queue and reach HBM
// with one reservation per QCAP items.
constexpr uint32_t NSLOT = 64;
constexpr uint32_t QCAP = 256;   // items per wavefront LDS queue
constexpr uint32_t BIG_NIT = 4;  // vertices with more items bypass the LDS queue

struct CtrSlot {
  uint32_t verts, cols;          // vertices emitted / with a non-empty column
  unsigned long long edges;      // entries behind the emitted items
  unsigned long long scanned;    // row entries read by pull steps
  uint32_t nmin_inv;             // ~(min distance in the next frontier), via atomicMax
  uint32_t touched;              // rows that received a message (measurement runs only)
};
struct Frontier {
  uint64_t* items;               // work items: (chunk << 32) | slot
  uint32_t* nitems;              // [0] items emitted
  CtrSlot* slots;                // [NSLOT]
};

// Per-wavefront emitter. cnt is wave-uniform; the counters are per lane.
struct Emitter {
  uint64_t* q;
  uint32_t cnt;
  uint32_t verts, cols;
  unsigned long long edges, scanned;
  uint32_t nmin_inv;  // ~(smallest distance this lane put in the next frontier)
  uint32_t touched;   // rows this lane saw receive their first message of the step
};

__device__ __forceinline__ void em_init(Emitter& em, uint64_t* lds) {
  em.q = lds + (threadIdx.x / WAVE) * QCAP;
  em.cnt = 0;
  em.verts = em.cols = 0;
  em.edges = em.scanned = 0;
  em.nmin_inv = 0;
  em.touched = 0;
}

__device__ __forceinline__ void em_flush(Emitter& em, const Frontier& f) {
  if (em.cnt == 0) return;
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(f.nitems, em.cnt);
  base = __shfl(base, 0, WAVE);
  __builtin_amdgcn_wave_barrier();
  for (uint32_t i = lane_id(); i < em.cnt; i += WAVE) f.items[base + i] = em.q[i];
  __builtin_amdgcn_wave_barrier();
  em.cnt = 0;
}

// Every lane of the wave must call this (nit may be 0).
__device__ __forceinline__ void em_push(Emitter& em, const Frontier& f, uint32_t slot, uint32_t nit,
                                        uint32_t nedges, bool counted) {
  em.verts += counted;
  em.cols += nit > 0;
  em.edges += nedges;
  const bool big = nit > BIG_NIT;
  const uint32_t small = big ? 0 : nit;
  const uint32_t incl = wave_incl_scan(small);
  const uint32_t total = __shfl(incl, WAVE - 1, WAVE);
  if (total) {
    if (em.cnt + total > QCAP) em_flush(em, f);
    const uint32_t o = em.cnt + incl - small;
    for (uint32_t k = 0; k < small; k++) em.q[o + k] = ((uint64_t)k << 32) | slot;
    em.cnt += total;
  }
  if (big) {  // hubs: reserve their own range
    const uint32_t base = atomicAdd(f.nitems, nit);
    for (uint32_t k = 0; k < nit; k++) f.items[base + k] = ((uint64_t)k << 32) | slot;
  }
}

// Flushes the queue and adds the block's counters into slot bid % NSLOT
// (bid = the block's index in its launch or in its share of a fused launch).
// Must be called once by every thread of the block (contains a block barrier).
__device__ __forceinline__ void em_finish(Emitter& em, const Frontier& f, uint32_t bid) {
  em_flush(em, f);
  __shared__ unsigned long long red[6][BLOCK / WAVE];
  const uint32_t w = threadIdx.x / WAVE;
  const unsigned long long v0 = wave_sum((unsigned long long)em.verts);
  const unsigned long long v1 = wave_sum((unsigned long long)em.cols);
  const unsigned long long v2 = wave_sum(em.edges);
  const unsigned long long v3 = wave_sum(em.scanned);
  const unsigned long long v5 = wave_sum((unsigned long long)em.touched);
  uint32_t mx = em.nmin_inv;
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, d, WAVE));
  if (lane_id() == 0) {
    red[0][w] = v0; red[1][w] = v1; red[2][w] = v2; red[3][w] = v3; red[4][w] = mx; red[5][w] = v5;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++)
      for (uint32_t j = 0; j < BLOCK / WAVE; j++) t[i] += red[i][j];
    for (uint32_t j = 0; j < BLOCK / WAVE; j++) t[4] = max(t[4], red[4][j]);
    for (uint32_t j = 0; j < BLOCK / WAVE; j++) t[5] += red[5][j];
    CtrSlot& cs = f.slots[bid % NSLOT];
    if (t[0]) atomicAdd(&cs.verts, (uint32_t)t[0]);
    if (t[1]) atomicAdd(&cs.cols, (uint32_t)t[1]);
    if (t[2]) atomicAdd(&cs.edges, t[2]);
    if (t[3]) atomicAdd(&cs.scanned, t[3]);
    if (t[4]) atomicMax(&cs.nmin_inv, (uint32_t)t[4]);
    if (t[5]) atomicAdd(&cs.touched, (uint32_t)t[5]);
  }
}

__device__ __forceinline__ void em_finish(Emitter& em, const Frontier& f) { em_finish(em, f, blockIdx.x); }

// Blocks of a launch, or of one task's share of a fused launch: block bid of
// nb, with the waves numbered bid * 4 + wave of nb * 4.
struct Part {
  uint32_t bid, nb;
  __device__ uint32_t wid() const { return bid * (BLOCK / WAVE) + threadIdx.x / WAVE; }
  __device__ uint32_t nwaves() const { return nb * (BLOCK / WAVE); }
  __device__ uint32_t tid() const { return bid * BLOCK + threadIdx.x; }
  __device__ uint32_t nthreads() const { return nb * BLOCK; }
};
__device__ __forceinline__ Part grid_part() { return Part{blockIdx.x, gridDim.x}; }

#define LA3_EMITTER(em)                                  \
  __shared__ uint64_t em##_lds[(BLOCK / WAVE) * QCAP];   \
  Emitter em;                                            \
  em_init(em, em##_lds)

// ------------------------------------------------------------- utilities ---
template <class T>
__global__ void k_fill(T* p, uint64_t n, T v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Plain copy (16-byte vectors, four loads in flight per lane); the
// known-size kernel for bandwidth checks.
__global__ void __launch_bounds__(BLOCK) la3_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                  uint64_t n) {
  constexpr int V = 4;
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  uint64_t i = blockIdx.x * (uint64_t)BLOCK + threadIdx.x;
  for (; i + (V - 1) * stride < n; i += V * stride) {
    uint4 t[V];
#pragma unroll
    for (int v = 0; v < V; v++) t[v] = src[i + v * stride];
#pragma unroll
    for (int v = 0; v < V; v++) dst[i + v * stride] = t[v];
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

// out[id] = v[slot(id)] for every original id (coalesced writes).
template <class T>
__global__ void k_gather_by_id(DG g, const T* v, T* out) {
  for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < g.nslots; id += gridDim.x * blockDim.x)
    out[id] = v[g.slot_of_id[id]];
}

// Degree: in-degree of the stored graph (DegVertex, degree.h:31-34).
__global__ void k_degree(DG g, uint32_t* out_by_id) {
  for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < g.nslots; id += gridDim.x * blockDim.x) {
    const uint32_t s = g.slot_of_id[id];
    out_by_id[id] = s < g.nrows ? g.row_ptr[s + 1] - g.row_ptr[s] : 0u;
  }
}

// Block 0 clears a queue's counter block (the queue the kernel reads from:
// its counts were already taken), so the next step finds it zeroed without
// a separate launch.
constexpr uint32_t CTR_WORDS = (64 + NSLOT * sizeof(CtrSlot)) / 4;
__device__ __forceinline__ void zero_ctr(uint32_t* zc) {
  if (zc && blockIdx.x == 0)
    for (uint32_t i = threadIdx.x; i < CTR_WORDS; i += blockDim.x) zc[i] = 0;
}

// Out-degree of each slot = length of its column (regular + sink rows).
__global__ void k_outdeg(DG g, uint32_t* deg) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x) {
    uint32_t d = 0;
    if (slot_has_col(g, s)) { uint32_t c = col_of_slot(g, s); d = g.col_ptr[c + 1] - g.col_ptr[c]; }
    deg[s] = d;
  }
}

// ------------------------------------------------------------------- BFS ---
// BfsVertex (bfs.h:27-49) on the non-optimized loop (vertex_program_execute.hpp:166-248):
// regular and sink rows every iteration. level[s] = iteration of discovery
// (u8, 255 = unreached). mark[s] = 1 + message-order key of the winning
// parent column (0 = unvisited, ~0 = root): among the frontier in-neighbours
// of a newly reached row, the reference keeps the last one combined
// (combine: y2 = y1, bfs.h:40), i.e. the largest message-order key.

// The root's own slot is written by the thread that resets that slot, so no
// other thread's reset can land after it.
__global__ void k_bfs_init(DG g, uint8_t* level, uint32_t* mark, uint32_t root, Frontier f, uint32_t* zc_other) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x) {
    level[s] = s == root ? 0 : (uint8_t)BFS_INF;
    mark[s] = s == root ? 0xffffffffu : 0u;
  }
  zero_ctr(zc_other);
  zero_ctr(f.nitems);
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    uint32_t nit = 0, ne = 0;
    if (slot_has_col(g, root)) {
      uint32_t c = col_of_slot(g, root);
      ne = g.col_ptr[c + 1] - g.col_ptr[c];
      nit = (ne + CHUNK - 1) / CHUNK;
    }
    for (uint32_t k = 0; k < nit; k++) f.items[k] = ((uint64_t)k << 32) | root;
    f.nitems[0] = nit;
    f.slots[0].verts = 1;
    f.slots[0].cols = nit > 0;
    f.slots[0].edges = ne;
  }
}

// Top-down (push) step over DCSC columns of the frontier; one work item
// (<= CHUNK entries of one column) per wavefront iteration.
__global__ void __launch_bounds__(BLOCK) k_bfs_push(DG g, const uint64_t* items, uint32_t nitems,
                                                    uint8_t* level, uint32_t* mark, uint32_t iter,
                                                    Frontier nf, uint32_t* zc) {
  zero_ctr(zc);
  const uint32_t nwaves = gridDim.x * (BLOCK / WAVE);
  const uint32_t wid = blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE;
  const uint8_t nxt = (uint8_t)(iter + 1);
  LA3_EMITTER(em);
  for (uint32_t it = wid; it < nitems; it += nwaves) {
    const uint64_t item = items[it];
    const uint32_t s = (uint32_t)item, ch = (uint32_t)(item >> 32);
    const uint32_t c = col_of_slot(g, s);
    const uint32_t beg = g.col_ptr[c] + ch * CHUNK;
    const uint32_t end = min(beg + CHUNK, g.col_ptr[c + 1]);
    const uint32_t key1 = msg_key(g, c) + 1;
    for (uint32_t e0 = beg; e0 < end; e0 += WAVE) {
      const uint32_t e = e0 + lane_id();
      bool disc = false;
      uint32_t r = 0;
      if (e < end) {
        r = g.col_rows[e];
        uint8_t lv = level[r];
        if (lv == (uint8_t)BFS_INF || lv == nxt) {
          uint32_t old = atomicMax(&mark[r], key1);
          if (old == 0) { level[r] = nxt; disc = true; }
        }
      }
      uint32_t nit = 0, ne = 0;
      if (disc && slot_has_col(g, r)) {
        uint32_t cr = col_of_slot(g, r);
        ne = g.col_ptr[cr + 1] - g.col_ptr[cr];
        nit = (ne + CHUNK - 1) / CHUNK;
      }
      em_push(em, nf, r, nit, ne, disc);
    }
  }
  em_finish(em, nf);
}

// Bottom-up (pull) step: each unreached row scans its in-entries from the
// end of the reference's message order; the first frontier member found is
// the last writer. Rows longer than LONG_ROW are left to k_bfs_pull_long.
__global__ void __launch_bounds__(BLOCK) k_bfs_pull(DG g, uint8_t* level, uint32_t* mark,
                                                    uint32_t iter, Frontier nf, uint32_t* zc) {
  zero_ctr(zc);
  const uint8_t cur = (uint8_t)iter, nxt = (uint8_t)(iter + 1);
  const uint32_t stride = gridDim.x * BLOCK;
  LA3_EMITTER(em);
  for (uint32_t base = blockIdx.x * BLOCK; base < g.nrows; base += stride) {
    const uint32_t r = base + threadIdx.x;
    bool disc = false;
    uint32_t scanned = 0;
    if (r < g.nrows && level[r] == (uint8_t)BFS_INF) {
      const uint32_t b = g.row_ptr[r], e = g.row_ptr[r + 1];
      if (e - b <= LONG_ROW) {
        for (uint32_t j = e; j > b; j--) {
          uint32_t src = g.row_cols[j - 1];
          scanned++;
          if (level[src] == cur) { mark[r] = msg_key(g, col_of_slot(g, src)) + 1; disc = true; break; }
        }
      }
    }
    em.scanned += scanned;
    uint32_t nit = 0, ne = 0;
    if (disc) {
      level[r] = nxt;
      if (slot_has_col(g, r)) {
        uint32_t cr = col_of_slot(g, r);
        ne = g.col_ptr[cr + 1] - g.col_ptr[cr];
        nit = (ne + CHUNK - 1) / CHUNK;
      }
    }
    em_push(em, nf, r, nit, ne, disc);
  }
  em_finish(em, nf);
}

// Bottom-up for long rows: one wavefront per row, scanning 64 entries at a
// time from the end; the highest lane with a frontier member wins.
__global__ void __launch_bounds__(BLOCK) k_bfs_pull_long(DG g, uint8_t* level, uint32_t* mark,
                                                         uint32_t iter, Frontier nf) {
  const uint8_t cur = (uint8_t)iter, nxt = (uint8_t)(iter + 1);
  const uint32_t nwaves = gridDim.x * (BLOCK / WAVE);
  const uint32_t wid = blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE;
  LA3_EMITTER(em);
  for (uint32_t i = wid; i < g.nlong; i += nwaves) {
    const uint32_t r = g.long_rows[i];
    if (level[r] != (uint8_t)BFS_INF) continue;  // wave-uniform
    const uint32_t b = g.row_ptr[r], e = g.row_ptr[r + 1];
    bool found = false;
    uint32_t scanned = 0;
    for (uint32_t top = e; top > b && !found; top = (top - b > WAVE) ? top - WAVE : b) {
      const uint32_t lo = (top - b > WAVE) ? top - WAVE : b;
      scanned += top - lo;
      const uint32_t j = lo + lane_id();
      bool hit = false;
      uint32_t src = 0;
      if (j < top) { src = g.row_cols[j]; hit = level[src] == cur; }
      uint64_t m = __ballot(hit);
      if (m) {
        const uint32_t win = 63 - __clzll(m);
        if (lane_id() == win) mark[r] = msg_key(g, col_of_slot(g, src)) + 1;
        found = true;
      }
    }
    if (lane_id() == 0) em.scanned += scanned;
    uint32_t nit = 0, ne = 0;
    bool disc = found && lane_id() == 0;
    if (disc) {
      level[r] = nxt;
      if (slot_has_col(g, r)) {
        uint32_t cr = col_of_slot(g, r);
        ne = g.col_ptr[cr + 1] - g.col_ptr[cr];
        nit = (ne + CHUNK - 1) / CHUNK;
      }
    }
    em_push(em, nf, r, nit, ne, disc);
  }
  em_finish(em, nf);
}

// hops / parent by original id. parent = hashed id of the winning column's
// vertex (Edge.src = csc.colidxs, vertex_program_execute.hpp:593).
__global__ void k_bfs_output(DG g, const uint8_t* level, const uint32_t* mark, uint8_t* hops,
                             uint32_t* parent) {
  for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < g.nslots; id += gridDim.x * blockDim.x) {
    const uint32_t s = g.slot_of_id[id];
    hops[id] = level[s];
    if (parent) {
      uint32_t m = mark[s], p = 0;
      if (m != 0 && m != 0xffffffffu) p = g.hid_of[slot_of_col(g, col_of_key(g, m - 1))];
      parent[id] = p;
    }
  }
}

// ------------------------------------------------------------------ SSSP ---
// SpVertex (sssp.h:26-46) on the optimized loop (vertex_program_execute.hpp:55-163):
// the loop relaxes regular rows only, then one pull pass computes the sinks.
// Relaxations are asynchronous (Bellman-Ford with a deduplicated work
// queue); min-plus distances are unique, so the result equals the
// reference's.

__global__ void k_sssp_init(DG g, uint32_t* dist, uint32_t* stamp, uint32_t root, Frontier f, uint32_t* zc_other) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x) {
    dist[s] = s == root ? 0u : SSSP_INF;
    stamp[s] = 0;
  }
  zero_ctr(zc_other);
  zero_ctr(f.nitems);
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    uint32_t nit = 0, ne = 0;
    if (slot_has_col(g, root)) {
      uint32_t c = col_of_slot(g, root);
      ne = g.col_split[c] - g.col_ptr[c];
      nit = (ne + CHUNK - 1) / CHUNK;
    }
    for (uint32_t k = 0; k < nit; k++) f.items[k] = ((uint64_t)k << 32) | root;
    f.nitems[0] = nit;
    f.slots[0].verts = 1;
    f.slots[0].cols = nit > 0;
    f.slots[0].edges = ne;
  }
}

// Near/far split: an item whose vertex distance is >= T is not relaxed but
// put back in the next frontier (chunk 0 re-adds the whole column once);
// the host raises T when nothing below it remains. T = ~0 is plain
// Bellman-Ford. Every improvement is still relaxed eventually, so the fixed
// point — the exact distances — is unchanged.
//
// Wavefront w takes items w, w + nw, ... of the work list, 64 at a time, one
// item per lane (its vertex distance and entry range load together); the
// entries of those items are then packed into passes of 64 lanes (lane l
// of a pass finds its item by a binary search over the lanes' exclusive
// offsets), so short columns do not leave most lanes idle.
//
// SYNC = true is the reference's synchronous iteration: a sender's message is
// msg[s], its distance when the previous step ended (k_sssp_snap), not a
// distance lowered earlier in this step; bounded runs use it so that m steps
// give exactly the reference's m iterations (T must then be ~0).
//
// nitems_dev: when set, the item count is read from the device (the count
// the previous step emitted) instead of `nitems`, so the host can enqueue
// steps without waiting for counts; zc must then be null (the host clears
// the consumed counters after the launch).
template <bool SYNC>
__global__ void __launch_bounds__(BLOCK) k_sssp_push(DG g, const uint64_t* items, uint32_t nitems,
                                                     uint32_t* dist, uint32_t* stamp, uint32_t round,
                                                     uint32_t T, Frontier nf, uint32_t* zc, const uint32_t* msg,
                                                     const uint32_t* nitems_dev) {
  if (nitems_dev) nitems = *nitems_dev;
  zero_ctr(zc);
  const uint32_t nw = gridDim.x * (BLOCK / WAVE);
  const uint32_t wid = blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE;
  const uint32_t lane = lane_id();
  const uint32_t tag = round + 1;
  LA3_EMITTER(em);
  const uint32_t mine = wid < nitems ? (nitems - wid + nw - 1) / nw : 0;
  for (uint32_t b0 = 0; b0 < mine; b0 += WAVE) {  // wave-uniform
    const uint32_t nb = min((uint32_t)WAVE, mine - b0);
    uint32_t s = 0, du = SSSP_INF, beg = 0, len = 0, fnit = 0, fne = 0;
    bool again = false;
    if (lane < nb) {
      const uint64_t item = items[wid + (uint64_t)(b0 + lane) * nw];
      s = (uint32_t)item;
      const uint32_t ch = (uint32_t)(item >> 32);
      const uint32_t c = col_of_slot(g, s);
      du = SYNC ? msg[s] : ld_relaxed(&dist[s]);
      const uint32_t cb = g.col_ptr[c], cs = g.col_split[c];
      if (du >= T) {  // far: keep it for a later round
        if (ch == 0 && atomicMax(&stamp[s], tag) < tag) {
          again = true;
          fne = cs - cb;
          fnit = (fne + CHUNK - 1) / CHUNK;
          em.nmin_inv = max(em.nmin_inv, ~du);
        }
      } else {
        beg = cb + ch * CHUNK;
        const uint32_t end = min(beg + CHUNK, cs);
        len = end > beg ? end - beg : 0;
      }
    }
    em_push(em, nf, s, fnit, fne, again);
    const uint32_t incl = wave_incl_scan(len);
    const uint32_t total = __shfl(incl, WAVE - 1, WAVE);
    const uint32_t excl = incl - len;
    for (uint32_t p0 = 0; p0 < total; p0 += WAVE) {  // wave-uniform
      const uint32_t gi = p0 + lane;
      // item of entry gi: the last lane x with excl[x] <= gi (excl[0] = 0)
      uint32_t x = 0;
#pragma unroll
      for (uint32_t step = WAVE / 2; step > 0; step >>= 1)
        if (__shfl(excl, x + step, WAVE) <= gi) x += step;
      const uint32_t ib = __shfl(beg, x, WAVE), ie = __shfl(excl, x, WAVE), idu = __shfl(du, x, WAVE);
      bool push = false;
      uint32_t r = 0;
      if (gi < total) {
        const uint32_t e = ib + (gi - ie);
        r = g.col_rows[e];
        const uint32_t nd = idu + g.col_w[e];
        if (nd < ld_relaxed(&dist[r])) {
          // Fire-and-forget: the comparison with the loaded value marks the
          // improvement (a stale load only re-activates r harmlessly).
          atomicMin(&dist[r], nd);
          em.nmin_inv = max(em.nmin_inv, ~nd);
          if (atomicMax(&stamp[r], tag) < tag) push = true;
        }
      }
      uint32_t nit = 0, ne = 0;
      if (push) {  // r is a regular row, hence a regular column with c == r
        ne = g.col_split[r] - g.col_ptr[r];
        nit = (ne + CHUNK - 1) / CHUNK;
      }
      em_push(em, nf, r, nit, ne, push);
    }
  }
  em_finish(em, nf);
}

// Messages of the next synchronous step: the distance each frontier vertex
// holds now that the step that queued it has ended.
__global__ void k_sssp_snap(const uint64_t* items, uint32_t nitems, const uint32_t* dist, uint32_t* msg) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nitems; i += gridDim.x * blockDim.x) {
    const uint32_t s = (uint32_t)items[i];
    msg[s] = dist[s];
  }
}

// Sink step (vertex_program_execute.hpp:131-148): every sink row takes the
// min over all its in-entries of dist + w (regular senders with their final
// distance; source senders other than the root are at INF and add nothing).
__global__ void k_sssp_sinks(DG g, uint32_t* dist) {
  const uint32_t n = g.nsink;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t r = g.nreg + i;
    uint32_t y = SSSP_INF;
    for (uint32_t j = g.row_ptr[r]; j < g.row_ptr[r + 1]; j++)
      y = min(y, dist[g.row_cols[j]] + g.row_w[j]);
    dist[r] = min(dist[r], y);
  }
}

// -------------------------------------------------------------- PageRank ---
// PrVertex (pr.h:34-45) / PPR: rank' = base(r) + (1 - alpha) * y with
// y = sum of x over the row's in-entries, x = rank / out-degree.
// base(r) = alpha for PR (seed == ~0), alpha at the seed only for PPR.
// Rows with <= LONG_ROW entries: one lane per row, sequential sum in the
// reference's message order (sources first). Longer rows: one wavefront,
// lane-strided partial sums and a fixed butterfly. Messages are stored by
// sender rank (DG::rank_of, decreasing out-degree) and gathered through
// DG::row_rank: the senders of most entries then share a few cached lines.
struct PrArgs {
  const double* x;       // messages, by rank_of[slot]
  double* xnext;         // next messages (regular rows), by rank_of[slot]
  double* rank;          // by slot
  double* srcsum;        // per regular row: sum over source entries (constant)
  const uint32_t* deg;   // out-degree by slot
  uint32_t* active;      // set when some |delta| > tol
  double alpha, damp, tol;
  uint32_t seed;         // slot of the PPR seed, ~0 for PR
  // A fixed-iteration run reads no rank between iterations (the new rank is
  // a function of the row sum; the next messages are formed in registers)
  // and needs no convergence flag: all but its last iteration skip the
  // rank's read and store. Zero: read, store and flag.
  uint32_t skip_rank;
};

__device__ __forceinline__ double pr_base(const PrArgs& a, uint32_t r) {
  return (a.seed == 0xffffffffu || a.seed == r) ? a.alpha : 0.0;
}

// mode 0: srcsum[r] = sum over source entries (regular rows)
// mode 1: loop step for regular rows
// mode 2: sink step for sink rows
template <int MODE>
__device__ __forceinline__ void pr_finish(const DG& g, const PrArgs& a, uint32_t r, double y) {
  if (MODE == 0) { a.srcsum[r] = y; return; }
  const double nr = __dadd_rn(pr_base(a, r), __dmul_rn(a.damp, y));
  if (MODE == 1) {
    if (!a.skip_rank) {
      const double old = a.rank[r];
      a.rank[r] = nr;
      if (fabs(nr - old) > a.tol) a.active[0] = 1;
    }
    a.xnext[g.rank_of[r]] = nr / (double)a.deg[r];
  } else {
    a.rank[r] = nr;
  }
}

// Entry range of row r summed in MODE: 0 = source columns [b, split),
// 1 = regular columns [split, e), 2 = all [b, e).
template <int MODE>
__device__ __forceinline__ void pr_range(const DG& g, uint32_t r, uint32_t& lo, uint32_t& hi) {
  const uint32_t b = g.row_ptr[r], e = g.row_ptr[r + 1], sp = g.row_split[r];
  lo = MODE == 1 ? sp : b;
  hi = MODE == 0 ? sp : e;
}

// y + x[cols[lo]] + ... + x[cols[hi - 1]], added in entry order; the loads
// of eight entries (the last < 8 with indices clamped to hi - 1) are issued
// before their adds.
__device__ __forceinline__ double seq_sum(const DG& g, const double* x, uint32_t j, uint32_t hi, double y) {
  for (; j + 8 <= hi; j += 8) {
    uint32_t cc[8];
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) cc[u] = g.row_rank[j + u];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = x[cc[u]];
#pragma unroll
    for (int u = 0; u < 8; u++) y = __dadd_rn(y, v[u]);
  }
  if (j < hi) {
    const uint32_t rem = hi - j;
    uint32_t cc[8];
    double v[8];
#pragma unroll
    for (uint32_t u = 0; u < 8; u++) cc[u] = g.row_rank[min(j + u, hi - 1)];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = x[cc[u]];
#pragma unroll
    for (uint32_t u = 0; u < 8; u++)
      if (u < rem) y = __dadd_rn(y, v[u]);
  }
  return y;
}

// Wavefront sum of x[cols[j]], j in [lo, hi): lane l adds entries lo + l,
// lo + l + 64, ... in order (eight, then four, loads in flight), then a
// fixed-shape wavefront reduction.
__device__ __forceinline__ double wave_range_sum(const DG& g, const double* x, uint32_t lo, uint32_t hi) {
  double p = 0.0;
  uint32_t j = lo + lane_id();
  for (; j + 7 * WAVE < hi; j += 8 * WAVE) {
    uint32_t cc[8];
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) cc[u] = g.row_rank[j + u * WAVE];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = x[cc[u]];
#pragma unroll
    for (int u = 0; u < 8; u++) p = __dadd_rn(p, v[u]);
  }
  for (; j + 3 * WAVE < hi; j += 4 * WAVE) {
    const uint32_t c0 = g.row_rank[j], c1 = g.row_rank[j + WAVE], c2 = g.row_rank[j + 2 * WAVE],
                   c3 = g.row_rank[j + 3 * WAVE];
    const double v0 = x[c0], v1 = x[c1], v2 = x[c2], v3 = x[c3];
    p = __dadd_rn(p, v0);
    p = __dadd_rn(p, v1);
    p = __dadd_rn(p, v2);
    p = __dadd_rn(p, v3);
  }
  for (; j < hi; j += WAVE) p = __dadd_rn(p, x[g.row_rank[j]]);
  return wave_sum(p);
}

// Block sum of x[cols[j]], j in [lo, hi): thread t adds entries lo + t,
// lo + t + BLOCK, ... in order (eight loads in flight), then the wavefront
// sums, then the wavefronts' sums in order. The result is on thread 0.
__device__ __forceinline__ double block_range_sum(const DG& g, const double* x, uint32_t lo, uint32_t hi,
                                                  double* red) {
  double p = 0.0;
  uint32_t j = lo + threadIdx.x;
  for (; j + 7 * BLOCK < hi; j += 8 * BLOCK) {
    uint32_t cc[8];
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) cc[u] = g.row_rank[j + u * BLOCK];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = x[cc[u]];
#pragma unroll
    for (int u = 0; u < 8; u++) p = __dadd_rn(p, v[u]);
  }
  for (; j < hi; j += BLOCK) p = __dadd_rn(p, x[g.row_rank[j]]);
  p = wave_sum(p);
  if (lane_id() == 0) red[threadIdx.x / WAVE] = p;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (uint32_t w = 0; w < BLOCK / WAVE; w++) s = __dadd_rn(s, red[w]);
  __syncthreads();  // red is reused by the next row
  return s;
}

// One PR step over rows [r0, r1) (regular rows, or sink rows for MODE 2):
//   blocks [0, nbr): rows with <= LONG_ROW entries, lane per row;
//   blocks [nbr, nbr + nbw): long rows with <= PSEG_ENTRIES entries,
//   wavefront per row; then the PSEG_ENTRIES-entry segments [s0, s1) of the
//   longer rows, wavefront per segment, partial sums into part[segment]
//   (k_pr_combine finishes those rows);
//   blocks [nbr + nbw, grid): the rows plong[p0, p1), longer than
//   PSEG_ENTRIES, a block per row (when the host passes no segments).
template <int MODE>
__global__ void __launch_bounds__(BLOCK) k_pr_step(DG g, PrArgs a, uint32_t r0, uint32_t r1, uint32_t nbr,
                                                   uint32_t s0, uint32_t s1, double* part, uint32_t nbw,
                                                   uint32_t p0, uint32_t p1) {
  if (blockIdx.x >= nbr + nbw) {
    __shared__ double red[BLOCK / WAVE];
    const uint32_t nbb = gridDim.x - nbr - nbw;
    for (uint32_t i = p0 + (blockIdx.x - nbr - nbw); i < p1; i += nbb) {  // block-uniform
      const uint32_t r = g.plong[i];
      uint32_t lo, hi;
      pr_range<MODE>(g, r, lo, hi);
      const double p = block_range_sum(g, a.x, lo, hi, red);
      if (threadIdx.x == 0) pr_finish<MODE>(g, a, r, MODE == 1 ? __dadd_rn(a.srcsum[r], p) : p);
    }
    return;
  }
  if (blockIdx.x < nbr) {
    for (uint32_t r = r0 + blockIdx.x * BLOCK + threadIdx.x; r < r1; r += nbr * BLOCK) {
      if (g.row_ptr[r + 1] - g.row_ptr[r] > LONG_ROW) continue;
      uint32_t lo, hi;
      pr_range<MODE>(g, r, lo, hi);
      pr_finish<MODE>(g, a, r, seq_sum(g, a.x, lo, hi, MODE == 1 ? a.srcsum[r] : 0.0));
    }
    return;
  }
  const uint32_t nwaves = nbw * (BLOCK / WAVE);
  const uint32_t wid = (blockIdx.x - nbr) * (BLOCK / WAVE) + threadIdx.x / WAVE;
  for (uint32_t i = wid; i < g.nlong; i += nwaves) {
    const uint32_t r = g.long_rows[i];
    if (r < r0 || r >= r1 || g.row_ptr[r + 1] - g.row_ptr[r] > PSEG_ENTRIES) continue;  // wave-uniform
    uint32_t lo, hi;
    pr_range<MODE>(g, r, lo, hi);
    const double p = wave_range_sum(g, a.x, lo, hi);
    if (lane_id() == 0) pr_finish<MODE>(g, a, r, MODE == 1 ? __dadd_rn(a.srcsum[r], p) : p);
  }
  for (uint32_t sg = s0 + wid; sg < s1; sg += nwaves) {
    const uint32_t r = g.pseg_row[sg];
    uint32_t lo, hi;
    pr_range<MODE>(g, r, lo, hi);
    lo = max(lo, g.pseg_beg[sg]);
    hi = min(hi, g.pseg_beg[sg] + PSEG_ENTRIES);
    const double p = wave_range_sum(g, a.x, lo, max(lo, hi));
    if (lane_id() == 0) part[sg] = p;
  }
}

// Rows plong[p0, p1) (rows longer than PSEG_ENTRIES): their segment partials
// in segment order, then the row's update.
template <int MODE>
__global__ void __launch_bounds__(BLOCK) k_pr_combine(DG g, PrArgs a, uint32_t p0, uint32_t p1,
                                                      const double* part) {
  for (uint32_t i = p0 + blockIdx.x * BLOCK + threadIdx.x; i < p1; i += gridDim.x * BLOCK) {
    const uint32_t r = g.plong[i];
    double y = MODE == 1 ? a.srcsum[r] : 0.0;
    // eight partials loaded before their adds (in segment order)
    const uint32_t f1 = g.pseg_first[i + 1];
    uint32_t sg = g.pseg_first[i];
    for (; sg + 8 <= f1; sg += 8) {
      double v[8];
#pragma unroll
      for (uint32_t u = 0; u < 8; u++) v[u] = part[sg + u];
#pragma unroll
      for (uint32_t u = 0; u < 8; u++) y = __dadd_rn(y, v[u]);
    }
    for (; sg < f1; sg++) y = __dadd_rn(y, part[sg]);
    pr_finish<MODE>(g, a, r, y);
  }
}

__global__ void k_pr_init(DG g, PrArgs a) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x) {
    const double r0 = pr_base(a, s);
    a.rank[s] = r0;
    double x = 0.0;
    if (slot_has_col(g, s)) x = r0 / (double)a.deg[s];
    ((double*)a.x)[g.rank_of[s]] = x;
    a.xnext[g.rank_of[s]] = x;
  }
}

// -------------------------------------------------------------------- CC ---
// CcVertex (cc.h:22-38) computes, per connected component, the minimum
// original id. Here: concurrent union-find over slots where a root is only
// ever hooked under a root with a smaller original id, so the surviving root
// of each component is its minimum-id vertex.
__device__ __forceinline__ uint32_t uf_find(uint32_t* p, uint32_t x) {
  for (;;) {
    const uint32_t y = ld_relaxed(&p[x]);
    if (y == x) return x;
    const uint32_t z = ld_relaxed(&p[y]);
    if (z == y) return y;
    atomicCAS(&p[x], y, z);  // path halving
    x = z;
  }
}

__device__ __forceinline__ void uf_union(const DG& g, uint32_t* p, uint32_t a, uint32_t b) {
  for (;;) {
    a = uf_find(p, a);
    b = uf_find(p, b);
    if (a == b) return;
    if (g.id_of[a] < g.id_of[b]) { uint32_t t = a; a = b; b = t; }
    if (atomicCAS(&p[a], a, b) == a) return;
  }
}

__global__ void k_cc_init(DG g, uint32_t* p) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x)
    p[s] = s;
}

// Sampling-based union-find (the Afforest scheme: link a few neighbours,
// compress, find the largest tree, then process the remaining edges of the
// vertices outside it). Edges of an undirected load are stored in both rows,
// so an edge with an endpoint outside the largest tree is seen from that
// endpoint; two endpoints inside it are connected already. The hooking rule
// (larger-id root under smaller-id root) is unchanged, so every root is its
// component's minimum-id vertex whatever order the unions run in.

// Every row r unions with the smallest-id vertex among its first `scan`
// entries when that id is below its own (lane per row). Each union then
// hooks the row's own root, so the threads do not compete for one word (a
// plain first-entry link hooks a hub's root under each smaller-id neighbour
// in turn, one compare-and-swap winner at a time).
__global__ void __launch_bounds__(BLOCK) k_cc_link(DG g, uint32_t* p, uint32_t scan) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < g.nrows; r += gridDim.x * blockDim.x) {
    const uint32_t b = g.row_ptr[r], e = min(g.row_ptr[r + 1], g.row_ptr[r] + scan);
    uint32_t best = r, bid = g.id_of[r];
    for (uint32_t j = b; j < e; j++) {
      const uint32_t s = g.row_cols[j], id = g.id_of[s];
      if (id < bid) { bid = id; best = s; }
    }
    if (best != r) uf_union(g, p, r, best);
  }
}

// p[s] = root of s (full compression; the unions have finished).
__global__ void k_cc_compress(DG g, uint32_t* p) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x) {
    const uint32_t x = uf_find(p, s);  // other threads only shorten paths to the same root
    p[s] = x;
  }
}

// Roots of n sampled rows (row i * stride mod nrows).
__global__ void k_cc_sample(DG g, const uint32_t* p, uint32_t n, uint32_t stride, uint32_t* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = p[(uint32_t)(((uint64_t)i * stride) % g.nrows)];
}

// Rows whose root is not `big`: union with entries from `from` on. Rows of at
// most LONG_ROW entries lane per row, longer rows a wavefront each.
__global__ void __launch_bounds__(BLOCK) k_cc_finish(DG g, uint32_t* p, uint32_t big, uint32_t from) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < g.nrows; r += gridDim.x * blockDim.x) {
    const uint32_t b = g.row_ptr[r], e = g.row_ptr[r + 1];
    if (e - b <= from || e - b > LONG_ROW || uf_find(p, r) == big) continue;
    for (uint32_t j = b + from; j < e; j++) uf_union(g, p, r, g.row_cols[j]);
  }
}

__global__ void __launch_bounds__(BLOCK) k_cc_finish_long(DG g, uint32_t* p, uint32_t big, uint32_t from) {
  const uint32_t nwaves = gridDim.x * (BLOCK / WAVE);
  const uint32_t wid = blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE;
  for (uint32_t i = wid; i < g.nlong; i += nwaves) {
    const uint32_t r = g.long_rows[i];
    if ((uint32_t)__builtin_amdgcn_readfirstlane((int)uf_find(p, r)) == big) continue;  // wave-uniform
    for (uint32_t j = g.row_ptr[r] + from + lane_id(); j < g.row_ptr[r + 1]; j += WAVE)
      uf_union(g, p, r, g.row_cols[j]);
  }
}

__global__ void k_cc_output(DG g, uint32_t* p, uint32_t* label_by_id) {
  for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < g.nslots; id += gridDim.x * blockDim.x)
    label_by_id[id] = g.id_of[uf_find(p, g.slot_of_id[id])];
}

// -------------------------------------------------------------------- TC ---
// GnVertex + CtVertex (tc.h:47-99) on the acyclic load (row = higher id,
// column = lower id): row u's in-entries are N-(u). For every entry (u, v)
// the reference adds |N-(u) n N-(v)| to v (tc.h:84-96). Both lists are sorted
// by message-order key, so one list is binary-searched in the other.
__device__ __forceinline__ uint32_t slot_key(const DG& g, uint32_t s) {
  return msg_key(g, col_of_slot(g, s));
}

__device__ uint32_t tc_intersect(const DG& g, uint32_t ab, uint32_t ae, uint32_t bb, uint32_t be) {
  if (ae - ab > be - bb) { uint32_t t = ab; ab = bb; bb = t; t = ae; ae = be; be = t; }
  uint32_t cnt = 0, lo = bb;
  for (uint32_t i = ab; i < ae && lo < be; i++) {
    const uint32_t k = slot_key(g, g.row_cols[i]);
    uint32_t l = lo, h = be;
    while (l < h) {
      const uint32_t m = (l + h) >> 1;
      if (slot_key(g, g.row_cols[m]) < k) l = m + 1; else h = m;
    }
    if (l < be && slot_key(g, g.row_cols[l]) == k) { cnt++; l++; }
    lo = l;
  }
  return cnt;
}

// mark[ptr[i]] = i for every non-empty owner i < n (a row or column whose
// entries are [ptr[i], ptr[i + 1])); an inclusive max-scan of the marks (0
// elsewhere) then gives every entry its owner.
__global__ void k_tc_mark_owner(const uint32_t* ptr, uint32_t n, uint32_t* mark) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t b = ptr[i];
    if (ptr[i + 1] > b) mark[b] = i;
  }
}

// Number of keys of list A = cols[ab, ae) found in list B = cols[bb, be)
// (both sorted by message-order key), with the 64 lanes of a wavefront each
// binary-searching one element of A in B. Every lane returns the count.
__device__ uint32_t tc_intersect_wave(const DG& g, uint32_t ab, uint32_t ae, uint32_t bb, uint32_t be) {
  uint32_t found = 0;
  for (uint32_t i0 = ab; i0 < ae; i0 += WAVE) {
    const uint32_t i = i0 + lane_id();
    bool hit = false;
    if (i < ae) {
      const uint32_t k = slot_key(g, g.row_cols[i]);
      uint32_t l = bb, h = be;
      while (l < h) {
        const uint32_t m = (l + h) >> 1;
        if (slot_key(g, g.row_cols[m]) < k) l = m + 1; else h = m;
      }
      hit = l < be && slot_key(g, g.row_cols[l]) == k;
    }
    found += __popcll(__ballot(hit));
  }
  return found;
}

// Number of keys common to A = cols[ab, ae) and B = cols[bb, be) (both
// sorted by message-order key, no repeats) by a merge split over the
// wavefront (merge path): lane l merges positions [l L, (l + 1) L) of the
// merged order, L = ceil((|A| + |B|) / 64), ties taking A first; an A
// element is counted when it is consumed and equals the next B element, so
// every common key is counted once. Every lane returns the count.
__device__ uint32_t tc_intersect_merge(const DG& g, uint32_t ab, uint32_t ae, uint32_t bb, uint32_t be) {
  const uint32_t a = ae - ab, b = be - bb, L = (a + b + WAVE - 1) / WAVE;
  const uint32_t d = min(lane_id() * L, a + b);
  // i = number of A elements among the first d merged ones
  uint32_t lo = d > b ? d - b : 0, hi = min(d, a);
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (slot_key(g, g.row_cols[ab + mid]) <= slot_key(g, g.row_cols[bb + d - mid - 1])) lo = mid + 1;
    else hi = mid;
  }
  uint32_t i = lo, j = d - lo, cnt = 0;
  const uint32_t stop = min(d + L, a + b);
  uint32_t ka = i < a ? slot_key(g, g.row_cols[ab + i]) : 0xffffffffu;
  uint32_t kb = j < b ? slot_key(g, g.row_cols[bb + j]) : 0xffffffffu;
  for (uint32_t pos = d; pos < stop; pos++) {
    if (i < a && (j >= b || ka <= kb)) {
      cnt += j < b && ka == kb;
      i++;
      ka = i < a ? slot_key(g, g.row_cols[ab + i]) : 0xffffffffu;
    } else {
      j++;
      kb = j < b ? slot_key(g, g.row_cols[bb + j]) : 0xffffffffu;
    }
  }
  return wave_sum(cnt);
}

// total[0] += triangles; total[1] += sum over entries of |N-(u)| + |N-(v)|
// (the list entries a merge intersection reads: the algorithmic bytes).
// Lane per entry (u, v); a pair whose shorter list has more than
// TC_WAVE_MIN entries, or whose searches would take one lane more than
// TC_LANE_STEPS dependent loads, is intersected by the whole wavefront
// instead, so a pair of hub lists is not one lane's serial chain of binary
// searches (128 measured best of 64, 128, 256 and no limit at scale 20).
constexpr uint32_t TC_WAVE_MIN = 32, TC_LANE_STEPS = 128;

// Entries go to the LDS hash passes (k_tc_hash) when the endpoint with the
// longer list N-(x), TC_HASH_MIN < d(x) <= TC_HASH_MAX, can hold it in a
// table that several entries share: x = u (the row) when d(v) <= d(u); x = v
// (the column) when d(u) < d(v) and column v has more than TC_HASH_MIN
// entries. The other endpoint's list is then probed into the table, d(y)
// probes instead of d(y) binary searches. Every other entry is counted here.
// Owner bands of the hash launches: (MIN, SPLIT] with 8-KB tables and
// (SPLIT, MAX] with 32 KB. Longer lists stay with k_tc_count: a third band
// to 16384 (128-KB tables, one block per CU) took 6.6 ms for pairs that the
// wavefront merges of k_tc_count count in about 0.4 ms.
constexpr uint32_t TC_HASH_MIN = 32, TC_HASH_SPLIT = 1024, TC_HASH_MAX = 4096;

__device__ __forceinline__ bool tc_hashed(uint32_t d) { return d > TC_HASH_MIN && d <= TC_HASH_MAX; }

__device__ __forceinline__ bool tc_by_hash(const DG& g, uint32_t v, uint32_t du, uint32_t dv) {
  if (dv <= du) return tc_hashed(du);
  if (!tc_hashed(dv)) return false;
  const uint32_t c = col_of_slot(g, v);
  return g.col_ptr[c + 1] - g.col_ptr[c] > TC_HASH_MIN;
}

__global__ void __launch_bounds__(BLOCK) k_tc_count(DG g, const uint32_t* row_of, uint32_t* per_slot,
                                                    unsigned long long* total) {
  unsigned long long acc = 0, work = 0, hwork[4] = {0, 0, 0, 0};
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < g.nnz; base += stride) {
    const uint64_t j = base + threadIdx.x;
    uint32_t u = 0, v = 0, ub = 0, ue = 0, vb = 0, ve = 0;
    bool pair = false;
    if (j < g.nnz) {
      u = row_of[j];
      v = g.row_cols[j];
      if (v < g.nrows) {
        ub = g.row_ptr[u]; ue = g.row_ptr[u + 1]; vb = g.row_ptr[v]; ve = g.row_ptr[v + 1];
        if (tc_by_hash(g, v, ue - ub, ve - vb)) {  // counted by a k_tc_hash launch
          const uint32_t du = ue - ub, dv = ve - vb, dx = max(du, dv);
          const uint32_t k = (dv <= du ? 0u : 2u) + (dx > TC_HASH_SPLIT ? 1u : 0u);
          hwork[k] += du + dv;
        } else {
          pair = true;
          work += (ue - ub) + (ve - vb);
        }
      }
    }
    // A lane's searches cost about min * log2(max) dependent loads; past
    // TC_LANE_STEPS the pair goes to the wavefront (log2(max) loads per 64
    // keys) even when its shorter list is short.
    const uint32_t mn = min(ue - ub, ve - vb), mx = max(ue - ub, ve - vb);
    const bool heavy = pair && (mn > TC_WAVE_MIN || mn * (32 - __clz(mx)) > TC_LANE_STEPS);
    if (pair && !heavy) {
      const uint32_t c = tc_intersect(g, ub, ue, vb, ve);
      if (c) { atomicAdd(&per_slot[v], c); acc += c; }
    }
    uint64_t hm = __ballot(heavy);
    while (hm) {  // wave-uniform
      const int src = __ffsll((long long)hm) - 1;
      hm &= hm - 1;
      const uint32_t hv = __shfl(v, src, WAVE);
      uint32_t ab = __shfl(ub, src, WAVE), ae = __shfl(ue, src, WAVE);
      uint32_t bb = __shfl(vb, src, WAVE), be = __shfl(ve, src, WAVE);
      if (ae - ab > be - bb) { uint32_t t = ab; ab = bb; bb = t; t = ae; ae = be; be = t; }
      // Merge split when its steps per lane are well below the binary
      // searches' (a merge step costs more: divergent, one dependent load;
      // the factor 3 is the fastest of 1, 2, 3, 5, 8 measured at scale 20).
      const uint32_t na = ae - ab, nb = be - bb;
      const uint32_t lg = 32 - __clz(nb);
      const bool merge = 3 * ((na + nb) / WAVE + lg) < ((na + WAVE - 1) / WAVE) * lg;
      const uint32_t c = merge ? tc_intersect_merge(g, ab, ae, bb, be) : tc_intersect_wave(g, ab, ae, bb, be);
      if (lane_id() == 0 && c) { atomicAdd(&per_slot[hv], c); acc += c; }
    }
  }
  acc = wave_sum(acc);
  work = wave_sum(work);
  if (lane_id() == 0 && acc) atomicAdd(&total[0], acc);
  if (lane_id() == 0 && work) atomicAdd(&total[1], work);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const unsigned long long h = wave_sum(hwork[k]);
    if (lane_id() == 0 && h) atomicAdd(&total[2 + k], h);
  }
}

// Multiplicative (Fibonacci) hash of a key to 32 - shift bits.
__device__ __forceinline__ uint32_t tc_slot_hash(uint32_t key, uint32_t shift) { return (key * 0x9E3779B1u) >> shift; }

// Exclusive prefix over the block of one value per thread; tot = the sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* ws, uint32_t& tot) {
  const uint32_t lane = lane_id(), w = threadIdx.x / WAVE;
  const uint32_t incl = wave_incl_scan(v);
  if (lane == WAVE - 1) ws[w] = incl;
  __syncthreads();
  uint32_t off = 0;
  tot = 0;
#pragma unroll
  for (uint32_t i = 0; i < BLOCK / WAVE; i++) {
    const uint32_t x = ws[i];
    off += i < w ? x : 0u;
    tot += x;
  }
  __syncthreads();
  return off + incl - v;
}

// LDS table of N-(x) (slots + 1, open addressing, at most half full);
// returns its size T through lg (T = 2^lg).
__device__ __forceinline__ void tc_build_table(const DG& g, uint32_t x, uint32_t* table, uint32_t& T,
                                               uint32_t& shift) {
  const uint32_t xb = g.row_ptr[x], xe = g.row_ptr[x + 1], dx = xe - xb;
  uint32_t lg = 6;
  while ((1u << lg) < 2 * dx) lg++;
  T = 1u << lg;
  shift = 32 - lg;
  for (uint32_t i = threadIdx.x; i < T; i += BLOCK) table[i] = 0;
  __syncthreads();
  for (uint32_t j = xb + threadIdx.x; j < xe; j += BLOCK) {  // keys are distinct
    const uint32_t key = g.row_cols[j] + 1;
    uint32_t h = tc_slot_hash(key, shift);
    while (atomicCAS(&table[h], 0u, key) != 0u) h = (h + 1) & (T - 1);
  }
  __syncthreads();
}

// Thread t brings list cols[lb, lb + len) (len = 0: none); the block probes
// every element of every list into the table, four positions per thread in
// flight, and thread t gets the number of its list's elements found.
__device__ uint32_t tc_probe_block(const DG& g, const uint32_t* table, uint32_t T, uint32_t shift,
                                   uint32_t* excl, uint32_t* lbeg, uint32_t* cnt, uint32_t* ws, uint32_t lb,
                                   uint32_t len) {
  const uint32_t t = threadIdx.x;
  uint32_t W;
  const uint32_t ex = block_excl_scan(len, ws, W);
  excl[t] = ex;
  lbeg[t] = lb;
  cnt[t] = 0;
  __syncthreads();
  constexpr uint32_t PQ = 4;
  for (uint32_t p0 = 0; p0 < W; p0 += PQ * BLOCK) {  // block-uniform
    uint32_t o[PQ], key[PQ];
#pragma unroll
    for (uint32_t q = 0; q < PQ; q++) {
      const uint32_t p = p0 + q * BLOCK + t;
      // owner: the last thread whose prefix is <= p (threads with empty lists
      // share their prefix with the next thread, which owns those positions)
      o[q] = 0;
#pragma unroll
      for (uint32_t step = BLOCK / 2; step; step >>= 1)
        if (excl[o[q] + step] <= p) o[q] += step;
      key[q] = p < W ? g.row_cols[lbeg[o[q]] + (p - excl[o[q]])] + 1 : 0u;
    }
#pragma unroll
    for (uint32_t q = 0; q < PQ; q++) {
      bool hit = false;
      if (key[q]) {
        uint32_t h = tc_slot_hash(key[q], shift);
        for (;;) {  // at most half the table is used: an empty slot ends every probe
          const uint32_t sv = table[h];
          if (sv == key[q]) { hit = true; break; }
          if (sv == 0) break;
          h = (h + 1) & (T - 1);
        }
      }
      // a wavefront's 64 consecutive positions mostly share one list: then
      // one add of their hit count
      const uint32_t o0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)o[q]);
      if (__ballot(o[q] != o0) == 0) {
        const uint64_t hb = __ballot(hit);
        if (hb && lane_id() == 0) atomicAdd(&cnt[o0], (uint32_t)__popcll(hb));
      } else if (hit) {
        atomicAdd(&cnt[o[q]], 1u);
      }
    }
  }
  __syncthreads();
  const uint32_t c = cnt[t];
  __syncthreads();  // excl / lbeg / cnt are rewritten by the next call
  return c;
}

// The hash passes over 256-entry ranges of the entries: SIDE 0 row-major
// (owner_of = row of each entry, x = u, y = v, counts to v), SIDE 1
// column-major (owner_of = column of each entry, x = v, y = u, counts to
// v). A block builds the table of each owner x of its range that passes the
// rule of k_tc_count's filter (at most 10 per range: such owners have more
// than 32 entries) and probes its eligible entries' N-(y).
// The launch takes the owners with DMIN < d(x) <= DMAX: smaller tables for
// the smaller owners leave room for more blocks per CU.
template <int SIDE, uint32_t DMIN, uint32_t DMAX>
__global__ void __launch_bounds__(BLOCK) k_tc_hash(DG g, const uint32_t* owner_of, uint32_t* per_slot,
                                                   unsigned long long* total) {
  __shared__ uint32_t table[2 * DMAX];
  __shared__ uint32_t excl[BLOCK], lbeg[BLOCK], cnt[BLOCK], xs[BLOCK];
  __shared__ uint32_t ws[BLOCK / WAVE];
  __shared__ uint32_t hx[BLOCK], hneed[BLOCK], hsum[BLOCK];
  __shared__ uint32_t nh;
  const uint32_t t = threadIdx.x;
  unsigned long long acc = 0;
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < g.nnz; base += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t j = base + t;
    uint32_t x = 0, y = 0, dx = 0, dy = 0, yb = 0;
    bool own = false;
    if (j < g.nnz) {
      // the owner's fields are shared by a run of entries (cached); the other
      // endpoint's list is read only for entries this launch owns
      x = SIDE == 0 ? owner_of[j] : slot_of_col(g, owner_of[j]);
      if (x < g.nrows) dx = g.row_ptr[x + 1] - g.row_ptr[x];
      own = tc_hashed(dx) && dx > DMIN && dx <= DMAX;
      if (SIDE == 1 && own) {
        const uint32_t c = owner_of[j];
        own = g.col_ptr[c + 1] - g.col_ptr[c] > TC_HASH_MIN;
      }
      if (own) {
        y = SIDE == 0 ? g.row_cols[j] : g.col_rows[j];
        if (y < g.nrows) { yb = g.row_ptr[y]; dy = g.row_ptr[y + 1] - yb; }
      }
    }
    const bool elig = own && (SIDE == 0 ? (y < g.nrows && dy <= dx) : dy < dx);
    if (t == 0) nh = 0;
    xs[t] = own ? x : 0xffffffffu;
    __syncthreads();
    if (own && (t == 0 || xs[t - 1] != x)) {  // first entry of an owner's run
      const uint32_t i = atomicAdd(&nh, 1u);
      hx[i] = x;
      hneed[i] = 0;
      hsum[i] = 0;
    }
    __syncthreads();
    const uint32_t n = nh;
    uint32_t mine = 0xffffffffu;
    if (elig)
      for (uint32_t i = 0; i < n; i++)
        if (hx[i] == x) mine = i;
    if (elig) atomicAdd(&hneed[mine], 1u);
    __syncthreads();
    for (uint32_t i = 0; i < n; i++) {  // block-uniform
      if (!hneed[i]) continue;
      uint32_t T, shift;
      tc_build_table(g, hx[i], table, T, shift);
      const bool go = mine == i;
      const uint32_t c = tc_probe_block(g, table, T, shift, excl, lbeg, cnt, ws, go ? yb : 0u, go ? dy : 0u);
      if (c) {
        if (SIDE == 0) { atomicAdd(&per_slot[y], c); acc += c; }
        else atomicAdd(&hsum[i], c);
      }
      __syncthreads();
      if (SIDE == 1 && t == 0 && hsum[i]) { atomicAdd(&per_slot[hx[i]], hsum[i]); acc += hsum[i]; }
    }
    __syncthreads();  // nh, xs, hx are rewritten for the next range
  }
  acc = wave_sum(acc);
  if (lane_id() == 0 && acc) atomicAdd(&total[0], acc);
}

}  // namespace la3
