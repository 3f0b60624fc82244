// Batched (multi-source) traversal kernels: k BFS queries in one traversal.
// Included once, by capi.hip, after kernels_dev.h.
//
// Layout, per vertex slot s (W = ceil(k / 64) words, K = 64 * W):
//   vis[s*W + w], fa[s*W + w], fb[s*W + w]   k-bit visited / frontier / next
//                                           frontier sets, bit q = query q
//   lev[s*K + q]                            u8 level of query q (255 = unreached)
// Every per-query result equals the single-query BfsVertex run from that
// root: a vertex joins query q's frontier at step i exactly when the
// single-query traversal discovers it at level i + 1.
#pragma once

#include "kernels_dev.h"

namespace la3 {

template <int W>
__global__ void k_msbfs_init(DG g, uint64_t* vis, uint64_t* fa, uint64_t* fb, uint64_t* lev64,
                             uint32_t* stamp) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x) {
#pragma unroll
    for (int w = 0; w < W; w++) {
      vis[(size_t)s * W + w] = 0;
      fa[(size_t)s * W + w] = 0;
      fb[(size_t)s * W + w] = 0;
    }
#pragma unroll
    for (int w = 0; w < 8 * W; w++) lev64[(size_t)s * 8 * W + w] = ~0ull;
    stamp[s] = 0;
  }
}

// One thread sets every root (roots may repeat, so no concurrent writers to a
// word), then the distinct root slots are turned into push work items.
template <int W>
__global__ void k_msbfs_roots(DG g, const uint32_t* root_slots, uint32_t k, uint64_t* vis,
                              uint64_t* fa, uint8_t* lev) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (uint32_t q = 0; q < k; q++) {
    const uint32_t s = root_slots[q];
    const uint64_t bit = 1ull << (q & 63);
    vis[(size_t)s * W + q / 64] |= bit;
    if (slot_has_col(g, s)) fa[(size_t)s * W + q / 64] |= bit;  // senders only
    lev[(size_t)s * 64 * W + q] = 0;
  }
}

__global__ void __launch_bounds__(BLOCK) k_msbfs_seed(DG g, const uint32_t* root_slots, uint32_t k,
                                                      uint32_t* stamp, Frontier f, int regular_only) {
  const uint32_t stride = gridDim.x * BLOCK;
  LA3_EMITTER(em);
  for (uint32_t base = blockIdx.x * BLOCK; base < k; base += stride) {
    const uint32_t q = base + threadIdx.x;
    bool first = false;
    uint32_t s = 0, nit = 0, ne = 0;
    if (q < k) {
      s = root_slots[q];
      first = atomicMax(&stamp[s], 1u) < 1u;
      if (first && slot_has_col(g, s)) {
        const uint32_t c = col_of_slot(g, s);
        ne = (regular_only ? g.col_split[c] : g.col_ptr[c + 1]) - g.col_ptr[c];
        nit = (ne + CHUNK - 1) / CHUNK;
      }
    }
    em_push(em, f, s, nit, ne, first);
  }
  em_finish(em, f);
}

// Top-down step: every entry (c -> r) of an active column ORs c's frontier
// bits that r has not yet seen into r's next-frontier words. The first
// toucher of r in this step (stamp) appends r to the touched list. With
// tst != nullptr (measurement runs) rows that receive a message are counted
// once per step (tst holds the last step tag ttag that counted the row).
template <int W>
__device__ __forceinline__ void msbfs_push_body(const DG& g, Part p, Emitter& em, const uint64_t* items,
                                                uint32_t nitems, const uint64_t* vis, const uint64_t* fa,
                                                uint64_t* fb, uint32_t* stamp, uint32_t tag,
                                                const Frontier& touched, uint32_t* tst = nullptr,
                                                uint32_t ttag = 0) {
  for (uint32_t it = p.wid(); it < nitems; it += p.nwaves()) {
    const uint64_t item = items[it];
    const uint32_t s = (uint32_t)item, ch = (uint32_t)(item >> 32);
    const uint32_t c = col_of_slot(g, s);
    const uint32_t beg = g.col_ptr[c] + ch * CHUNK;
    const uint32_t end = min(beg + CHUNK, g.col_ptr[c + 1]);
    uint64_t f[W];
#pragma unroll
    for (int w = 0; w < W; w++) f[w] = fa[(size_t)s * W + w];
    for (uint32_t e0 = beg; e0 < end; e0 += WAVE) {
      const uint32_t e = e0 + lane_id();
      bool first = false;
      uint32_t r = 0;
      if (e < end) {
        r = g.col_rows[e];
        bool any = false;
#pragma unroll
        for (int w = 0; w < W; w++) {
          const uint64_t m = f[w] & ~vis[(size_t)r * W + w];
          if (m) { atomicOr((unsigned long long*)&fb[(size_t)r * W + w], (unsigned long long)m); any = true; }
        }
        if (any) first = atomicMax(&stamp[r], tag) < tag;
        if (tst && atomicMax(&tst[r], ttag) < ttag) em.touched++;
      }
      em_push(em, touched, r, first ? 1u : 0u, 0, first);
    }
  }
}

template <int W>
__global__ void __launch_bounds__(BLOCK) k_msbfs_push(DG g, const uint64_t* items, uint32_t nitems,
                                                      const uint64_t* vis, const uint64_t* fa,
                                                      uint64_t* fb, uint32_t* stamp, uint32_t tag,
                                                      Frontier touched) {
  LA3_EMITTER(em);
  msbfs_push_body<W>(g, grid_part(), em, items, nitems, vis, fa, fb, stamp, tag, touched);
  em_finish(em, touched);
}

// Sets level byte q of a 64-query level row to v for every bit q of m,
// eight bytes at a time (read-modify-write of the 8-byte groups holding set
// bits; one lane owns the row in the step).
__device__ __forceinline__ void set_levels(uint8_t* row, uint64_t m, uint8_t v) {
  uint64_t* w8 = (uint64_t*)row;  // level rows are 64-byte aligned
  const uint64_t vv = 0x0101010101010101ull * v;
#pragma unroll
  for (int g = 0; g < 8; g++) {
    uint64_t x = (m >> (8 * g)) & 0xffu;
    if (!x) continue;
    x = (x | (x << 28)) & 0x0000000f0000000full;  // bit i -> byte i
    x = (x | (x << 14)) & 0x0003000300030003ull;
    x = (x | (x << 7)) & 0x0101010101010101ull;
    const uint64_t sel = x * 0xffu;
    w8[g] = (w8[g] & ~sel) | (vv & sel);
  }
}

// Marks the new bits of each touched vertex visited, records their level and
// emits the vertex's column as push work for the next step. New bits are the
// next-frontier bits not yet visited; a sink's next-frontier words are
// cleared here (no step reads a sink's frontier bits), so after the step the
// only non-zero frontier words are those of emitted vertices.
template <int W>
__device__ __forceinline__ void msbfs_apply_body(const DG& g, Part p, Emitter& em, const uint64_t* touched,
                                                 uint32_t ntouched, uint64_t* vis, uint64_t* fb,
                                                 uint8_t* lev, uint32_t iter, const Frontier& nf) {
  const uint8_t nxt = (uint8_t)(iter + 1);
  for (uint32_t base = p.bid * BLOCK; base < ntouched; base += p.nthreads()) {
    const uint32_t i = base + threadIdx.x;
    bool act = false;
    uint32_t r = 0, nit = 0, ne = 0;
    if (i < ntouched) {
      r = (uint32_t)touched[i];
      const bool has_col = slot_has_col(g, r);
#pragma unroll
      for (int w = 0; w < W; w++) {
        const uint64_t v = vis[(size_t)r * W + w];
        uint64_t m = fb[(size_t)r * W + w] & ~v;
        if (!has_col) fb[(size_t)r * W + w] = 0;
        if (!m) continue;
        act = true;
        vis[(size_t)r * W + w] = v | m;
        set_levels(lev + (size_t)r * 64 * W + w * 64, m, nxt);
      }
      if (act && has_col) {
        const uint32_t c = col_of_slot(g, r);
        ne = g.col_ptr[c + 1] - g.col_ptr[c];
        nit = (ne + CHUNK - 1) / CHUNK;
      }
    }
    em_push(em, nf, r, nit, ne, act);
  }
}

template <int W>
__global__ void __launch_bounds__(BLOCK) k_msbfs_apply(DG g, const uint64_t* touched,
                                                       uint32_t ntouched, uint64_t* vis,
                                                       uint64_t* fb, uint8_t* lev,
                                                       uint32_t iter, Frontier nf) {
  LA3_EMITTER(em);
  msbfs_apply_body<W>(g, grid_part(), em, touched, ntouched, vis, fb, lev, iter, nf);
  em_finish(em, nf);
}

// Clears the W words of `bits` of every vertex in the work list (chunk 0
// items: one per vertex). Used on the frontier words a step has consumed.
template <int W>
__device__ __forceinline__ void clear_items_body(Part p, const uint64_t* items, uint32_t nitems,
                                                 uint64_t* bits) {
  for (uint32_t i = p.tid(); i < nitems; i += p.nthreads()) {
    const uint64_t item = items[i];
    if ((uint32_t)(item >> 32) != 0) continue;
    const uint32_t s = (uint32_t)item;
#pragma unroll
    for (int w = 0; w < W; w++) bits[(size_t)s * W + w] = 0;
  }
}

template <int W>
__global__ void __launch_bounds__(BLOCK) k_clear_items(const uint64_t* items, uint32_t nitems, uint64_t* bits) {
  clear_items_body<W>(grid_part(), items, nitems, bits);
}

// Bottom-up step: each row ORs the frontier words of its in-neighbours,
// restricted to the queries that have not reached it, stopping once all
// of them are covered. Rows longer than LONG_ROW go to msbfs_pull_long_body.
// Only rows with a column (senders of the next step) keep frontier words.
template <int W>
__device__ __forceinline__ void msbfs_pull_body(const DG& g, Part p, Emitter& em, uint64_t* vis,
                                                const uint64_t* fa, uint64_t* fb, uint8_t* lev,
                                                uint32_t iter, const uint64_t* valid, const Frontier& nf) {
  const uint8_t nxt = (uint8_t)(iter + 1);
  for (uint32_t base = p.bid * BLOCK; base < g.nrows; base += p.nthreads()) {
    const uint32_t r = base + threadIdx.x;
    bool act = false;
    uint32_t nit = 0, ne = 0, scanned = 0;
    if (r < g.nrows) {
      const uint32_t b = g.row_ptr[r], e = g.row_ptr[r + 1];
      uint64_t need[W], acc[W];
      bool any = false;
#pragma unroll
      for (int w = 0; w < W; w++) {
        need[w] = ~vis[(size_t)r * W + w] & valid[w];
        acc[w] = 0;
        any |= need[w] != 0;
      }
      if (any && e - b <= LONG_ROW) {
        // UB in-neighbours at a time (their frontier words load together),
        // stopping once every needed query is covered: further entries
        // cannot change acc then, so the result is that of the serial scan.
        constexpr uint32_t UB = W <= 2 ? 4 : 2;
        bool done = false;
        uint32_t j = b;
        for (; j + UB <= e && !done; j += UB) {
          uint32_t src[UB];
#pragma unroll
          for (uint32_t t = 0; t < UB; t++) src[t] = g.row_cols[j + t];
          done = true;
#pragma unroll
          for (int w = 0; w < W; w++) {
            uint64_t f = 0;
#pragma unroll
            for (uint32_t t = 0; t < UB; t++) f |= fa[(size_t)src[t] * W + w];
            acc[w] |= f & need[w];
            done &= acc[w] == need[w];
          }
          scanned += UB;
        }
        for (; j < e && !done; j++) {
          const uint32_t src = g.row_cols[j];
          done = true;
          scanned++;
#pragma unroll
          for (int w = 0; w < W; w++) {
            acc[w] |= fa[(size_t)src * W + w] & need[w];
            done &= acc[w] == need[w];
          }
        }
        const bool has_col = slot_has_col(g, r);
#pragma unroll
        for (int w = 0; w < W; w++) {
          const uint64_t m = acc[w];
          if (!m) continue;
          act = true;
          vis[(size_t)r * W + w] |= m;
          if (has_col) fb[(size_t)r * W + w] = m;
          set_levels(lev + (size_t)r * 64 * W + w * 64, m, nxt);
        }
      }
      if (act && slot_has_col(g, r)) {
        const uint32_t c = col_of_slot(g, r);
        ne = g.col_ptr[c + 1] - g.col_ptr[c];
        nit = (ne + CHUNK - 1) / CHUNK;
      }
    }
    em.scanned += scanned;
    em_push(em, nf, r, nit, ne, act);
  }
}

template <int W>
__global__ void __launch_bounds__(BLOCK) k_msbfs_pull(DG g, uint64_t* vis, const uint64_t* fa,
                                                      uint64_t* fb, uint8_t* lev, uint32_t iter,
                                                      const uint64_t* valid, Frontier nf) {
  LA3_EMITTER(em);
  msbfs_pull_body<W>(g, grid_part(), em, vis, fa, fb, lev, iter, valid, nf);
  em_finish(em, nf);
}

template <int W>
__device__ __forceinline__ void msbfs_pull_long_body(const DG& g, Part p, Emitter& em, uint64_t* vis,
                                                     const uint64_t* fa, uint64_t* fb, uint8_t* lev,
                                                     uint32_t iter, const uint64_t* valid,
                                                     const Frontier& nf) {
  const uint8_t nxt = (uint8_t)(iter + 1);
  for (uint32_t i = p.wid(); i < g.nlong; i += p.nwaves()) {
    const uint32_t r = g.long_rows[i];
    uint64_t need[W], acc[W];
    bool any = false;
#pragma unroll
    for (int w = 0; w < W; w++) {
      need[w] = ~vis[(size_t)r * W + w] & valid[w];
      acc[w] = 0;
      any |= need[w] != 0;
    }
    if (!any) continue;  // wave-uniform: every lane read the same words
    const uint32_t b = g.row_ptr[r], e = g.row_ptr[r + 1];
    uint32_t scanned = 0;
    for (uint32_t j0 = b; j0 < e; j0 += WAVE) {
      const uint32_t j = j0 + lane_id();
      if (j < e) {
        const uint32_t src = g.row_cols[j];
        scanned++;
#pragma unroll
        for (int w = 0; w < W; w++) acc[w] |= fa[(size_t)src * W + w] & need[w];
      }
      bool done = true;
#pragma unroll
      for (int w = 0; w < W; w++) {
        uint64_t x = acc[w];
#pragma unroll
        for (int d = WAVE / 2; d > 0; d >>= 1) x |= __shfl_xor(x, d, WAVE);
        acc[w] = x;
        done &= x == need[w];
      }
      if (done) break;
    }
    scanned = wave_sum(scanned);
    bool act = false;
    uint32_t nit = 0, ne = 0;
    if (lane_id() == 0) {
      em.scanned += scanned;
      const bool has_col = slot_has_col(g, r);
#pragma unroll
      for (int w = 0; w < W; w++) {
        uint64_t m = acc[w];
        if (!m) continue;
        act = true;
        vis[(size_t)r * W + w] |= m;
        if (has_col) fb[(size_t)r * W + w] = m;
        set_levels(lev + (size_t)r * 64 * W + w * 64, m, nxt);
      }
      if (act && has_col) {
        const uint32_t c = col_of_slot(g, r);
        ne = g.col_ptr[c + 1] - g.col_ptr[c];
        nit = (ne + CHUNK - 1) / CHUNK;
      }
    }
    em_push(em, nf, r, nit, ne, act);
  }
}

template <int W>
__global__ void __launch_bounds__(BLOCK) k_msbfs_pull_long(DG g, uint64_t* vis, const uint64_t* fa,
                                                           uint64_t* fb, uint8_t* lev, uint32_t iter,
                                                           const uint64_t* valid, Frontier nf) {
  LA3_EMITTER(em);
  msbfs_pull_long_body<W>(g, grid_part(), em, vis, fa, fb, lev, iter, valid, nf);
  em_finish(em, nf);
}

// ------------------------------------------------------------ batched SSSP ---
// k SSSP queries (SpVertex, sssp.h:26-46) in one Bellman-Ford traversal.
// dist[s*K + q] (K = 64*QW); lane l of a wavefront owns queries l, l+64, ...
// cm / nm hold, per vertex, the queries whose distance improved in the last /
// current step; only those queries are relaxed from it. As in the single
// query, the loop relaxes regular rows and one pull pass computes the sinks.

// Appends `slot` of every lane whose `flag` is set (in lane order).
__device__ __forceinline__ void em_push_lanes(Emitter& em, const Frontier& f, uint32_t slot, bool flag) {
  const uint64_t m = __ballot(flag);
  if (!m) return;  // wave-uniform
  const uint32_t tot = __popcll(m);
  if (em.cnt + tot > QCAP) em_flush(em, f);
  if (flag) {
    em.q[em.cnt + __popcll(m & ((1ull << lane_id()) - 1))] = slot;
    em.verts++;
    em.cols++;
  }
  em.cnt += tot;
}

// Appends `slot` from lane 0 when `flag` (wave-uniform) is set.
__device__ __forceinline__ void em_push_one(Emitter& em, const Frontier& f, uint32_t slot, bool flag) {
  if (!flag) return;
  if (em.cnt + 1 > QCAP) em_flush(em, f);
  if (lane_id() == 0) { em.q[em.cnt] = slot; em.verts++; em.cols++; }
  em.cnt++;
}

template <int QW>
__global__ void k_mssp_roots(DG g, const uint32_t* root_slots, uint32_t k, uint32_t* dist, uint64_t* cm) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (uint32_t q = 0; q < k; q++) {
    const uint32_t s = root_slots[q];
    dist[(size_t)s * 64 * QW + q] = 0;
    cm[(size_t)s * QW + q / 64] |= 1ull << (q & 63);
  }
}

__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ uint64_t rdlane64(uint64_t v, uint32_t l) {
  return ((uint64_t)rdlane((uint32_t)(v >> 32), l) << 32) | rdlane((uint32_t)v, l);
}

// Near/far split as in k_sssp_push, per (vertex, query) pair with one
// threshold T for the batch: far pairs keep their bit for the next step and
// their vertex is re-added (chunk 0) without scanning the column. Improved
// rows are queued without deduplication, so no relaxation waits on an atomic.
//
// Wavefront w takes items w, w + nw, w + 2 nw, ... of the work list (the
// chunks of a hub's column, adjacent in the list, go to different
// wavefronts). The items' slots and entry ranges are read 64 at a time, one
// item per lane, and handed to the whole wavefront with readlane; a column's
// entries are read 64 at a time the same way. Per item, the vertex's
// distance row, its query mask and its first 64 entries are loaded
// together, instead of behind a chain of dependent loads (item, column
// pointers, mask, distances, entries), and the entry indices live in one
// register, not U.
// DIRECT (one mask word per row only): queued rows' work items go straight
// into `touched`, which is then the next step's work list (no items pass).
template <int QW, bool DIRECT = false>
__device__ __forceinline__ void mssp_push_body(const DG& g, Part p, Emitter& em, const uint64_t* items,
                                               uint32_t nitems, uint32_t* dist, const uint64_t* cm,
                                               uint64_t* nm, uint32_t T, const Frontier& touched,
                                               uint32_t* tst = nullptr, uint32_t ttag = 0) {
  static_assert(!DIRECT || QW == 1, "direct emission needs one mask word per row");
  constexpr uint32_t K = 64 * QW;
  // Edges in groups of U: the U distance rows are loaded before any is
  // used, so a wavefront keeps U row loads in flight.
  constexpr uint32_t U = QW == 1 ? 8 : (QW <= 2 ? 4 : 2);
  const uint32_t lane = lane_id();
  const uint32_t nw = p.nwaves(), wid = p.wid();
  const uint32_t mine = wid < nitems ? (nitems - wid + nw - 1) / nw : 0;
  for (uint32_t b0 = 0; b0 < mine; b0 += WAVE) {  // wave-uniform
    const uint32_t nb = min((uint32_t)WAVE, mine - b0);
    uint32_t ms = 0, mbeg = 0, mend = 0, mfirst = 0;
    if (lane < nb) {
      const uint64_t item = items[wid + (uint64_t)(b0 + lane) * nw];
      ms = (uint32_t)item;
      const uint32_t ch = (uint32_t)(item >> 32);
      const uint32_t c = col_of_slot(g, ms);
      mbeg = g.col_ptr[c] + ch * CHUNK;
      mend = max(mbeg, min(mbeg + CHUNK, g.col_split[c]));
      mfirst = ch == 0;  // first chunk of the column
    }
    for (uint32_t x = 0; x < nb; x++) {
      const uint32_t s = rdlane(ms, x), beg = rdlane(mbeg, x), end = rdlane(mend, x);
      const bool first = rdlane(mfirst, x) != 0;
      // distance row, query mask and first entries: all in flight together
      uint32_t ds[QW];
      uint64_t qm[QW];
#pragma unroll
      for (int j = 0; j < QW; j++) {
        ds[j] = ld_relaxed(&dist[(size_t)s * K + 64 * j + lane]);
        qm[j] = cm[(size_t)s * QW + j];
      }
      uint32_t ne = min((uint32_t)WAVE, end - beg);
      uint32_t rv = lane < ne ? g.col_rows[beg + lane] : 0;
      uint32_t wv = lane < ne ? g.col_w[beg + lane] : 0;
      bool act[QW];
      bool any_near = false, requeue = false;  // requeue: wave-uniform
#pragma unroll
      for (int j = 0; j < QW; j++) {
        const bool a0 = (qm[j] >> lane) & 1;
        if (!a0) ds[j] = SSSP_INF;
        const bool far = a0 && ds[j] >= T;
        act[j] = a0 && !far;
        any_near |= act[j];
        if (far && first) em.nmin_inv = max(em.nmin_inv, ~ds[j]);
        const uint64_t fb = __ballot(far);
        if (fb && first) {
          unsigned long long old = 0;
          if (lane == 0) old = atomicOr((unsigned long long*)&nm[(size_t)s * QW + j], (unsigned long long)fb);
          // one mask word: queued only by the step's first OR into it (as
          // the improvements below), so the touched list has no repeats
          requeue |= QW == 1 ? __ballot(lane == 0 && old == 0) != 0 : true;
        }
      }
      // far pairs of this vertex: queue it again
      if (DIRECT) {
        if (requeue) {
          const uint32_t ne = g.col_split[s] - g.col_ptr[s];
          em_push(em, touched, s, lane == 0 ? (ne + CHUNK - 1) / CHUNK : 0u, lane == 0 ? ne : 0u, lane == 0);
        }
      } else {
        em_push_one(em, touched, s, requeue);
      }
      if (__ballot(any_near) == 0) continue;  // nothing near: no column scan
      bool actw[QW];  // wave-uniform: some lane relaxes in word j
#pragma unroll
      for (int j = 0; j < QW; j++) actw[j] = __ballot(act[j]) != 0;
      for (uint32_t e0 = beg; e0 < end; e0 += WAVE) {  // wave-uniform
        if (e0 != beg) {
          ne = min((uint32_t)WAVE, end - e0);
          rv = lane < ne ? g.col_rows[e0 + lane] : 0;
          wv = lane < ne ? g.col_w[e0 + lane] : 0;
        }
        for (uint32_t u0 = 0; u0 < ne; u0 += U) {
          const uint32_t n = min(U, ne - u0);
          // Row bases are wave-uniform (scalar registers); lanes add their
          // query offset. A word of a row is loaded whole when any lane
          // relaxes in it (the lines are fetched either way), so the loads
          // need no per-lane branches.
          uint32_t r[U], cur[U][QW];
#pragma unroll
          for (uint32_t u = 0; u < U; u++) {
            r[u] = rdlane(rv, min(u0 + u, (uint32_t)WAVE - 1));
            const uint32_t* row = dist + (size_t)r[u] * K;
#pragma unroll
            for (int j = 0; j < QW; j++) cur[u][j] = (u < n && actw[j]) ? ld_relaxed(row + 64 * j + lane) : 0u;
          }
          // Lane u collects entry u's row and improved-query masks, so the
          // mask updates and the queue appends of the whole group are one
          // lane-parallel atomic and one compaction.
          uint32_t vrow = 0;
          uint64_t vbits[QW];
#pragma unroll
          for (int j = 0; j < QW; j++) vbits[j] = 0;
#pragma unroll
          for (uint32_t u = 0; u < U; u++) {
            if (u >= n) continue;  // wave-uniform; no break, so the loop unrolls
            const uint32_t w = rdlane(wv, u0 + u);
            vrow = lane == u ? r[u] : vrow;
#pragma unroll
            for (int j = 0; j < QW; j++) {
              // The pre-loaded value decides "improved"; the atomic's result
              // is not waited for. A stale value can only over-report (the
              // pair is re-activated and relaxes with its current, smaller
              // distance); a real improvement is never missed.
              const uint32_t nd = ds[j] + w;
              const bool imp = act[j] && nd < cur[u][j];
              if (imp) {
                atomicMin(dist + (size_t)r[u] * K + 64 * j + lane, nd);
                em.nmin_inv = max(em.nmin_inv, ~nd);
              }
              const uint64_t b = __ballot(imp);
              vbits[j] = lane == u ? b : vbits[j];
            }
          }
          // Measurement runs: lane u counts entry u's row once per step (every
          // entry of a column with a near query carries a message).
          if (tst && lane < n && atomicMax(&tst[vrow], ttag) < ttag) em.touched++;
          // With one mask word per row, the row is queued by the lane whose
          // atomicOr finds the word still zero: the first improvement of the
          // row in this step (nm starts the step zeroed, far bits included).
          // Wider rows are queued on every improvement (waiting on QW
          // returned words costs more than it saves) and k_mssp_items drops
          // the repeats; with one word the touched list has none.
          bool any = false;
#pragma unroll
          for (int j = 0; j < QW; j++) {
            if (!vbits[j]) continue;
            unsigned long long* w = (unsigned long long*)&nm[(size_t)vrow * QW + j];
            if (QW == 1) any = atomicOr(w, (unsigned long long)vbits[j]) == 0;
            else { atomicOr(w, (unsigned long long)vbits[j]); any = true; }
          }
          if (DIRECT) {
            if (__ballot(any)) {  // wave-uniform; a regular row's column index is its slot
              uint32_t ne = 0;
              if (any) ne = g.col_split[vrow] - g.col_ptr[vrow];
              em_push(em, touched, vrow, (ne + CHUNK - 1) / CHUNK, ne, any);
            }
          } else {
            em_push_lanes(em, touched, vrow, any);
          }
        }
      }
    }
  }
}

template <int QW>
__global__ void __launch_bounds__(BLOCK) k_mssp_push(DG g, const uint64_t* items, uint32_t nitems,
                                                     uint32_t* dist, const uint64_t* cm, uint64_t* nm,
                                                     uint32_t T, Frontier touched, uint32_t* zc) {
  zero_ctr(zc);
  LA3_EMITTER(em);
  mssp_push_body<QW, QW == 1>(g, grid_part(), em, items, nitems, dist, cm, nm, T, touched);
  em_finish(em, touched);
}

// ------------------------------------------- batched SSSP, narrow rows ---
// Batches of at most 32 queries: a distance row holds LPR (1, 2, 4, 8, 16
// or 32) queries, lane l works on query l % LPR of entry group l / LPR, and a
// wavefront relaxes 64 / LPR entries of a column at once. A small batch
// then reads LPR x 4 B per entry instead of a 256-B row, and its lanes are
// busy. The relaxation rule, the near/far split and the queueing are those
// of mssp_push_body (one mask word per row); distances are exact either way.
template <int LPR>
__device__ __forceinline__ void mssp_sub_push_body(const DG& g, Part p, Emitter& em, const uint64_t* items,
                                                   uint32_t nitems, uint32_t* dist, const uint64_t* cm,
                                                   uint64_t* nm, uint32_t T, const Frontier& next) {
  constexpr uint32_t EPP = WAVE / LPR;       // entries per pass
  constexpr uint32_t U = LPR >= 4 ? 4 : LPR;  // passes in flight (U * EPP <= 64)
  constexpr uint64_t QMASK = LPR == 64 ? ~0ull : ((1ull << LPR) - 1);
  const uint32_t lane = lane_id(), grp = lane / LPR, q = lane % LPR;
  const uint32_t nw = p.nwaves(), wid = p.wid();
  const uint32_t mine = wid < nitems ? (nitems - wid + nw - 1) / nw : 0;
  for (uint32_t b0 = 0; b0 < mine; b0 += WAVE) {  // wave-uniform
    const uint32_t nb = min((uint32_t)WAVE, mine - b0);
    uint32_t ms = 0, mbeg = 0, mend = 0, mfirst = 0;
    if (lane < nb) {
      const uint64_t item = items[wid + (uint64_t)(b0 + lane) * nw];
      ms = (uint32_t)item;
      const uint32_t ch = (uint32_t)(item >> 32);
      const uint32_t c = col_of_slot(g, ms);
      mbeg = g.col_ptr[c] + ch * CHUNK;
      mend = max(mbeg, min(mbeg + CHUNK, g.col_split[c]));
      mfirst = ch == 0;
    }
    for (uint32_t x = 0; x < nb; x++) {
      const uint32_t s = rdlane(ms, x), beg = rdlane(mbeg, x), end = rdlane(mend, x);
      const bool first = rdlane(mfirst, x) != 0;
      uint32_t ds = ld_relaxed(&dist[(size_t)s * LPR + q]);
      const uint64_t qm = cm[s];
      uint32_t ne = min((uint32_t)WAVE, end - beg);
      uint32_t rv = lane < ne ? g.col_rows[beg + lane] : 0;
      uint32_t wv = lane < ne ? g.col_w[beg + lane] : 0;
      const bool a0 = (qm >> q) & 1;
      if (!a0) ds = SSSP_INF;
      const bool far = a0 && ds >= T;
      const bool act = a0 && !far;
      if (far && first) em.nmin_inv = max(em.nmin_inv, ~ds);
      const uint64_t fb = __ballot(far) & QMASK;  // group 0 holds every query
      bool requeue = false;  // wave-uniform: the step's first OR into nm[s]
      if (fb && first) {
        unsigned long long old = 0;
        if (lane == 0) old = atomicOr((unsigned long long*)&nm[s], (unsigned long long)fb);
        requeue = __ballot(lane == 0 && old == 0) != 0;
      }
      if (requeue) {  // s's work items go straight into the next work list
        const uint32_t ne = g.col_split[s] - g.col_ptr[s];
        em_push(em, next, s, lane == 0 ? (ne + CHUNK - 1) / CHUNK : 0u, lane == 0 ? ne : 0u, lane == 0);
      }
      if ((__ballot(act) & QMASK) == 0) continue;  // nothing near: no column scan
      for (uint32_t e0 = beg; e0 < end; e0 += WAVE) {  // wave-uniform
        if (e0 != beg) {
          ne = min((uint32_t)WAVE, end - e0);
          rv = lane < ne ? g.col_rows[e0 + lane] : 0;
          wv = lane < ne ? g.col_w[e0 + lane] : 0;
        }
        for (uint32_t u0 = 0; u0 < ne; u0 += U * EPP) {
          const uint32_t n = min(U * EPP, ne - u0);
          uint32_t r[U], w[U], cur[U];
#pragma unroll
          for (uint32_t u = 0; u < U; u++) {
            const uint32_t e = u * EPP + grp;  // this lane's entry in the group
            const uint32_t src = min(u0 + e, (uint32_t)WAVE - 1);
            r[u] = __shfl(rv, src, WAVE);
            w[u] = __shfl(wv, src, WAVE);
            cur[u] = (u * EPP < n && e < n) ? ld_relaxed(&dist[(size_t)r[u] * LPR + q]) : 0u;
          }
          uint32_t vrow = 0;
          uint64_t vbits = 0;
#pragma unroll
          for (uint32_t u = 0; u < U; u++) {
            if (u * EPP >= n) continue;  // wave-uniform
            const uint32_t e = u * EPP + grp;
            const uint32_t nd = ds + w[u];
            const bool imp = act && e < n && nd < cur[u];
            if (imp) {
              atomicMin(&dist[(size_t)r[u] * LPR + q], nd);
              em.nmin_inv = max(em.nmin_inv, ~nd);
            }
            if constexpr (LPR == 1) {  // one pass; every lane collects its own entry
              vrow = r[u];
              vbits = imp ? 1 : 0;
            } else {
              const uint64_t b = __ballot(imp);
#pragma unroll
              for (uint32_t gg = 0; gg < EPP; gg++) {
                const uint32_t at = u * EPP + gg;  // collecting lane of entry group gg
                vrow = lane == at ? rdlane(r[u], gg * LPR) : vrow;
                vbits = lane == at ? (b >> (gg * LPR)) & QMASK : vbits;
              }
            }
          }
          // The lane whose atomicOr finds the row's word zero queues the row
          // (its first improvement in this step), so each row is queued
          // once: its work items (the regular-row part of its column, a
          // row's column index being its slot) go straight into the next
          // work list, with no touched list and no items pass.
          bool any = false;
          if (vbits) any = atomicOr((unsigned long long*)&nm[vrow], (unsigned long long)vbits) == 0;
          if (__ballot(any)) {  // wave-uniform
            uint32_t ne = 0;
            if (any) ne = g.col_split[vrow] - g.col_ptr[vrow];
            em_push(em, next, vrow, (ne + CHUNK - 1) / CHUNK, ne, any);
          }
        }
      }
    }
  }
}

// next: the next step's work list (its counters, the smallest queued
// distance included, are read by the host after the launch).
template <int LPR>
__global__ void __launch_bounds__(BLOCK) k_mssp_sub_push(DG g, const uint64_t* items, uint32_t nitems,
                                                         uint32_t* dist, const uint64_t* cm, uint64_t* nm,
                                                         uint32_t T, Frontier next, uint32_t* zc) {
  zero_ctr(zc);
  LA3_EMITTER(em);
  mssp_sub_push_body<LPR>(g, grid_part(), em, items, nitems, dist, cm, nm, T, next);
  em_finish(em, next);
}

// Roots: dist[slot * K + q] = 0 and the query's mask bit (K lanes per row).
__global__ void k_mssp_roots_k(DG g, const uint32_t* root_slots, uint32_t k, uint32_t K, uint32_t* dist,
                               uint64_t* cm) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (uint32_t q = 0; q < k; q++) {
    const uint32_t s = root_slots[q];
    dist[(size_t)s * K + q] = 0;
    cm[(size_t)s * (K / 64 ? K / 64 : 1) + q / 64] |= 1ull << (q & 63);
  }
}

// Sink rows from their final in-neighbours, LPR lanes per sink row.
template <int LPR>
__global__ void k_mssp_sub_sinks(DG g, uint32_t* dist) {
  constexpr uint32_t RPW = WAVE / LPR;
  const uint32_t nwaves = gridDim.x * (BLOCK / WAVE);
  const uint32_t wid = blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE;
  const uint32_t lane = lane_id(), grp = lane / LPR, q = lane % LPR;
  for (uint64_t base = (uint64_t)wid * RPW; base < g.nsink; base += (uint64_t)nwaves * RPW) {
    if (base + grp >= g.nsink) continue;
    const uint32_t r = g.nreg + (uint32_t)base + grp;
    uint32_t y = SSSP_INF;
    const uint32_t e1 = g.row_ptr[r + 1];
    uint32_t e = g.row_ptr[r];
    for (; e + 8 <= e1; e += 8) {  // eight in-neighbours in flight (min: any order)
      uint32_t v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = dist[(size_t)g.row_cols[e + u] * LPR + q] + g.row_w[e + u];
#pragma unroll
      for (int u = 0; u < 8; u++) y = min(y, v[u]);
    }
    for (; e < e1; e++) y = min(y, dist[(size_t)g.row_cols[e] * LPR + q] + g.row_w[e]);
    uint32_t* pd = &dist[(size_t)r * LPR + q];
    *pd = min(*pd, y);
  }
}

// Turns the touched list into push work items for the regular-row part of
// each column. With `dedup` a row may repeat in the list and the first
// occurrence in this step is kept (stamp); without, the list holds each row
// once (rows of one mask word: the push queues a row only on the step's
// first OR into its word, which starts the step zero at every row that has
// items, since the finished step's masks are cleared at its work list).
__device__ __forceinline__ void mssp_items_body(const DG& g, Part p, Emitter& em, const uint64_t* touched,
                                                uint32_t ntouched, uint32_t* stamp, uint32_t tag, bool dedup,
                                                const Frontier& nf) {
  for (uint32_t base = p.bid * BLOCK; base < ntouched; base += p.nthreads()) {
    const uint32_t i = base + threadIdx.x;
    uint32_t r = 0, nit = 0, ne = 0;
    bool first = false;
    if (i < ntouched) {
      r = (uint32_t)touched[i];
      first = !dedup || atomicMax(&stamp[r], tag) < tag;
      if (first) {
        ne = g.col_split[r] - g.col_ptr[r];  // touched rows are regular: column index == slot
        nit = (ne + CHUNK - 1) / CHUNK;
      }
    }
    em_push(em, nf, r, nit, ne, first);
  }
}

__global__ void __launch_bounds__(BLOCK) k_mssp_items(DG g, const uint64_t* touched, uint32_t ntouched,
                                                      uint32_t* stamp, uint32_t tag, int dedup, Frontier nf,
                                                      uint32_t* zc) {
  zero_ctr(zc);
  LA3_EMITTER(em);
  mssp_items_body(g, grid_part(), em, touched, ntouched, stamp, tag, dedup != 0, nf);
  em_finish(em, nf);
}

template <int QW>
__global__ void __launch_bounds__(BLOCK) k_mssp_sinks(DG g, uint32_t* dist) {
  constexpr uint32_t K = 64 * QW;
  const uint32_t nwaves = gridDim.x * (BLOCK / WAVE);
  const uint32_t wid = blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE;
  const uint32_t lane = lane_id();
  for (uint32_t i = wid; i < g.nsink; i += nwaves) {
    const uint32_t r = g.nreg + i;
    uint32_t y[QW];
#pragma unroll
    for (int j = 0; j < QW; j++) y[j] = SSSP_INF;
    const uint32_t e1 = g.row_ptr[r + 1];
    uint32_t e = g.row_ptr[r];
    for (; e + 8 <= e1; e += 8) {  // eight in-neighbours' rows in flight (min: any order)
      uint32_t src[8], w[8];
#pragma unroll
      for (int u = 0; u < 8; u++) { src[u] = g.row_cols[e + u]; w[u] = g.row_w[e + u]; }
#pragma unroll
      for (int j = 0; j < QW; j++) {
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = dist[(size_t)src[u] * K + 64 * j + lane];
#pragma unroll
        for (int u = 0; u < 8; u++) y[j] = min(y[j], v[u] + w[u]);
      }
    }
    for (; e < e1; e++) {
      const uint32_t src = g.row_cols[e], w = g.row_w[e];
#pragma unroll
      for (int j = 0; j < QW; j++) y[j] = min(y[j], dist[(size_t)src * K + 64 * j + lane] + w);
    }
#pragma unroll
    for (int j = 0; j < QW; j++) {
      uint32_t* p = &dist[(size_t)r * K + 64 * j + lane];
      *p = min(*p, y[j]);
    }
  }
}

// out[q * nslots + id] = v[slot(id) * K + q] for q < k, through a 64 x 64
// LDS tile: rows are read whole (lane = query, coalesced) and written per
// query as runs of 64 consecutive ids (lane = id, coalesced). The tile's 64
// slots are read once, one per lane, and each wavefront keeps its 16 row
// loads (and then its 16 stores) in flight together.
template <class T>
__device__ __forceinline__ void k_batch_output_body(const DG& g, const T* v, uint32_t K, uint32_t k, T* out) {
  __shared__ T tile[64][65];
  const uint32_t lane = lane_id(), wv = threadIdx.x / WAVE;
  constexpr uint32_t PER = 64 / (BLOCK / WAVE);  // rows (then queries) per wavefront
  const uint32_t ntile = (g.nslots + 63) / 64;
  for (uint32_t t = blockIdx.x; t < ntile; t += gridDim.x) {
    const uint32_t id0 = t * 64;
    const uint32_t nid = min(64u, g.nslots - id0);
    const uint32_t msl = lane < nid ? g.slot_of_id[id0 + lane] : 0;
    for (uint32_t q0 = 0; q0 < k; q0 += 64) {
      T val[PER];
#pragma unroll
      for (uint32_t x = 0; x < PER; x++) {
        const uint32_t i = wv + (BLOCK / WAVE) * x;
        const uint32_t sl = rdlane(msl, i);
        val[x] = (i < nid && q0 + lane < k) ? v[(size_t)sl * K + q0 + lane] : T();
      }
#pragma unroll
      for (uint32_t x = 0; x < PER; x++) tile[wv + (BLOCK / WAVE) * x][lane] = val[x];
      __syncthreads();
#pragma unroll
      for (uint32_t x = 0; x < PER; x++) {
        const uint32_t qi = wv + (BLOCK / WAVE) * x;
        if (q0 + qi < k && lane < nid) out[(size_t)(q0 + qi) * g.nslots + id0 + lane] = tile[lane][qi];
      }
      __syncthreads();
    }
  }
}

template <class T>
__global__ void __launch_bounds__(BLOCK) k_batch_output(DG g, const T* v, uint32_t K, uint32_t k, T* out) {
  k_batch_output_body(g, v, K, k, out);
}

// hops[q * nslots + id] = lev[slot(id) * K + q] for q < k (same tiling).
__global__ void k_msbfs_output(DG g, const uint8_t* lev, uint32_t K, uint32_t k, uint8_t* hops) {
  k_batch_output_body(g, lev, K, k, hops);
}

}  // namespace la3

namespace la3 {

// ------------------------------------------------------------- batched PPR ---
// k personalized-PageRank queries (the PR iteration of pr.h:34-45 with the
// teleport alpha at each query's seed) in FP32 with compensation:
//   state per (vertex, query): rank hi (f32) + its compensation lo (f32);
//   message x = f32((hi + lo) / degree);
//   row sum y = Kahan sum of the messages in the reference's per-row order
//   (sources first, then regular columns), one wavefront per row, lane = query;
//   rank' = tele + (1 - alpha) * y in error-free transforms, renormalized to
//   (hi, lo).
// Rows longer than PSEG entries are summed in PSEG-entry segments whose
// (sum, compensation) pairs are combined in segment order by k_mppr_combine.
constexpr uint32_t PSEG = PSEG_ENTRIES;

struct PprBatchArgs {
  const float* x;        // messages [slot * K + q]; narrow rows: [rank_of[slot] * K + q]
  float* xnext;
  float* hi;             // rank [slot * K + q]
  float* lo;             // compensation
  const uint32_t* deg;   // out-degree by slot
  const uint32_t* seeds; // seed slot per query [k]
  uint32_t k;
  float alpha_hi, alpha_lo;  // alpha as an unevaluated f32 pair (teleport at the seed)
  float damp_hi, damp_lo;    // (1 - alpha) as an unevaluated f32 pair
  double* resid;            // per block x K partial |delta| sums (FP64)
  float* part;              // long-row segment partials [seg * 2K] (sum, comp)
  const uint32_t* seg_row;  // segment -> row
  const uint32_t* seg_beg;  // segment -> first entry
  const uint32_t* long_ppr; // rows with > PSEG entries
  uint32_t nlong_ppr;
  const uint32_t* seg_first;  // long row i -> first segment index (nlong_ppr + 1 entries)
  // An iteration's new rank is a function of its row sum alone, and the
  // next iteration's messages are formed from it in registers: only the
  // last iteration's ranks and residual are ever read. Earlier iterations
  // skip storing (hi, lo) (all but the last two) and reading the old pair
  // for the residual (all but the last). Zero: store and compute both.
  uint32_t skip_rank, skip_resid;
};

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// (hi, lo) of (th + tl) + (dh + dl) * (s - c), all error-free transforms.
__device__ __forceinline__ void ppr_apply(float th, float tl, float dh, float dl, float s, float c,
                                          float& nh, float& nl) {
  const float p = dh * s;
  const float pe = fmaf(dh, s, -p) + fmaf(dl, s, -dh * c);   // product error + low-order terms
  const float t = th + p;                                     // TwoSum(th, p)
  const float bv = t - th;
  const float te = (th - (t - bv)) + (p - bv);
  const float e = te + pe + tl;
  nh = t + e;                                                 // Fast2Sum renormalization
  nl = e - (nh - t);
}

__device__ __forceinline__ float ppr_msg(float h, float l, uint32_t d) {
  const float fd = (float)d;
  const float q = h / fd;
  const float r = fmaf(-q, fd, h);   // exact remainder of h / d
  return q + (r + l) / fd;
}

// seed: this lane's queries' seed slots (~0 past k); deg: out-degree of r.
template <int QW>
__device__ __forceinline__ void ppr_finish(const PprBatchArgs& a, uint32_t r, const float* s, const float* c,
                                           bool loop, double* res, const uint32_t* seed, uint32_t deg) {
  constexpr uint32_t K = 64 * QW;
  const uint32_t lane = lane_id();
#pragma unroll
  for (int w = 0; w < QW; w++) {
    const uint32_t q = 64 * w + lane;
    if (q >= a.k) continue;
    const size_t o = (size_t)r * K + q;
    const bool sd = seed[w] == r;
    float nh, nl;
    ppr_apply(sd ? a.alpha_hi : 0.f, sd ? a.alpha_lo : 0.f, a.damp_hi, a.damp_lo, s[w], c[w], nh, nl);
    if (!a.skip_resid) res[w] += fabs(((double)nh + nl) - ((double)a.hi[o] + a.lo[o]));
    if (!a.skip_rank) {
      a.hi[o] = nh;
      a.lo[o] = nl;
    }
    if (loop) a.xnext[o] = ppr_msg(nh, nl, deg);
  }
}

template <int QW>
__device__ __forceinline__ void ppr_seeds(const PprBatchArgs& a, uint32_t* seed) {
#pragma unroll
  for (int w = 0; w < QW; w++) {
    const uint32_t q = 64 * w + lane_id();
    seed[w] = q < a.k ? a.seeds[q] : 0xffffffffu;
  }
}

// Per-block, per-query reduction of the residual partials (FP64, fixed order).
// red: LDS of (BLOCK / WAVE) x K doubles.
template <int QW>
__device__ __forceinline__ void ppr_resid_flush(const PprBatchArgs& a, double* res, uint32_t bid, double* red) {
  constexpr uint32_t K = 64 * QW;
  const uint32_t w0 = threadIdx.x / WAVE, lane = lane_id();
#pragma unroll
  for (int w = 0; w < QW; w++) red[w0 * K + 64 * w + lane] = res[w];
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < K; q += BLOCK) {
    double t = 0;
    for (uint32_t j = 0; j < BLOCK / WAVE; j++) t += red[j * K + q];
    a.resid[(size_t)bid * K + q] = t;
  }
}

// Kahan sums of the messages of entries [b, e), lane = query. Lanes of
// queries beyond k load nothing (their vector-load lanes are off), so a
// small batch reads only the cache lines of its live queries.
template <int QW>
__device__ __forceinline__ void ppr_row_sum(const DG& g, const PprBatchArgs& a, uint32_t b, uint32_t e,
                                            float* s, float* c) {
  constexpr uint32_t K = 64 * QW;
  // messages of UR entries loaded before any is added (adds stay in entry
  // order): UR k-wide rows in flight per wavefront
  constexpr uint32_t UR = QW <= 2 ? 8 : 4;
  const uint32_t lane = lane_id();
  bool live[QW];
#pragma unroll
  for (int w = 0; w < QW; w++) live[w] = 64 * w + lane < a.k;
  uint32_t j = b;
  for (; j + UR <= e; j += UR) {
    uint32_t sc[UR];
#pragma unroll
    for (uint32_t u = 0; u < UR; u++) sc[u] = g.row_cols[j + u];
#pragma unroll
    for (int w = 0; w < QW; w++) {
      if (!live[w]) continue;
      float v[UR];
#pragma unroll
      for (uint32_t u = 0; u < UR; u++) v[u] = a.x[(size_t)sc[u] * K + 64 * w + lane];
#pragma unroll
      for (uint32_t u = 0; u < UR; u++) kahan_add(s[w], c[w], v[u]);
    }
  }
  if (j < e) {  // the last < UR entries: loaded together (clamped to e - 1), added in order
    const uint32_t rem = e - j;
    uint32_t sc[UR];
#pragma unroll
    for (uint32_t u = 0; u < UR; u++) sc[u] = g.row_cols[min(j + u, e - 1)];
#pragma unroll
    for (int w = 0; w < QW; w++) {
      if (!live[w]) continue;
      float v[UR];
#pragma unroll
      for (uint32_t u = 0; u < UR; u++) v[u] = a.x[(size_t)sc[u] * K + 64 * w + lane];
#pragma unroll
      for (uint32_t u = 0; u < UR; u++)
        if (u < rem) kahan_add(s[w], c[w], v[u]);
    }
  }
}

// Rows [r0, r1) with at most PSEG entries: one wavefront per row (rows
// r0 + w, r0 + w + nw, ... for wavefront w). The rows' entry ranges and
// out-degrees are read 64 at a time, one row per lane, and handed out with
// readlane, so a row's first message loads do not wait behind its pointers.
template <int QW>
__device__ __forceinline__ void mppr_rows_body(const DG& g, Part p, const PprBatchArgs& a, uint32_t r0,
                                               uint32_t r1, bool loop, double* red) {
  double res[QW];
  uint32_t seed[QW];
#pragma unroll
  for (int w = 0; w < QW; w++) res[w] = 0.0;
  ppr_seeds<QW>(a, seed);
  const uint32_t lane = lane_id();
  const uint32_t nw = p.nwaves(), wid = p.wid();
  const uint32_t nrow = r1 > r0 ? r1 - r0 : 0;
  const uint32_t mine = wid < nrow ? (nrow - wid + nw - 1) / nw : 0;
  for (uint32_t b0 = 0; b0 < mine; b0 += WAVE) {  // wave-uniform
    const uint32_t nb = min((uint32_t)WAVE, mine - b0);
    uint32_t mr = 0, mb = 0, me = 0, md = 1;
    if (lane < nb) {
      mr = r0 + wid + (b0 + lane) * nw;
      mb = g.row_ptr[mr];
      me = g.row_ptr[mr + 1];
      md = a.deg[mr];
    }
    for (uint32_t x = 0; x < nb; x++) {
      const uint32_t r = rdlane(mr, x), b = rdlane(mb, x), e = rdlane(me, x);
      if (e - b > PSEG) continue;  // wave-uniform; segmented path
      float s[QW], c[QW];
#pragma unroll
      for (int w = 0; w < QW; w++) { s[w] = 0.f; c[w] = 0.f; }
      ppr_row_sum<QW>(g, a, b, e, s, c);
      ppr_finish<QW>(a, r, s, c, loop, res, seed, rdlane(md, x));
    }
  }
  ppr_resid_flush<QW>(a, res, p.bid, red);
}

template <int QW>
__global__ void __launch_bounds__(BLOCK) k_mppr_rows(DG g, PprBatchArgs a, uint32_t r0, uint32_t r1,
                                                     int loop) {
  __shared__ double red[(BLOCK / WAVE) * 64 * QW];
  mppr_rows_body<QW>(g, grid_part(), a, r0, r1, loop != 0, red);
}

// One wavefront per PSEG-entry segment of a long row: partial (sum, comp).
// Segments [0, nseg) of a.seg_row / a.seg_beg / a.part.
template <int QW>
__device__ __forceinline__ void mppr_segs_body(const DG& g, Part p, const PprBatchArgs& a, uint32_t nseg) {
  constexpr uint32_t K = 64 * QW;
  const uint32_t lane = lane_id();
  for (uint32_t sg = p.wid(); sg < nseg; sg += p.nwaves()) {
    const uint32_t r = a.seg_row[sg], b = a.seg_beg[sg];
    const uint32_t e = min(b + PSEG, g.row_ptr[r + 1]);
    float s[QW], c[QW];
#pragma unroll
    for (int w = 0; w < QW; w++) { s[w] = 0.f; c[w] = 0.f; }
    ppr_row_sum<QW>(g, a, b, e, s, c);
#pragma unroll
    for (int w = 0; w < QW; w++) {
      a.part[(size_t)sg * 2 * K + 64 * w + lane] = s[w];
      a.part[(size_t)sg * 2 * K + K + 64 * w + lane] = c[w];
    }
  }
}

template <int QW>
__global__ void __launch_bounds__(BLOCK) k_mppr_segs(DG g, PprBatchArgs a, uint32_t nseg) {
  mppr_segs_body<QW>(g, grid_part(), a, nseg);
}

// One iteration's row sums in one launch: the short rows, then the segments
// of the long rows (`as`: the segment arrays from the first segment on), so
// the two parts' tails overlap. The residual partials are those of the rows
// part, one per block.
template <int QW>
__global__ void __launch_bounds__(BLOCK) k_mppr_rowsegs(DG g, PprBatchArgs a, uint32_t r0, uint32_t r1, int loop,
                                                        PprBatchArgs as, uint32_t nseg) {
  __shared__ double red[(BLOCK / WAVE) * 64 * QW];
  mppr_rows_body<QW>(g, grid_part(), a, r0, r1, loop != 0, red);
  mppr_segs_body<QW>(g, grid_part(), as, nseg);
}

// Combines each long row's segment partials in segment order, then applies.
template <int QW>
__device__ __forceinline__ void mppr_combine_body(const DG& g, Part p, const PprBatchArgs& a, uint32_t r0,
                                                  uint32_t r1, bool loop, double* red) {
  constexpr uint32_t K = 64 * QW;
  const uint32_t lane = lane_id();
  double res[QW];
  uint32_t seed[QW];
#pragma unroll
  for (int w = 0; w < QW; w++) res[w] = 0.0;
  ppr_seeds<QW>(a, seed);
  for (uint32_t i = p.wid(); i < a.nlong_ppr; i += p.nwaves()) {
    const uint32_t r = a.long_ppr[i];
    if (r < r0 || r >= r1) continue;
    float s[QW], c[QW];
#pragma unroll
    for (int w = 0; w < QW; w++) { s[w] = 0.f; c[w] = 0.f; }
    // the partials of four segments loaded before their adds (in order)
    const uint32_t f1 = a.seg_first[i + 1];
    uint32_t sg = a.seg_first[i];
    for (; sg + 4 <= f1; sg += 4) {
#pragma unroll
      for (int w = 0; w < QW; w++) {
        if (64 * w + lane >= a.k) continue;
        float ps[4], pc[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; u++) {
          ps[u] = a.part[(size_t)(sg + u) * 2 * K + 64 * w + lane];
          pc[u] = a.part[(size_t)(sg + u) * 2 * K + K + 64 * w + lane];
        }
#pragma unroll
        for (uint32_t u = 0; u < 4; u++) {
          kahan_add(s[w], c[w], ps[u]);
          kahan_add(s[w], c[w], -pc[u]);
        }
      }
    }
    for (; sg < f1; sg++) {
#pragma unroll
      for (int w = 0; w < QW; w++) {
        if (64 * w + lane >= a.k) continue;
        kahan_add(s[w], c[w], a.part[(size_t)sg * 2 * K + 64 * w + lane]);
        kahan_add(s[w], c[w], -a.part[(size_t)sg * 2 * K + K + 64 * w + lane]);
      }
    }
    ppr_finish<QW>(a, r, s, c, loop, res, seed, a.deg[r]);
  }
  ppr_resid_flush<QW>(a, res, p.bid, red);
}

template <int QW>
__global__ void __launch_bounds__(BLOCK) k_mppr_combine(DG g, PprBatchArgs a, uint32_t r0, uint32_t r1,
                                                        int loop) {
  __shared__ double red[(BLOCK / WAVE) * 64 * QW];
  mppr_combine_body<QW>(g, grid_part(), a, r0, r1, loop != 0, red);
}

// ------------------------------------------- batched PPR, narrow rows ---
// Batches of at most 32 queries: LPR (1, 2, 4, 8, 16 or 32) lanes per row,
// one per query, and 64 / LPR rows per wavefront, so a small batch does not leave
// most lanes of every row idle. Per query the arithmetic is that of the
// 64-lane kernels above (same entries, same order, same updates), so a
// query's ranks do not depend on the width of the batch it ran in; the FP64
// residual partials are grouped differently. Message rows are stored by
// sender rank (DG::rank_of, decreasing out-degree) and gathered through
// DG::row_rank, so the senders of most entries share a few cached lines.
template <int LPR, uint32_t UR = 8>
__device__ __forceinline__ void ppr_sub_row_sum(const DG& g, const PprBatchArgs& a, uint32_t b, uint32_t e,
                                                bool live, float& s, float& c) {
  const uint32_t q = lane_id() % LPR;
  uint32_t j = b;
  for (; j + UR <= e; j += UR) {
    uint32_t sc[UR];
#pragma unroll
    for (uint32_t u = 0; u < UR; u++) sc[u] = g.row_rank[j + u];
    if (!live) continue;
    float v[UR];
#pragma unroll
    for (uint32_t u = 0; u < UR; u++) v[u] = a.x[(size_t)sc[u] * LPR + q];
#pragma unroll
    for (uint32_t u = 0; u < UR; u++) kahan_add(s, c, v[u]);
  }
  if (j < e) {  // the last < UR entries: loaded together (clamped to e - 1), added in order
    const uint32_t rem = e - j;
    uint32_t sc[UR];
#pragma unroll
    for (uint32_t u = 0; u < UR; u++) sc[u] = g.row_rank[min(j + u, e - 1)];
    if (live) {
      float v[UR];
#pragma unroll
      for (uint32_t u = 0; u < UR; u++) v[u] = a.x[(size_t)sc[u] * LPR + q];
#pragma unroll
      for (uint32_t u = 0; u < UR; u++)
        if (u < rem) kahan_add(s, c, v[u]);
    }
  }
}

// Lane `q` of a row group: the row's update for query q (q < k); xr is the
// row's message row (rank_of[r]; read only when the messages advance).
template <int LPR>
__device__ __forceinline__ void ppr_sub_finish(const PprBatchArgs& a, uint32_t r, uint32_t xr, float s, float c,
                                               bool loop, uint32_t seed, double& res) {
  const uint32_t q = lane_id() % LPR;
  const size_t o = (size_t)r * LPR + q;
  const bool sd = seed == r;
  float nh, nl;
  ppr_apply(sd ? a.alpha_hi : 0.f, sd ? a.alpha_lo : 0.f, a.damp_hi, a.damp_lo, s, c, nh, nl);
  if (!a.skip_resid) res += fabs(((double)nh + nl) - ((double)a.hi[o] + a.lo[o]));
  if (!a.skip_rank) {
    a.hi[o] = nh;
    a.lo[o] = nl;
  }
  if (loop) a.xnext[(size_t)xr * LPR + q] = ppr_msg(nh, nl, a.deg[r]);
}

// Per-block, per-query FP64 residual: the row groups of a wavefront by a
// fixed butterfly, then the wavefronts in order. red: (BLOCK / WAVE) x LPR.
template <int LPR>
__device__ __forceinline__ void ppr_sub_resid_flush(const PprBatchArgs& a, double res, uint32_t bid, double* red) {
#pragma unroll
  for (int d = LPR; d < WAVE; d <<= 1) res += __shfl_xor(res, d, WAVE);
  const uint32_t w0 = threadIdx.x / WAVE, lane = lane_id();
  if (lane < LPR) red[w0 * LPR + lane] = res;
  __syncthreads();
  if (threadIdx.x < LPR) {
    double t = 0;
    for (uint32_t j = 0; j < BLOCK / WAVE; j++) t += red[j * LPR + threadIdx.x];
    a.resid[(size_t)bid * LPR + threadIdx.x] = t;
  }
}

// Chunks of RPW consecutive items (segments, long rows) over the wavefronts
// of a partition. With `spread` and fewer chunks than the first CHUNK_SPAN
// wavefronts (2048 blocks, one full residency of the device), every
// stride-th of those takes one, so the work spreads over all CUs instead of
// the first few blocks; otherwise chunk c goes to wavefront c mod nw, as a
// plain grid-stride loop. Wave-uniform. Spreading pays with one lane per
// row, where a chunk is 64 long segments (R-MAT, one-query PPR, 30
// iterations: scale 18 2.32 -> 1.48 ms, scale 20 5.7 -> 4.5 ms, 22 and 24
// unchanged); with 4 lanes per row it cost 5-7% at scales 18-20.
constexpr uint32_t CHUNK_SPAN = 8192;
struct ChunkMap {
  uint32_t stride, nwe;  // wavefronts wid % stride == 0 take chunks wid / stride + t * nwe
  __device__ ChunkMap(uint64_t nchunks, uint32_t nw, bool spread) {
    const uint32_t span = min(nw, CHUNK_SPAN);
    stride = spread && nchunks < span ? span / (uint32_t)max(nchunks, (uint64_t)1) : 1u;
    nwe = (nw + stride - 1) / stride;
  }
  __device__ bool takes(uint32_t wid) const { return wid % stride == 0; }
  __device__ uint64_t first(uint32_t wid) const { return wid / stride; }
};

// Rows [r0, r1) with at most PSEG entries; row group g of wavefront w takes
// rows r0 + (w + t * nw) * (64 / LPR) + g. red: LDS of (BLOCK / WAVE) x LPR
// doubles (all threads of the block call).
template <int LPR>
__device__ __forceinline__ void mppr_sub_rows_body(const DG& g, Part p, const PprBatchArgs& a, uint32_t r0,
                                                   uint32_t r1, bool loop, double* red) {
  constexpr uint32_t RPW = WAVE / LPR;
  const uint32_t lane = lane_id(), grp = lane / LPR, q = lane % LPR;
  const bool live = q < a.k;
  const uint32_t seed = live ? a.seeds[q] : 0xffffffffu;
  const uint32_t nw = p.nwaves(), wid = p.wid();
  double res = 0.0;
  for (uint64_t base = r0 + (uint64_t)wid * RPW; base < r1; base += (uint64_t)nw * RPW) {
    const uint32_t r = (uint32_t)base + grp;
    if (r >= r1) continue;
    const uint32_t b = g.row_ptr[r], e = g.row_ptr[r + 1];
    if (e - b > PSEG) continue;  // segmented path
    float s = 0.f, c = 0.f;
    ppr_sub_row_sum<LPR>(g, a, b, e, live, s, c);
    if (live) ppr_sub_finish<LPR>(a, r, loop ? g.rank_of[r] : 0u, s, c, loop, seed, res);
  }
  ppr_sub_resid_flush<LPR>(a, res, p.bid, red);
}

template <int LPR>
__global__ void __launch_bounds__(BLOCK) k_mppr_sub_rows(DG g, PprBatchArgs a, uint32_t r0, uint32_t r1, int loop) {
  __shared__ double red[(BLOCK / WAVE) * LPR];
  mppr_sub_rows_body<LPR>(g, grid_part(), a, r0, r1, loop != 0, red);
}

// PSEG-entry segments [0, nseg) of the long rows: partial (sum, comp).
// UR: entries in flight per lane along a segment (0: 32 for LPR <= 8, else
// 8); the sums are the same for any UR (loads batched, adds in order).
template <int LPR, uint32_t UR = 0>
__device__ __forceinline__ void mppr_sub_segs_body(const DG& g, Part p, const PprBatchArgs& a, uint32_t nseg) {
  constexpr uint32_t SEG_UR = UR ? UR : LPR <= 8 ? 32 : 8;
  constexpr uint32_t RPW = WAVE / LPR;
  const uint32_t lane = lane_id(), grp = lane / LPR, q = lane % LPR;
  const bool live = q < a.k;
  const uint32_t nw = p.nwaves(), wid = p.wid();
  const uint64_t nch = ((uint64_t)nseg + RPW - 1) / RPW;
  const ChunkMap cm(nch, nw, LPR == 1);
  if (!cm.takes(wid)) return;
  for (uint64_t ch = cm.first(wid); ch < nch; ch += cm.nwe) {  // wave-uniform
    const uint64_t si = ch * RPW + grp;
    if (si >= nseg) continue;
    const uint32_t sg = (uint32_t)si;
    const uint32_t r = a.seg_row[sg], b = a.seg_beg[sg];
    const uint32_t e = min(b + PSEG, g.row_ptr[r + 1]);
    float s = 0.f, c = 0.f;
    ppr_sub_row_sum<LPR, SEG_UR>(g, a, b, e, live, s, c);
    a.part[(size_t)sg * 2 * LPR + q] = s;
    a.part[(size_t)sg * 2 * LPR + LPR + q] = c;
  }
}

template <int LPR>
__global__ void __launch_bounds__(BLOCK) k_mppr_sub_segs(DG g, PprBatchArgs a, uint32_t nseg) {
  mppr_sub_segs_body<LPR>(g, grid_part(), a, nseg);
}

// Short rows and long-row segments in one launch, as k_mppr_rowsegs. The
// segment part keeps 32 entries in flight per lane only with one lane per
// row: in a launch that also holds the row part, 32 costs three of eight
// waves per SIMD, which pays only there.
template <int LPR>
__global__ void __launch_bounds__(BLOCK) k_mppr_sub_rowsegs(DG g, PprBatchArgs a, uint32_t r0, uint32_t r1,
                                                            int loop, PprBatchArgs as, uint32_t nseg) {
  __shared__ double red[(BLOCK / WAVE) * LPR];
  mppr_sub_rows_body<LPR>(g, grid_part(), a, r0, r1, loop != 0, red);
  mppr_sub_segs_body<LPR, LPR == 1 ? 32 : 8>(g, grid_part(), as, nseg);
}

// Each long row's segment partials in segment order, then its update.
template <int LPR>
__device__ __forceinline__ void mppr_sub_combine_body(const DG& g, Part p, const PprBatchArgs& a, uint32_t r0,
                                                      uint32_t r1, bool loop, double* red) {
  constexpr uint32_t RPW = WAVE / LPR;
  const uint32_t lane = lane_id(), grp = lane / LPR, q = lane % LPR;
  const bool live = q < a.k;
  const uint32_t seed = live ? a.seeds[q] : 0xffffffffu;
  const uint32_t nw = p.nwaves(), wid = p.wid();
  double res = 0.0;
  // long rows spread over the grid as the segments are
  const uint64_t nch = ((uint64_t)a.nlong_ppr + RPW - 1) / RPW;
  const ChunkMap cm(nch, nw, LPR == 1);
  for (uint64_t ch = cm.takes(wid) ? cm.first(wid) : nch; ch < nch; ch += cm.nwe) {  // wave-uniform
    const uint64_t li = ch * RPW + grp;
    if (li >= a.nlong_ppr) continue;
    const uint32_t i = (uint32_t)li;
    const uint32_t r = a.long_ppr[i];
    if (r < r0 || r >= r1 || !live) continue;
    float s = 0.f, c = 0.f;
    // the partials of eight segments loaded before their adds (in order)
    const uint32_t f1 = a.seg_first[i + 1];
    uint32_t sg = a.seg_first[i];
    for (; sg + 8 <= f1; sg += 8) {
      float ps[8], pc[8];
#pragma unroll
      for (uint32_t u = 0; u < 8; u++) {
        ps[u] = a.part[(size_t)(sg + u) * 2 * LPR + q];
        pc[u] = a.part[(size_t)(sg + u) * 2 * LPR + LPR + q];
      }
#pragma unroll
      for (uint32_t u = 0; u < 8; u++) {
        kahan_add(s, c, ps[u]);
        kahan_add(s, c, -pc[u]);
      }
    }
    for (; sg < f1; sg++) {
      kahan_add(s, c, a.part[(size_t)sg * 2 * LPR + q]);
      kahan_add(s, c, -a.part[(size_t)sg * 2 * LPR + LPR + q]);
    }
    ppr_sub_finish<LPR>(a, r, loop ? g.rank_of[r] : 0u, s, c, loop, seed, res);
  }
  ppr_sub_resid_flush<LPR>(a, res, p.bid, red);
}

template <int LPR>
__global__ void __launch_bounds__(BLOCK) k_mppr_sub_combine(DG g, PprBatchArgs a, uint32_t r0, uint32_t r1,
                                                            int loop) {
  __shared__ double red[(BLOCK / WAVE) * LPR];
  mppr_sub_combine_body<LPR>(g, grid_part(), a, r0, r1, loop != 0, red);
}

// Initial state for K lanes per vertex (K = 64 * QW, or LPR; narrow rows,
// K < 64, store the messages by rank), for a run of `iters` iterations.
// Only what a later read needs is written: the messages of the senders
// (regular and source slots), the next-step messages of the sources (no
// row updates them), and the ranks of the slots no iteration writes
// (sources, isolated; regular rows too when one iteration reads them back
// for its residual). Sink ranks come from the sink step.
template <uint32_t K>
__device__ __forceinline__ void mppr_init_body(const DG& g, const PprBatchArgs& a, uint32_t iters) {
  const uint64_t total = (uint64_t)g.nslots * K;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = (uint32_t)(i / K), q = (uint32_t)(i % K);
    const bool seed = q < a.k && a.seeds[q] == s;
    const float th = seed ? a.alpha_hi : 0.f, tl = seed ? a.alpha_lo : 0.f;
    const bool regular = s < g.nreg, sink = !regular && s < g.nrows;
    const bool source = s >= g.nrows && s < g.nrows + g.nsrc;
    if (!sink && !(regular && iters >= 2)) {
      a.hi[i] = th;
      a.lo[i] = tl;
    }
    if (regular || source) {
      const float x = ppr_msg(th, tl, a.deg[s]);
      const uint64_t xi = K < 64 ? (uint64_t)g.rank_of[s] * K + q : i;
      ((float*)a.x)[xi] = x;
      if (source) a.xnext[xi] = x;
    }
  }
}

template <uint32_t K>
__global__ void k_mppr_init_k(DG g, PprBatchArgs a, uint32_t iters) {
  mppr_init_body<K>(g, a, iters);
}

template <int QW>
__global__ void k_mppr_init(DG g, PprBatchArgs a, uint32_t iters) {
  mppr_init_body<64 * QW>(g, a, iters);
}

}  // namespace la3
