// Fused, device-driven steps for a mix of batched queries: any number of
// batches (BFS k-bit, SSSP k-wide, PPR FP32+compensation), on one or several
// resident graphs, advance together, one step per iteration of:
//   k_fused_a  the traversal / row-sum work of every batch (one launch per
//              width class of 64, 128 or 256 queries present);
//   k_fused_b  the work that needs A's results (BFS apply, SSSP work items,
//              PPR long-row combine) and the clearing of consumed masks;
//   k_fused_c  one block: reads each batch's counters on the device, picks
//              the next BFS direction and SSSP threshold, flips buffers,
//              zeroes counters and splits the next launches' blocks among
//              the batches in proportion to their work.
// No kernel takes a per-step host argument, so a run of steps is captured
// once in a hipGraph and replayed; the host reads one word per replay.
// Included once, by capi.hip, after kernels_batch.h.
#pragma once

#include "kernels_batch.h"

namespace la3 {

enum : uint32_t { FK_BFS = 0, FK_SSSP = 1, FK_PPR = 2 };
constexpr uint32_t FMAX_TASKS = 64;   // batches of <= 256 queries in one run
constexpr uint32_t FMAX_WIDTH = 4;    // 64-query words per batch
// Width of the narrow tasks: SSSP and PPR tasks of at most 32 queries hold
// LPR (1, 2, 4, 8, 16 or 32) lanes per vertex row, as the batched calls'
// narrow-row kernels do, instead of a 64-lane row.
constexpr uint32_t FW_NARROW = 3;

// Launch classes: the wide tasks of 1, 2 and 4 query words (classes 0-2),
// and one class per (program, LPR) of the narrow tasks (3-8 SSSP, 9-14 PPR),
// each launched by a kernel of its own, so that each launch carries the
// registers of one body only.
constexpr uint32_t NCLS = 15;
__host__ __device__ constexpr uint32_t lpr_log2(uint32_t lpr) {
  return lpr <= 1 ? 0 : lpr == 2 ? 1 : lpr == 4 ? 2 : lpr == 8 ? 3 : lpr == 16 ? 4 : 5;
}
__host__ __device__ constexpr uint32_t wide_cls(uint32_t width) { return width == 1 ? 0 : width == 2 ? 1 : 2; }
__host__ __device__ constexpr uint32_t narrow_cls(uint32_t kind, uint32_t lpr) {
  return 3 + (kind == FK_PPR ? 6 : 0) + lpr_log2(lpr);
}

// One batch ("task"): immutable during a run.
struct FTask {
  DG g;
  uint32_t kind, width, k, iters;  // width = ceil(k / 64), or FW_NARROW; iters: PPR iterations
  uint32_t lpr;                    // narrow tasks: lanes per vertex row (1..32); else 0
  uint32_t cls;                    // launch class (wide_cls / narrow_cls)
  uint32_t delta;                  // SSSP near/far delta (0 = Bellman-Ford)
  uint32_t nslots_pad;
  uint32_t* stamp;
  uint32_t* tstamp;                // measurement runs: last step that counted a row as touched
  Frontier q[2];                   // work lists, ping-pong by FCtl::cur
  Frontier touched;
  // BFS
  uint64_t* vis;
  uint64_t* fr[2];                 // frontier words (current = fr[cur])
  uint8_t* lev;
  const uint64_t* valid;
  // SSSP
  uint32_t* dist;
  uint64_t* m[2];                  // improved-query masks (current = m[cur])
  // PPR
  PprBatchArgs pa;                 // x / xnext / resid are set per phase
  float* xs[2];                    // messages (current = xs[cur])
  double* res_rows;                // per A-block partial residuals [block * K]
  double* res_comb;                // per B-block partial residuals
  double* resid;                   // [K] FP64 residual of the last iteration
  double step_bytes;               // PPR: algorithmic bytes per iteration
  unsigned long long step_edges;   // PPR: entries read per iteration
};

// Mutable per-task control word block, written by k_fused_c.
enum : uint32_t { FS_FREE = 0, FS_NEW = 1, FS_RUN = 2 };

struct FCtl {
  uint32_t iter, cur, done, bottom_up;
  uint32_t thr, nitems, a_beg, a_end;
  uint32_t b_beg, b_end, p_beg, p_end;   // block shares: A, B and pull launches
  uint32_t fcols, state, useful, pad;    // state: FS_*; useful: blocks the next A/pull can use
  double est;                  // estimated bytes of the next step (block split)
  unsigned long long fedges;   // entries behind the current work list
  unsigned long long edges;    // entries scanned over the run
  double bytes;                // algorithmic bytes over the run
  double pbytes;               // of which in bottom-up BFS steps (pull launch)
};

struct FHdr {
  uint32_t ntasks, ga, gb, active;  // active: tasks not done
  uint32_t steps, pad0;
  uint64_t sub_mask;                // tasks that also run the sub-steps (FC_ADV_SUB, FC_SPLIT_SUB)
  uint32_t pad[8];
};

// Control modes (flags): FC_INIT sets up new tasks only; FC_ADV_SUB
// advances only the tasks of hdr->sub_mask (a sub-step ran: the others kept
// their state); FC_SPLIT_SUB gives the next launches' blocks to those tasks
// only (a sub-step comes next; the others get none). With neither, every
// running task advances and shares the next full step. So small tasks take
// several iterations per step of the large ones.
enum : uint32_t { FC_ALL = 0, FC_INIT = 1, FC_ADV_SUB = 2, FC_SPLIT_SUB = 4 };

// The task of launch class CLS whose share of this launch holds blockIdx.x.
enum : uint32_t { PH_A = 0, PH_B = 1, PH_PULL = 2 };

template <uint32_t CLS>
__device__ __forceinline__ const FTask* find_task(const FTask* tasks, const FHdr* h, const FCtl* ctl,
                                                  uint32_t phase, Part& p, uint32_t& t) {
  const uint32_t nt = h->ntasks;
  for (t = 0; t < nt; t++) {
    if (tasks[t].cls != CLS) continue;
    const uint32_t b = phase == PH_A ? ctl[t].a_beg : phase == PH_B ? ctl[t].b_beg : ctl[t].p_beg;
    const uint32_t e = phase == PH_A ? ctl[t].a_end : phase == PH_B ? ctl[t].b_end : ctl[t].p_end;
    if (blockIdx.x >= b && blockIdx.x < e) {
      p.bid = blockIdx.x - b;
      p.nb = e - b;
      return &tasks[t];
    }
  }
  return nullptr;
}

// Top-down BFS step (bottom-up steps run in k_fused_pull: the W-bit pull
// bodies need more registers than any other body, and kept in their own
// launch they do not force spills into the SSSP / PPR loops of k_fused_a).
template <int W>
__device__ __forceinline__ void fused_bfs_a(const FTask& T, const FCtl& C, Part p, Emitter& em) {
  const uint32_t cur = C.cur;
  msbfs_push_body<W>(T.g, p, em, T.q[cur].items, C.nitems, T.vis, T.fr[cur], T.fr[cur ^ 1], T.stamp,
                     C.iter + 2, T.touched, T.tstamp, C.iter + 1);
  em_finish(em, T.touched, p.bid);
}

template <int W>
__device__ __forceinline__ void fused_bfs_b(const FTask& T, const FCtl& C, Part p, Emitter& em) {
  const uint32_t cur = C.cur;
  if (!C.bottom_up) {
    const DG g = T.g;
    const Frontier nf = T.q[cur ^ 1];
    msbfs_apply_body<W>(g, p, em, T.touched.items, *T.touched.nitems, T.vis, T.fr[cur ^ 1], T.lev, C.iter, nf);
    em_finish(em, nf, p.bid);
  }
  clear_items_body<W>(p, T.q[cur].items, C.nitems, T.fr[cur]);
}

template <int QW>
__device__ __forceinline__ void fused_sssp_a(const FTask& T, const FCtl& C, Part p, Emitter& em) {
  const uint32_t cur = C.cur;
  mssp_push_body<QW>(T.g, p, em, T.q[cur].items, C.nitems, T.dist, T.m[cur], T.m[cur ^ 1], C.thr, T.touched,
                     T.tstamp, C.iter + 1);
  em_finish(em, T.touched, p.bid);
}

template <int QW>
__device__ __forceinline__ void fused_sssp_b(const FTask& T, const FCtl& C, Part p, Emitter& em) {
  const uint32_t cur = C.cur;
  const DG g = T.g;
  const Frontier nf = T.q[cur ^ 1];
  mssp_items_body(g, p, em, T.touched.items, *T.touched.nitems, T.stamp, C.iter + 2, QW > 1, nf);
  em_finish(em, nf, p.bid);
  clear_items_body<QW>(p, T.q[cur].items, C.nitems, T.m[cur]);
}

// PPR steps reduce their residual partials in the emitter's LDS (unused by
// PPR): (BLOCK / WAVE) x 64 * QW doubles <= the emitter's 4 x QCAP words.
static_assert((BLOCK / WAVE) * 64 * FMAX_WIDTH <= (BLOCK / WAVE) * QCAP, "LDS reuse");

template <int QW>
__device__ __forceinline__ void fused_ppr_a(const FTask& T, const FCtl& C, Part p, double* red) {
  PprBatchArgs a = T.pa;
  a.x = T.xs[C.cur];
  a.xnext = T.xs[C.cur ^ 1];
  a.skip_rank = C.iter + 2 < T.iters;
  a.skip_resid = C.iter + 1 < T.iters;
  a.resid = T.res_rows;
  mppr_rows_body<QW>(T.g, p, a, 0, T.g.nreg, true, red);
  mppr_segs_body<QW>(T.g, p, a, T.g.npseg_reg);
}

template <int QW>
__device__ __forceinline__ void fused_ppr_b(const FTask& T, const FCtl& C, Part p, double* red) {
  PprBatchArgs a = T.pa;
  a.x = T.xs[C.cur];
  a.xnext = T.xs[C.cur ^ 1];
  a.skip_rank = C.iter + 2 < T.iters;
  a.skip_resid = C.iter + 1 < T.iters;
  a.resid = T.res_comb;
  const DG g = T.g;
  mppr_combine_body<QW>(g, p, a, 0, g.nreg, true, red);
}

// Narrow tasks: the narrow-row bodies of the batched calls, so a task of a
// few queries reads LPR x 4 B per entry with busy lanes instead of a 64-lane
// row. The SSSP push queues each improved row once and emits its work items
// into the next list itself (no touched list, no items pass; the smallest
// queued distance lands in that list's counters); B clears the consumed
// masks (one word per row).
template <int LPR>
__device__ __forceinline__ void fused_sssp_an(const FTask& T, const FCtl& C, Part p, Emitter& em) {
  const uint32_t cur = C.cur;
  mssp_sub_push_body<LPR>(T.g, p, em, T.q[cur].items, C.nitems, T.dist, T.m[cur], T.m[cur ^ 1], C.thr,
                          T.q[cur ^ 1]);
  em_finish(em, T.q[cur ^ 1], p.bid);
}

template <int LPR>
__device__ __forceinline__ void fused_ppr_an(const FTask& T, const FCtl& C, Part p, double* red) {
  PprBatchArgs a = T.pa;
  a.x = T.xs[C.cur];
  a.xnext = T.xs[C.cur ^ 1];
  a.skip_rank = C.iter + 2 < T.iters;
  a.skip_resid = C.iter + 1 < T.iters;
  a.resid = T.res_rows;
  mppr_sub_rows_body<LPR>(T.g, p, a, 0, T.g.nreg, true, red);
  // entries in flight per lane along a segment: the launch holds both
  // bodies, and 32 costs it three of its eight waves per SIMD, which pays
  // only with one lane per row (scale 22, 30 iterations: 1 query 16.8 ms with
  // 32 against 20.3 with 8; 8 queries 31.4 against 26.4)
  mppr_sub_segs_body<LPR, LPR == 1 ? 32 : 8>(T.g, p, a, T.g.npseg_reg);
}

template <int LPR>
__device__ __forceinline__ void fused_ppr_bn(const FTask& T, const FCtl& C, Part p, double* red) {
  PprBatchArgs a = T.pa;
  a.x = T.xs[C.cur];
  a.xnext = T.xs[C.cur ^ 1];
  a.skip_rank = C.iter + 2 < T.iters;
  a.skip_resid = C.iter + 1 < T.iters;
  a.resid = T.res_comb;
  mppr_sub_combine_body<LPR>(T.g, p, a, 0, T.g.nreg, true, red);
}

// One narrow class per kernel: program KIND with LPR lanes per row.
template <uint32_t KIND, int LPR>
__global__ void __launch_bounds__(BLOCK) k_fused_an(const FTask* __restrict__ tasks, const FHdr* hdr,
                                                 const FCtl* ctl) {
  Part p;
  uint32_t t;
  const FTask* T = find_task<narrow_cls(KIND, LPR)>(tasks, hdr, ctl, PH_A, p, t);
  if (!T) return;  // block-uniform
  const FCtl& C = ctl[t];
  if constexpr (KIND == FK_SSSP) {
    LA3_EMITTER(em);
    fused_sssp_an<LPR>(*T, C, p, em);
  } else {
    __shared__ double red[(BLOCK / WAVE) * LPR];
    fused_ppr_an<LPR>(*T, C, p, red);
  }
}

template <uint32_t KIND, int LPR>
__global__ void __launch_bounds__(BLOCK) k_fused_bn(const FTask* __restrict__ tasks, const FHdr* hdr,
                                                 const FCtl* ctl) {
  Part p;
  uint32_t t;
  const FTask* T = find_task<narrow_cls(KIND, LPR)>(tasks, hdr, ctl, PH_B, p, t);
  if (!T) return;  // block-uniform
  const FCtl& C = ctl[t];
  if constexpr (KIND == FK_SSSP) {
    clear_items_body<1>(p, T->q[C.cur].items, C.nitems, T->m[C.cur]);
  } else {
    __shared__ double red[(BLOCK / WAVE) * LPR];
    fused_ppr_bn<LPR>(*T, C, p, red);
  }
}

// One launch per width class W present in the run (1, 2 or 4 query words):
// each holds only the W-wide bodies, so its registers are those of its own
// widest body.
template <int W>
__global__ void __launch_bounds__(BLOCK, 8) k_fused_a(const FTask* __restrict__ tasks, const FHdr* hdr,
                                                   const FCtl* ctl) {
  Part p;
  uint32_t t;
  const FTask* T = find_task<wide_cls(W)>(tasks, hdr, ctl, PH_A, p, t);
  if (!T) return;  // block-uniform
  const FCtl& C = ctl[t];
  LA3_EMITTER(em);
  switch (T->kind) {
    case FK_BFS: fused_bfs_a<W>(*T, C, p, em); break;
    case FK_SSSP: fused_sssp_a<W>(*T, C, p, em); break;
    case FK_PPR: fused_ppr_a<W>(*T, C, p, (double*)em_lds); break;
    default: break;
  }
}

// Bottom-up BFS steps of width W.
template <int W>
__global__ void __launch_bounds__(BLOCK) k_fused_pull(const FTask* __restrict__ tasks, const FHdr* hdr,
                                                      const FCtl* ctl) {
  Part p;
  uint32_t t;
  const FTask* T = find_task<wide_cls(W)>(tasks, hdr, ctl, PH_PULL, p, t);
  if (!T) return;  // block-uniform
  const FCtl& C = ctl[t];
  const uint32_t cur = C.cur;
  const DG g = T->g;
  const Frontier nf = T->q[cur ^ 1];
  LA3_EMITTER(em);
  msbfs_pull_body<W>(g, p, em, T->vis, T->fr[cur], T->fr[cur ^ 1], T->lev, C.iter, T->valid, nf);
  msbfs_pull_long_body<W>(g, p, em, T->vis, T->fr[cur], T->fr[cur ^ 1], T->lev, C.iter, T->valid, nf);
  em_finish(em, nf, p.bid);
}

template <int W>
__global__ void __launch_bounds__(BLOCK) k_fused_b(const FTask* __restrict__ tasks, const FHdr* hdr,
                                                   const FCtl* ctl) {
  Part p;
  uint32_t t;
  const FTask* T = find_task<wide_cls(W)>(tasks, hdr, ctl, PH_B, p, t);
  if (!T) return;  // block-uniform
  const FCtl& C = ctl[t];
  LA3_EMITTER(em);
  switch (T->kind) {
    case FK_BFS: fused_bfs_b<W>(*T, C, p, em); break;
    case FK_SSSP: fused_sssp_b<W>(*T, C, p, em); break;
    case FK_PPR: fused_ppr_b<W>(*T, C, p, (double*)em_lds); break;
    default: break;
  }
}

// Sums of one queue's counter slots (one wavefront; NSLOT == WAVE).
struct FSums {
  uint32_t verts, cols, nmin, touched;
  unsigned long long edges, scanned;
};
__device__ __forceinline__ FSums fsum(const Frontier& f) {
  const CtrSlot& s = f.slots[lane_id()];
  FSums o;
  o.verts = wave_sum(s.verts);
  o.cols = wave_sum(s.cols);
  o.edges = wave_sum(s.edges);
  o.scanned = wave_sum(s.scanned);
  o.touched = wave_sum(s.touched);
  uint32_t mx = s.nmin_inv;
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, d, WAVE));
  o.nmin = ~mx;
  return o;
}

// Zeroes a queue's item count and counter slots (one wavefront).
__device__ __forceinline__ void fzero(const Frontier& f) {
  uint32_t* w = f.nitems;
  for (uint32_t i = lane_id(); i < CTR_WORDS; i += WAVE) w[i] = 0;
}

// r + p[0] + p[stride] + ... + p[(n - 1) * stride], added in that order;
// eight loads are issued before their adds (one thread sums thousands of
// block partials).
__device__ __forceinline__ double ordered_sum(const double* p, uint32_t n, uint32_t stride, double r) {
  uint32_t b = 0;
  for (; b + 8 <= n; b += 8) {
    double v[8];
#pragma unroll
    for (uint32_t u = 0; u < 8; u++) v[u] = p[(size_t)(b + u) * stride];
#pragma unroll
    for (uint32_t u = 0; u < 8; u++) r += v[u];
  }
  for (; b < n; b++) r += p[(size_t)b * stride];
  return r;
}

// Control step (one block of BLOCK threads). Tasks in state FS_NEW are set
// up from their seeded work lists and become FS_RUN; running tasks account
// for the step just run and advance (done tasks and FS_FREE slots are
// skipped; with FC_ADV_SUB only the tasks of hdr->sub_mask, with FC_INIT
// none). Then the blocks of the next launches are split among the running
// tasks (with FC_SPLIT_SUB among those of the mask only).
__global__ void __launch_bounds__(BLOCK) k_fused_c(const FTask* __restrict__ tasks, FHdr* hdr, FCtl* ctl,
                                                   uint32_t mode) {
  __shared__ uint32_t fin[FMAX_TASKS];   // PPR tasks whose last iteration just ran
  const uint32_t nt = hdr->ntasks;
  const uint32_t lane = lane_id();
  const uint64_t sub = (mode & FC_ADV_SUB) ? hdr->sub_mask : ~0ull;     // tasks that advance
  const uint64_t next = (mode & FC_SPLIT_SUB) ? hdr->sub_mask : ~0ull;  // tasks of the next launches
  const bool init_only = mode & FC_INIT;
  for (uint32_t t = threadIdx.x / WAVE; t < nt; t += BLOCK / WAVE) {
    const FTask& T = tasks[t];
    FCtl& C = ctl[t];
    const DG& g = T.g;
    const uint32_t K = T.lpr ? T.lpr : 64 * T.width;  // lanes per vertex row
    if (lane == 0) fin[t] = 0;
    const uint32_t state = C.state;
    const bool init = state == FS_NEW;
    if (!(sub >> t & 1)) continue;
    if (state == FS_FREE || (state == FS_RUN && (C.done || init_only))) continue;
    if (T.kind == FK_PPR) {
      if (lane == 0) {
        if (!init) {
          C.bytes += T.step_bytes;
          C.edges += T.step_edges;
          C.iter++;
          C.cur ^= 1;
        } else {
          C.iter = 0; C.cur = 0; C.bytes = 0; C.pbytes = 0; C.edges = 0; C.bottom_up = 0;
        }
        C.done = C.iter >= T.iters;
        C.state = FS_RUN;
        fin[t] = !init && C.done;
        C.est = C.done ? 0.0 : (4.0 + 4.0 * T.k) * (double)g.nnz + 20.0 * T.k * (double)g.nreg;
        C.useful = C.done ? 0 : (g.nreg + 3) / 4 + (g.npseg_reg + 3) / 4;
      }
      continue;
    }
    // Traversals: the work list just produced is q[cur ^ 1] (q[0] at init).
    const uint32_t nxt = init ? 0 : (C.cur ^ 1);
    const Frontier& nf = T.q[nxt];
    const FSums ns = fsum(nf);
    const uint32_t nitems = *nf.nitems;
    if (!init) {
      const FSums ts = fsum(T.touched);  // push counters: smallest queued distance
      if (lane == 0) {
        // Algorithmic bytes of the step (the compulsory traffic of one
        // vertex-program iteration with perfect reuse of the vectors):
        //   (4 + w) E_scan + (8 + m) C_act + (a + s_r + s_w) R_touch + m R_act,
        // per-vertex terms per live query: BFS bit sets (m = a = s_r = s_w =
        // k/8 B), SSSP u32 distances (w = 4, m = a = s_r = s_w = 4k B).
        // R_touch (rows that received a message) is counted by the push in
        // measurement runs (tstamp); a bottom-up step touches exactly the
        // rows it activates.
        const double kb = T.k / 8.0;
        double b = 0;
        unsigned long long e = 0;
        if (T.kind == FK_BFS) {
          if (!C.bottom_up) {
            b = 4.0 * C.fedges + (8.0 + kb) * C.fcols + 3 * kb * ts.touched + kb * ns.verts;
            e = C.fedges;
          } else {
            b = 4.0 * ns.scanned + (8.0 + kb) * C.fcols + 4 * kb * ns.verts;
            e = ns.scanned;
            C.pbytes += b;
          }
        } else {
          // narrow tasks queue the improved rows directly: R_touch taken as
          // the rows queued (a lower bound)
          const double k4 = 4.0 * T.k;
          const uint32_t rt = T.lpr ? ns.verts : ts.touched;
          b = 8.0 * C.fedges + (8.0 + k4) * C.fcols + 3 * k4 * rt + k4 * ns.verts;
          e = C.fedges;
        }
        C.bytes += b;
        C.edges += e;
        C.iter++;
        if (T.kind == FK_BFS) {
          if (!C.bottom_up && ns.edges * 20 > g.nnz) C.bottom_up = 1;
          else if (C.bottom_up && (uint64_t)ns.verts * 24 < g.nslots) C.bottom_up = 0;
        } else if (T.delta) {
          // Threshold unchanged while the new work list holds pairs below it
          // (the push's smallest queued distance: in the touched list, or for
          // a narrow task in the next work list).
          const uint32_t nmin = T.lpr ? ns.nmin : ts.nmin;
          if (nmin >= C.thr) C.thr = (uint32_t)min(0xffffffffull, (unsigned long long)nmin + T.delta);
        }
        C.cur = nxt;
      }
    } else if (lane == 0) {
      C.iter = 0; C.cur = 0; C.bottom_up = 0; C.bytes = 0; C.pbytes = 0; C.edges = 0;
      C.thr = T.delta ? T.delta : 0xffffffffu;
    }
    // The list just consumed (and the touched list) are written next step.
    __builtin_amdgcn_wave_barrier();
    fzero(T.q[nxt ^ 1]);
    fzero(T.touched);
    if (lane == 0) {
      C.nitems = nitems;
      C.fedges = ns.edges;
      C.fcols = ns.cols;
      C.done = nitems == 0;
      C.state = FS_RUN;
      const double K8 = T.kind == FK_BFS ? 4.0 + T.k / 8.0 : 8.0 + 4.0 * K;
      const bool pull = T.kind == FK_BFS && C.bottom_up;
      C.est = C.done ? 0.0 : (pull ? K8 * (double)g.nnz / 4 : K8 * (double)ns.edges + 64.0 * nitems);
      C.useful = C.done ? 0 : (pull ? (g.nrows + BLOCK - 1) / BLOCK + (g.nlong + 3) / 4 : (nitems + 3) / 4);
    }
  }
  __syncthreads();
  // FP64 residual of a PPR task's last iteration, per query, in a fixed
  // order: the A-block partials, then the B-block partials of that step.
  for (uint32_t t = 0; t < nt; t++) {
    if (!fin[t]) continue;  // block-uniform
    const FTask& T = tasks[t];
    const FCtl& C = ctl[t];
    const uint32_t K = T.lpr ? T.lpr : 64 * T.width;
    const uint32_t na = C.a_end - C.a_beg, nb = C.b_end - C.b_beg;
    for (uint32_t q = threadIdx.x; q < K; q += BLOCK) {
      double r = ordered_sum(T.res_rows + q, na, K, 0.0);
      if (T.g.npseg_reg) r = ordered_sum(T.res_comb + q, nb, K, r);
      T.resid[q] = r;
    }
  }
  __syncthreads();  // the partitions read above are rewritten below
  if (threadIdx.x == 0) {
    // Blocks of each launch class's launches, split among its running tasks:
    // A (and, for bottom-up BFS steps, the pull launch) in proportion to the
    // estimated bytes (at least 1, at most the useful count), B evenly.
    double tot[NCLS] = {}, totp[NCLS] = {};
    uint32_t act[NCLS] = {}, a[NCLS] = {}, b[NCLS] = {}, pp[NCLS] = {};
    uint32_t active = 0;
    for (uint32_t t = 0; t < nt; t++) {
      const FCtl& C = ctl[t];
      if (C.state != FS_RUN || C.done || !(next >> t & 1)) continue;
      const uint32_t w = tasks[t].cls;
      const bool pull = tasks[t].kind == FK_BFS && C.bottom_up;
      (pull ? totp : tot)[w] += C.est;
      act[w]++;
      active++;
    }
    const uint32_t ga = hdr->ga, gb = hdr->gb;
    for (uint32_t t = 0; t < nt; t++) {
      FCtl& C = ctl[t];
      const bool run = C.state == FS_RUN && !C.done && (next >> t & 1);
      const uint32_t w = run ? tasks[t].cls : 0;
      const bool pull = run && tasks[t].kind == FK_BFS && C.bottom_up;
      uint32_t na = 0, nb = 0, np = 0;
      if (run) {
        const double tw = pull ? totp[w] : tot[w];
        uint32_t n = tw > 0 ? (uint32_t)(ga * (C.est / tw)) : 1;
        n = max(1u, min(n, max(1u, C.useful)));
        (pull ? np : na) = n;
        nb = max(1u, gb / max(act[w], 1u));
      }
      C.a_beg = a[w]; C.a_end = a[w] + na; a[w] += na;
      C.p_beg = pp[w]; C.p_end = pp[w] + np; pp[w] += np;
      C.b_beg = b[w]; C.b_end = b[w] + nb; b[w] += nb;
    }
    if (next == ~0ull) hdr->active = active;
    if (!init_only && sub == ~0ull && active) hdr->steps++;
  }
}

}  // namespace la3
