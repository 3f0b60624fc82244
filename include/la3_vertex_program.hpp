// la3_vertex_program.hpp — the reference's vertex-program interface
// VertexProgram<W, M, A, S> (src/vprogram/vertex_program.h:21-220) as a
// header-only HIP template for gfx950. A program is one struct of __device__
// operations, the counterparts of the reference's overridable virtuals
// (vertex_program.h:56-137):
//
//   struct MyProgram {
//     using W = uint32_t;  // edge weight (la3::Empty for pairs)
//     using M = ...;       // message
//     using A = ...;       // accumulator; A() is the combine identity
//     using S = ...;       // vertex state; S() is the initial state
//     __device__ bool init(uint32_t vid, S& s) const;                    // default: return stationary
//     __device__ bool init(uint32_t vid, const OS& other, S& s) const;   // for initialize(other)
//     __device__ M scatter(const S& s) const;                            // default: M()
//     __device__ A gather(const la3::Edge<W>& e, const M& m) const;      // default: A()
//     __device__ A gather(const la3::Edge<W>& e, const M& m, const S& s) const;  // reads the row's state
//     __device__ void combine(const A& y1, A& y2) const;                 // default: nothing
//     __device__ bool apply(const A& y, S& s) const;                     // default: false
//     __device__ bool apply(const A& y, S& s, uint32_t iter) const;      // reads the iteration
//   };
//   la3::VertexProgram<MyProgram> vp(&G, stationary);   // G: la3::Graph<W>
//   vp.initialize(); vp.execute(max_iters); vp.states();
//
// Any operation may be left out; it then behaves as the reference's default.
// The three-argument gather and apply are detected at compile time, as the
// reference detects their overrides at run time (initialize_flags,
// vertex_program.hpp:66-81): either disables the optimized execution, and a
// state-reading gather sees the row vertex's state as initialize() left it
// (the reference mirrors states once, vertex_program.hpp:159-173).
// M, A and S must be trivially copyable (they live in device memory).
//
// Execution follows the reference exactly (vertex_program_execute.hpp:11-345):
// execute(1) is execute_single; a directed graph with neither three-argument
// operation and `optimizable` set runs execute_ (regular rows in the loop,
// then every regular vertex re-scatters its final state and the sink rows
// take one step, :55-163); otherwise execute_non_opt (regular and sink rows in
// every iteration, :166-248). max_iters = 0 runs until an iteration activates
// no regular vertex; max_iters = m > 1 runs exactly m iterations. Source
// messages reach regular rows only when stationary or in iteration 0, sink
// rows always (:464-483). A row applies only when it received a message, with
// y = combine(partial, A()) (:678-691), and a regular row scatters when
// apply returned true or the program is stationary (:711-716).
//
// Each row folds its messages in the reference's order: source columns
// first, then regular columns, each in increasing hashed id — the graph's
// row-major companion holds a row's senders in exactly that order. So
// results equal the reference's even for order-dependent combines
// (last-writer-wins, floating-point sums). Rows of at most 64 entries are
// folded by one lane; longer rows are gathered in parallel and folded in
// entry order by a wavefront, or, for a program that declares its combine
// associative or gives a segment type (seg_add / seg_join / seg_value),
// folded as 256-entry runs by many wavefronts whose partials are joined in
// order.
//
// Device memory: the states, two message buffers and two activity-flag
// buffers (by vertex slot), plus a mirror of the states when gather reads
// them. The program runs on its context's stream; states() synchronizes.
#ifndef LA3_VERTEX_PROGRAM_HPP
#define LA3_VERTEX_PROGRAM_HPP

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "la3.hpp"

namespace la3 {

// Edge<W> (src/vprogram/types.h:29-40): src = the sender's hashed id (the
// column), dst = the receiving row's hashed id, weight = the entry's weight.
template <class W>
struct Edge {
  uint32_t src, dst;
  W weight;
};
template <>
struct Edge<Empty> {
  uint32_t src, dst;
  Empty weight;
};

namespace vp_detail {

constexpr uint32_t VP_BLOCK = 256;
constexpr uint32_t VP_WAVE = 64;
constexpr uint32_t VP_LONG = 64;  // rows with more entries are folded by a wavefront

// ---- detection of the overridable operations ----
template <class P, class = void> struct has_init : std::false_type {};
template <class P>
struct has_init<P, std::void_t<decltype(std::declval<const P&>().init(uint32_t(), std::declval<typename P::S&>()))>>
    : std::true_type {};
template <class P, class OS, class = void> struct has_init_other : std::false_type {};
template <class P, class OS>
struct has_init_other<P, OS, std::void_t<decltype(std::declval<const P&>().init(
                                 uint32_t(), std::declval<const OS&>(), std::declval<typename P::S&>()))>>
    : std::true_type {};
template <class P, class = void> struct has_scatter : std::false_type {};
template <class P>
struct has_scatter<P, std::void_t<decltype(std::declval<const P&>().scatter(std::declval<const typename P::S&>()))>>
    : std::true_type {};
template <class P, class = void> struct has_gather2 : std::false_type {};
template <class P>
struct has_gather2<P, std::void_t<decltype(std::declval<const P&>().gather(
                          std::declval<const Edge<typename P::W>&>(), std::declval<const typename P::M&>()))>>
    : std::true_type {};
template <class P, class = void> struct has_gather3 : std::false_type {};
template <class P>
struct has_gather3<P, std::void_t<decltype(std::declval<const P&>().gather(
                          std::declval<const Edge<typename P::W>&>(), std::declval<const typename P::M&>(),
                          std::declval<const typename P::S&>()))>> : std::true_type {};
template <class P, class = void> struct has_combine : std::false_type {};
template <class P>
struct has_combine<P, std::void_t<decltype(std::declval<const P&>().combine(std::declval<const typename P::A&>(),
                                                                             std::declval<typename P::A&>()))>>
    : std::true_type {};
template <class P, class = void> struct has_apply2 : std::false_type {};
template <class P>
struct has_apply2<P, std::void_t<decltype(std::declval<const P&>().apply(std::declval<const typename P::A&>(),
                                                                          std::declval<typename P::S&>()))>>
    : std::true_type {};
template <class P, class = void> struct has_apply3 : std::false_type {};
template <class P>
struct has_apply3<P, std::void_t<decltype(std::declval<const P&>().apply(
                         std::declval<const typename P::A&>(), std::declval<typename P::S&>(), uint32_t()))>>
    : std::true_type {};

// Optional segmented fold, for long rows: a program whose combine is
// associative (combine(a, combine(b, y)) == combine(combine(a, b), y)) may
// declare `static constexpr bool associative = true`; any program may instead
// give a segment type that reproduces its in-order fold:
//   using Seg = ...;                                        // Seg() = no message
//   __device__ void seg_add(const A& y, Seg& s) const;      // s := s then y
//   __device__ void seg_join(const Seg& later, Seg& s) const; // s := s then later
//   __device__ A seg_value(const Seg& s) const;             // the fold from A()
// Rows of more than VP_LONG entries are then folded as 256-entry runs, by
// many wavefronts at once, whose partial folds are joined in entry order.
template <class P, class = void> struct has_seg : std::false_type {};
template <class P>
struct has_seg<P, std::void_t<typename P::Seg,
                              decltype(std::declval<const P&>().seg_add(std::declval<const typename P::A&>(),
                                                                        std::declval<typename P::Seg&>())),
                              decltype(std::declval<const P&>().seg_join(std::declval<const typename P::Seg&>(),
                                                                         std::declval<typename P::Seg&>())),
                              decltype(std::declval<const P&>().seg_value(std::declval<const typename P::Seg&>()))>>
    : std::true_type {};
template <class P, class = void> struct is_assoc : std::false_type {};
template <class P> struct is_assoc<P, std::enable_if_t<P::associative>> : std::true_type {};

// ---- the operations with the reference's defaults ----
template <class P>
__device__ __forceinline__ bool op_init(const P& p, bool stationary, uint32_t vid, typename P::S& s) {
  if constexpr (has_init<P>::value) return p.init(vid, s);
  else return stationary;
}
template <class P, class OS>
__device__ __forceinline__ bool op_init_other(const P& p, bool stationary, uint32_t vid, const OS& o,
                                              typename P::S& s) {
  if constexpr (has_init_other<P, OS>::value) return p.init(vid, o, s);
  else return stationary;
}
template <class P>
__device__ __forceinline__ typename P::M op_scatter(const P& p, const typename P::S& s) {
  if constexpr (has_scatter<P>::value) return p.scatter(s);
  else return typename P::M();
}
template <class P>
__device__ __forceinline__ typename P::A op_gather(const P& p, const Edge<typename P::W>& e,
                                                   const typename P::M& m, const typename P::S* mirror) {
  if constexpr (has_gather3<P>::value) return p.gather(e, m, *mirror);
  else if constexpr (has_gather2<P>::value) return p.gather(e, m);
  else return typename P::A();
}
template <class P>
__device__ __forceinline__ void op_combine(const P& p, const typename P::A& y1, typename P::A& y2) {
  if constexpr (has_combine<P>::value) p.combine(y1, y2);
}
template <class P>
__device__ __forceinline__ bool op_apply(const P& p, const typename P::A& y, typename P::S& s, uint32_t iter) {
  if constexpr (has_apply3<P>::value) return p.apply(y, s, iter);
  else if constexpr (has_apply2<P>::value) return p.apply(y, s);
  else return false;
}

// The segmented fold a program allows (on = false: none).
template <class P, bool = has_seg<P>::value, bool = is_assoc<P>::value>
struct SegFold {
  static constexpr bool on = false;
  using Seg = typename P::A;
};
template <class P, bool X>
struct SegFold<P, true, X> {
  static constexpr bool on = true;
  using Seg = typename P::Seg;
  __device__ static void add(const P& p, const typename P::A& y, Seg& s) { p.seg_add(y, s); }
  __device__ static void join(const P& p, const Seg& later, Seg& s) { p.seg_join(later, s); }
  __device__ static typename P::A value(const P& p, const Seg& s) { return p.seg_value(s); }
};
template <class P>
struct SegFold<P, false, true> {  // associative combine: a run's fold is its accumulator
  static constexpr bool on = true;
  using Seg = typename P::A;
  __device__ static void add(const P& p, const typename P::A& y, Seg& s) { op_combine(p, y, s); }
  __device__ static void join(const P& p, const Seg& later, Seg& s) { op_combine(p, later, s); }
  __device__ static typename P::A value(const P&, const Seg& s) { return s; }
};

template <class W>
__device__ __forceinline__ Edge<W> make_edge(const la3_device_view& g, uint32_t sender, uint32_t row, uint32_t e) {
  Edge<W> ed;
  ed.src = g.hid_of[sender];
  ed.dst = g.hid_of[row];
  if constexpr (std::is_same<W, Empty>::value) ed.weight = Empty();
  else ed.weight = (W)g.row_w[e];
  return ed;
}

// Device state of one program.
template <class P>
struct Bufs {
  typename P::S* v = nullptr;   // states by slot
  typename P::S* vm = nullptr;  // mirror (state-reading gather only)
  typename P::M* x[2] = {nullptr, nullptr};  // messages by slot (regular and source slots)
  uint8_t* act[2] = {nullptr, nullptr};      // message flags by slot
  // [0] some regular vertex was activated; [4 + k] active regular senders
  // counted for a pass (two slots, used alternately); [6] active sources
  uint32_t* ctr = nullptr;
  uint32_t* stamp = nullptr;    // by row: the last sparse pass that reached it
  // Vertex activity by slot (the state vector's activity set,
  // random_access_array.h:14): set by apply for a regular vertex that was
  // activated or is stationary and a sink that was activated
  // (vertex_program_execute.hpp:693-713), kept across executes, cleared by
  // reset_activity() and after the first mirroring; read by active_only
  // reduce / topk and by the mirroring of a non-stationary program.
  uint8_t* vact = nullptr;
};

// One pass over rows [r0, r1): read the messages of buffer `cur` (source
// slots only when use_src), apply, and for regular rows write the next
// messages into buffer `cur ^ 1` (unless single).
struct PassArgs {
  uint32_t r0, r1, cur, iter;
  uint32_t use_src, single, stationary;
  uint32_t tag, slot, dense_only;  // sparse-pass stamp, count slot, no sparse pass
};

// Dense or sparse pass, by the reference's density rule
// (serializable_bitvector.h:54-55): dense when more than
// (columns / 32) * 2 / 3 columns send. A sparse pass first stamps the rows
// that the active senders' columns reach (k_vp_mark); the row kernels then
// skip the other rows, which would receive nothing. Either way a processed
// row folds all its messages in the same order, so results are the same.
__device__ __forceinline__ bool vp_sparse(const la3_device_view& g, const uint32_t* ctr, const PassArgs& a) {
  if (a.dense_only) return false;
  const uint32_t n = ctr[4 + a.slot] + (a.use_src ? ctr[6] : 0);
  return n <= (g.ncols / 32) * 2 / 3;
}

// Adds every thread's n into *dst, one atomic per block (all threads call).
__device__ __forceinline__ void block_add(uint32_t n, uint32_t* dst) {
  __shared__ uint32_t part;
  if (threadIdx.x == 0) part = 0;
  __syncthreads();
  if (n) atomicAdd(&part, n);
  __syncthreads();
  if (threadIdx.x == 0 && part) atomicAdd(dst, part);
}

template <class P>
__device__ __forceinline__ void vp_finish_row(const P& p, const la3_device_view& g, const Bufs<P>& b,
                                              const PassArgs& a, uint32_t r, const typename P::A& y) {
  using A = typename P::A;
  using S = typename P::S;
  A fin = A();
  op_combine(p, y, fin);
  S st = b.v[r];
  const bool act = op_apply(p, fin, st, a.iter);
  b.v[r] = st;
  if (r < g.nreg) {
    if (act) b.ctr[0] = 1;
    if (act || a.stationary) {
      b.vact[r] = 1;
      if (!a.single) {
        b.x[a.cur ^ 1][r] = op_scatter(p, st);
        b.act[a.cur ^ 1][r] = 1;
      }
    }
  } else if (act) {
    b.vact[r] = 1;
  }
}

// Short rows (at most VP_LONG entries): lane per row. Entries are read four
// at a time (columns, then flags, then messages: independent loads in
// flight together) and combined in entry order.
template <class P>
__global__ void __launch_bounds__(VP_BLOCK) k_vp_pull(la3_device_view g, P p, Bufs<P> b, PassArgs a) {
  using A = typename P::A;
  using M = typename P::M;
  const M* x = b.x[a.cur];
  const uint8_t* act = b.act[a.cur];
  const bool sparse = vp_sparse(g, b.ctr, a);
  for (uint32_t r = a.r0 + blockIdx.x * blockDim.x + threadIdx.x; r < a.r1; r += gridDim.x * blockDim.x) {
    if (sparse && b.stamp[r] != a.tag) continue;
    const uint32_t e0 = g.row_ptr[r], e1 = g.row_ptr[r + 1];
    if (e1 - e0 > VP_LONG) continue;
    A y = A();
    bool touched = false;
    const typename P::S* mir = b.vm ? b.vm + r : nullptr;
    uint32_t e = a.use_src ? e0 : g.row_split[r];
    for (; e + 4 <= e1; e += 4) {
      uint32_t u[4];
      bool on[4];
#pragma unroll
      for (int j = 0; j < 4; j++) u[j] = g.row_cols[e + j];
#pragma unroll
      for (int j = 0; j < 4; j++) on[j] = act[u[j]] != 0;
      M m[4];
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (on[j]) m[j] = x[u[j]];
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (on[j]) {
          op_combine(p, op_gather(p, make_edge<typename P::W>(g, u[j], r, e + j), m[j], mir), y);
          touched = true;
        }
    }
    for (; e < e1; e++) {
      const uint32_t u = g.row_cols[e];
      if (!act[u]) continue;
      op_combine(p, op_gather(p, make_edge<typename P::W>(g, u, r, e), x[u], mir), y);
      touched = true;
    }
    if (touched) vp_finish_row(p, g, b, a, r, y);
  }
}

// Long rows (more than VP_LONG entries) in two phases, exact for any
// combine: (A) every entry's gather, in parallel (a wavefront per
// VP_SEG-entry segment), into a staging area: the value, and per 64-entry
// chunk a mask of the entries whose sender was active; (B) a wavefront per
// row folds the chunks' active values in entry order. Phase B's fold is the
// only serial part; its loads are issued VP_DEPTH chunks ahead.
constexpr uint32_t VP_SEG = 256;  // entries per phase-A segment: four chunks
struct LongRow {
  uint32_t r, e0, e1, cbase;  // row slot, entries [e0, e1), first staging chunk
};
template <class A>
struct LongTabs {
  const LongRow* rows = nullptr;   // by long-row index
  const uint32_t* sbase = nullptr; // first segment of each long row
  const uint32_t* seg = nullptr;   // long-row index of each segment
  const uint32_t* nseg = nullptr;  // device: number of segments
  void* part = nullptr;            // segmented fold: one partial (SegFold::Seg) per segment
  uint8_t* pany = nullptr;         // segmented fold: the segment held a message
  A* val = nullptr;                // 64 values per chunk
  uint64_t* mask = nullptr;        // one mask per chunk
};

// The long-row tables (once per program): chunk and segment offsets of each
// long row, exclusive prefix sums in long-row order, by a three-kernel scan
// over blocks of VP_TAB rows; totals[0] = chunks, totals[1] = segments.
constexpr uint32_t VP_TAB = 1024;
__device__ __forceinline__ void long_counts(const la3_device_view& g, uint32_t i, uint32_t& c, uint32_t& s) {
  const uint32_t r = g.long_rows[i], len = g.row_ptr[r + 1] - g.row_ptr[r];
  c = (len + VP_WAVE - 1) / VP_WAVE;
  s = (len + VP_SEG - 1) / VP_SEG;
}

// In-block exclusive scan of two counters (VP_TAB threads).
__device__ __forceinline__ void block_scan2(uint32_t& c, uint32_t& s, uint32_t& tc, uint32_t& ts) {
  __shared__ uint32_t sc[VP_TAB], ss[VP_TAB];
  const uint32_t t = threadIdx.x;
  sc[t] = c;
  ss[t] = s;
  __syncthreads();
  for (uint32_t off = 1; off < VP_TAB; off <<= 1) {  // inclusive (Hillis-Steele)
    const uint32_t ac = t >= off ? sc[t - off] : 0, as = t >= off ? ss[t - off] : 0;
    __syncthreads();
    sc[t] += ac;
    ss[t] += as;
    __syncthreads();
  }
  tc = sc[VP_TAB - 1];
  ts = ss[VP_TAB - 1];
  c = sc[t] - c;
  s = ss[t] - s;
}

static __global__ void __launch_bounds__(VP_TAB) k_vp_long_sums(la3_device_view g, uint32_t* bsum) {
  const uint32_t i = blockIdx.x * VP_TAB + threadIdx.x;
  uint32_t c = 0, s = 0, tc, ts;
  if (i < g.nlong) long_counts(g, i, c, s);
  block_scan2(c, s, tc, ts);
  if (threadIdx.x == 0) {
    bsum[2 * blockIdx.x] = tc;
    bsum[2 * blockIdx.x + 1] = ts;
  }
}

// One block: exclusive scan of the nb block sums (in place), totals.
static __global__ void __launch_bounds__(VP_TAB) k_vp_long_scan(uint32_t* bsum, uint32_t nb, uint32_t* totals) {
  uint32_t carry_c = 0, carry_s = 0;
  for (uint32_t base = 0; base < nb; base += VP_TAB) {  // block-uniform
    const uint32_t i = base + threadIdx.x;
    uint32_t c = i < nb ? bsum[2 * i] : 0, s = i < nb ? bsum[2 * i + 1] : 0, tc, ts;
    block_scan2(c, s, tc, ts);
    if (i < nb) {
      bsum[2 * i] = carry_c + c;
      bsum[2 * i + 1] = carry_s + s;
    }
    carry_c += tc;
    carry_s += ts;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    totals[0] = carry_c;
    totals[1] = carry_s;
  }
}

static __global__ void __launch_bounds__(VP_TAB) k_vp_long_table(la3_device_view g, const uint32_t* bsum, LongRow* rows,
                                                         uint32_t* sbase) {
  const uint32_t i = blockIdx.x * VP_TAB + threadIdx.x;
  uint32_t c = 0, s = 0, tc, ts;
  if (i < g.nlong) long_counts(g, i, c, s);
  block_scan2(c, s, tc, ts);
  if (i < g.nlong) {
    const uint32_t r = g.long_rows[i];
    rows[i] = LongRow{r, g.row_ptr[r], g.row_ptr[r + 1], bsum[2 * blockIdx.x] + c};
    sbase[i] = bsum[2 * blockIdx.x + 1] + s;
  }
}

static __global__ void k_vp_long_segs(la3_device_view g, const LongRow* rows, const uint32_t* sbase, uint32_t* seg) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < g.nlong; i += gridDim.x * blockDim.x) {
    const uint32_t ns = (rows[i].e1 - rows[i].e0 + VP_SEG - 1) / VP_SEG;
    for (uint32_t k = 0; k < ns; k++) seg[sbase[i] + k] = i;
  }
}

__device__ __forceinline__ uint32_t vp_lane() { return threadIdx.x % VP_WAVE; }

// A long-row record read by a whole wavefront, held in scalar registers.
__device__ __forceinline__ LongRow uniform_row(const LongRow* rows, uint32_t li) {
  const LongRow R = rows[li];
  return LongRow{(uint32_t)__builtin_amdgcn_readfirstlane(R.r), (uint32_t)__builtin_amdgcn_readfirstlane(R.e0),
                 (uint32_t)__builtin_amdgcn_readfirstlane(R.e1), (uint32_t)__builtin_amdgcn_readfirstlane(R.cbase)};
}

// Lane l's copy of v from lane l + off (off < 64; lanes past 63 read their own).
template <class T>
__device__ __forceinline__ T wave_down(const T& v, uint32_t off) {
  constexpr int NW = (sizeof(T) + 3) / 4;
  int w[NW];
  w[NW - 1] = 0;
  memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < NW; i++) w[i] = __shfl_down(w[i], off, VP_WAVE);
  T out;
  memcpy(&out, w, sizeof(T));
  return out;
}

// Ordered reduction of one partial fold per lane (lane order = entry order;
// `ne` = the lane's run held a message): lane 0 ends with the fold of all
// 64 runs. Empty runs are skipped, not joined (a join with Seg() is not the
// identity for every combine, e.g. last-writer).
template <class P>
__device__ __forceinline__ void wave_fold(const P& p, typename SegFold<P>::Seg& s, bool& ne) {
  using SF = SegFold<P>;
  const uint32_t lane = threadIdx.x % VP_WAVE;
#pragma unroll
  for (uint32_t off = 1; off < VP_WAVE; off <<= 1) {
    const typename SF::Seg o = wave_down(s, off);
    const bool one = __shfl_down((int)ne, off, VP_WAVE) != 0;
    if ((lane & (2 * off - 1)) == 0 && lane + off < VP_WAVE && one) {
      if (ne) SF::join(p, o, s);
      else s = o;
      ne = true;
    }
  }
}

template <class P>
__global__ void __launch_bounds__(VP_BLOCK) k_vp_gather_long(la3_device_view g, P p, Bufs<P> b, PassArgs a,
                                                            LongTabs<typename P::A> t) {
  using M = typename P::M;
  const M* x = b.x[a.cur];
  const uint8_t* act = b.act[a.cur];
  const uint32_t lane = vp_lane();
  const uint32_t nseg = *t.nseg;
  const uint32_t nw = gridDim.x * (VP_BLOCK / VP_WAVE);
  const bool sparse = vp_sparse(g, b.ctr, a);
  for (uint32_t s = blockIdx.x * (VP_BLOCK / VP_WAVE) + threadIdx.x / VP_WAVE; s < nseg; s += nw) {
    const uint32_t li = (uint32_t)__builtin_amdgcn_readfirstlane(t.seg[s]);
    const LongRow R = uniform_row(t.rows, li);
    if (R.r < a.r0 || R.r >= a.r1) continue;  // wave-uniform
    if (sparse && (uint32_t)__builtin_amdgcn_readfirstlane(b.stamp[R.r]) != a.tag) continue;
    const uint32_t k = s - t.sbase[li];
    const uint32_t lo = a.use_src ? R.e0 : g.row_split[R.r];
    const uint32_t eb = R.e0 + k * VP_SEG;
    const typename P::S* mir = b.vm ? b.vm + R.r : nullptr;
    uint32_t u[4];
    bool on[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t e = eb + j * VP_WAVE + lane;
      u[j] = e < R.e1 ? g.row_cols[e] : 0u;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t e = eb + j * VP_WAVE + lane;
      on[j] = e < R.e1 && e >= lo && act[u[j]] != 0;
    }
    M m[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (on[j]) m[j] = x[u[j]];
    if constexpr (SegFold<P>::on) {
      // the segment's partial fold: per 64-entry chunk an ordered wave
      // reduction, the chunks joined in order by lane 0
      using SF = SegFold<P>;
      typename SF::Seg acc = typename SF::Seg();
      bool any = false;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint32_t e = eb + j * VP_WAVE + lane;
        typename SF::Seg sj = typename SF::Seg();
        bool ne = on[j];
        if (on[j]) SF::add(p, op_gather(p, make_edge<typename P::W>(g, u[j], R.r, e), m[j], mir), sj);
        wave_fold(p, sj, ne);
        if (lane == 0 && ne) {
          if (any) SF::join(p, sj, acc);
          else acc = sj;
          any = true;
        }
      }
      if (lane == 0) {
        ((typename SF::Seg*)t.part)[s] = acc;
        t.pany[s] = any;
      }
    } else {
      const uint64_t c0 = (uint64_t)R.cbase + k * (VP_SEG / VP_WAVE);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint32_t e = eb + j * VP_WAVE + lane;
        if (on[j]) t.val[(c0 + j) * VP_WAVE + lane] = op_gather(p, make_edge<typename P::W>(g, u[j], R.r, e), m[j], mir);
        const uint64_t bm = __ballot(on[j]);
        if (lane == 0 && eb + j * VP_WAVE < R.e1) t.mask[c0 + j] = bm;
      }
    }
  }
}

// Long rows of a program with a segmented fold: a wavefront per row joins
// its segments' partials in order (64 at a time, by an ordered wave
// reduction), then applies.
template <class P>
__global__ void __launch_bounds__(VP_BLOCK) k_vp_join_long(la3_device_view g, P p, Bufs<P> b, PassArgs a,
                                                          LongTabs<typename P::A> t) {
  using SF = SegFold<P>;
  const uint32_t lane = threadIdx.x % VP_WAVE;
  const uint32_t nw = gridDim.x * (VP_BLOCK / VP_WAVE);
  const bool sparse = vp_sparse(g, b.ctr, a);
  const typename SF::Seg* part = (const typename SF::Seg*)t.part;
  for (uint32_t li = blockIdx.x * (VP_BLOCK / VP_WAVE) + threadIdx.x / VP_WAVE; li < g.nlong; li += nw) {
    const LongRow R = t.rows[li];
    const uint32_t r = (uint32_t)__builtin_amdgcn_readfirstlane(R.r);
    if (r < a.r0 || r >= a.r1) continue;  // wave-uniform
    if (sparse && (uint32_t)__builtin_amdgcn_readfirstlane(b.stamp[r]) != a.tag) continue;
    const uint32_t s0 = (uint32_t)__builtin_amdgcn_readfirstlane(t.sbase[li]);
    const uint32_t ns = ((uint32_t)__builtin_amdgcn_readfirstlane(R.e1) - (uint32_t)__builtin_amdgcn_readfirstlane(R.e0) +
                         VP_SEG - 1) / VP_SEG;
    typename SF::Seg acc = typename SF::Seg();
    bool any = false;
    for (uint32_t c = 0; c < ns; c += VP_WAVE) {  // wave-uniform
      typename SF::Seg sc = typename SF::Seg();
      bool ne = false;
      if (c + lane < ns) {
        ne = t.pany[s0 + c + lane] != 0;
        if (ne) sc = part[s0 + c + lane];
      }
      wave_fold(p, sc, ne);
      if (lane == 0 && ne) {
        if (any) SF::join(p, sc, acc);
        else acc = sc;
        any = true;
      }
    }
    if (lane == 0 && any) vp_finish_row(p, g, b, a, r, SF::value(p, acc));
  }
}

// One lane's copy of v from lane `src` (src uniform across the wavefront).
template <class T>
__device__ __forceinline__ T wave_read(const T& v, uint32_t src) {
  constexpr int NW = (sizeof(T) + 3) / 4;
  uint32_t w[NW];
  w[NW - 1] = 0;
  memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < NW; i++) w[i] = __builtin_amdgcn_readlane(w[i], src);
  T out;
  memcpy(&out, w, sizeof(T));
  return out;
}

template <class P>
__global__ void __launch_bounds__(VP_BLOCK) k_vp_fold_long(la3_device_view g, P p, Bufs<P> b, PassArgs a,
                                                          LongTabs<typename P::A> t) {
  using A = typename P::A;
  constexpr uint32_t D = sizeof(A) <= 8 ? 8 : 2;  // chunks loaded ahead of the fold
  const uint32_t lane = vp_lane();
  const uint32_t nw = gridDim.x * (VP_BLOCK / VP_WAVE);
  const bool sparse = vp_sparse(g, b.ctr, a);
  for (uint32_t li = blockIdx.x * (VP_BLOCK / VP_WAVE) + threadIdx.x / VP_WAVE; li < g.nlong; li += nw) {
    const LongRow R = uniform_row(t.rows, li);
    if (R.r < a.r0 || R.r >= a.r1) continue;  // wave-uniform
    if (sparse && (uint32_t)__builtin_amdgcn_readfirstlane(b.stamp[R.r]) != a.tag) continue;
    const uint32_t nch = (R.e1 - R.e0 + VP_WAVE - 1) / VP_WAVE;
    A y = A();
    bool touched = false;
    // Loads of chunks past the row's last read its last chunk again (and are
    // not folded): unconditional loads keep the wait counts exact, so the
    // next group's loads stay in flight while this group is folded.
    uint64_t m[D];
    A v[D];
#pragma unroll
    for (uint32_t j = 0; j < D; j++) {
      const uint64_t ch = (uint64_t)R.cbase + min(j, nch - 1);
      m[j] = t.mask[ch];
      v[j] = t.val[ch * VP_WAVE + lane];
    }
    for (uint32_t c = 0; c < nch; c += D) {
      uint64_t mn[D];
      A vn[D];
#pragma unroll
      for (uint32_t j = 0; j < D; j++) {
        const uint64_t ch = (uint64_t)R.cbase + min(c + D + j, nch - 1);
        mn[j] = t.mask[ch];
        vn[j] = t.val[ch * VP_WAVE + lane];
      }
#pragma unroll
      for (uint32_t j = 0; j < D; j++)
        if (c + j < nch) {
          // the mask is the same in every lane: in scalar registers the
          // fold loop and a uniform accumulator run on the scalar unit
          uint64_t mm = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(m[j] >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)m[j]);
          touched |= mm != 0;
          if (mm == ~0ull) {  // every entry active: a straight-line fold, lanes in order
#pragma unroll
            for (uint32_t l = 0; l < VP_WAVE; l++) op_combine(p, wave_read(v[j], l), y);
          } else {
            while (mm) {
              const uint32_t src = (uint32_t)__builtin_ctzll(mm);
              mm &= mm - 1;
              op_combine(p, wave_read(v[j], src), y);
            }
          }
        }
#pragma unroll
      for (uint32_t j = 0; j < D; j++) {
        m[j] = mn[j];
        v[j] = vn[j];
      }
    }
    if (lane == 0 && touched) vp_finish_row(p, g, b, a, R.r, y);
  }
}

// Sparse passes: stamps every row of [r0, r1) that an active sender's column
// reaches (columns [regular | source]: column i is sender slot i, or
// i + nsink for a source). Columns of at most VP_WAVE entries are walked by
// their lane, longer ones by the whole wavefront.
template <class P>
__global__ void __launch_bounds__(VP_BLOCK) k_vp_mark(la3_device_view g, Bufs<P> b, PassArgs a) {
  if (!vp_sparse(g, b.ctr, a)) return;
  const uint8_t* act = b.act[a.cur];
  const bool sink_rows = a.r0 >= g.nreg;
  const uint32_t nc = g.nreg + (a.use_src ? g.nsrc : 0);
  const uint32_t lane = vp_lane();
  for (uint32_t base = blockIdx.x * blockDim.x; base < nc; base += gridDim.x * blockDim.x) {  // block-uniform
    const uint32_t i = base + threadIdx.x;
    uint32_t lo = 0, hi = 0;
    if (i < nc && act[i < g.nreg ? i : i + g.nsink]) {
      lo = sink_rows ? g.col_split[i] : g.col_ptr[i];
      hi = sink_rows ? g.col_ptr[i + 1] : g.col_split[i];
    }
    if (hi - lo <= VP_WAVE) {
      for (uint32_t e = lo; e < hi; e++) b.stamp[g.col_rows[e]] = a.tag;
      lo = hi = 0;
    }
    uint64_t wide = __ballot(hi > lo);
    while (wide) {
      const uint32_t src = (uint32_t)__builtin_ctzll(wide);
      wide &= wide - 1;
      const uint32_t l0 = __builtin_amdgcn_readlane(lo, src), l1 = __builtin_amdgcn_readlane(hi, src);
      for (uint32_t e = l0 + lane; e < l1; e += VP_WAVE) b.stamp[g.col_rows[e]] = a.tag;
    }
  }
}

// Zeroes the counters (before initialize).
template <class P>
__global__ void k_vp_reset_ctr(Bufs<P> b) {
  if (threadIdx.x < 16) b.ctr[threadIdx.x] = 0;
}

// States value-initialized, as the reference's vertex vector is on
// allocation (random_access_array.h:26-28).
template <class S>
__global__ void k_vp_fill_states(S* v, uint32_t n) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) v[s] = S();
}

// initialize() / initialize(vids) (vertex_program.hpp:84-157, 177-236): init
// on each vertex's current state (all vertices, or the slots flagged in
// `only`; the others keep theirs); regular and, on a directed graph, source
// vertices whose init returns true scatter into message buffer 0; no other
// vertex has a message pending. Flags and messages of buffer 1 start clear.
// With `mirror` (the first initialize of the program's states), the mirrors
// take the active states — every state when stationary (activate_all), else
// those whose activity is set — and the activity is then cleared
// (vertex_program.hpp:159-173).
template <class P>
__global__ void k_vp_init(la3_device_view g, P p, Bufs<P> b, uint32_t stationary, const uint8_t* only,
                          uint32_t mirror) {
  using S = typename P::S;
  const uint32_t src0 = g.nreg + g.nsink, src1 = src0 + g.nsrc;
  uint32_t nsrc_act = 0;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x) {
    S st = b.v[s];
    bool a = false;
    if (!only || only[s]) a = op_init(p, stationary != 0, g.id_of[s], st);
    b.v[s] = st;
    if (b.vm && mirror) {
      b.vm[s] = stationary || b.vact[s] ? st : S();
      b.vact[s] = 0;
    }
    const bool sends = s < g.nreg || (g.directed && s >= src0 && s < src1);
    b.act[0][s] = sends && a;
    b.act[1][s] = 0;
    if (sends && a) {
      const typename P::M m = op_scatter(p, st);
      b.x[0][s] = m;
      if (s >= g.nreg) b.x[1][s] = m;  // source messages are read from either buffer
    }
    if (sends && a && s >= g.nreg) {
      b.act[1][s] = 1;
      nsrc_act++;
    }
  }
  block_add(nsrc_act, b.ctr + 6);
}

// initialize(other) (vertex_program.hpp:410-508): init(vid, other's state of
// the same vertex, s). The reference pairs positions of G and the reversed
// graph (sinks of one are sources of the other); pairing by vertex id gives
// the same state to the same vertex.
template <class P, class OS>
__global__ void k_vp_init_other(la3_device_view g, P p, Bufs<P> b, uint32_t stationary, const OS* other,
                                const uint32_t* other_slot_of_id, uint32_t mirror) {
  using S = typename P::S;
  const uint32_t src0 = g.nreg + g.nsink, src1 = src0 + g.nsrc;
  uint32_t nsrc_act = 0;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x) {
    S st = b.v[s];
    const uint32_t vid = g.id_of[s];
    const bool a = op_init_other(p, stationary != 0, vid, other[other_slot_of_id[vid]], st);
    b.v[s] = st;
    if (b.vm && mirror) {  // vertex_program.hpp:490-505
      b.vm[s] = stationary || b.vact[s] ? st : S();
      b.vact[s] = 0;
    }
    const bool sends = s < g.nreg || (g.directed && s >= src0 && s < src1);
    b.act[0][s] = sends && a;
    b.act[1][s] = 0;
    if (sends && a) {
      const typename P::M m = op_scatter(p, st);
      b.x[0][s] = m;
      if (s >= g.nreg) b.x[1][s] = m;
    }
    if (sends && a && s >= g.nreg) {
      b.act[1][s] = 1;
      nsrc_act++;
    }
  }
  block_add(nsrc_act, b.ctr + 6);
}

// Before a pass that reads buffer which ^ 1 and writes buffer `which`:
// clears the regular flags of `which` and the activation flag, counts the
// active regular senders of which ^ 1 into count slot `slot` (zeroed by the
// previous call, or at initialize) and zeroes the other slot.
template <class P>
__global__ void k_vp_clear(la3_device_view g, Bufs<P> b, uint32_t which, uint32_t slot) {
  uint32_t n = 0;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nreg; s += gridDim.x * blockDim.x) {
    b.act[which][s] = 0;
    n += b.act[which ^ 1][s];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    b.ctr[0] = 0;
    b.ctr[4 + (slot ^ 1)] = 0;
  }
  block_add(n, b.ctr + 4 + slot);
}

static __global__ void k_vp_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Every regular vertex scatters its final state into buffer `which` (the
// sink step of execute_, vertex_program_execute.hpp:137-143).
template <class P>
__global__ void k_vp_scatter_all(la3_device_view g, P p, Bufs<P> b, uint32_t which) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nreg; s += gridDim.x * blockDim.x) {
    b.x[which][s] = op_scatter(p, b.v[s]);
    b.act[which][s] = 1;
  }
}

// States by original id (the reference's states by unhashed id).
template <class S>
__global__ void k_vp_by_id(la3_device_view g, const S* v, S* out) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x)
    out[g.id_of[s]] = v[s];
}

inline uint32_t vp_blocks(uint64_t n) {
  uint64_t b = (n + VP_BLOCK - 1) / VP_BLOCK;
  if (b == 0) b = 1;
  if (b > 8192) b = 8192;
  return (uint32_t)b;
}

// Device memory for one program: allocate / free callbacks (hipMalloc by
// default; the library passes its context scratch).
struct Allocator {
  void* self = nullptr;
  void* (*alloc)(void* self, const char* name, size_t bytes) = nullptr;
  void (*release)(void* self, void* p) = nullptr;
};

inline void* hip_alloc(void*, const char*, size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  return p;
}
inline void hip_release(void*, void* p) {
  if (p) (void)hipFree(p);
}

static __global__ void k_vp_fill_u8(uint8_t* p, uint64_t n, uint8_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// The vertex side of a program — states, their activity and the mirrors —
// which a program constructed from another borrows (the reference's
// VectorV, shared when owns_vertices is false, vertex_program.hpp:22-30).
// Allocated by the first engine that needs it, with that engine's
// allocator; freed when the last engine holding it lets it go.
template <class S>
struct VStore {
  S* v = nullptr;
  S* vm = nullptr;
  uint8_t* vact = nullptr;
  bool mirrored = false;  // v->mirrors_allocated (vertex_program.hpp:159-173)
  uint32_t n = 0;
  Allocator al;
  VStore() = default;
  VStore(const VStore&) = delete;
  VStore& operator=(const VStore&) = delete;
  ~VStore() {
    if (!al.release) return;
    al.release(al.self, v);
    al.release(al.self, vm);
    al.release(al.self, vact);
  }
  // States value-initialized and activity clear (random_access_array.h:26-28).
  hipError_t ensure(const Allocator& a, uint32_t nslots, hipStream_t st) {
    if (v) return hipSuccess;
    al = a;
    n = nslots ? nslots : 1;
    v = (S*)al.alloc(al.self, "vp.v", sizeof(S) * (size_t)n);
    vact = (uint8_t*)al.alloc(al.self, "vp.vact", (size_t)n + 16);
    if (!v || !vact) return hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_vp_fill_states<S>, dim3(vp_blocks(n)), dim3(VP_BLOCK), 0, st, v, n);
    hipLaunchKernelGGL(k_vp_fill_u8, dim3(vp_blocks(n)), dim3(VP_BLOCK), 0, st, vact, (uint64_t)n, (uint8_t)0);
    return hipGetLastError();
  }
  hipError_t ensure_mirrors() {
    if (vm) return hipSuccess;
    vm = (S*)al.alloc(al.self, "vp.vm", sizeof(S) * (size_t)n);
    return vm ? hipSuccess : hipErrorOutOfMemory;
  }
};

// The execution engine (vertex_program_execute.hpp:11-345) over one graph's
// device layout on one stream. Returns a hipError_t; `err` says what failed.
template <class P>
class Engine {
 public:
  using W = typename P::W;
  using M = typename P::M;
  using A = typename P::A;
  using S = typename P::S;
  static_assert(std::is_trivially_copyable<M>::value && std::is_trivially_copyable<A>::value &&
                    std::is_trivially_copyable<S>::value,
                "messages, accumulators and states must be trivially copyable");
  static_assert(std::is_same<W, Empty>::value || std::is_same<W, uint32_t>::value,
                "edge weights are la3::Empty (pairs) or uint32_t (LA3 weighted triples)");
  static constexpr bool gather_with_state = has_gather3<P>::value;
  static constexpr bool apply_with_iter = has_apply3<P>::value;

  // store: the vertex side of another program on the same graph to borrow
  // (null: this program owns its own).
  Engine(const la3_device_view& g, hipStream_t stream, bool stationary, const P& ops, Allocator al = Allocator(),
         std::shared_ptr<VStore<S>> store = nullptr)
      : g_(g), stream_(stream), stationary_(stationary), p_(ops), al_(al), vs_(std::move(store)) {
    if (!al_.alloc) {
      al_.alloc = hip_alloc;
      al_.release = hip_release;
    }
    if (!vs_) vs_ = std::make_shared<VStore<S>>();
  }
  ~Engine() { release(); }
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  std::string err;
  bool optimizable = true;
  uint32_t iters_run = 0;
  const Bufs<P>& bufs() const { return b_; }
  const la3_device_view& view() const { return g_; }
  const std::shared_ptr<VStore<S>>& store() const { return vs_; }
  // The states, allocated (and value-initialized) on first use.
  const S* states_ptr() {
    if (!vs_) vs_ = std::make_shared<VStore<S>>();
    return vs_->ensure(al_, g_.nslots, stream_) == hipSuccess ? vs_->v : nullptr;
  }

  hipError_t initialize(const uint8_t* only_slots = nullptr) {
    hipError_t e = ensure();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_vp_reset_ctr<P>, dim3(1), dim3(64), 0, stream_, b_);
    hipLaunchKernelGGL(k_vp_init<P>, dim3(vp_blocks(g_.nslots)), dim3(VP_BLOCK), 0, stream_, g_, p_, b_,
                       (uint32_t)stationary_, only_slots, (uint32_t)mirror_now());
    initialized_ = true;
    cur_ = 0;
    return hipGetLastError();
  }

  template <class OS>
  hipError_t initialize_other(const OS* other_states, const uint32_t* other_slot_of_id) {
    hipError_t e = ensure();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_vp_reset_ctr<P>, dim3(1), dim3(64), 0, stream_, b_);
    hipLaunchKernelGGL((k_vp_init_other<P, OS>), dim3(vp_blocks(g_.nslots)), dim3(VP_BLOCK), 0, stream_, g_, p_,
                       b_, (uint32_t)stationary_, other_states, other_slot_of_id, (uint32_t)mirror_now());
    initialized_ = true;
    cur_ = 0;
    return hipGetLastError();
  }

  // reset_activity() / activate_all() (vertex_program.hpp:549-555): clear or
  // set the activity of every vertex state.
  hipError_t set_activity(bool on) {
    hipError_t e = ensure();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_vp_fill_u8, dim3(vp_blocks(g_.nslots)), dim3(VP_BLOCK), 0, stream_, b_.vact,
                       (uint64_t)g_.nslots, (uint8_t)(on ? 1 : 0));
    return sync_error();
  }

  // States and activity in slot order (the reference's state positions).
  hipError_t by_slot(S* states, uint8_t* activity) {
    hipError_t e = ensure();
    if (e != hipSuccess) return e;
    if (states) e = hipMemcpyAsync(states, b_.v, sizeof(S) * (size_t)g_.nslots, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess && activity)
      e = hipMemcpyAsync(activity, b_.vact, (size_t)g_.nslots, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    return e == hipSuccess ? hipSuccess : fail(e);
  }

  // execute(max_iters) (vertex_program_execute.hpp:11-52).
  hipError_t execute(uint32_t max_iters) {
    hipError_t e;
    if (!initialized_ && (e = initialize()) != hipSuccess) return e;
    const bool directed = g_.directed != 0;
    const bool opt = optimizable && directed && !gather_with_state && !apply_with_iter;
    const uint32_t nrows = g_.nreg + g_.nsink;
    // A second execute() without reset() continues from the messages the
    // first one left (the reference initializes only once, :15-16).
    if (max_iters == 1) {  // execute_single / execute_single_undirected (:251-345)
      clear(cur_ ^ 1);
      pass(0, g_.nreg, 1, 0, true);
      if (directed) pass(g_.nreg, nrows, 1, 0, true);
      iters_run = 1;
      return sync_error();
    }
    const bool until = max_iters == 0;
    uint32_t iter = 0;
    for (;;) {
      if (!until && iter >= max_iters) break;
      clear(cur_ ^ 1);
      pass(0, g_.nreg, stationary_ || iter == 0, iter, false);
      if (!opt && directed) pass(g_.nreg, nrows, 1, iter, false);
      cur_ ^= 1;
      iter++;
      if (until) {
        uint32_t any = 0;
        if ((e = hipMemcpyAsync(&any, b_.ctr, 4, hipMemcpyDeviceToHost, stream_)) != hipSuccess) return fail(e);
        if ((e = hipStreamSynchronize(stream_)) != hipSuccess) return fail(e);
        if (!any) break;
      }
    }
    iters_run = iter;
    if (opt && g_.nsink) {
      // Sink step: every regular vertex re-scatters its final state; the
      // sink rows take source and regular messages once (:131-148).
      hipLaunchKernelGGL(k_vp_scatter_all<P>, dim3(vp_blocks(g_.nreg)), dim3(VP_BLOCK), 0, stream_, g_, p_, b_, cur_);
      pass(g_.nreg, nrows, 1, iter, false, true);
    }
    return sync_error();
  }

  hipError_t states_by_id(S* host_out) {
    hipError_t e0 = ensure();
    if (e0 != hipSuccess) return e0;
    S* tmp = (S*)al_.alloc(al_.self, "vp.out", sizeof(S) * (size_t)g_.nslots);
    if (!tmp) return fail(hipErrorOutOfMemory);
    hipLaunchKernelGGL(k_vp_by_id<S>, dim3(vp_blocks(g_.nslots)), dim3(VP_BLOCK), 0, stream_, g_,
                       (const S*)b_.v, tmp);
    hipError_t e = hipMemcpyAsync(host_out, tmp, sizeof(S) * (size_t)g_.nslots, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    al_.release(al_.self, tmp);
    return e == hipSuccess ? hipSuccess : fail(e);
  }

  // reset() (vertex_program.hpp:51-63): the next execute() initializes again
  // (pending messages dropped); the states stay.
  void reset() { initialized_ = false; }

  // free() (vertex_program.hpp:38-48): this program's buffers, and the
  // states unless another program still holds them.
  void release() {
    if (!allocated_) {
      if (vs_ && vs_->v) (void)hipStreamSynchronize(stream_);
      vs_.reset();
      return;
    }
    (void)hipStreamSynchronize(stream_);
    vs_.reset();  // the states go with their last holder
    al_.release(al_.self, b_.x[0]);
    al_.release(al_.self, b_.x[1]);
    al_.release(al_.self, b_.act[0]);
    al_.release(al_.self, b_.act[1]);
    al_.release(al_.self, b_.ctr);
    al_.release(al_.self, b_.stamp);
    al_.release(al_.self, (void*)lt_.rows);
    al_.release(al_.self, (void*)lt_.sbase);
    al_.release(al_.self, (void*)ltot_);
    al_.release(al_.self, (void*)lt_.seg);
    al_.release(al_.self, lt_.val);
    al_.release(al_.self, lt_.mask);
    al_.release(al_.self, lt_.part);
    al_.release(al_.self, lt_.pany);
    lt_ = LongTabs<A>();
    ltot_ = nullptr;
    nseg_ = 0;
    b_ = Bufs<P>();
    allocated_ = false;
    initialized_ = false;
  }

 private:
  la3_device_view g_;
  hipStream_t stream_;
  bool stationary_;
  P p_;
  Allocator al_;
  Bufs<P> b_;
  std::shared_ptr<VStore<S>> vs_;
  LongTabs<A> lt_;
  uint32_t* ltot_ = nullptr;
  uint32_t nseg_ = 0;
  uint32_t tag_ = 0;                   // the last pass's stamp
  uint32_t ccount_ = 0, cslot_ = 0;    // count slots used alternately
  bool allocated_ = false, initialized_ = false;
  uint32_t cur_ = 0;

  // The first initialize of the (possibly shared) states of a state-reading
  // program mirrors them; later ones do not.
  bool mirror_now() {
    if (!gather_with_state || vs_->mirrored) return false;
    vs_->mirrored = true;
    return true;
  }

  hipError_t fail(hipError_t e) {
    if (err.empty()) err = hipGetErrorString(e);
    return e;
  }

  hipError_t ensure() {
    if (allocated_) return hipSuccess;
    if (!vs_) vs_ = std::make_shared<VStore<S>>();  // after free(): fresh states
    const size_t S_ = g_.nslots ? g_.nslots : 1;
    hipError_t es = vs_->ensure(al_, g_.nslots, stream_);
    if (es == hipSuccess && gather_with_state) es = vs_->ensure_mirrors();
    if (es != hipSuccess) {
      err = "device allocation failed";
      return es;
    }
    b_.v = vs_->v;
    b_.vm = vs_->vm;
    b_.vact = vs_->vact;
    b_.x[0] = (M*)al_.alloc(al_.self, "vp.x0", sizeof(M) * S_ + 16);
    b_.x[1] = (M*)al_.alloc(al_.self, "vp.x1", sizeof(M) * S_ + 16);
    b_.act[0] = (uint8_t*)al_.alloc(al_.self, "vp.a0", S_ + 16);
    b_.act[1] = (uint8_t*)al_.alloc(al_.self, "vp.a1", S_ + 16);
    b_.ctr = (uint32_t*)al_.alloc(al_.self, "vp.ctr", 64);
    b_.stamp = (uint32_t*)al_.alloc(al_.self, "vp.stamp", 4 * S_);
    allocated_ = true;
    if (!b_.x[0] || !b_.x[1] || !b_.act[0] || !b_.act[1] || !b_.ctr || !b_.stamp) {
      release();
      err = "device allocation failed";
      return hipErrorOutOfMemory;
    }
    tag_ = 0;
    hipLaunchKernelGGL(k_vp_fill_u32, dim3(vp_blocks(S_)), dim3(VP_BLOCK), 0, stream_, b_.stamp, (uint64_t)S_, 0u);
    const hipError_t e = ensure_long();
    if (e != hipSuccess) {
      release();
      err = e == hipErrorOutOfMemory ? "device allocation failed" : hipGetErrorString(e);
      return e;
    }
    return hipSuccess;
  }

  void clear(uint32_t which) {
    cslot_ = (ccount_++) & 1;
    hipLaunchKernelGGL(k_vp_clear<P>, dim3(std::min<uint32_t>(vp_blocks(g_.nreg), 1024)), dim3(VP_BLOCK), 0, stream_,
                       g_, b_, which, cslot_);
  }

  // One pass over rows [r0, r1): with dense_only, every row (the sink step
  // after every regular vertex re-scattered); otherwise dense or sparse by
  // the counted senders.
  void pass(uint32_t r0, uint32_t r1, bool use_src, uint32_t iter, bool single, bool dense_only = false) {
    if (r1 <= r0) return;
    if (++tag_ == 0) {  // stamps wrapped: clear them
      hipLaunchKernelGGL(k_vp_fill_u32, dim3(vp_blocks(g_.nslots)), dim3(VP_BLOCK), 0, stream_, b_.stamp,
                         (uint64_t)g_.nslots, 0u);
      tag_ = 1;
    }
    PassArgs a{r0, r1, cur_, iter, (uint32_t)use_src, (uint32_t)single, (uint32_t)stationary_,
               tag_, cslot_, (uint32_t)dense_only};
    if (!dense_only) {
      const uint32_t nc = g_.nreg + (use_src ? g_.nsrc : 0);
      hipLaunchKernelGGL(k_vp_mark<P>, dim3(vp_blocks(nc)), dim3(VP_BLOCK), 0, stream_, g_, b_, a);
    }
    hipLaunchKernelGGL(k_vp_pull<P>, dim3(vp_blocks(r1 - r0)), dim3(VP_BLOCK), 0, stream_, g_, p_, b_, a);
    if (g_.nlong) {
      hipLaunchKernelGGL(k_vp_gather_long<P>, dim3(vp_blocks((uint64_t)nseg_ * VP_WAVE)), dim3(VP_BLOCK), 0,
                         stream_, g_, p_, b_, a, lt_);
      if constexpr (SegFold<P>::on)
        hipLaunchKernelGGL(k_vp_join_long<P>, dim3(vp_blocks((uint64_t)g_.nlong * VP_WAVE)), dim3(VP_BLOCK), 0,
                           stream_, g_, p_, b_, a, lt_);
      else
        hipLaunchKernelGGL(k_vp_fold_long<P>, dim3(vp_blocks((uint64_t)g_.nlong * VP_WAVE)), dim3(VP_BLOCK), 0,
                           stream_, g_, p_, b_, a, lt_);
    }
  }

  // The long-row tables and staging (once per program; one synchronization
  // to size the staging).
  hipError_t ensure_long() {
    if (!g_.nlong) return hipSuccess;
    const uint32_t nb = (g_.nlong + VP_TAB - 1) / VP_TAB;
    LongRow* rows = (LongRow*)al_.alloc(al_.self, "vp.lrows", sizeof(LongRow) * (size_t)g_.nlong);
    uint32_t* sbase = (uint32_t*)al_.alloc(al_.self, "vp.lsbase", 4ull * g_.nlong);
    uint32_t* totals = (uint32_t*)al_.alloc(al_.self, "vp.ltot", 16 + 8ull * nb);  // totals, block sums
    lt_.rows = rows;
    lt_.sbase = sbase;
    lt_.nseg = totals ? totals + 1 : nullptr;
    ltot_ = totals;
    if (!rows || !sbase || !totals) return hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_vp_long_sums, dim3(nb), dim3(VP_TAB), 0, stream_, g_, totals + 4);
    hipLaunchKernelGGL(k_vp_long_scan, dim3(1), dim3(VP_TAB), 0, stream_, totals + 4, nb, totals);
    hipLaunchKernelGGL(k_vp_long_table, dim3(nb), dim3(VP_TAB), 0, stream_, g_, (const uint32_t*)(totals + 4),
                       rows, sbase);
    uint32_t tot[2] = {0, 0};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(tot, totals, 8, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return e;
    nseg_ = tot[1];
    uint32_t* seg = (uint32_t*)al_.alloc(al_.self, "vp.lseg", 4ull * (nseg_ ? nseg_ : 1));
    lt_.seg = seg;
    if constexpr (SegFold<P>::on) {  // one partial per segment
      lt_.part = al_.alloc(al_.self, "vp.lpart", sizeof(typename SegFold<P>::Seg) * (size_t)(nseg_ ? nseg_ : 1));
      lt_.pany = (uint8_t*)al_.alloc(al_.self, "vp.lpany", nseg_ ? nseg_ : 1);
      if (!seg || !lt_.part || !lt_.pany) return hipErrorOutOfMemory;
    } else {  // every entry's value and per-chunk masks
      lt_.val = (A*)al_.alloc(al_.self, "vp.lval", sizeof(A) * VP_WAVE * (size_t)(tot[0] ? tot[0] : 1));
      lt_.mask = (uint64_t*)al_.alloc(al_.self, "vp.lmask", 8ull * (tot[0] ? tot[0] : 1));
      if (!seg || !lt_.val || !lt_.mask) return hipErrorOutOfMemory;
    }
    hipLaunchKernelGGL(k_vp_long_segs, dim3(vp_blocks(g_.nlong)), dim3(VP_BLOCK), 0, stream_, g_,
                       (const LongRow*)rows, (const uint32_t*)sbase, seg);
    return hipGetLastError();
  }

  hipError_t sync_error() {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    return e == hipSuccess ? hipSuccess : fail(e);
  }
};

}  // namespace vp_detail

// VertexProgram<W, M, A, S> (vertex_program.h:21-220) over a resident graph:
// the program's operations are the members of P.
template <class P>
class VertexProgram {
 public:
  using W = typename P::W;
  using M = typename P::M;
  using A = typename P::A;
  using S = typename P::S;

  // VertexProgram(const Graph<W>* G, bool stationary) (vertex_program.h:33).
  explicit VertexProgram(const Graph<W>* G, bool stationary = false, const P& ops = P())
      : VertexProgram(G->context().get(), G->handle(), stationary, ops) {}
  VertexProgram(la3_ctx* ctx, la3_graph* g, bool stationary = false, const P& ops = P())
      : ctx_(ctx), g_(g), view_(view_of(ctx, g)),
        e_(view_, (hipStream_t)la3_ctx_stream(ctx), stationary, ops) {}

  // VertexProgram(const VertexProgram<W, M2, A2, S>& other, bool stationary)
  // (vertex_program.h:41, vertex_program.hpp:22-30): the same graph, and the
  // other program's vertex states, their activity and mirrors, shared (the
  // reference's borrowed VectorV); messages and accumulators of its own.
  // The states stay allocated while either program holds them.
  template <class P2>
  VertexProgram(const VertexProgram<P2>& other, bool stationary = false, const P& ops = P())
      : ctx_(other.context_handle()), g_(other.graph_handle()), view_(other.view()),
        e_(view_, (hipStream_t)la3_ctx_stream(ctx_), stationary, ops, vp_detail::Allocator(),
           other.shared_states()) {
    static_assert(std::is_same<typename P2::S, S>::value, "a borrowing program has the other's state type");
    static_assert(std::is_same<typename P2::W, W>::value, "a borrowing program has the other's weight type");
  }

  // LA3-optimizable execution (vertex_program.h:140-148); cleared to force
  // execute_non_opt. Every initialize ANDs in the graph's and the program's
  // conditions, as initialize_flags does (vertex_program.hpp:66-81).
  bool optimizable = true;

  // initialize() (vertex_program.hpp:84-174): init on every vertex's
  // current state (S() until the first execute).
  void initialize() {
    and_flags();
    check(e_.initialize(), "initialize");
  }

  // initialize(vids) (vertex_program.hpp:177-259): init only the listed
  // vertices; the others keep S() and send nothing.
  void initialize(const std::vector<uint32_t>& vids) {
    and_flags();
    std::vector<uint32_t> slot(view_.nslots);
    check(hipMemcpy(slot.data(), view_.slot_of_id, 4ull * view_.nslots, hipMemcpyDeviceToHost), "initialize");
    std::vector<uint8_t> only(view_.nslots, 0);
    for (uint32_t v : vids) {
      if (v >= view_.nslots) throw Error(LA3_EINVAL, "initialize: vertex id out of range");
      only[slot[v]] = 1;
    }
    uint8_t* d = nullptr;
    check(hipMalloc(&d, only.size()), "initialize");
    hipError_t e = hipMemcpy(d, only.data(), only.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = e_.initialize(d);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)la3_ctx_stream(ctx_));
    (void)hipFree(d);
    check(e, "initialize");
  }

  // initialize(other) (vertex_program.hpp:410-508): init(vid, state in
  // `other`, s) for every vertex, other being a program on the same vertex
  // set (typically the reversed graph).
  template <class P2>
  void initialize(const VertexProgram<P2>& other) {
    and_flags();
    check(e_.initialize_other(other.device_states(), other.view().slot_of_id), "initialize(other)");
  }

  // execute(max_iters) (vertex_program_execute.hpp:11-52); without a prior
  // initialize it initializes, and a second execute continues from the
  // messages the first left.
  void execute(uint32_t max_iters = 0) {
    and_flags();
    e_.optimizable = optimizable;
    check(e_.execute(max_iters), "execute");
  }

  // reset() (vertex_program.hpp:51-63): pending messages are dropped and the
  // next execute initializes again; the states stay. free() (:38-48).
  void reset() { e_.reset(); }
  void free() { e_.release(); }

  uint32_t iterations() const { return e_.iters_run; }

  // reset_activity() / activate_all() (vertex_program.hpp:549-555): clear or
  // set the activity of every vertex state. Apply sets it for a regular
  // vertex that was activated or is stationary and for a sink that was
  // activated (vertex_program_execute.hpp:693-713); it is kept across
  // executes and resets, and cleared after the first mirroring of a
  // state-reading program, which mirrors the active states.
  void reset_activity() { check(e_.set_activity(false), "reset_activity"); }
  void activate_all() { check(e_.set_activity(true), "activate_all"); }

  // States by original vertex id (nslots = n + 1 entries).
  std::vector<S> states() {
    std::vector<S> out(view_.nslots);
    check(e_.states_by_id(out.data()), "states");
    return out;
  }
  // Vertex activity by original vertex id (1 = active).
  std::vector<uint8_t> activity() {
    std::vector<uint8_t> a(view_.nslots), out(view_.nslots);
    check(e_.by_slot(nullptr, a.data()), "activity");
    const std::vector<uint32_t>& ids = ids_by_slot();
    for (size_t p = 0; p < a.size(); p++) out[ids[p]] = a[p];
    return out;
  }
  const S* device_states() const {
    const S* v = const_cast<vp_detail::Engine<P>&>(e_).states_ptr();
    if (!v) throw Error(LA3_ENOMEM, "device allocation failed");
    return v;
  }
  const la3_device_view& view() const { return view_; }
  la3_ctx* context_handle() const { return ctx_; }
  la3_graph* graph_handle() const { return g_; }
  std::shared_ptr<vp_detail::VStore<S>> shared_states() const {
    device_states();  // allocated, so that both programs hold the same states
    return e_.store();
  }

  // reduce (vertex_program.hpp:558-604) at one rank: map every vertex (or
  // every active one) in the reference's slot order, with its original id,
  // and fold with reduce into Value().
  template <class Value, class Mapper, class Reducer>
  Value reduce(Mapper map, Reducer red, bool active_only = false) {
    std::vector<S> st;
    std::vector<uint8_t> act;
    slot_states(st, act, active_only);
    const std::vector<uint32_t>& ids = ids_by_slot();
    Value acc = Value();
    for (size_t p = 0; p < st.size(); p++)
      if (!active_only || act[p]) red(acc, map(ids[p], st[p]));
    return acc;
  }

  // topk (vertex_program.hpp:607-680) at one rank, step for step:
  // (id, map(id, state)) pairs in slot order (active vertices only if asked),
  // std::partial_sort of the first min(k, count) by cmp, resized to k (pairs
  // I(), V() appended when fewer), partial_sort of those k, resized to k.
  template <class I, class V, class Mapper, class Comparator>
  void topk(uint32_t k, std::vector<std::pair<I, V>>& out, Mapper map, Comparator cmp, bool active_only = false) {
    std::vector<S> st;
    std::vector<uint8_t> act;
    slot_states(st, act, active_only);
    const std::vector<uint32_t>& ids = ids_by_slot();
    std::vector<std::pair<I, V>> ivs;
    for (size_t p = 0; p < st.size(); p++)
      if (!active_only || act[p]) ivs.push_back(std::pair<I, V>((I)ids[p], map(ids[p], st[p])));
    out.clear();
    select_k(ivs, k, cmp);
    out = ivs;
  }

  // btopk<BATCH_SIZE> (vertex_program.hpp:683-767) at one rank: map returns
  // a pointer to BATCH_SIZE values of a vertex; topk's steps per batch b
  // into topk[b].
  template <uint32_t BATCH_SIZE, class I, class V, class Mapper, class Comparator>
  void btopk(uint32_t k, std::vector<std::pair<I, V>>* out, Mapper map, Comparator cmp, bool active_only = false) {
    std::vector<S> st;
    std::vector<uint8_t> act;
    slot_states(st, act, active_only);
    const std::vector<uint32_t>& ids = ids_by_slot();
    std::vector<std::pair<I, V>> ivs[BATCH_SIZE];
    for (uint32_t b = 0; b < BATCH_SIZE; b++) out[b].clear();
    for (size_t p = 0; p < st.size(); p++) {
      if (active_only && !act[p]) continue;
      const V* vals = map(ids[p], st[p]);
      for (uint32_t b = 0; b < BATCH_SIZE; b++) ivs[b].push_back(std::pair<I, V>((I)ids[p], vals[b]));
    }
    for (uint32_t b = 0; b < BATCH_SIZE; b++) {
      select_k(ivs[b], k, cmp);
      out[b] = ivs[b];
    }
  }

  // display<values_only>(nvertices) (vertex_program.hpp:511-546): in slot
  // order, every vertex whose original id j is at most min(nvertices, n) as
  // "j: " + state.to_string() (values_only: the string alone, and one empty
  // line after the list), one line each, to `os` (the reference logs them).
  // S must have `std::string to_string() const`, as the reference's states do.
  template <bool values_only = false>
  void display(uint32_t nvertices = 30, std::ostream& os = std::cout) {
    static_assert(has_to_string<S>::value, "display() needs S::to_string() const (as the reference's states)");
    std::vector<S> st;
    std::vector<uint8_t> act;
    slot_states(st, act, false);
    const std::vector<uint32_t>& ids = ids_by_slot();
    const uint32_t nmax = std::min<uint32_t>(nvertices, view_.nslots ? view_.nslots - 1 : 0);
    for (size_t p = 0; p < st.size(); p++) {
      if (ids[p] > nmax) continue;
      if (values_only) os << st[p].to_string() << "\n";
      else os << ids[p] << ": " << st[p].to_string() << "\n";
    }
    if (values_only) os << "\n";
  }

 private:
  template <class T, class = void> struct has_to_string : std::false_type {};
  template <class T>
  struct has_to_string<T, std::void_t<decltype(std::declval<const T&>().to_string())>> : std::true_type {};

  std::vector<uint32_t> ids_;  // original id by slot (the layout does not change)
  const std::vector<uint32_t>& ids_by_slot() {
    if (ids_.size() != view_.nslots) {
      std::vector<uint32_t> hid(view_.nslots);
      if (la3_graph_slots(g_, hid.data()) != LA3_OK) throw Error(LA3_EINVAL, "slot order");
      ids_.resize(view_.nslots);
      for (uint32_t p = 0; p < view_.nslots; p++) ids_[p] = la3_graph_unhash(g_, hid[p]);
    }
    return ids_;
  }
  void slot_states(std::vector<S>& st, std::vector<uint8_t>& act, bool with_activity) {
    st.resize(view_.nslots);
    if (with_activity) act.resize(view_.nslots);
    check(e_.by_slot(st.data(), with_activity ? act.data() : nullptr), "states");
  }
  template <class IV, class Comparator>
  static void select_k(std::vector<IV>& ivs, uint32_t k, Comparator cmp) {
    std::partial_sort(ivs.begin(), ivs.begin() + std::min<size_t>(k, ivs.size()), ivs.end(), cmp);
    ivs.resize(k);
    std::partial_sort(ivs.begin(), ivs.begin() + std::min<size_t>(k, ivs.size()), ivs.end(), cmp);
    ivs.resize(k);
  }

  la3_ctx* ctx_;
  la3_graph* g_;
  la3_device_view view_;
  vp_detail::Engine<P> e_;

  void and_flags() {
    optimizable = optimizable && view_.directed && !vp_detail::Engine<P>::gather_with_state &&
                  !vp_detail::Engine<P>::apply_with_iter;
  }

  static la3_device_view view_of(la3_ctx* ctx, la3_graph* g) {
    la3_device_view v;
    std::memset(&v, 0, sizeof(v));
    la3_status s = la3_graph_device_view(g, &v);
    if (s) throw Error(s, std::string("la3_graph_device_view: ") + la3_last_error(ctx));
    if (std::is_same<W, uint32_t>::value && !v.weighted)
      throw Error(LA3_EINVAL, "a weighted program (W = uint32_t) needs a weighted graph");
    return v;
  }
  void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(LA3_EDEVICE, std::string(what) + ": " + (e_.err.empty() ? hipGetErrorString(e) : e_.err));
  }
};

}  // namespace la3

#endif  // LA3_VERTEX_PROGRAM_HPP
