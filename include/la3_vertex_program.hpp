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
// folded by one lane; longer rows by a wavefront that gathers 64 messages at
// a time in parallel and folds them in order from LDS.
//
// Device memory: the states, two message buffers and two activity-flag
// buffers (by vertex slot), plus a mirror of the states when gather reads
// them. The program runs on its context's stream; states() synchronizes.
#ifndef LA3_VERTEX_PROGRAM_HPP
#define LA3_VERTEX_PROGRAM_HPP

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
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
  uint32_t* ctr = nullptr;      // [0] = some regular vertex was activated
};

// One pass over rows [r0, r1): read the messages of buffer `cur` (source
// slots only when use_src), apply, and for regular rows write the next
// messages into buffer `cur ^ 1` (unless single). Lane per row; rows longer
// than VP_LONG are left to k_vp_pull_long.
struct PassArgs {
  uint32_t r0, r1, cur, iter;
  uint32_t use_src, single, stationary;
};

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
    if (!a.single && (act || a.stationary)) {
      b.x[a.cur ^ 1][r] = op_scatter(p, st);
      b.act[a.cur ^ 1][r] = 1;
    }
  }
}

template <class P>
__global__ void __launch_bounds__(VP_BLOCK) k_vp_pull(la3_device_view g, P p, Bufs<P> b, PassArgs a) {
  using A = typename P::A;
  const typename P::M* x = b.x[a.cur];
  const uint8_t* act = b.act[a.cur];
  for (uint32_t r = a.r0 + blockIdx.x * blockDim.x + threadIdx.x; r < a.r1; r += gridDim.x * blockDim.x) {
    const uint32_t e0 = g.row_ptr[r], es = g.row_split[r], e1 = g.row_ptr[r + 1];
    if (e1 - e0 > VP_LONG) continue;
    A y = A();
    bool touched = false;
    const typename P::S* mir = b.vm ? b.vm + r : nullptr;
    for (uint32_t e = a.use_src ? e0 : es; e < e1; e++) {
      const uint32_t u = g.row_cols[e];
      if (!act[u]) continue;
      op_combine(p, op_gather(p, make_edge<typename P::W>(g, u, r, e), x[u], mir), y);
      touched = true;
    }
    if (touched) vp_finish_row(p, g, b, a, r, y);
  }
}

// Rows longer than VP_LONG: one wavefront per row. Lanes gather 64 entries'
// values at a time (in parallel: flags, messages, gathers); lane 0 then
// combines them in entry order from LDS.
// Orders the LDS writes of one wavefront's lanes before its lanes' reads.
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class P>
__global__ void __launch_bounds__(VP_BLOCK) k_vp_pull_long(la3_device_view g, P p, Bufs<P> b, PassArgs a) {
  using A = typename P::A;
  __shared__ alignas(16) unsigned char raw[VP_BLOCK * sizeof(A)];
  __shared__ uint8_t ok[VP_BLOCK];
  A* val = reinterpret_cast<A*>(raw);
  const uint32_t lane = threadIdx.x % VP_WAVE;
  const uint32_t wv = threadIdx.x / VP_WAVE;
  const uint32_t nw = gridDim.x * (VP_BLOCK / VP_WAVE);
  const typename P::M* x = b.x[a.cur];
  const uint8_t* act = b.act[a.cur];
  for (uint32_t li = blockIdx.x * (VP_BLOCK / VP_WAVE) + wv; li < g.nlong; li += nw) {
    const uint32_t r = g.long_rows[li];
    if (r < a.r0 || r >= a.r1) continue;  // wave-uniform
    const uint32_t e0 = g.row_ptr[r], es = g.row_split[r], e1 = g.row_ptr[r + 1];
    const typename P::S* mir = b.vm ? b.vm + r : nullptr;
    A y = A();
    bool touched = false;
    for (uint32_t base = a.use_src ? e0 : es; base < e1; base += VP_WAVE) {  // wave-uniform
      const uint32_t e = base + lane;
      bool have = false;
      if (e < e1) {
        const uint32_t u = g.row_cols[e];
        if (act[u]) {
          val[threadIdx.x] = op_gather(p, make_edge<typename P::W>(g, u, r, e), x[u], mir);
          have = true;
        }
      }
      ok[threadIdx.x] = have;
      wave_sync_lds();
      if (lane == 0) {
        const uint32_t n = min((uint32_t)VP_WAVE, e1 - base);
        for (uint32_t j = 0; j < n; j++)
          if (ok[wv * VP_WAVE + j]) {
            op_combine(p, val[wv * VP_WAVE + j], y);
            touched = true;
          }
      }
      wave_sync_lds();
    }
    if (lane == 0 && touched) vp_finish_row(p, g, b, a, r, y);
  }
}

// initialize() / initialize(vids) (vertex_program.hpp:84-157, 177-236): init
// per vertex (all, or the slots flagged in `only`); regular and, on a
// directed graph, source vertices whose init returns true scatter into
// message buffer 0. Flags and messages of buffer 1 start clear.
template <class P>
__global__ void k_vp_init(la3_device_view g, P p, Bufs<P> b, uint32_t stationary, const uint8_t* only) {
  using S = typename P::S;
  const uint32_t src0 = g.nreg + g.nsink, src1 = src0 + g.nsrc;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x) {
    S st = S();
    bool a = false;
    if (!only || only[s]) a = op_init(p, stationary != 0, g.id_of[s], st);
    b.v[s] = st;
    if (b.vm) b.vm[s] = stationary ? st : S();  // the active states (vertex_program.hpp:159-173)
    const bool sends = s < g.nreg || (g.directed && s >= src0 && s < src1);
    b.act[0][s] = sends && a;
    b.act[1][s] = 0;
    if (sends && a) {
      const typename P::M m = op_scatter(p, st);
      b.x[0][s] = m;
      if (s >= g.nreg) b.x[1][s] = m;  // source messages are read from either buffer
    }
    if (sends && a && s >= g.nreg) b.act[1][s] = 1;
  }
}

// initialize(other) (vertex_program.hpp:410-508): init(vid, other's state of
// the same vertex, s). The reference pairs positions of G and the reversed
// graph (sinks of one are sources of the other); pairing by vertex id gives
// the same state to the same vertex.
template <class P, class OS>
__global__ void k_vp_init_other(la3_device_view g, P p, Bufs<P> b, uint32_t stationary, const OS* other,
                                const uint32_t* other_slot_of_id) {
  using S = typename P::S;
  const uint32_t src0 = g.nreg + g.nsink, src1 = src0 + g.nsrc;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nslots; s += gridDim.x * blockDim.x) {
    S st = S();
    const uint32_t vid = g.id_of[s];
    const bool a = op_init_other(p, stationary != 0, vid, other[other_slot_of_id[vid]], st);
    b.v[s] = st;
    if (b.vm) b.vm[s] = stationary ? st : S();  // the active states (vertex_program.hpp:159-173)
    const bool sends = s < g.nreg || (g.directed && s >= src0 && s < src1);
    b.act[0][s] = sends && a;
    b.act[1][s] = 0;
    if (sends && a) {
      const typename P::M m = op_scatter(p, st);
      b.x[0][s] = m;
      if (s >= g.nreg) b.x[1][s] = m;
    }
    if (sends && a && s >= g.nreg) b.act[1][s] = 1;
  }
}

// Clears the regular flags of buffer `which` and the activation counter.
template <class P>
__global__ void k_vp_clear(la3_device_view g, Bufs<P> b, uint32_t which) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < g.nreg; s += gridDim.x * blockDim.x)
    b.act[which][s] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) b.ctr[0] = 0;
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

  Engine(const la3_device_view& g, hipStream_t stream, bool stationary, const P& ops, Allocator al = Allocator())
      : g_(g), stream_(stream), stationary_(stationary), p_(ops), al_(al) {
    if (!al_.alloc) {
      al_.alloc = hip_alloc;
      al_.release = hip_release;
    }
  }
  ~Engine() { release(); }
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  std::string err;
  bool optimizable = true;
  uint32_t iters_run = 0;
  const Bufs<P>& bufs() const { return b_; }
  const la3_device_view& view() const { return g_; }

  hipError_t initialize(const uint8_t* only_slots = nullptr) {
    hipError_t e = ensure();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_vp_init<P>, dim3(vp_blocks(g_.nslots)), dim3(VP_BLOCK), 0, stream_, g_, p_, b_,
                       (uint32_t)stationary_, only_slots);
    initialized_ = true;
    return hipGetLastError();
  }

  template <class OS>
  hipError_t initialize_other(const OS* other_states, const uint32_t* other_slot_of_id) {
    hipError_t e = ensure();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_vp_init_other<P, OS>), dim3(vp_blocks(g_.nslots)), dim3(VP_BLOCK), 0, stream_, g_, p_,
                       b_, (uint32_t)stationary_, other_states, other_slot_of_id);
    initialized_ = true;
    return hipGetLastError();
  }

  // execute(max_iters) (vertex_program_execute.hpp:11-52).
  hipError_t execute(uint32_t max_iters) {
    hipError_t e;
    if (!initialized_ && (e = initialize()) != hipSuccess) return e;
    const bool directed = g_.directed != 0;
    const bool opt = optimizable && directed && !gather_with_state && !apply_with_iter;
    const uint32_t nrows = g_.nreg + g_.nsink;
    cur_ = 0;
    if (max_iters == 1) {  // execute_single / execute_single_undirected (:251-345)
      clear(1);
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
      pass(g_.nreg, nrows, 1, iter, false);
    }
    return sync_error();
  }

  hipError_t states_by_id(S* host_out) {
    S* tmp = (S*)al_.alloc(al_.self, "vp.out", sizeof(S) * (size_t)g_.nslots);
    if (!tmp) return fail(hipErrorOutOfMemory);
    hipLaunchKernelGGL(k_vp_by_id<S>, dim3(vp_blocks(g_.nslots)), dim3(VP_BLOCK), 0, stream_, g_,
                       (const S*)b_.v, tmp);
    hipError_t e = hipMemcpyAsync(host_out, tmp, sizeof(S) * (size_t)g_.nslots, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    al_.release(al_.self, tmp);
    return e == hipSuccess ? hipSuccess : fail(e);
  }

  // reset() (vertex_program.hpp:51-63): the next execute() initializes again.
  void reset() { initialized_ = false; }

  void release() {
    if (!allocated_) return;
    (void)hipStreamSynchronize(stream_);
    al_.release(al_.self, b_.v);
    al_.release(al_.self, b_.vm);
    al_.release(al_.self, b_.x[0]);
    al_.release(al_.self, b_.x[1]);
    al_.release(al_.self, b_.act[0]);
    al_.release(al_.self, b_.act[1]);
    al_.release(al_.self, b_.ctr);
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
  bool allocated_ = false, initialized_ = false;
  uint32_t cur_ = 0;

  hipError_t fail(hipError_t e) {
    if (err.empty()) err = hipGetErrorString(e);
    return e;
  }

  hipError_t ensure() {
    if (allocated_) return hipSuccess;
    const size_t S_ = g_.nslots ? g_.nslots : 1;
    b_.v = (S*)al_.alloc(al_.self, "vp.v", sizeof(S) * S_);
    if (gather_with_state) b_.vm = (S*)al_.alloc(al_.self, "vp.vm", sizeof(S) * S_);
    b_.x[0] = (M*)al_.alloc(al_.self, "vp.x0", sizeof(M) * S_ + 16);
    b_.x[1] = (M*)al_.alloc(al_.self, "vp.x1", sizeof(M) * S_ + 16);
    b_.act[0] = (uint8_t*)al_.alloc(al_.self, "vp.a0", S_ + 16);
    b_.act[1] = (uint8_t*)al_.alloc(al_.self, "vp.a1", S_ + 16);
    b_.ctr = (uint32_t*)al_.alloc(al_.self, "vp.ctr", 64);
    allocated_ = true;
    if (!b_.v || (gather_with_state && !b_.vm) || !b_.x[0] || !b_.x[1] || !b_.act[0] || !b_.act[1] || !b_.ctr) {
      release();
      err = "device allocation failed";
      return hipErrorOutOfMemory;
    }
    return hipSuccess;
  }

  void clear(uint32_t which) {
    hipLaunchKernelGGL(k_vp_clear<P>, dim3(vp_blocks(g_.nreg)), dim3(VP_BLOCK), 0, stream_, g_, b_, which);
  }

  void pass(uint32_t r0, uint32_t r1, bool use_src, uint32_t iter, bool single) {
    if (r1 <= r0) return;
    PassArgs a{r0, r1, cur_, iter, (uint32_t)use_src, (uint32_t)single, (uint32_t)stationary_};
    hipLaunchKernelGGL(k_vp_pull<P>, dim3(vp_blocks(r1 - r0)), dim3(VP_BLOCK), 0, stream_, g_, p_, b_, a);
    if (g_.nlong)
      hipLaunchKernelGGL(k_vp_pull_long<P>, dim3(vp_blocks((uint64_t)g_.nlong * VP_WAVE)), dim3(VP_BLOCK), 0,
                         stream_, g_, p_, b_, a);
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

  // LA3-optimizable execution (vertex_program.h:140-148); cleared to force
  // execute_non_opt.
  bool optimizable = true;

  // initialize() (vertex_program.hpp:84-174).
  void initialize() { check(e_.initialize(), "initialize"); }

  // initialize(vids) (vertex_program.hpp:177-259): init only the listed
  // vertices; the others keep S() and send nothing.
  void initialize(const std::vector<uint32_t>& vids) {
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
    check(e_.initialize_other(other.device_states(), other.view().slot_of_id), "initialize(other)");
  }

  // execute(max_iters) (vertex_program_execute.hpp:11-52).
  void execute(uint32_t max_iters = 0) {
    e_.optimizable = optimizable;
    check(e_.execute(max_iters), "execute");
  }

  // reset() (vertex_program.hpp:51-63) and free() (:38-48).
  void reset() { e_.reset(); }
  void free() { e_.release(); }

  uint32_t iterations() const { return e_.iters_run; }

  // States by original vertex id (nslots = n + 1 entries).
  std::vector<S> states() {
    std::vector<S> out(view_.nslots);
    check(e_.states_by_id(out.data()), "states");
    return out;
  }
  const S* device_states() const { return e_.bufs().v; }
  const la3_device_view& view() const { return view_; }

  // reduce (vertex_program.hpp:558-604): map every vertex in the reference's
  // slot order and fold with reduce.
  template <class Value, class Mapper, class Reducer>
  Value reduce(Mapper map, Reducer red) {
    std::vector<S> st = states();
    std::vector<uint32_t> hid(view_.nslots);
    if (la3_graph_slots(g_, hid.data()) != LA3_OK) throw Error(LA3_EINVAL, "reduce: slots");
    Value acc = Value();
    for (uint32_t h : hid) {
      const uint32_t vid = la3_graph_unhash(g_, h);
      red(acc, map(vid, st[vid]));
    }
    return acc;
  }

 private:
  la3_ctx* ctx_;
  la3_graph* g_;
  la3_device_view view_;
  vp_detail::Engine<P> e_;

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
