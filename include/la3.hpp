// la3.hpp — C++ drop-in over the C ABI (include/la3.h).
//
// Keeps the reference's host-side names so that an LA3 app driver ports by
// changing includes and handing a Context to the Graph:
//   Graph<W>::load_directed / load_undirected / free / get_nvertices
//       (src/matrix/graph.h:38-109)
//   DegVertex, PrVertex, BfsVertex, SpVertex, CcVertex, GnVertex, CtVertex
//       (src/apps/graph_analytics/*.h) with initialize(), initialize(other),
//       execute(max_iters), reset(), free(), reduce / topk (active_only too),
//       reset_activity(), activate_all(), display() and a public `root`
//       where the reference has one
//       (src/vprogram/vertex_program.h:158-214).
// Programs with other operations are written against la3::VertexProgram<P>
// in la3_vertex_program.hpp (HIP device code).
// States are exposed per vertex with the reference's field names (degree,
// rank, parent, hops, distance, label, ntriangles). Errors throw la3::Error
// (the reference exits or asserts; a C++ caller can catch here).
#ifndef LA3_HPP
#define LA3_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "la3.h"

namespace la3 {

struct Error : std::runtime_error {
  la3_status status;
  Error(la3_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

enum Hashing { NONE = LA3_HASH_NONE, BUCKET = LA3_HASH_BUCKET };
struct Empty {};

class Context {
 public:
  explicit Context(int device = 0) {
    la3_status s = la3_ctx_create(device, &c_);
    if (s) throw Error(s, "la3_ctx_create failed");
  }
  ~Context() { la3_ctx_destroy(c_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  la3_ctx* get() const { return c_; }
  void check(la3_status s, const char* what) const {
    if (s) throw Error(s, std::string(what) + ": " + la3_last_error(c_));
  }

 private:
  la3_ctx* c_ = nullptr;
};

// Graph<W> at one rank (np = 1), resident on the device. W = uint32_t loads
// weighted triples, anything else loads pairs.
template <class W = Empty>
class Graph {
 public:
  explicit Graph(Context& ctx) : ctx_(&ctx) {}
  ~Graph() { free(); }
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;

  void load_directed(bool binary, std::string filepath, uint32_t nvertices = 0,
                     bool reverse_edges = false, bool remove_cycles = false,
                     Hashing hashing = BUCKET) {
    la3_mode m = reverse_edges ? (remove_cycles ? LA3_REVERSED_ACYCLIC : LA3_REVERSED)
                               : (remove_cycles ? LA3_ACYCLIC : LA3_DIRECTED);
    load(binary, filepath, nvertices, m, hashing);
  }
  void load_undirected(bool binary, std::string filepath, uint32_t nvertices = 0,
                       Hashing hashing = BUCKET) {
    load(binary, filepath, nvertices, LA3_UNDIRECTED, hashing);
  }
  void free() {
    if (g_) la3_graph_free(g_);
    g_ = nullptr;
  }

  uint32_t get_nvertices() const { return info_.n; }
  uint64_t get_nedges() const { return info_.nrecords; }
  bool is_directed() const { return info_.directed != 0; }
  const la3_graph_info& info() const { return info_; }
  la3_graph* handle() const { return g_; }
  Context& context() const { return *ctx_; }
  uint32_t unhash(uint32_t v) const { return la3_graph_unhash(g_, v); }
  // Slot order of the reference's vertex-state vector (for reduce order).
  std::vector<uint32_t> slot_ids() const {
    std::vector<uint32_t> hid(info_.nslots), ids(info_.nslots);
    ctx_->check(la3_graph_slots(g_, hid.data()), "graph_slots");
    for (uint32_t p = 0; p < info_.nslots; p++) ids[p] = la3_graph_unhash(g_, hid[p]);
    return ids;
  }

 private:
  void load(bool binary, const std::string& path, uint32_t n, la3_mode m, Hashing h) {
    if (!binary) throw Error(LA3_ENOTSUP, "text input is not implemented (as in graph.hpp:216-221)");
    free();
    ctx_->check(la3_graph_load_file(ctx_->get(), path.c_str(), n, weighted, m, (la3_hashing)h, &g_),
                "graph_load_file");
    ctx_->check(la3_graph_info_get(g_, &info_), "graph_info");
  }
  static constexpr int weighted = std::is_same<W, uint32_t>::value ? 1 : 0;
  Context* ctx_;
  la3_graph* g_ = nullptr;
  la3_graph_info info_{};
};

// Common part of the program wrappers: per-vertex states by original id,
// vertex activity, and reduce / topk / display in the reference's slot order
// (vertex_program.hpp:511-680).
template <class G, class S>
class ProgramBase {
 public:
  using State = S;
  explicit ProgramBase(const G* graph) : G_(graph) {}
  const G* get_graph() const { return G_; }
  const std::vector<S>& states() const { return states_; }
  const S& state(uint32_t vid) const { return states_.at(vid); }

  // reset() (vertex_program.hpp:51-63): the next execute() starts from
  // initialize() again — every execute() here does. free() (:38-48)
  // releases the states.
  void reset() {}
  void free() {
    std::vector<S>().swap(states_);
    std::vector<uint8_t>().swap(active_);
  }

  // Vertex activity by original id (vertex_program.hpp:549-555): apply sets
  // it for a regular vertex that was activated or is stationary and for a
  // sink that was activated (vertex_program_execute.hpp:693-713); executes
  // add to it, reset_activity() clears it, activate_all() sets every vertex.
  void reset_activity() { active_.assign(nslots(), 0); }
  void activate_all() { active_.assign(nslots(), 1); }
  const std::vector<uint8_t>& activity() const {
    if (!tracks_activity_) throw Error(LA3_ENOTSUP, "this program's vertex activity is not tracked");
    return active_;
  }

  // reduce (vertex_program.hpp:558-604) at one rank: every vertex (or every
  // active one) in slot order, folded into Value().
  template <class Value, class Mapper, class Reducer>
  Value reduce(Mapper map, Reducer red, bool active_only = false) const {
    have_states(active_only);
    Value acc = Value();
    for (uint32_t vid : G_->slot_ids())
      if (!active_only || active_[vid]) red(acc, map(vid, states_[vid]));
    return acc;
  }

  // topk (vertex_program.hpp:607-680) at one rank, step for step: (id,
  // map(id, state)) pairs in slot order (active vertices only if asked),
  // std::partial_sort of the first min(k, count) by cmp, resized to k (pairs
  // I(), V() appended when fewer), partial_sort of those k, resized to k.
  // Same input order and algorithm, so equal values come out as in the
  // reference. For large outputs left on the device, la3_result_topk selects
  // on the GPU.
  template <class I, class V, class Mapper, class Comparator>
  void topk(uint32_t k, std::vector<std::pair<I, V>>& out, Mapper map, Comparator cmp,
            bool active_only = false) const {
    have_states(active_only);
    std::vector<std::pair<I, V>> ivs;
    for (uint32_t vid : G_->slot_ids())
      if (!active_only || active_[vid]) ivs.push_back(std::pair<I, V>((I)vid, map(vid, states_[vid])));
    select_k(ivs, k, cmp);
    out = ivs;
  }

  // display<values_only>(nvertices) (vertex_program.hpp:511-546): in slot
  // order, every vertex whose original id j is at most min(nvertices, n), as
  // "j: {...}" in the reference state's to_string format (values_only: the
  // braces alone, then one empty line), one line each, to `os`.
  template <bool values_only = false>
  void display(uint32_t nvertices = 30, std::ostream& os = std::cout) const {
    have_states(false);
    const uint32_t nmax = std::min<uint32_t>(nvertices, G_->get_nvertices());
    for (uint32_t vid : G_->slot_ids()) {
      if (vid > nmax) continue;
      if (values_only) os << to_string(states_[vid]) << "\n";
      else os << vid << ": " << to_string(states_[vid]) << "\n";
    }
    if (values_only) os << "\n";
  }

  template <class IV, class Comparator>
  static void select_k(std::vector<IV>& ivs, uint32_t k, Comparator cmp) {
    std::partial_sort(ivs.begin(), ivs.begin() + std::min<size_t>(k, ivs.size()), ivs.end(), cmp);
    ivs.resize(k);
    std::partial_sort(ivs.begin(), ivs.begin() + std::min<size_t>(k, ivs.size()), ivs.end(), cmp);
    ivs.resize(k);
  }

 protected:
  const G* G_;
  std::vector<S> states_;
  std::vector<uint8_t> active_;    // by original id
  bool tracks_activity_ = true;
  uint32_t nslots() const { return G_->info().nslots; }
  Context& ctx() const { return G_->context(); }
  void have_states(bool active_only) const {
    if (states_.size() != nslots()) throw Error(LA3_EINVAL, "no vertex states: execute() first");
    if (active_only) {
      activity();
      if (active_.size() != states_.size()) throw Error(LA3_EINVAL, "no vertex activity");
    }
  }
  // Adds the vertices for which active(vid, state) holds to the activity.
  template <class F>
  void add_active(F active) {
    if (active_.size() != states_.size()) active_.assign(states_.size(), 0);
    for (uint32_t v = 0; v < states_.size(); v++)
      if (active(v, states_[v])) active_[v] = 1;
  }
};

struct DegState { uint32_t degree = 0; };
struct PrState : DegState { double rank = 0.15; };
struct BfsState { uint32_t parent = 0; uint8_t hops = 255; };
struct SpState { uint32_t distance = 0x7fffffffu; };
struct CcState { uint32_t label = 0; };
struct CtState { uint32_t ntriangles = 0; };

// The states' to_string formats (degree.h:20, pr.h:23-24, bfs.h:22-23,
// sssp.h:21-22, cc.h:18). CtState keeps the count only (its neighbour list
// is not kept), so TC programs have no display.
inline std::string to_string(const DegState& s) { return "{degree: " + std::to_string(s.degree) + "}"; }
inline std::string to_string(const PrState& s) {
  return "{rank: " + std::to_string(s.rank) + ", degree: " + std::to_string(s.degree) + "}";
}
inline std::string to_string(const BfsState& s) {
  return "{parent: " + std::to_string(s.parent) + ", hops: " + std::to_string(s.hops) + "}";
}
inline std::string to_string(const SpState& s) { return "{distance: " + std::to_string(s.distance) + "}"; }
inline std::string to_string(const CcState& s) { return "{label: " + std::to_string(s.label) + "}"; }

// DegVertex (degree.h:25-35): in-degree of the stored graph.
template <class W = Empty>
class DegVertex : public ProgramBase<Graph<W>, DegState> {
 public:
  DegVertex(const Graph<W>* G, bool /*stationary*/ = true) : ProgramBase<Graph<W>, DegState>(G) {}
  void initialize() {}
  // Every max_iters >= 1 gives the same degrees; 0 (until convergence)
  // never ends in the reference, whose apply always activates.
  void execute(uint32_t max_iters = 1) {
    if (max_iters == 0) throw Error(LA3_EINVAL, "DegVertex::execute(0) does not converge");
    std::vector<uint32_t> d(this->nslots());
    this->ctx().check(la3_degree(this->G_->handle(), d.data()), "degree");
    this->states_.assign(d.size(), DegState());
    for (size_t i = 0; i < d.size(); i++) this->states_[i].degree = d[i];
    // stationary, apply always activates: every vertex that received a
    // message, i.e. every vertex with an in-entry
    this->add_active([](uint32_t, const DegState& s) { return s.degree > 0; });
  }
};

// PrVertex (pr.h:28-46). initialize(vp_degree) is accepted for drop-in: the
// out-degrees the reference gets from DegVertex on the reversed load are the
// column lengths of this graph, which the device reads directly.
class PrVertex : public ProgramBase<Graph<Empty>, PrState> {
 public:
  PrVertex(const Graph<Empty>* G, bool /*stationary*/ = true) : ProgramBase(G) {}
  template <class Other>
  void initialize(const Other&) {}
  void initialize() {}
  void execute(uint32_t max_iters = 0) {
    std::vector<double> r(nslots());
    std::vector<uint32_t> d(nslots());
    ctx().check(la3_pagerank(G_->handle(), max_iters, r.data(), d.data(), &iters_), "pagerank");
    states_.assign(r.size(), PrState());
    for (size_t i = 0; i < r.size(); i++) { states_[i].rank = r[i]; states_[i].degree = d[i]; }
    // stationary: every regular vertex (each receives in every iteration);
    // a sink when its one apply moved the rank from 0.15 by more than the
    // tolerance (pr.h:38-43)
    std::vector<uint8_t> reg(states_.size(), 0), sink(states_.size(), 0);
    const std::vector<uint32_t> ids = G_->slot_ids();
    const uint32_t nreg = G_->info().nregular, nsink = G_->info().nsink;
    for (uint32_t p = 0; p < nreg + nsink && p < ids.size(); p++) (p < nreg ? reg : sink)[ids[p]] = 1;
    add_active([&](uint32_t v, const PrState& st) {
      return reg[v] || (sink[v] && std::fabs(st.rank - 0.15) > 1e-5);
    });
  }
  uint32_t iterations() const { return iters_; }

 private:
  uint32_t iters_ = 0;
};

// BfsVertex (bfs.h:27-49).
class BfsVertex : public ProgramBase<Graph<Empty>, BfsState> {
 public:
  BfsVertex(const Graph<Empty>* G, bool = false) : ProgramBase(G) {}
  uint32_t root = 0;
  void initialize() {}
  void execute(uint32_t max_iters = 0) {
    std::vector<uint8_t> h(nslots());
    std::vector<uint32_t> p(nslots());
    ctx().check(la3_bfs(G_->handle(), root, max_iters, h.data(), p.data(), &iters_), "bfs");
    states_.assign(h.size(), BfsState());
    for (size_t i = 0; i < h.size(); i++) { states_[i].hops = h[i]; states_[i].parent = p[i]; }
    // apply activates a vertex once, when it is reached (bfs.h:40-47); the
    // root was set by init
    add_active([this](uint32_t v, const BfsState& st) { return st.hops != 255 && v != root; });
  }
  uint32_t iterations() const { return iters_; }

 private:
  uint32_t iters_ = 0;
};

// SpVertex (sssp.h:26-46).
class SpVertex : public ProgramBase<Graph<uint32_t>, SpState> {
 public:
  SpVertex(const Graph<uint32_t>* G, bool = false) : ProgramBase(G) {}
  uint32_t root = 0;
  void initialize() {}
  void execute(uint32_t max_iters = 0) {
    std::vector<uint32_t> d(nslots());
    ctx().check(la3_sssp(G_->handle(), root, max_iters, d.data(), &iters_), "sssp");
    states_.assign(d.size(), SpState());
    for (size_t i = 0; i < d.size(); i++) states_[i].distance = d[i];
    // apply activates a vertex whenever its distance drops (sssp.h:38-44):
    // every reached vertex but the root, which init set to 0
    add_active([this](uint32_t v, const SpState& st) { return st.distance != 0x7fffffffu && v != root; });
  }
  uint32_t iterations() const { return iters_; }

 private:
  uint32_t iters_ = 0;
};

// CcVertex (cc.h:22-38).
class CcVertex : public ProgramBase<Graph<Empty>, CcState> {
 public:
  CcVertex(const Graph<Empty>* G, bool = false) : ProgramBase(G) {}
  void initialize() {}
  void execute(uint32_t max_iters = 0) {
    std::vector<uint32_t> l(nslots());
    ctx().check(la3_cc(G_->handle(), max_iters, l.data(), &iters_), "cc");
    states_.assign(l.size(), CcState());
    for (size_t i = 0; i < l.size(); i++) states_[i].label = l[i];
    // apply activates a vertex whenever its label drops (cc.h:32-37); labels
    // start at the vertex id and only drop
    add_active([](uint32_t v, const CcState& st) { return st.label != v; });
  }
  uint32_t iterations() const { return iters_; }

 private:
  uint32_t iters_ = 0;
};

// GnVertex + CtVertex (tc.h:47-99): the pair runs as one masked SpGEMM on the
// acyclic load GnVertex was constructed with; CtVertex(GR).initialize(gn)
// takes that graph over, execute(1) fills ntriangles (middle-id counts).
class GnVertex : public ProgramBase<Graph<Empty>, CtState> {
 public:
  GnVertex(const Graph<Empty>* G, bool = true) : ProgramBase(G) { tracks_activity_ = false; }
  void initialize() {}
  void execute(uint32_t = 1) {}
};

class CtVertex : public ProgramBase<Graph<Empty>, CtState> {
 public:
  CtVertex(const Graph<Empty>* GR, bool = true) : ProgramBase(GR) { tracks_activity_ = false; }
  void initialize(const GnVertex& gn) { acyclic_ = gn.get_graph(); }
  void execute(uint32_t = 1) {
    if (!acyclic_) throw Error(LA3_EINVAL, "CtVertex::initialize(GnVertex) was not called");
    std::vector<uint32_t> per(acyclic_->info().nslots);
    acyclic_->context().check(la3_tc(acyclic_->handle(), &total_, per.data()), "tc");
    states_.assign(per.size(), CtState());
    for (size_t i = 0; i < per.size(); i++) states_[i].ntriangles = per[i];
  }
  uint64_t total() const { return total_; }

 private:
  const Graph<Empty>* acyclic_ = nullptr;
  uint64_t total_ = 0;
};

// ------------------------------------------------------------- batched ---
// B queries of one program as one batched run (la3_*_batch). The state of
// vertex vid holds the B per-query values. The batched programs have no
// counterpart program in the reference; btopk keeps VertexProgram::btopk's
// signature and algorithm (vertex_program.hpp:683-767) at one rank.
template <uint32_t B, class V>
struct BatchState {
  V v[B];
};

template <class G, uint32_t B, class V>
class BatchProgramBase : public ProgramBase<G, BatchState<B, V>> {
 public:
  explicit BatchProgramBase(const G* graph) : ProgramBase<G, BatchState<B, V>>(graph) {
    this->tracks_activity_ = false;
  }
  uint32_t roots[B] = {};  // query q starts at roots[q] (its seed for PPR)

  // Per query b < BATCH_SIZE: (id, map(id, state)[b]) for every id in slot
  // order, std::partial_sort by cmp, the first k, into topk[b].
  // Batched programs have no counterpart in the reference and track no
  // vertex activity: active_only is refused.
  template <uint32_t BATCH_SIZE, class I, class VV, class Mapper, class Comparator>
  void btopk(uint32_t k, std::vector<std::pair<I, VV>>* topk, Mapper map, Comparator cmp,
             bool active_only = false) const {
    static_assert(BATCH_SIZE <= B, "btopk: BATCH_SIZE exceeds the batch");
    if (active_only) throw Error(LA3_ENOTSUP, "btopk: batched programs track no vertex activity");
    std::vector<std::pair<I, VV>> ivs[BATCH_SIZE];
    for (uint32_t vid : this->G_->slot_ids()) {
      const VV* vals = map(vid, this->states_[vid]);
      for (uint32_t b = 0; b < BATCH_SIZE; b++) ivs[b].push_back(std::pair<I, VV>((I)vid, vals[b]));
    }
    for (uint32_t b = 0; b < BATCH_SIZE; b++) {
      ProgramBase<G, BatchState<B, V>>::select_k(ivs[b], k, cmp);
      topk[b] = ivs[b];
    }
  }
  uint32_t iterations() const { return iters_; }

 protected:
  uint32_t iters_ = 0;
  // rows[q * nslots + id], query-major as the batch calls return them
  void fill(const std::vector<V>& rows) {
    const uint32_t S = this->nslots();
    this->states_.assign(S, BatchState<B, V>());
    for (uint32_t q = 0; q < B; q++)
      for (uint32_t i = 0; i < S; i++) this->states_[i].v[q] = rows[(size_t)q * S + i];
  }
};

// B BFS queries (levels, 255 = unreached), one k-bit traversal.
template <uint32_t B>
class BfsBatchVertex : public BatchProgramBase<Graph<Empty>, B, uint8_t> {
 public:
  BfsBatchVertex(const Graph<Empty>* G, bool = false) : BatchProgramBase<Graph<Empty>, B, uint8_t>(G) {}
  void initialize() {}
  void execute(uint32_t max_iters = 0) {
    std::vector<uint8_t> h((size_t)B * this->nslots());
    this->ctx().check(la3_bfs_batch(this->G_->handle(), this->roots, B, max_iters, h.data(), &this->iters_),
                      "bfs_batch");
    this->fill(h);
  }
};

// B SSSP queries (distances, 2^31 - 1 = unreachable), one k-wide traversal.
template <uint32_t B>
class SpBatchVertex : public BatchProgramBase<Graph<uint32_t>, B, uint32_t> {
 public:
  SpBatchVertex(const Graph<uint32_t>* G, bool = false) : BatchProgramBase<Graph<uint32_t>, B, uint32_t>(G) {}
  void initialize() {}
  void execute(uint32_t max_iters = 0) {
    std::vector<uint32_t> d((size_t)B * this->nslots());
    this->ctx().check(la3_sssp_batch(this->G_->handle(), this->roots, B, max_iters, d.data(), &this->iters_),
                      "sssp_batch");
    this->fill(d);
  }
};

// B personalized-PageRank queries (FP32 with compensation), max_iters
// iterations of the PR step with teleport alpha at each query's seed.
template <uint32_t B>
class PprBatchVertex : public BatchProgramBase<Graph<Empty>, B, float> {
 public:
  PprBatchVertex(const Graph<Empty>* G, bool = true) : BatchProgramBase<Graph<Empty>, B, float>(G) {}
  double alpha = 0.15;
  void initialize() {}
  void execute(uint32_t max_iters) {
    std::vector<float> r((size_t)B * this->nslots());
    this->ctx().check(la3_ppr_batch(this->G_->handle(), this->roots, B, max_iters, alpha, r.data(), nullptr, nullptr),
                      "ppr_batch");
    this->iters_ = max_iters;
    this->fill(r);
  }
};

}  // namespace la3

#endif  // LA3_HPP
