// LA3 vertex-program oracle: a CPU restatement of the reference's execution path
// at one rank (np=1, one tile).
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product (la3_amd/, include/) links,
// loads or calls this file. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, and only as the checker.
//
// What it restates (reference paths are relative to the LA3 tree):
//   ingest            src/matrix/graph.hpp:38-214   (self-loop drop, transpose,
//                                                     reverse, acyclic orientation,
//                                                     bucket hashing, mirroring)
//   hashing           src/matrix/hashers.h:39-67   (SimpleBucketHasher, nranks=1)
//   tiling            src/matrix/matrix2d.hpp:6-41   (np=1: one (n+1)x(n+1) tile)
//   vertex classes    src/matrix/processed_matrix2d.hpp:182-229
//   locators          src/utils/locator.h:84-155 and
//                     src/matrix/processed_matrix2d_details.hpp:156-216
//   DCSC build        src/matrix/csc_matrix2d.hpp:20-88, src/utils/csc.h:76-130
//                     (first-inserted-wins dedup, regular/sink split, entries
//                      sorted by global_idx)
//   initialize        src/vprogram/vertex_program.hpp:84-174, 417-508
//   execute variants  src/vprogram/vertex_program_execute.hpp:11-345
//   SpMV              src/vprogram/vertex_program_execute.hpp:564-603
//   apply/scatter     src/vprogram/vertex_program_execute.hpp:606-724
//   containers        src/structures/random_access_array.h:76-82 (pop resets to
//                     Value()), src/structures/streaming_array.h:92-108 (pushes
//                     in increasing index)
//   programs          src/apps/graph_analytics/{degree,pr,bfs,sssp,cc,tc}.h
//
// PPR is not in the reference. It is defined here as PR with a per-query
// teleport vector (alpha at the seed, 0 elsewhere), same iteration structure.
//
// Pinning (every function here is checked against answers that do not come
// from this file):
//   hashing            hand-worked hashed ids for n = 8, 300, 5000, 2^16+3 and
//                      the bucket construction over every id, inverse and
//                      identity above max_range (tests/test_oracle_layout.py)
//   classes, slot order, both DCSCs (colptrs, colidxs, global_idx, idx,
//                      first-record weights)   brute-force recomputation from
//                      the records for all five load modes, bucket and no
//                      hashing, ids 0 and >= max_range (test_oracle_layout.py)
//   programs           g1 known answers (tests/golden/g1_kat.json); scipy BFS,
//                      Dijkstra and connected components, brute-force and trace
//                      triangle counts, degree counts, a dense PR/PPR iteration
//                      with the post-loop sink step, BFS parents recomputed
//                      from classes and the bucket construction
//                      (tests/test_oracle.py)
//   execute(max_iters) hop-bounded BFS levels and Bellman-Ford distances,
//                      m-round min-label propagation and its round count,
//                      single-iteration PR (tests/test_oracle_bounded.py)
// Not pinned: nothing on the np = 1 path; the reference's own tests hold no
// vectors (it has none), so these closed forms are the pins.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <parallel/algorithm>
#endif

namespace {

// ---------------------------------------------------------------- hashing ---
// hashers.h:39-67. nbuckets = nranks = 1 here.
struct Hasher {
  bool bucket;
  int64_t nparts = 128, height = 0, max_range = 0;
  Hasher(int64_t max_domain, bool bucket_) : bucket(bucket_) {
    height = max_domain / nparts;
    max_range = height * nparts;
  }
  int64_t hash(int64_t v) const {
    if (!bucket || v >= max_range) return v;
    int64_t col = (uint32_t)v % nparts;
    int64_t row = v / nparts;
    return row + col * height;
  }
  int64_t unhash(int64_t v) const {
    if (!bucket || v >= max_range) return v;
    int64_t col = v / height;
    int64_t row = v % height;
    return col + row * nparts;
  }
};

struct Entry {
  uint32_t global_idx;  // position among regular rows or among sink rows
  uint32_t idx;         // hashed row id
  uint32_t w;
};

struct OCSC {
  uint32_t ncols = 0;
  std::vector<uint32_t> colptrs, colidxs;
  std::vector<Entry> entries;
};

struct Rec {
  uint32_t row, col, w;
};

enum Mode { DIRECTED = 0, REVERSED = 1, UNDIRECTED = 2, ACYCLIC = 3, REVERSED_ACYCLIC = 4 };

struct OGraph {
  uint32_t n = 0, nslots = 0;
  bool directed = true, weighted = false;
  Hasher h{0, true};
  uint32_t nreg = 0, nsink = 0, nsrc = 0;
  std::vector<uint32_t> dash;    // hid -> dashboard position [reg|sink|src|rest]
  std::vector<uint32_t> hid_of;  // dashboard position -> hid
  std::vector<uint32_t> rowloc;  // hid -> row-group position [reg|sink|rest]
  std::vector<uint32_t> colloc;  // hid -> col-group position [reg|src|rest]
  OCSC reg, sink;
};

// csc.h:76-130 for one of the two row sets of the tile.
void build_csc(OCSC& c, uint32_t ncols, const std::vector<Rec>& triples, const OGraph& G) {
  c.ncols = ncols;
  c.colptrs.assign(ncols + 1, 0);
  c.colidxs.assign(ncols + 1, 0);
  c.entries.resize(triples.size());
  for (auto& t : triples) {
    ++c.colptrs[G.colloc[t.col]];
    c.colidxs[G.colloc[t.col]] = t.col;
  }
  std::partial_sum(c.colptrs.begin(), c.colptrs.end(), c.colptrs.begin());
  for (auto& t : triples) {
    uint32_t loc = G.rowloc[t.row];
    uint32_t off = loc < G.nreg ? 0 : G.nreg;
    uint32_t col = G.colloc[t.col];
    c.entries[--c.colptrs[col]] = Entry{loc - off, t.row, t.w};
  }
  // Each column's entries by global_idx (unique within a column), one
  // column per task.
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint32_t i = 0; i < ncols; i++)
    std::sort(c.entries.begin() + c.colptrs[i], c.entries.begin() + c.colptrs[i + 1],
              [](const Entry& a, const Entry& b) { return a.global_idx < b.global_idx; });
}

// graph.hpp:38-214 (np=1), processed_matrix2d*.hpp, csc_matrix2d.hpp.
// Returns nullptr if an id exceeds n (the reference writes out of bounds there).
OGraph* build_graph(const uint32_t* data, uint64_t nrec, uint32_t n, bool weighted, int mode,
                    bool bucket) {
  auto* G = new OGraph();
  G->n = n;
  G->nslots = n + 1;  // matrix2d.hpp:8 tile_height = n/1 + 1
  G->weighted = weighted;
  G->h = Hasher(n, bucket);
  const bool directed = mode != UNDIRECTED;
  const bool reverse = mode == REVERSED || mode == REVERSED_ACYCLIC;
  const bool acyclic = mode == ACYCLIC || mode == REVERSED_ACYCLIC;
  G->directed = directed;
  const uint32_t rw = weighted ? 3 : 2;

  std::vector<Rec> tri;
  tri.reserve(directed ? nrec : 2 * nrec);
  for (uint64_t k = 0; k < nrec; k++) {
    Rec t{data[k * rw], data[k * rw + 1], weighted ? data[k * rw + 2] : 0u};
    if (t.row > n || t.col > n) { delete G; return nullptr; }
    if (t.row == t.col) continue;                      // graph.hpp:165-167
    if (directed && !reverse) std::swap(t.row, t.col);  // graph.hpp:171-172
    if (acyclic) {                                      // graph.hpp:174-179
      if ((!reverse && t.col > t.row) || (reverse && t.col < t.row)) std::swap(t.row, t.col);
    }
    t.row = (uint32_t)G->h.hash(t.row);
    t.col = (uint32_t)G->h.hash(t.col);
    tri.push_back(t);
    if (!directed) { std::swap(t.row, t.col); tri.push_back(t); }
  }

  // Classes over the stored matrix (processed_matrix2d.hpp:182-229).
  const uint32_t S = G->nslots;
  std::vector<uint8_t> rows(S, 0), cols(S, 0);
  for (auto& t : tri) { rows[t.row] = 1; cols[t.col] = 1; }
  G->dash.assign(S, 0);
  G->hid_of.assign(S, 0);
  G->rowloc.assign(S, 0);
  G->colloc.assign(S, 0);
  uint32_t pos = 0;
  for (uint32_t i = 0; i < S; i++) if (rows[i] && cols[i]) { G->dash[i] = pos; G->hid_of[pos++] = i; }
  G->nreg = pos;
  for (uint32_t i = 0; i < S; i++) if (rows[i] && !cols[i]) { G->dash[i] = pos; G->hid_of[pos++] = i; }
  G->nsink = pos - G->nreg;
  for (uint32_t i = 0; i < S; i++) if (!rows[i] && cols[i]) { G->dash[i] = pos; G->hid_of[pos++] = i; }
  G->nsrc = pos - G->nreg - G->nsink;
  for (uint32_t i = 0; i < S; i++) if (!rows[i] && !cols[i]) { G->dash[i] = pos; G->hid_of[pos++] = i; }
  // Row-group locator [regular | sink | rest]; at np=1 the global locator is equal.
  pos = 0;
  for (uint32_t i = 0; i < S; i++) if (rows[i] && cols[i]) G->rowloc[i] = pos++;
  for (uint32_t i = 0; i < S; i++) if (rows[i] && !cols[i]) G->rowloc[i] = pos++;
  for (uint32_t i = 0; i < S; i++) if (!rows[i]) G->rowloc[i] = pos++;
  // Col-group locator [regular | source | rest].
  pos = 0;
  for (uint32_t i = 0; i < S; i++) if (rows[i] && cols[i]) G->colloc[i] = pos++;
  for (uint32_t i = 0; i < S; i++) if (!rows[i] && cols[i]) G->colloc[i] = pos++;
  const uint32_t ncols = pos;
  for (uint32_t i = 0; i < S; i++) if (!cols[i]) G->colloc[i] = pos++;

  // Dedup with first-inserted-wins (unordered_set::insert keeps the first):
  // ordering by (row, col, insertion index) puts the first occurrence of each
  // key first. The pairs are unique, so any sort gives this one order.
  std::vector<std::pair<uint64_t, uint64_t>> order(tri.size());
  for (uint64_t k = 0; k < tri.size(); k++) order[k] = {((uint64_t)tri[k].row << 32) | tri[k].col, k};
#ifdef _OPENMP
  __gnu_parallel::sort(order.begin(), order.end());
#else
  std::sort(order.begin(), order.end());
#endif
  std::vector<Rec> regset, sinkset;
  for (uint64_t k = 0; k < order.size(); k++) {
    if (k > 0 && order[k - 1].first == order[k].first) continue;
    const Rec& t = tri[order[k].second];
    if (G->rowloc[t.row] < G->nreg) regset.push_back(t); else sinkset.push_back(t);
  }
  build_csc(G->reg, ncols, regset, *G);
  build_csc(G->sink, ncols, sinkset, *G);
  return G;
}

// ------------------------------------------------------------- programs ---
struct Empty {};

struct EdgeV {
  uint32_t src, dst, w;
};

constexpr double PR_ALPHA = 0.15;  // pr.h:16
constexpr double PR_TOL = 1e-5;    // pr.h:17

// degree.h:25-35
struct DegProg {
  using M = Empty; using A = uint32_t;
  struct S { uint32_t degree = 0; };
  static constexpr bool stationary = true, gather_state = false, apply_iter = false;
  A identity() const { return 0; }
  bool init(uint32_t, S&) { return stationary; }
  M scatter(const S&) { return M(); }
  A gather(const EdgeV&, const M&) { return 1; }
  void combine(const A& y1, A& y2) { y2 += y1; }
  bool apply(const A& y, S& s) { s.degree = y; return true; }
};

// pr.h:20-46
struct PrProg {
  using M = double; using A = double;
  struct S { uint32_t degree = 0; double rank = PR_ALPHA; };
  static constexpr bool stationary = true, gather_state = false, apply_iter = false;
  A identity() const { return 0.0; }
  bool init(uint32_t, S&) { return stationary; }
  bool init_other(uint32_t, const DegProg::S& o, S& s) { s.degree = o.degree; return true; }
  M scatter(const S& s) { return s.rank / s.degree; }
  A gather(const EdgeV&, const M& m) { return m; }
  void combine(const A& y1, A& y2) { y2 += y1; }
  bool apply(const A& y, S& s) {
    double tmp = s.rank;
    s.rank = PR_ALPHA + (1.0 - PR_ALPHA) * y;
    return std::fabs(s.rank - tmp) > PR_TOL;
  }
};

// Personalized PageRank (not in the reference): PR with teleport alpha at the
// seed only; initial rank = teleport. Same structure as pr.h:34-45.
struct PprProg {
  using M = double; using A = double;
  struct S { uint32_t degree = 0; double rank = 0.0; double tele = 0.0; };
  static constexpr bool stationary = true, gather_state = false, apply_iter = false;
  uint32_t seed = 0;
  double alpha = PR_ALPHA;
  A identity() const { return 0.0; }
  bool init(uint32_t, S&) { return stationary; }
  bool init_other(uint32_t vid, const DegProg::S& o, S& s) {
    s.degree = o.degree;
    s.tele = (vid == seed) ? alpha : 0.0;
    s.rank = s.tele;
    return true;
  }
  M scatter(const S& s) { return s.rank / s.degree; }
  A gather(const EdgeV&, const M& m) { return m; }
  void combine(const A& y1, A& y2) { y2 += y1; }
  bool apply(const A& y, S& s) {
    double tmp = s.rank;
    s.rank = s.tele + (1.0 - alpha) * y;
    return std::fabs(s.rank - tmp) > PR_TOL;
  }
};

// bfs.h:11-49
constexpr uint8_t BFS_INF = 255;
struct BfsProg {
  using M = Empty; using A = uint32_t;
  struct S { uint32_t parent = 0; uint8_t hops = BFS_INF; };
  static constexpr bool stationary = false, gather_state = false, apply_iter = true;
  uint32_t root = 0;
  A identity() const { return 0; }
  bool init(uint32_t vid, S& s) { if (vid == root) { s.hops = 0; return true; } return false; }
  M scatter(const S&) { return M(); }
  A gather(const EdgeV& e, const M&) { return e.src; }
  void combine(const A& y1, A& y2) { y2 = y1; }
  bool apply(const A& y, S& s, uint32_t iter) {
    if (s.hops != BFS_INF) return false;
    s.hops = (uint8_t)(iter + 1);
    s.parent = y;
    return true;
  }
};

// sssp.h:11-46 (IntegerWrapper<INF> default-constructs to INF, common.h:106-127)
constexpr uint32_t SSSP_INF = UINT32_MAX / 2;
struct SsspProg {
  using M = uint32_t; using A = uint32_t;
  struct S { uint32_t dist = SSSP_INF; };
  static constexpr bool stationary = false, gather_state = false, apply_iter = false;
  uint32_t root = 0;
  A identity() const { return SSSP_INF; }
  bool init(uint32_t vid, S& s) { if (vid == root) { s.dist = 0; return true; } return false; }
  M scatter(const S& s) { return s.dist; }
  A gather(const EdgeV& e, const M& m) { return m + e.w; }
  void combine(const A& y1, A& y2) { y2 = std::min(y1, y2); }
  bool apply(const A& y, S& s) { uint32_t t = s.dist; s.dist = std::min(s.dist, y); return t != s.dist; }
};

// cc.h:11-38
struct CcProg {
  using M = uint32_t; using A = uint32_t;
  struct S { uint32_t label = 0; };
  static constexpr bool stationary = false, gather_state = false, apply_iter = false;
  A identity() const { return 0; }
  bool init(uint32_t vid, S& s) { s.label = vid; return true; }
  M scatter(const S& s) { return s.label; }
  A gather(const EdgeV&, const M& m) { return m; }
  void combine(const A& y1, A& y2) { y2 = (y2 == 0) ? y1 : std::min(y1, y2); }
  bool apply(const A& y, S& s) { uint32_t t = s.label; s.label = std::min(s.label, y); return t != s.label; }
};

// tc.h:47-65 (GnVertex)
struct GnProg {
  using M = Empty; using A = std::vector<uint32_t>;
  struct S { std::vector<uint32_t> neighbors; };
  static constexpr bool stationary = true, gather_state = false, apply_iter = false;
  A identity() const { return A(); }
  bool init(uint32_t, S&) { return stationary; }
  M scatter(const S&) { return M(); }
  A gather(const EdgeV& e, const M&) { return A{e.src}; }
  void combine(const A& y1, A& y2) { y2.insert(y2.end(), y1.begin(), y1.end()); }
  bool apply(const A& y, S& s) {
    if (!y.empty()) { s.neighbors = y; std::sort(s.neighbors.begin(), s.neighbors.end()); }
    return false;
  }
};

// tc.h:68-99 (CtVertex); gather reads the row vertex's state (mirrors).
struct CtProg {
  using M = std::vector<uint32_t>; using A = uint32_t;
  struct S { std::vector<uint32_t> list; uint32_t ntriangles = 0; };
  static constexpr bool stationary = true, gather_state = true, apply_iter = false;
  A identity() const { return 0; }
  bool init(uint32_t, S&) { return stationary; }
  bool init_other(uint32_t, const GnProg::S& o, S& s) {
    if (o.neighbors.empty()) return false;
    s.list = o.neighbors;
    return true;
  }
  M scatter(const S& s) { return s.list; }
  A gather(const EdgeV&, const M& msg, const S& s) {
    A t = 0;
    size_t i = 0, j = 0;
    while (i != s.list.size() && j != msg.size()) {
      if (s.list[i] == msg[j]) { t++; ++i; ++j; }
      else if (s.list[i] < msg[j]) ++i;
      else ++j;
    }
    return t;
  }
  void combine(const A& y1, A& y2) { y2 += y1; }
  bool apply(const A& y, S& s) { s.ntriangles = y; return false; }
};

// Programs written only against the vertex-program interface, with no app
// of the reference behind them; the same operations are compiled against
// include/la3_vertex_program.hpp in tests/vp_programs.hip. They exercise the
// engine's paths: weights and execute_single (WDeg), execute_ with the sink
// step and execute_non_opt (MaxLabel), apply with the iteration and an
// order-dependent combine (Level), floating-point sums and
// initialize(other) (PrAlpha), and a state-reading gather of a stationary
// program (CountGreater: mirrors hold the initial states) and of a
// non-stationary one (CountNonzero: mirrors hold S()).
struct WDegProg {
  using M = Empty; using A = uint32_t;
  struct S { uint32_t wdeg = 0; };
  static constexpr bool stationary = true, gather_state = false, apply_iter = false;
  A identity() const { return 0; }
  bool init(uint32_t, S&) { return stationary; }
  M scatter(const S&) { return M(); }
  A gather(const EdgeV& e, const M&) { return e.w; }
  void combine(const A& y1, A& y2) { y2 += y1; }
  bool apply(const A& y, S& s) { s.wdeg = y; return true; }
};

struct MaxLabelProg {
  using M = uint32_t; using A = uint32_t;
  struct S { uint32_t label = 0; };
  static constexpr bool stationary = false, gather_state = false, apply_iter = false;
  A identity() const { return 0; }
  bool init(uint32_t vid, S& s) { s.label = vid; return true; }
  M scatter(const S& s) { return s.label; }
  A gather(const EdgeV&, const M& m) { return m; }
  void combine(const A& y1, A& y2) { y2 = std::max(y1, y2); }
  bool apply(const A& y, S& s) { uint32_t t = s.label; s.label = std::max(s.label, y); return t != s.label; }
};

struct LevelProg {
  using M = Empty; using A = uint32_t;
  struct S { uint32_t level = 0xffffffffu; uint32_t via = 0; };
  static constexpr bool stationary = false, gather_state = false, apply_iter = true;
  uint32_t root = 0;
  A identity() const { return 0; }
  bool init(uint32_t vid, S& s) { if (vid == root) { s.level = 0; return true; } return false; }
  M scatter(const S&) { return M(); }
  A gather(const EdgeV& e, const M&) { return e.src; }
  void combine(const A& y1, A& y2) { y2 = y1; }
  bool apply(const A& y, S& s, uint32_t iter) {
    if (s.level != 0xffffffffu) return false;
    s.level = iter + 1;
    s.via = y;
    return true;
  }
};

struct PrAlphaProg {
  using M = double; using A = double;
  struct S { uint32_t degree = 0; double rank = 0.2; };
  static constexpr bool stationary = true, gather_state = false, apply_iter = false;
  A identity() const { return 0.0; }
  bool init(uint32_t, S&) { return stationary; }
  bool init_other(uint32_t, const DegProg::S& o, S& s) { s.degree = o.degree; return true; }
  M scatter(const S& s) { return s.rank / s.degree; }
  A gather(const EdgeV&, const M& m) { return m; }
  void combine(const A& y1, A& y2) { y2 += y1; }
  bool apply(const A& y, S& s) {
    double t = s.rank;
    s.rank = 0.2 + 0.8 * y;
    return std::fabs(s.rank - t) > 1e-9;
  }
};

struct CountGreaterProg {
  using M = uint32_t; using A = uint32_t;
  struct S { uint32_t val = 0; uint32_t cnt = 0; };
  static constexpr bool stationary = true, gather_state = true, apply_iter = false;
  A identity() const { return 0; }
  bool init(uint32_t vid, S& s) { s.val = vid % 7; return true; }
  M scatter(const S& s) { return s.val; }
  A gather(const EdgeV&, const M& m, const S& s) { return m > s.val ? 1u : 0u; }
  void combine(const A& y1, A& y2) { y2 += y1; }
  bool apply(const A& y, S& s) { s.cnt += y; s.val = y % 7; return false; }
};

// Not stationary, gather reads the row's state: the mirrors keep S() (val
// 0), so a row counts the messages above 0, not above its own value.
struct CountNonzeroProg {
  using M = uint32_t; using A = uint32_t;
  struct S { uint32_t val = 0; uint32_t cnt = 0; };
  static constexpr bool stationary = false, gather_state = true, apply_iter = false;
  A identity() const { return 0; }
  bool init(uint32_t vid, S& s) { s.val = vid % 7; return vid % 3 == 0; }
  M scatter(const S& s) { return s.val; }
  A gather(const EdgeV&, const M& m, const S& s) { return m > s.val ? 1u : 0u; }
  void combine(const A& y1, A& y2) { y2 += y1; }
  bool apply(const A& y, S& s) { s.cnt += y; return y > 0; }
};

// Stationary, gather reads the row's state, and init adds to the state it
// finds: a second initialize (after reset) changes the states but not the
// mirrors, which keep the first initialize's values.
struct AccGreaterProg {
  using M = uint32_t; using A = uint32_t;
  struct S { uint32_t val = 0; uint32_t cnt = 0; };
  static constexpr bool stationary = true, gather_state = true, apply_iter = false;
  A identity() const { return 0; }
  bool init(uint32_t vid, S& s) { s.val += vid % 5 + 1; return true; }
  M scatter(const S& s) { return s.val; }
  A gather(const EdgeV&, const M& m, const S& s) { return m > s.val ? 1u : 0u; }
  void combine(const A& y1, A& y2) { y2 += y1; }
  bool apply(const A& y, S& s) { s.cnt += y; s.val = y % 7; return false; }
};

// ---------------------------------------------------------------- engine ---
template <class P>
struct Engine {
  using M = typename P::M; using A = typename P::A; using S = typename P::S;
  P& p;
  const OGraph& G;
  std::vector<S> v;                                  // dashboard positions
  std::vector<S> vm;  // mirrors read by a state-reading gather (vertex_program.hpp:159-173)
  std::vector<std::pair<uint32_t, M>> xreg, xsrc;    // StreamingArray (pushes ascending)
  std::vector<A> yreg, ysink;
  std::vector<uint8_t> treg, tsink;                  // activity bitvectors
  // The state vector's activity set (random_access_array.h:14), by
  // position: pushed by apply (vertex_program_execute.hpp:693-713), never
  // cleared by execute or reset, cleared by reset_activity() and after the
  // first mirroring (vertex_program.hpp:166-172).
  std::vector<uint8_t> vact;
  uint32_t iters_run = 0;

  Engine(P& p_, const OGraph& G_) : p(p_), G(G_), v(G_.nslots), vact(G_.nslots, 0) {
    yreg.assign(G.nreg + 1, p.identity());
    ysink.assign(G.nsink + 1, p.identity());
    treg.assign(G.nreg + 1, 0);
    tsink.assign(G.nsink + 1, 0);
  }

  uint32_t vid_of(uint32_t pos) const { return (uint32_t)G.h.unhash(G.hid_of[pos]); }

  // reset_activity() / activate_all() (vertex_program.hpp:549-555).
  void reset_activity() { vact.assign(G.nslots, 0); }
  void activate_all() { vact.assign(G.nslots, 1); }

  // Mirrors are filled once, at the first initialize (`not
  // v->mirrors_allocated`, vertex_program.hpp:159-173): a stationary program
  // first activates every vertex; the active states are broadcast (init
  // itself activates none, vertex_program.hpp:103-104, so a non-stationary
  // program mirrors only what an earlier execute or activate_all left
  // active; the other mirrors keep S()); then the activity is cleared.
  bool mirrored = false;
  void mirror() {
    if constexpr (P::gather_state) {
      if (mirrored) return;
      if (P::stationary) activate_all();
      vm.assign(G.nslots, S());
      for (uint32_t i = 0; i < G.nslots; i++)
        if (vact[i]) vm[i] = v[i];
      reset_activity();
    }
    mirrored = true;
  }

  // reset() (vertex_program.hpp:51-63): messages and accumulators are
  // dropped; the states stay.
  void reset() {
    xreg.clear();
    xsrc.clear();
  }

  // initialize(vids) (vertex_program.hpp:177-259): init on the listed
  // vertices only (flags by position), by class; regular and source
  // vertices whose init returns true send, in increasing position.
  void initialize_vids(const std::vector<uint8_t>& only) {
    xreg.clear(); xsrc.clear();
    for (uint32_t i = 0; i < G.nreg; i++)
      if (only[i] && p.init(vid_of(i), v[i])) xreg.emplace_back(i, p.scatter(v[i]));
    if (G.directed) {
      for (uint32_t k = 0; k < G.nsink; k++) {
        uint32_t i = G.nreg + k;
        if (only[i]) p.init(vid_of(i), v[i]);
      }
      for (uint32_t k = 0; k < G.nsrc; k++) {
        uint32_t i = G.nreg + G.nsink + k;
        if (only[i] && p.init(vid_of(i), v[i])) xsrc.emplace_back(k, p.scatter(v[i]));
      }
    }
    for (uint32_t i = G.nreg + G.nsink + G.nsrc; i < G.nslots; i++)
      if (only[i]) p.init(vid_of(i), v[i]);
    mirror();
  }

  // vertex_program.hpp:84-157
  void initialize() {
    xreg.clear(); xsrc.clear();
    for (uint32_t i = 0; i < G.nreg; i++)
      if (p.init(vid_of(i), v[i])) xreg.emplace_back(i, p.scatter(v[i]));
    if (G.directed) {
      for (uint32_t k = 0; k < G.nsink; k++) { uint32_t i = G.nreg + k; p.init(vid_of(i), v[i]); }
      for (uint32_t k = 0; k < G.nsrc; k++) {
        uint32_t i = G.nreg + G.nsink + k;
        if (p.init(vid_of(i), v[i])) xsrc.emplace_back(k, p.scatter(v[i]));
      }
    }
    for (uint32_t i = G.nreg + G.nsink + G.nsrc; i < G.nslots; i++) p.init(vid_of(i), v[i]);
    mirror();
  }

  // vertex_program.hpp:417-491: positions swap sink(G) <-> source(G reversed).
  template <class OS>
  void initialize_other(const std::vector<OS>& o) {
    xreg.clear(); xsrc.clear();
    for (uint32_t i = 0; i < G.nreg; i++)
      if (p.init_other(vid_of(i), o[i], v[i])) xreg.emplace_back(i, p.scatter(v[i]));
    if (G.directed) {
      for (uint32_t k = 0; k < G.nsink; k++) {
        uint32_t i = G.nreg + k, j = G.nreg + G.nsrc + k;
        p.init_other(vid_of(i), o[j], v[i]);
      }
      for (uint32_t k = 0; k < G.nsrc; k++) {
        uint32_t j = G.nreg + k, i = G.nreg + G.nsink + k;
        if (p.init_other(vid_of(i), o[j], v[i])) xsrc.emplace_back(k, p.scatter(v[i]));
      }
    }
    for (uint32_t i = G.nreg + G.nsink + G.nsrc; i < G.nslots; i++) p.init_other(vid_of(i), o[i], v[i]);
    mirror();
  }

  // vertex_program_execute.hpp:564-603
  void spmv(const OCSC& csc, const std::vector<std::pair<uint32_t, M>>& xs, std::vector<A>& y,
            std::vector<uint8_t>& t, uint32_t offset, uint32_t mirror_base) {
    for (auto& im : xs) {
      uint32_t c = offset + im.first;
      for (uint32_t j = csc.colptrs[c]; j < csc.colptrs[c + 1]; j++) {
        const Entry& e = csc.entries[j];
        EdgeV ev{csc.colidxs[c], e.idx, e.w};
        if constexpr (P::gather_state)
          p.combine(p.gather(ev, im.second, vm[mirror_base + e.global_idx]), y[e.global_idx]);
        else
          p.combine(p.gather(ev, im.second), y[e.global_idx]);
        t[e.global_idx] = 1;
      }
    }
  }

  // process_ready_messages at np=1 (vertex_program_execute.hpp:433-492, 498-561)
  void process(bool sink, uint32_t iter) {
    if (sink) {
      spmv(G.sink, xsrc, ysink, tsink, G.nreg, G.nreg);
      spmv(G.sink, xreg, ysink, tsink, 0, G.nreg);
    } else {
      if (P::stationary || iter == 0) spmv(G.reg, xsrc, yreg, treg, G.nreg, 0);
      spmv(G.reg, xreg, yreg, treg, 0, 0);
    }
  }

  bool do_apply(const A& y, S& s, uint32_t iter) {
    if constexpr (P::apply_iter) return p.apply(y, s, iter);
    else return p.apply(y, s);
  }

  // produce_messages + combine_accumulators + apply_and_scatter_messages
  // (vertex_program_execute.hpp:606-724). Entries are popped in increasing
  // index and reset to Value(); the final accumulator starts at Value().
  bool produce(bool sink, bool single_iter, uint32_t iter) {
    bool any = false;
    if (sink) {
      for (uint32_t idx = 0; idx < G.nsink; idx++) {
        if (!tsink[idx]) continue;
        A part = ysink[idx]; ysink[idx] = p.identity(); tsink[idx] = 0;
        A fin = p.identity(); p.combine(part, fin);
        if (do_apply(fin, v[G.nreg + idx], iter)) vact[G.nreg + idx] = 1;
      }
      return false;
    }
    std::vector<std::pair<uint32_t, M>> out;
    for (uint32_t idx = 0; idx < G.nreg; idx++) {
      if (!treg[idx]) continue;
      A part = yreg[idx]; yreg[idx] = p.identity(); treg[idx] = 0;
      A fin = p.identity(); p.combine(part, fin);
      bool act = do_apply(fin, v[idx], iter);
      any |= act;
      if (act || P::stationary) {
        vact[idx] = 1;
        if (!single_iter) out.emplace_back(idx, p.scatter(v[idx]));
      }
    }
    if (!single_iter) xreg.swap(out);
    return any;
  }

  bool optimizable() const { return G.directed && !P::gather_state && !P::apply_iter; }

  // vertex_program_execute.hpp:11-52
  void execute(uint32_t max_iters) {
    if (max_iters == 1) {
      if (G.directed) { process(false, 0); process(true, 0); produce(false, true, 0); produce(true, true, 0); }
      else { process(false, 0); produce(false, true, 0); }
      iters_run = 1;
      return;
    }
    const bool until = max_iters == 0;
    uint32_t iter = 0;
    bool converged = false;
    if (optimizable()) {
      // execute_ (55-163)
      while (until ? !converged : iter < max_iters) {
        process(false, iter);
        converged = !produce(false, false, iter);
        iter++;
      }
      // Sink processing: every regular vertex re-scatters its final state.
      xreg.clear();
      for (uint32_t i = 0; i < G.nreg; i++) xreg.emplace_back(i, p.scatter(v[i]));
      process(true, iter);
      produce(true, false, iter);
    } else {
      // execute_non_opt (166-248)
      while (until ? !converged : iter < max_iters) {
        process(false, iter);
        if (G.directed) process(true, iter);
        converged = !produce(false, false, iter);
        if (G.directed) converged &= !produce(true, false, iter);
        iter++;
      }
    }
    iters_run = iter;
  }
};

struct Handle {
  OGraph* G;
};

}  // namespace

template <class P>
void run_pr_like(P& p, const OGraph& G, const OGraph& GR, uint32_t iters, double* rank,
                        uint32_t* deg, uint32_t* iters_run) {
  DegProg dp;
  Engine<DegProg> de(dp, GR);
  de.initialize();
  de.execute(1);
  Engine<P> e(p, G);
  e.initialize_other(de.v);
  e.execute(iters);
  for (uint32_t i = 0; i < G.nslots; i++) {
    uint32_t id = e.vid_of(i);
    if (rank) rank[id] = e.v[i].rank;
    if (deg) deg[id] = e.v[i].degree;
  }
  if (iters_run) *iters_run = e.iters_run;
}

// One run of an interface program for ora_vp_activity.
template <class P, class F>
static void act_run(P& p, const OGraph& G, uint32_t max_iters, uint32_t flags, F field, uint32_t* out,
                    uint8_t* act) {
  Engine<P> e(p, G);
  if (flags & 1) e.activate_all();
  e.initialize();
  e.execute(max_iters);
  if (flags & 2) e.reset_activity();
  if (flags & 4) e.activate_all();
  for (uint32_t i = 0; i < G.nslots; i++) {
    out[e.vid_of(i)] = field(e.v[i]);
    act[e.vid_of(i)] = e.vact[i];
  }
}

// ----------------------------------------------------------------- C API ---
extern "C" {

// Build a graph from the records of an LA3 binary file (u32 pairs, or u32
// triples when weighted). mode: 0 directed, 1 reversed, 2 undirected,
// 3 acyclic, 4 reversed acyclic. hashing: 1 bucket (default), 0 none.
void* ora_graph_new(const uint32_t* data, uint64_t nrec, uint32_t n, int weighted, int mode,
                    int hashing) {
  OGraph* G = build_graph(data, nrec, n, weighted != 0, mode, hashing != 0);
  if (!G) return nullptr;
  return new Handle{G};
}

void ora_graph_free(void* h) {
  if (!h) return;
  delete ((Handle*)h)->G;
  delete (Handle*)h;
}

// info: [nslots, nreg, nsink, nsrc, ncols, nnz_reg, nnz_sink, directed]
void ora_graph_info(void* h, uint64_t* info) {
  const OGraph& G = *((Handle*)h)->G;
  info[0] = G.nslots; info[1] = G.nreg; info[2] = G.nsink; info[3] = G.nsrc;
  info[4] = G.reg.ncols; info[5] = G.reg.entries.size(); info[6] = G.sink.entries.size();
  info[7] = G.directed;
}

// Layout export: hid_of[nslots] (dashboard position -> hashed id), and for the
// regular (which=0) or sink (which=1) DCSC: colptrs[ncols+1], colidxs[ncols+1],
// entry global_idx / idx / weight [nnz].
void ora_graph_layout(void* h, uint32_t* hid_of) {
  const OGraph& G = *((Handle*)h)->G;
  std::memcpy(hid_of, G.hid_of.data(), sizeof(uint32_t) * G.nslots);
}

void ora_graph_csc(void* h, int which, uint32_t* colptrs, uint32_t* colidxs, uint32_t* gidx,
                   uint32_t* idx, uint32_t* w) {
  const OGraph& G = *((Handle*)h)->G;
  const OCSC& c = which ? G.sink : G.reg;
  std::memcpy(colptrs, c.colptrs.data(), sizeof(uint32_t) * (c.ncols + 1));
  std::memcpy(colidxs, c.colidxs.data(), sizeof(uint32_t) * (c.ncols + 1));
  for (size_t k = 0; k < c.entries.size(); k++) {
    gidx[k] = c.entries[k].global_idx;
    idx[k] = c.entries[k].idx;
    w[k] = c.entries[k].w;
  }
}

uint32_t ora_hash(void* h, uint32_t v) { return (uint32_t)((Handle*)h)->G->h.hash(v); }
uint32_t ora_unhash(void* h, uint32_t v) { return (uint32_t)((Handle*)h)->G->h.unhash(v); }

// degree.cpp: in-degree of the stored graph, by original id [nslots].
int ora_degree(void* h, uint32_t* deg) {
  const OGraph& G = *((Handle*)h)->G;
  DegProg p;
  Engine<DegProg> e(p, G);
  e.initialize();
  e.execute(1);
  for (uint32_t i = 0; i < G.nslots; i++) deg[e.vid_of(i)] = e.v[i].degree;
  return 0;
}

// pr.cpp:11-55. g = directed graph, gr = same file loaded reversed.
int ora_pr(void* g, void* gr, uint32_t iters, double* rank, uint32_t* deg, uint32_t* iters_run) {
  PrProg p;
  run_pr_like(p, *((Handle*)g)->G, *((Handle*)gr)->G, iters, rank, deg, iters_run);
  return 0;
}

int ora_ppr(void* g, void* gr, uint32_t seed, uint32_t iters, double alpha, double* rank,
            uint32_t* iters_run) {
  PprProg p;
  p.seed = seed;
  p.alpha = alpha;
  run_pr_like(p, *((Handle*)g)->G, *((Handle*)gr)->G, iters, rank, nullptr, iters_run);
  return 0;
}

// bfs.cpp:11-37 (on whatever load mode g was built with). parent is the
// hashed id of the last in-neighbour processed, as the reference stores it.
int ora_bfs(void* h, uint32_t root, uint32_t max_iters, uint8_t* hops, uint32_t* parent,
            uint32_t* iters_run) {
  const OGraph& G = *((Handle*)h)->G;
  BfsProg p;
  p.root = root;
  Engine<BfsProg> e(p, G);
  e.initialize();
  e.execute(max_iters);
  for (uint32_t i = 0; i < G.nslots; i++) {
    uint32_t id = e.vid_of(i);
    hops[id] = e.v[i].hops;
    if (parent) parent[id] = e.v[i].parent;
  }
  if (iters_run) *iters_run = e.iters_run;
  return 0;
}

// sssp.cpp:11-37
int ora_sssp(void* h, uint32_t root, uint32_t max_iters, uint32_t* dist, uint32_t* iters_run) {
  const OGraph& G = *((Handle*)h)->G;
  SsspProg p;
  p.root = root;
  Engine<SsspProg> e(p, G);
  e.initialize();
  e.execute(max_iters);
  for (uint32_t i = 0; i < G.nslots; i++) dist[e.vid_of(i)] = e.v[i].dist;
  if (iters_run) *iters_run = e.iters_run;
  return 0;
}

// cc.cpp:11-31
int ora_cc(void* h, uint32_t max_iters, uint32_t* label, uint32_t* iters_run) {
  const OGraph& G = *((Handle*)h)->G;
  CcProg p;
  Engine<CcProg> e(p, G);
  e.initialize();
  e.execute(max_iters);
  for (uint32_t i = 0; i < G.nslots; i++) label[e.vid_of(i)] = e.v[i].label;
  if (iters_run) *iters_run = e.iters_run;
  return 0;
}

// tc.cpp:11-52. g = acyclic load, gr = reversed acyclic load.
int ora_tc(void* g, void* gr, uint64_t* total, uint32_t* per_vertex) {
  const OGraph& G = *((Handle*)g)->G;
  const OGraph& GR = *((Handle*)gr)->G;
  GnProg gp;
  Engine<GnProg> ge(gp, G);
  ge.initialize();
  ge.execute(1);
  CtProg cp;
  Engine<CtProg> ce(cp, GR);
  ce.initialize_other(ge.v);
  ce.execute(1);
  uint64_t t = 0;
  for (uint32_t i = 0; i < GR.nslots; i++) {
    t += ce.v[i].ntriangles;
    if (per_vertex) per_vertex[ce.vid_of(i)] = ce.v[i].ntriangles;
  }
  *total = t;
  return 0;
}

// The interface-only programs above; per-vertex outputs by original id.
int ora_vp_wdeg(void* h, uint32_t max_iters, uint32_t* wdeg, uint32_t* iters_run) {
  const OGraph& G = *((Handle*)h)->G;
  WDegProg p;
  Engine<WDegProg> e(p, G);
  e.initialize();
  e.execute(max_iters);
  for (uint32_t i = 0; i < G.nslots; i++) wdeg[e.vid_of(i)] = e.v[i].wdeg;
  if (iters_run) *iters_run = e.iters_run;
  return 0;
}

int ora_vp_maxlabel(void* h, uint32_t max_iters, uint32_t* label, uint32_t* iters_run) {
  const OGraph& G = *((Handle*)h)->G;
  MaxLabelProg p;
  Engine<MaxLabelProg> e(p, G);
  e.initialize();
  e.execute(max_iters);
  for (uint32_t i = 0; i < G.nslots; i++) label[e.vid_of(i)] = e.v[i].label;
  if (iters_run) *iters_run = e.iters_run;
  return 0;
}

int ora_vp_level(void* h, uint32_t root, uint32_t max_iters, uint32_t* level, uint32_t* via, uint32_t* iters_run) {
  const OGraph& G = *((Handle*)h)->G;
  LevelProg p;
  p.root = root;
  Engine<LevelProg> e(p, G);
  e.initialize();
  e.execute(max_iters);
  for (uint32_t i = 0; i < G.nslots; i++) {
    level[e.vid_of(i)] = e.v[i].level;
    via[e.vid_of(i)] = e.v[i].via;
  }
  if (iters_run) *iters_run = e.iters_run;
  return 0;
}

int ora_vp_pralpha(void* g, void* gr, uint32_t max_iters, double* rank, uint32_t* iters_run) {
  const OGraph& G = *((Handle*)g)->G;
  const OGraph& GR = *((Handle*)gr)->G;
  DegProg dp;
  Engine<DegProg> de(dp, GR);
  de.initialize();
  de.execute(1);
  PrAlphaProg p;
  Engine<PrAlphaProg> e(p, G);
  e.initialize_other(de.v);
  e.execute(max_iters);
  for (uint32_t i = 0; i < G.nslots; i++) rank[e.vid_of(i)] = e.v[i].rank;
  if (iters_run) *iters_run = e.iters_run;
  return 0;
}

int ora_vp_countgreater(void* h, uint32_t max_iters, uint32_t* val, uint32_t* cnt, uint32_t* iters_run) {
  const OGraph& G = *((Handle*)h)->G;
  CountGreaterProg p;
  Engine<CountGreaterProg> e(p, G);
  e.initialize();
  e.execute(max_iters);
  for (uint32_t i = 0; i < G.nslots; i++) {
    val[e.vid_of(i)] = e.v[i].val;
    cnt[e.vid_of(i)] = e.v[i].cnt;
  }
  if (iters_run) *iters_run = e.iters_run;
  return 0;
}

// MaxLabel run twice: initialize(); execute(m1); then either execute(m2)
// (mode 0: the second run continues from the messages the first left) or
// reset(); initialize(vids) (all vertices when vids is NULL); execute(m2)
// (mode 1). vids are original ids.
int ora_vp_maxlabel_seq(void* h, uint32_t m1, int mode, const uint32_t* vids, uint32_t nvids, uint32_t m2,
                        uint32_t* label, uint32_t* iters_run) {
  const OGraph& G = *((Handle*)h)->G;
  MaxLabelProg p;
  Engine<MaxLabelProg> e(p, G);
  e.initialize();
  e.execute(m1);
  if (mode == 1) {
    e.reset();
    if (vids) {
      std::vector<uint8_t> only(G.nslots, 0);
      for (uint32_t i = 0; i < G.nslots; i++) {
        const uint32_t id = e.vid_of(i);
        for (uint32_t j = 0; j < nvids; j++)
          if (vids[j] == id) only[i] = 1;
      }
      e.initialize_vids(only);
    } else {
      e.initialize();
    }
  }
  e.execute(m2);
  for (uint32_t i = 0; i < G.nslots; i++) label[e.vid_of(i)] = e.v[i].label;
  if (iters_run) *iters_run = e.iters_run;
  return 0;
}

// AccGreater: initialize(); execute(1); reset(); initialize(); execute(1).
int ora_vp_accgreater_twice(void* h, uint32_t* val, uint32_t* cnt) {
  const OGraph& G = *((Handle*)h)->G;
  AccGreaterProg p;
  Engine<AccGreaterProg> e(p, G);
  e.initialize();
  e.execute(1);
  e.reset();
  e.initialize();
  e.execute(1);
  for (uint32_t i = 0; i < G.nslots; i++) {
    val[e.vid_of(i)] = e.v[i].val;
    cnt[e.vid_of(i)] = e.v[i].cnt;
  }
  return 0;
}

// Vertex activity after a run of an interface program: prog 0 WDeg
// (weighted load), 1 MaxLabel, 2 Level from root, 3 CountNonzero, 4
// CountGreater. flags: 1 = activate_all() before initialize(), 2 =
// reset_activity() after execute, 4 = activate_all() after execute. out: the
// program's first state field by original id (wdeg, label, level, cnt,
// cnt); act: the activity by original id.
int ora_vp_activity(void* h, int prog, uint32_t root, uint32_t max_iters, uint32_t flags, uint32_t* out,
                    uint8_t* act) {
  const OGraph& G = *((Handle*)h)->G;
  switch (prog) {
    case 0: { WDegProg p; act_run(p, G, max_iters, flags, [](const WDegProg::S& s) { return s.wdeg; }, out, act); break; }
    case 1: { MaxLabelProg p; act_run(p, G, max_iters, flags, [](const MaxLabelProg::S& s) { return s.label; }, out, act); break; }
    case 2: {
      LevelProg p;
      p.root = root;
      act_run(p, G, max_iters, flags, [](const LevelProg::S& s) { return s.level; }, out, act);
      break;
    }
    case 3: { CountNonzeroProg p; act_run(p, G, max_iters, flags, [](const CountNonzeroProg::S& s) { return s.cnt; }, out, act); break; }
    case 4: { CountGreaterProg p; act_run(p, G, max_iters, flags, [](const CountGreaterProg::S& s) { return s.cnt; }, out, act); break; }
    default: return 1;
  }
  return 0;
}

int ora_vp_countnonzero(void* h, uint32_t max_iters, uint32_t* cnt, uint32_t* iters_run) {
  const OGraph& G = *((Handle*)h)->G;
  CountNonzeroProg p;
  Engine<CountNonzeroProg> e(p, G);
  e.initialize();
  e.execute(max_iters);
  for (uint32_t i = 0; i < G.nslots; i++) cnt[e.vid_of(i)] = e.v[i].cnt;
  if (iters_run) *iters_run = e.iters_run;
  return 0;
}

}  // extern "C"
