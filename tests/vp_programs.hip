// Vertex programs written only against include/la3_vertex_program.hpp (the
// reference's VertexProgram<W, M, A, S> interface, vertex_program.h:21-137),
// compiled outside the library as a caller would, for tests/test_gpu_vp.py.
// None of them is one of the library's programs; the oracle holds the same
// operations (oracle/la3_oracle.cpp: WDegProg, MaxLabelProg, LevelProg,
// PrAlphaProg, CountGreaterProg, CountNonzeroProg). Each entry point returns 0 or a la3_status
// and writes per-vertex outputs by original id.
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>

#include "la3_vertex_program.hpp"

namespace {

thread_local std::string g_err;

// Weighted in-degree: stationary, weights, execute_single. The three programs
// below declare their combine associative (integer sum, max, last writer),
// so their long rows take the engine's segmented fold; the others keep the
// in-order fold of every message.
struct WDeg {
  static constexpr bool associative = true;
  using W = uint32_t;
  using M = la3::Empty;
  using A = uint32_t;
  struct S {
    uint32_t wdeg = 0;
  };
  __device__ A gather(const la3::Edge<W>& e, const M&) const { return e.weight; }
  __device__ void combine(const A& y1, A& y2) const { y2 += y1; }
  __device__ bool apply(const A& y, S& s) const {
    s.wdeg = y;
    return true;
  }
};

// Max-label propagation: execute_ with the sink step on directed loads,
// execute_non_opt on undirected ones.
struct MaxLabel {
  static constexpr bool associative = true;
  using W = la3::Empty;
  using M = uint32_t;
  using A = uint32_t;
  struct S {
    uint32_t label = 0;
    std::string to_string() const { return "{label: " + std::to_string(label) + "}"; }
  };
  __device__ bool init(uint32_t vid, S& s) const {
    s.label = vid;
    return true;
  }
  __device__ M scatter(const S& s) const { return s.label; }
  __device__ A gather(const la3::Edge<W>&, const M& m) const { return m; }
  __device__ void combine(const A& y1, A& y2) const { y2 = max(y1, y2); }
  __device__ bool apply(const A& y, S& s) const {
    const uint32_t t = s.label;
    s.label = max(s.label, y);
    return t != s.label;
  }
};

// First-reach level with the last combined sender: apply reads the
// iteration, and the combine keeps the last message, so the result depends
// on the order in which a row combines its messages.
struct Level {
  static constexpr bool associative = true;
  using W = la3::Empty;
  using M = la3::Empty;
  using A = uint32_t;
  struct S {
    uint32_t level = 0xffffffffu;
    uint32_t via = 0;
  };
  uint32_t root = 0;
  __device__ bool init(uint32_t vid, S& s) const {
    if (vid != root) return false;
    s.level = 0;
    return true;
  }
  __device__ A gather(const la3::Edge<W>& e, const M&) const { return e.src; }
  __device__ void combine(const A& y1, A& y2) const { y2 = y1; }
  __device__ bool apply(const A& y, S& s, uint32_t iter) const {
    if (s.level != 0xffffffffu) return false;
    s.level = iter + 1;
    s.via = y;
    return true;
  }
};

// Degrees on the reversed load, handed to PrAlpha through initialize(other).
struct Deg {
  using W = la3::Empty;
  using M = la3::Empty;
  using A = uint32_t;
  struct S {
    uint32_t degree = 0;
  };
  __device__ A gather(const la3::Edge<W>&, const M&) const { return 1; }
  __device__ void combine(const A& y1, A& y2) const { y2 += y1; }
  __device__ bool apply(const A& y, S& s) const {
    s.degree = y;
    return true;
  }
};

struct PrAlpha {
  using W = la3::Empty;
  using M = double;
  using A = double;
  struct S {
    uint32_t degree = 0;
    double rank = 0.2;
  };
  __device__ bool init(uint32_t, const Deg::S& o, S& s) const {
    s.degree = o.degree;
    return true;
  }
  __device__ M scatter(const S& s) const { return s.rank / s.degree; }
  __device__ A gather(const la3::Edge<W>&, const M& m) const { return m; }
  __device__ void combine(const A& y1, A& y2) const { y2 += y1; }
  __device__ bool apply(const A& y, S& s) const {
    const double t = s.rank;
    s.rank = 0.2 + 0.8 * y;
    return fabs(s.rank - t) > 1e-9;
  }
};

// A gather that reads the row's state (its mirror as initialize left it).
struct CountGreater {
  using W = la3::Empty;
  using M = uint32_t;
  using A = uint32_t;
  struct S {
    uint32_t val = 0;
    uint32_t cnt = 0;
  };
  __device__ bool init(uint32_t vid, S& s) const {
    s.val = vid % 7;
    return true;
  }
  __device__ M scatter(const S& s) const { return s.val; }
  __device__ A gather(const la3::Edge<W>&, const M& m, const S& s) const { return m > s.val ? 1u : 0u; }
  __device__ void combine(const A& y1, A& y2) const { y2 += y1; }
  __device__ bool apply(const A& y, S& s) const {
    s.cnt += y;
    s.val = y % 7;
    return false;
  }
};

// Not stationary, state-reading gather: the mirrors keep S() (val 0).
struct CountNonzero {
  using W = la3::Empty;
  using M = uint32_t;
  using A = uint32_t;
  struct S {
    uint32_t val = 0;
    uint32_t cnt = 0;
  };
  __device__ bool init(uint32_t vid, S& s) const {
    s.val = vid % 7;
    return vid % 3 == 0;
  }
  __device__ M scatter(const S& s) const { return s.val; }
  __device__ A gather(const la3::Edge<W>&, const M& m, const S& s) const { return m > s.val ? 1u : 0u; }
  __device__ void combine(const A& y1, A& y2) const { y2 += y1; }
  __device__ bool apply(const A& y, S& s) const {
    s.cnt += y;
    return y > 0;
  }
};

// Stationary, state-reading gather, and an init that adds to the state it
// finds: re-initialized after reset(), its mirrors keep the first values.
struct AccGreater {
  using W = la3::Empty;
  using M = uint32_t;
  using A = uint32_t;
  struct S {
    uint32_t val = 0;
    uint32_t cnt = 0;
  };
  __device__ bool init(uint32_t vid, S& s) const {
    s.val += vid % 5 + 1;
    return true;
  }
  __device__ M scatter(const S& s) const { return s.val; }
  __device__ A gather(const la3::Edge<W>&, const M& m, const S& s) const { return m > s.val ? 1u : 0u; }
  __device__ void combine(const A& y1, A& y2) const { y2 += y1; }
  __device__ bool apply(const A& y, S& s) const {
    s.cnt += y;
    s.val = y % 7;
    return false;
  }
};

// Shares MaxLabel's states (the borrowing constructor): init adds 1000 to
// the label it finds and activates nobody.
struct Bump {
  using W = la3::Empty;
  using M = uint32_t;
  using A = uint32_t;
  using S = MaxLabel::S;
  __device__ bool init(uint32_t, S& s) const {
    s.label += 1000;
    return false;
  }
};

template <class F>
int guarded(F f) {
  try {
    f();
    return 0;
  } catch (const la3::Error& e) {
    g_err = e.what();
    return (int)e.status;
  } catch (const std::exception& e) {
    g_err = e.what();
    return (int)LA3_EDEVICE;
  }
}

}  // namespace

extern "C" {

const char* vp_error() { return g_err.c_str(); }

int vp_wdeg(la3_ctx* c, la3_graph* g, uint32_t max_iters, uint32_t* wdeg, uint32_t* iters) {
  return guarded([&] {
    la3::VertexProgram<WDeg> vp(c, g, true);
    vp.initialize();
    vp.execute(max_iters);
    auto st = vp.states();
    for (size_t i = 0; i < st.size(); i++) wdeg[i] = st[i].wdeg;
    *iters = vp.iterations();
  });
}

int vp_maxlabel(la3_ctx* c, la3_graph* g, uint32_t max_iters, uint32_t* label, uint32_t* iters) {
  return guarded([&] {
    la3::VertexProgram<MaxLabel> vp(c, g);
    vp.initialize();
    vp.execute(max_iters);
    auto st = vp.states();
    for (size_t i = 0; i < st.size(); i++) label[i] = st[i].label;
    *iters = vp.iterations();
  });
}

int vp_level(la3_ctx* c, la3_graph* g, uint32_t root, uint32_t max_iters, uint32_t* level, uint32_t* via,
             uint32_t* iters) {
  return guarded([&] {
    Level ops;
    ops.root = root;
    la3::VertexProgram<Level> vp(c, g, false, ops);
    vp.initialize();
    vp.execute(max_iters);
    auto st = vp.states();
    for (size_t i = 0; i < st.size(); i++) {
      level[i] = st[i].level;
      via[i] = st[i].via;
    }
    *iters = vp.iterations();
  });
}

// g: directed load, gr: the same file reversed (as pr.cpp:15-27 loads them).
int vp_pralpha(la3_ctx* c, la3_graph* g, la3_graph* gr, uint32_t max_iters, double* rank, uint32_t* iters) {
  return guarded([&] {
    la3::VertexProgram<Deg> deg(c, gr, true);
    deg.initialize();
    deg.execute(1);
    la3::VertexProgram<PrAlpha> vp(c, g, true);
    vp.initialize(deg);
    vp.execute(max_iters);
    auto st = vp.states();
    for (size_t i = 0; i < st.size(); i++) rank[i] = st[i].rank;
    *iters = vp.iterations();
  });
}

int vp_countgreater(la3_ctx* c, la3_graph* g, uint32_t max_iters, uint32_t* val, uint32_t* cnt, uint32_t* iters) {
  return guarded([&] {
    la3::VertexProgram<CountGreater> vp(c, g, true);
    vp.initialize();
    vp.execute(max_iters);
    auto st = vp.states();
    for (size_t i = 0; i < st.size(); i++) {
      val[i] = st[i].val;
      cnt[i] = st[i].cnt;
    }
    *iters = vp.iterations();
  });
}

int vp_countnonzero(la3_ctx* c, la3_graph* g, uint32_t max_iters, uint32_t* cnt, uint32_t* iters) {
  return guarded([&] {
    la3::VertexProgram<CountNonzero> vp(c, g);
    vp.initialize();
    vp.execute(max_iters);
    auto st = vp.states();
    for (size_t i = 0; i < st.size(); i++) cnt[i] = st[i].cnt;
    *iters = vp.iterations();
  });
}

// initialize(); execute(m1); then execute(m2) (mode 0), or reset();
// initialize(vids) (all when vids is NULL); execute(m2) (mode 1).
int vp_maxlabel_seq(la3_ctx* c, la3_graph* g, uint32_t m1, int mode, const uint32_t* vids, uint32_t nvids,
                    uint32_t m2, uint32_t* label, uint32_t* iters) {
  return guarded([&] {
    la3::VertexProgram<MaxLabel> vp(c, g);
    vp.initialize();
    vp.execute(m1);
    if (mode == 1) {
      vp.reset();
      if (vids) vp.initialize(std::vector<uint32_t>(vids, vids + nvids));
      else vp.initialize();
    }
    vp.execute(m2);
    auto st = vp.states();
    for (size_t i = 0; i < st.size(); i++) label[i] = st[i].label;
    *iters = vp.iterations();
  });
}

int vp_accgreater_twice(la3_ctx* c, la3_graph* g, uint32_t* val, uint32_t* cnt) {
  return guarded([&] {
    la3::VertexProgram<AccGreater> vp(c, g, true);
    vp.initialize();
    vp.execute(1);
    vp.reset();
    vp.initialize();
    vp.execute(1);
    auto st = vp.states();
    for (size_t i = 0; i < st.size(); i++) {
      val[i] = st[i].val;
      cnt[i] = st[i].cnt;
    }
  });
}

// Vertex activity after a run (the oracle's ora_vp_activity): prog 0 WDeg,
// 1 MaxLabel, 2 Level from root, 3 CountNonzero, 4 CountGreater; flags 1 =
// activate_all() before initialize(), 2 = reset_activity() after execute,
// 4 = activate_all() after execute. out: the first state field by id.
int vp_activity(la3_ctx* c, la3_graph* g, int prog, uint32_t root, uint32_t max_iters, uint32_t flags,
                uint32_t* out, uint8_t* act) {
  return guarded([&] {
    auto run = [&](auto& vp, auto field) {
      if (flags & 1) vp.activate_all();
      vp.initialize();
      vp.execute(max_iters);
      if (flags & 2) vp.reset_activity();
      if (flags & 4) vp.activate_all();
      auto st = vp.states();
      auto a = vp.activity();
      for (size_t i = 0; i < st.size(); i++) {
        out[i] = field(st[i]);
        act[i] = a[i];
      }
    };
    if (prog == 0) {
      la3::VertexProgram<WDeg> vp(c, g, true);
      run(vp, [](const WDeg::S& s) { return s.wdeg; });
    } else if (prog == 1) {
      la3::VertexProgram<MaxLabel> vp(c, g);
      run(vp, [](const MaxLabel::S& s) { return s.label; });
    } else if (prog == 2) {
      Level ops;
      ops.root = root;
      la3::VertexProgram<Level> vp(c, g, false, ops);
      run(vp, [](const Level::S& s) { return s.level; });
    } else if (prog == 3) {
      la3::VertexProgram<CountNonzero> vp(c, g);
      run(vp, [](const CountNonzero::S& s) { return s.cnt; });
    } else if (prog == 4) {
      la3::VertexProgram<CountGreater> vp(c, g, true);
      run(vp, [](const CountGreater::S& s) { return s.cnt; });
    } else {
      throw la3::Error(LA3_EINVAL, "unknown program");
    }
  });
}

// MaxLabel until convergence, then: topk(k) of the labels (largest first;
// active_only as asked) into ids/vals; btopk<2> of (label, id) into bids /
// bvals (2 x k, batch-major); reduce of the labels (all and active) into
// sums[0..1]; display<values_only>(nvertices) into text (cap bytes, NUL
// terminated).
int vp_extract(la3_ctx* c, la3_graph* g, uint32_t k, int active_only, uint32_t* ids, uint32_t* vals, uint32_t* bids,
               uint32_t* bvals, uint64_t* sums, uint32_t nvertices, int values_only, char* text, uint64_t cap) {
  return guarded([&] {
    la3::VertexProgram<MaxLabel> vp(c, g);
    vp.initialize();
    vp.execute(0);
    using IV = std::pair<uint32_t, uint32_t>;
    auto greater = [](const IV& a, const IV& b) { return a.second > b.second; };
    std::vector<IV> top;
    vp.topk<uint32_t, uint32_t>(k, top, [](uint32_t, const MaxLabel::S& s) { return s.label; }, greater,
                                active_only != 0);
    if (top.size() != k) throw la3::Error(LA3_EDEVICE, "topk returned a wrong count");
    for (uint32_t i = 0; i < k; i++) {
      ids[i] = top[i].first;
      vals[i] = top[i].second;
    }
    std::vector<IV> bt[2];
    thread_local uint32_t pair[2];
    vp.btopk<2, uint32_t, uint32_t>(
        k, bt,
        [](uint32_t vid, const MaxLabel::S& s) {
          pair[0] = s.label;
          pair[1] = vid;
          return (const uint32_t*)pair;
        },
        greater, active_only != 0);
    for (uint32_t b = 0; b < 2; b++) {
      if (bt[b].size() != k) throw la3::Error(LA3_EDEVICE, "btopk returned a wrong count");
      for (uint32_t i = 0; i < k; i++) {
        bids[b * k + i] = bt[b][i].first;
        bvals[b * k + i] = bt[b][i].second;
      }
    }
    auto lab = [](uint32_t, const MaxLabel::S& s) { return (uint64_t)s.label; };
    auto add = [](uint64_t& a, const uint64_t& b) { a += b; };
    sums[0] = vp.reduce<uint64_t>(lab, add);
    sums[1] = vp.reduce<uint64_t>(lab, add, true);
    std::ostringstream os;
    if (values_only) vp.display<true>(nvertices, os);
    else vp.display(nvertices, os);
    const std::string t = os.str();
    if (t.size() + 1 > cap) throw la3::Error(LA3_EINVAL, "display text exceeds the buffer");
    std::memcpy(text, t.c_str(), t.size() + 1);
  });
}

// The borrowing constructor: MaxLabel runs to convergence into label0; a
// Bump program built from it initializes (adding 1000 to every shared label)
// and MaxLabel's states then read label1; Bump is freed, MaxLabel's states
// still read label2; MaxLabel is freed and Bump (constructed before)
// re-runs initialize on fresh states into label3 (every label 1000).
int vp_borrow(la3_ctx* c, la3_graph* g, uint32_t* label0, uint32_t* label1, uint32_t* label2, uint32_t* label3) {
  return guarded([&] {
    auto copy = [](const std::vector<MaxLabel::S>& st, uint32_t* out) {
      for (size_t i = 0; i < st.size(); i++) out[i] = st[i].label;
    };
    la3::VertexProgram<MaxLabel> ml(c, g);
    ml.initialize();
    ml.execute(0);
    copy(ml.states(), label0);
    la3::VertexProgram<Bump> bump(ml);
    bump.initialize();
    bump.execute(1);
    copy(ml.states(), label1);
    bump.free();
    copy(ml.states(), label2);
    ml.free();
    bump.initialize();
    copy(bump.states(), label3);
  });
}

}  // extern "C"
