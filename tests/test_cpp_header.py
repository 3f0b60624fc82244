"""The C++ drop-in header (include/la3.hpp) compiles with every program
wrapper instantiated, the batched ones and btopk included (host compiler
only; no device calls)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include "la3.hpp"
using la3::Empty;
int main(int argc, char**) {
  if (argc > 99) {  // never runs: instantiation only
    la3::Context ctx(0);
    la3::Graph<Empty> G(ctx);
    la3::Graph<uint32_t> GW(ctx);
    la3::DegVertex<Empty> deg(&G); deg.execute(1);
    la3::PrVertex pr(&G); pr.initialize(deg); pr.execute(20);
    la3::BfsVertex bfs(&G); bfs.execute();
    la3::SpVertex sp(&GW); sp.execute();
    la3::CcVertex cc(&G); cc.execute();
    la3::GnVertex gn(&G); la3::CtVertex ct(&G); ct.initialize(gn); ct.execute(1);
    la3::BfsBatchVertex<8> bb(&G); bb.execute();
    la3::SpBatchVertex<16> sb(&GW); sb.execute();
    la3::PprBatchVertex<4> pb(&G); pb.execute(30);
    using iv8 = std::pair<uint32_t, uint8_t>;
    std::vector<iv8> t8[8];
    bb.btopk<8, uint32_t, uint8_t>(3, t8, [](uint32_t, const la3::BatchState<8, uint8_t>& s) { return s.v; },
                                   [](const iv8& a, const iv8& b) { return a.second < b.second; });
    using ivf = std::pair<uint32_t, float>;
    std::vector<ivf> tf[2];
    pb.btopk<2, uint32_t, float>(5, tf, [](uint32_t, const la3::BatchState<4, float>& s) { return s.v; },
                                 [](const ivf& a, const ivf& b) { return a.second > b.second; });
    std::vector<std::pair<uint32_t, double>> top;
    pr.topk<uint32_t, double>(3, top, [](uint32_t, const la3::PrState& s) { return s.rank; },
                              [](const std::pair<uint32_t, double>& a, const std::pair<uint32_t, double>& b) {
                                return a.second > b.second;
                              }, true);
    deg.display(30, std::cout);
    pr.display<true>(5);
    bfs.display();
    sp.display();
    cc.reset_activity();
    cc.activate_all();
    long s = cc.reduce<long>([](uint32_t, const la3::CcState& st) { return (long)st.label; },
                             [](long& a, const long& b) { a += b; }, true);
    (void)s;
  }
  return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_header_instantiates_every_wrapper(tmp_path):
    src = tmp_path / "inst.cpp"
    src.write_text(SRC)
    out = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I",
                          os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr


RUN_SRC = r'''
#include <cstdio>
#include <sstream>
#include "la3.hpp"
// A stand-in for la3::Graph with a fixed slot order, so ProgramBase's host
// logic runs without a device.
struct FakeGraph {
  std::vector<uint32_t> ids;
  la3_graph_info inf{};
  std::vector<uint32_t> slot_ids() const { return ids; }
  const la3_graph_info& info() const { return inf; }
  uint32_t get_nvertices() const { return inf.n; }
  la3::Context& context() const { throw la3::Error(LA3_EINVAL, "no context"); }
};
struct Labels : la3::ProgramBase<FakeGraph, la3::CcState> {
  explicit Labels(const FakeGraph* g) : ProgramBase(g) {}
  void run(const std::vector<uint32_t>& lab) {
    states_.assign(lab.size(), la3::CcState());
    for (size_t i = 0; i < lab.size(); i++) states_[i].label = lab[i];
    add_active([](uint32_t v, const la3::CcState& s) { return s.label != v; });
  }
};
int main() {
  FakeGraph g;
  g.ids = {3, 1, 4, 0, 5, 9, 2, 6, 8, 7};
  g.inf.n = 9;
  g.inf.nslots = 10;
  Labels p(&g);
  p.run({0, 1, 1, 3, 3, 3, 6, 6, 8, 8});
  using IV = std::pair<uint32_t, uint32_t>;
  auto greater = [](const IV& a, const IV& b) { return a.second > b.second; };
  auto lab = [](uint32_t, const la3::CcState& s) { return s.label; };
  for (int ao = 0; ao < 2; ao++)
    for (uint32_t k : {3u, 12u}) {
      std::vector<IV> top;
      p.topk<uint32_t, uint32_t>(k, top, lab, greater, ao != 0);
      printf("topk %d %u:", ao, k);
      for (auto& t : top) printf(" %u/%u", t.first, t.second);
      printf("\n");
    }
  auto add = [](long& a, const long& b) { a += b; };
  auto lab_l = [](uint32_t, const la3::CcState& s) { return (long)s.label; };
  printf("sum %ld %ld\n", p.reduce<long>(lab_l, add), p.reduce<long>(lab_l, add, true));
  std::ostringstream a, b;
  p.display(4, a);
  p.display<true>(2, b);
  printf("display[%s]\nvalues[%s]\n", a.str().c_str(), b.str().c_str());
  p.reset_activity();
  printf("reset %ld\n", p.reduce<long>(lab_l, add, true));
  p.activate_all();
  printf("all %ld\n", p.reduce<long>(lab_l, add, true));
  la3::BfsState bs;
  bs.parent = 7;
  bs.hops = 2;
  la3::PrState ps;
  ps.rank = 0.3061875;
  ps.degree = 1;
  printf("fmt %s %s %s\n", la3::to_string(bs).c_str(), la3::to_string(ps).c_str(),
         la3::to_string(la3::SpState()).c_str());
  return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_program_wrapper_topk_activity_display(tmp_path):
    """ProgramBase's host logic on fixed states: topk follows the reference's
    steps (vertex_program.hpp:607-680: partial_sort of the first min(k, n),
    resize to k with (0, 0) pairs, partial_sort of the k), active_only keeps
    the vertices whose apply activated them (here: label != id, as CcVertex),
    display prints ids <= min(nvertices, n) in slot order in the states'
    to_string format (vertex_program.hpp:511-546; bfs.h:22-23, pr.h:23-24)."""
    src = tmp_path / "run.cpp"
    exe = tmp_path / "run"
    src.write_text(RUN_SRC)
    out = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                          str(src), "-o", str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    got = res.stdout
    lab = [0, 1, 1, 3, 3, 3, 6, 6, 8, 8]
    act = [v for v in range(10) if lab[v] != v]
    lines = {ln.split(":")[0]: ln for ln in got.splitlines() if ln.startswith("topk")}
    for ao in (0, 1):
        pool = act if ao else list(range(10))
        for k in (3, 12):
            pairs = [tuple(map(int, t.split("/"))) for t in lines[f"topk {ao} {k}"].split(":")[1].split()]
            assert len(pairs) == k
            vals = sorted([lab[v] for v in pool] + [0] * max(0, k - len(pool)), reverse=True)[:k]
            assert [p[1] for p in pairs] == vals, (ao, k, pairs)
            for i, v in pairs:
                assert v == lab[i] and (i in pool or (i, v) == (0, 0))
    assert f"sum {sum(lab)} {sum(lab[v] for v in act)}" in got
    slot = [3, 1, 4, 0, 5, 9, 2, 6, 8, 7]
    disp = "".join(f"{i}: {{label: {lab[i]}}}\n" for i in slot if i <= 4)
    vals = "".join(f"{{label: {lab[i]}}}\n" for i in slot if i <= 2) + "\n"
    assert f"display[{disp}]" in got and f"values[{vals}]" in got
    assert "reset 0\n" in got and f"all {sum(lab)}\n" in got
    assert "fmt {parent: 7, hops: 2} {rank: 0.306188, degree: 1} {distance: 2147483647}" in got
