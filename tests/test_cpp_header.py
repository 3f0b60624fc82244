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
                              });
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
