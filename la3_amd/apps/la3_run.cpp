// la3_run — drivers of the graph-analytics apps over the C++ drop-in
// (include/la3.hpp), with the reference's positional arguments and output
// lines, display() of ids up to 30 included, without the logger's prefix
// (src/apps/graph_analytics/{degree,pr,bfs,sssp,cc,tc}.cpp):
//   la3_run pr     <filepath> <num_vertices> [<iterations>]
//   la3_run bfs    <filepath> <num_vertices> <root>       (undirected, bfs.cpp:14)
//   la3_run sssp   <filepath> <num_vertices> <root>       (weighted triples)
//   la3_run cc     <filepath> <num_vertices>
//   la3_run tc     <filepath> <num_vertices>
//   la3_run degree <filepath> <num_vertices> [-o]
//   la3_run ppr_topk <filepath> <num_vertices> <iterations> <k> <seed0..seed3>
//     (four personalized-PageRank queries as one batch, per-query top-k by
//      btopk, vertex_program.hpp:683-767; no reference driver)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../include/la3.hpp"

using la3::Empty;

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void timer(const char* name, double t0) {
  const double s = now() - t0;
  // Same layout as DistTimer::report at one rank (dist_timer.h:43-63).
  printf("Timer %s stats: %f secs on average (0.000000 stdev) (%f [on rank 0] -> %f [on rank 0])\n",
         name, s, s, s);
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "Usage: %s <pr|bfs|sssp|cc|tc|degree|ppr_topk> <filepath> <num_vertices> [args]\n", argv[0]);
    return 1;
  }
  const std::string app = argv[1], path = argv[2];
  const uint32_t n = (uint32_t)std::atol(argv[3]);
  try {
    la3::Context ctx(0);
    if (app == "pr") {
      const uint32_t niters = argc > 4 ? (uint32_t)atoi(argv[4]) : 0;
      la3::Graph<Empty> GR(ctx);
      GR.load_directed(true, path, n, true);
      la3::DegVertex<Empty> vp_degree(&GR, true);
      double t = now();
      vp_degree.execute(1);
      timer("Degree Execution", t);
      GR.free();
      la3::Graph<Empty> G(ctx);
      G.load_directed(true, path, n);
      la3::PrVertex vp(&G, true);
      vp.initialize(vp_degree);
      t = now();
      vp.execute(niters);
      vp.display();  // pr.cpp:42
      timer("Pagerank Execution", t);
      long deg = vp.reduce<long>([](uint32_t, const la3::PrState& s) -> long { return s.degree; },
                                 [](long& a, const long& b) { a += b; });
      printf("Degree Checksum = %lu \n", deg);
      double pr = vp.reduce<double>([](uint32_t, const la3::PrState& s) -> double { return s.rank; },
                                    [](double& a, const double& b) { a += b; });
      printf("Pagerank Checksum = %lf \n", pr);
    } else if (app == "bfs" || app == "sssp") {
      if (argc < 5) { fprintf(stderr, "missing <root>\n"); return 1; }
      const uint32_t root = (uint32_t)std::atol(argv[4]);
      long reach = 0, sum = 0;
      if (app == "bfs") {
        la3::Graph<Empty> G(ctx);
        G.load_undirected(true, path, n);
        la3::BfsVertex vp(&G);
        vp.root = root;
        vp.initialize();
        double t = now();
        vp.execute();
        vp.display();  // bfs.cpp:25
        timer("BFS Execution", t);
        reach = vp.reduce<long>([](uint32_t, const la3::BfsState& s) -> long { return s.hops != 255; },
                                [](long& a, const long& b) { a += b; });
        sum = vp.reduce<long>([](uint32_t, const la3::BfsState& s) -> long { return (long)s.hops * s.parent; },
                              [](long& a, const long& b) { a += b; });
      } else {
        la3::Graph<uint32_t> G(ctx);
        G.load_directed(true, path, n);
        la3::SpVertex vp(&G);
        vp.root = root;
        vp.initialize();
        double t = now();
        vp.execute();
        vp.display();  // sssp.cpp:25
        timer("SSSP Execution", t);
        reach = vp.reduce<long>([](uint32_t, const la3::SpState& s) -> long { return s.distance != 0x7fffffffu; },
                                [](long& a, const long& b) { a += b; });
        sum = vp.reduce<long>([](uint32_t, const la3::SpState& s) -> long { return s.distance; },
                              [](long& a, const long& b) { a += b; });
      }
      printf("Reachable Vertices = %lu \n", reach);
      printf("Checksum = %lu \n", sum);
    } else if (app == "cc") {
      la3::Graph<Empty> G(ctx);
      G.load_undirected(true, path, n);
      la3::CcVertex vp(&G);
      vp.initialize();
      double t = now();
      vp.execute();
      vp.display();  // cc.cpp:24
      timer("CC Execution", t);
      long sum = vp.reduce<long>([](uint32_t, const la3::CcState& s) -> long { return s.label; },
                                 [](long& a, const long& b) { a += b; });
      printf("Checksum = %lu \n", sum);
    } else if (app == "tc") {
      la3::Graph<Empty> G(ctx);
      G.load_directed(true, path, n, false, true);
      la3::Graph<Empty> GR(ctx);
      GR.load_directed(true, path, n, true, true);
      la3::GnVertex gn(&G, true);
      gn.initialize();
      double t = now();
      gn.execute(1);
      la3::CtVertex ct(&GR, true);
      ct.initialize(gn);
      ct.execute(1);
      timer("Triangle Counting", t);
      long tri = ct.reduce<long>([](uint32_t, const la3::CtState& s) -> long { return s.ntriangles; },
                                 [](long& a, const long& b) { a += b; });
      printf("Triangles = %lu \n", tri);
    } else if (app == "degree") {
      const bool out = argc > 4 && std::string(argv[4]) == "-o";
      la3::Graph<Empty> G(ctx);
      G.load_directed(true, path, n, out);
      la3::DegVertex<Empty> vp(&G, true);
      double t = now();
      vp.execute(1);
      printf(out ? "Out-Degrees: \n" : "In-Degrees: \n");  // degree.cpp:28-32
      vp.display();
      timer("Degree Execution", t);
      long sum = vp.reduce<long>([](uint32_t, const la3::DegState& s) -> long { return s.degree; },
                                 [](long& a, const long& b) { a += b; });
      printf("%s-Degree Checksum = %lu \n", out ? "Out" : "In", sum);
      // degree.cpp:34-43
      constexpr int k = 3;
      using iv_t = std::pair<uint32_t, uint32_t>;
      std::vector<iv_t> top;
      vp.topk<uint32_t, uint32_t>(k, top, [](uint32_t, const la3::DegState& s) { return s.degree; },
                                  [](const iv_t& a, const iv_t& b) { return a.second > b.second; });
      printf("Top-%i: \n", k);
      for (const auto& iv : top) printf("idx %u: degree %u \n", iv.first, iv.second);
    } else if (app == "ppr_topk") {
      constexpr uint32_t B = 4;
      if (argc < 6 + (int)B) { fprintf(stderr, "missing <iterations> <k> <seed0..seed3>\n"); return 1; }
      const uint32_t niters = (uint32_t)atoi(argv[4]), k = (uint32_t)atoi(argv[5]);
      la3::Graph<Empty> G(ctx);
      G.load_directed(true, path, n);
      la3::PprBatchVertex<B> vp(&G);
      for (uint32_t q = 0; q < B; q++) vp.roots[q] = (uint32_t)std::atol(argv[6 + q]);
      vp.initialize();
      double t = now();
      vp.execute(niters);
      timer("PPR Batch Execution", t);
      using iv_t = std::pair<uint32_t, float>;
      std::vector<iv_t> top[B];
      vp.btopk<B, uint32_t, float>(k, top, [](uint32_t, const la3::BatchState<B, float>& s) { return s.v; },
                                   [](const iv_t& a, const iv_t& b) { return a.second > b.second; });
      for (uint32_t q = 0; q < B; q++) {
        printf("Query %u (seed %u) top-%u: \n", q, vp.roots[q], k);
        for (const auto& iv : top[q]) printf("idx %u: rank %.9g \n", iv.first, (double)iv.second);
      }
    } else {
      fprintf(stderr, "unknown app %s\n", app.c_str());
      return 1;
    }
  } catch (const la3::Error& e) {
    fprintf(stderr, "la3 error %d: %s\n", (int)e.status, e.what());
    return 2;
  }
  return 0;
}
