// Serving loop over resident graphs: continuous batching by worker threads,
// admission by per-query state bytes. Uses only the public C ABI
// (include/la3.h); each worker owns a context (stream + scratch).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/la3.h"

namespace {

struct Query {
  uint64_t qid;
  la3_query_kind kind;
  uint32_t graph, root, iters;
  double submit_s, start_s = 0, done_s = 0;
  uint32_t batch = 0;
  int32_t status = 0;
};

struct GraphEntry {
  la3_graph* g;
  uint32_t nslots;
  bool weighted, directed;
  uint64_t nnz;
};

}  // namespace

struct la3_server {
  la3_server_config cfg{};
  int device = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::vector<GraphEntry> graphs;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Query> pending;
  std::deque<la3_query_info> done;
  std::map<uint64_t, std::vector<uint8_t>> results;
  uint64_t next_qid = 1, submitted = 0, completed = 0, in_flight_bytes = 0;
  uint32_t running_batches = 0;
  // (kind, graph, iters) groups with a batch running on a worker
  std::set<std::tuple<int, uint32_t, uint32_t>> running_groups;
  bool stop = false;
  std::vector<std::thread> workers;
  uint64_t warm_gen = 0, warm_done = 0;  // la3_server_warm requests / worker completions

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

// Per-query device state of a batched run, per vertex slot (matches what
// the batched drivers allocate per query lane, including the result buffer).
double per_vertex_bytes(la3_query_kind k) {
  switch (k) {
    case LA3_Q_BFS: return 3.0 / 8 + 1 + 1;        // visited/frontier/next bits, level, result
    case LA3_Q_SSSP: return 4 + 2.0 / 8 + 4;       // distance, masks, result
    default: return 4 * 4 + 4;                     // two message buffers, rank, compensation, result
  }
}

uint64_t batch_bytes(const la3_server* s, la3_query_kind k, uint32_t graph, uint32_t nq) {
  const uint64_t lanes = 64ull * ((nq + 63) / 64);
  return (uint64_t)(per_vertex_bytes(k) * s->graphs[graph].nslots * lanes);
}

size_t elem_bytes(la3_query_kind k) { return k == LA3_Q_BFS ? 1 : 4; }

std::tuple<int, uint32_t, uint32_t> group_of(const Query& q) { return {(int)q.kind, q.graph, q.iters}; }

// Expected work of a query: the graph's entries times the passes over them
// (BFS 1, SSSP about 4 with its re-scans, PPR one per iteration).
double expected_work(const la3_server* s, const Query& q) {
  const double f = q.kind == LA3_Q_BFS ? 1.0 : q.kind == LA3_Q_SSSP ? 4.0 : (double)std::max(1u, q.iters);
  return f * (double)std::max<uint64_t>(s->graphs[q.graph].nnz, 1);
}

// Takes, with the lock held, a head query and the later pending queries with
// its (kind, graph, iters), up to `cap`, while the admitted state bytes fit
// the budget (a lone query is admitted when nothing runs). The head is the
// oldest pending query, or with max_wait_s > 0 (and `exclusive`) the one of
// least expected work unless the oldest has waited longer than max_wait_s:
// small graphs' queries then do not queue behind large PPR batches (at 400
// QPS on the six-graph stream p50 8.2 -> 5.1-5.5 ms, p99 unchanged), and the
// age bound keeps the large ones from starving. With `exclusive`, queries of
// a group that has a batch running are left pending: they go together in
// that group's next batch instead of a second sweep of the same graph beside
// it. Returns the picked queries (arrival order) and their state bytes.
std::vector<Query> pick_group(la3_server* s, uint32_t cap, uint64_t& bytes, bool exclusive) {
  std::vector<Query> batch;
  bytes = 0;
  if (s->pending.empty() || cap == 0) return batch;
  size_t h = 0;
  if (exclusive) {
    const size_t n = s->pending.size();
    size_t oldest = n, least = n;
    double lw = 0;
    for (size_t i = 0; i < n; i++) {
      const Query& q = s->pending[i];
      if (s->running_groups.count(group_of(q))) continue;
      if (oldest == n) {
        oldest = i;
        if (s->cfg.max_wait_s <= 0) break;
      }
      const double w = expected_work(s, q);
      if (least == n || w < lw) { least = i; lw = w; }
    }
    h = oldest;
    if (s->cfg.max_wait_s > 0 && oldest < n && s->now() - s->pending[oldest].submit_s <= s->cfg.max_wait_s)
      h = least;
  }
  if (h == s->pending.size()) return batch;
  const Query head = s->pending[h];
  std::vector<size_t> pick;
  for (size_t i = h; i < s->pending.size() && pick.size() < cap; i++) {
    const Query& q = s->pending[i];
    if (q.kind != head.kind || q.graph != head.graph || q.iters != head.iters) continue;
    const uint64_t b = batch_bytes(s, head.kind, head.graph, (uint32_t)pick.size() + 1);
    if (s->in_flight_bytes + b > s->cfg.state_budget && !(pick.empty() && s->running_batches == 0)) break;
    pick.push_back(i);
  }
  if (pick.empty()) return batch;
  bytes = batch_bytes(s, head.kind, head.graph, (uint32_t)pick.size());
  for (size_t j = pick.size(); j-- > 0;) {
    batch.push_back(s->pending[pick[j]]);
    s->pending.erase(s->pending.begin() + pick[j]);
  }
  std::reverse(batch.begin(), batch.end());
  s->in_flight_bytes += bytes;
  s->running_batches++;
  if (exclusive) s->running_groups.insert(group_of(head));
  return batch;
}

// Records the completion of one batch's queries (lock held).
void complete(la3_server* s, const std::vector<Query>& qs, uint32_t batch_size, la3_status st, double start,
              double fin, const uint8_t* out, uint64_t bytes) {
  for (uint32_t j = 0; j < qs.size(); j++) {
    const Query& q = qs[j];
    la3_query_info info{};
    info.qid = q.qid; info.kind = q.kind; info.graph = q.graph; info.root = q.root;
    info.batch_size = batch_size; info.status = st; info.submit_s = q.submit_s;
    info.start_s = start; info.done_s = fin;
    s->done.push_back(info);
    if (s->cfg.keep_results && st == LA3_OK && out) {
      const size_t row = (size_t)s->graphs[q.graph].nslots * elem_bytes(q.kind);
      s->results[q.qid] = std::vector<uint8_t>(out + (size_t)j * row, out + (size_t)(j + 1) * row);
    }
  }
  s->completed += qs.size();
  s->in_flight_bytes -= bytes;
  s->running_batches--;
}

// Fused mode: the worker keeps one session (a fused run on its context).
// Between steps it admits pending queries as new batches —
// grouped by (kind, graph, iters), any number of groups, while task slots
// and budget allow — and retires the batches that finished, so queries of
// every program and graph share each fused launch and new ones join as
// slots and memory free up.
void worker_fused(la3_server* s, la3_ctx* ctx) {
  struct Running {
    std::vector<Query> qs;
    GraphEntry ge{};  // copied under the lock: add_graph may grow the table
    uint64_t bytes = 0;
    std::vector<uint8_t> out;
    double start = 0;
  };
  la3_session* sess = nullptr;
  const la3_status cst = la3_session_create(ctx, s->cfg.ppr_alpha, &sess);
  std::map<uint64_t, Running> running;
  for (;;) {
    std::vector<Running> admit;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop) {
          lk.unlock();
          if (sess) la3_session_destroy(sess);
          la3_ctx_destroy(ctx);
          return;
        }
        uint32_t free = sess ? la3_session_free_slots(sess) : 0;
        while (free > 0 || !sess) {
          uint64_t bytes = 0;
          const uint32_t cap = sess ? std::min<uint32_t>(s->cfg.max_batch, 256 * free) : s->cfg.max_batch;
          std::vector<Query> qs = pick_group(s, cap, bytes, false);
          if (qs.empty()) break;
          if (!sess) {  // no session: every query fails with the creation status
            complete(s, qs, (uint32_t)qs.size(), cst, s->now(), s->now(), nullptr, bytes);
            continue;
          }
          free -= ((uint32_t)qs.size() + 255) / 256;
          Running r;
          r.ge = s->graphs[qs.front().graph];
          r.qs = std::move(qs);
          r.bytes = bytes;
          admit.push_back(std::move(r));
        }
        if (!sess) s->cv.notify_all();
        if (!admit.empty() || !running.empty()) break;
        s->cv.wait(lk);
      }
    }
    const double now = s->now();
    std::vector<std::pair<Running, la3_status>> failed;
    for (Running& r : admit) {
      const Query& h = r.qs.front();
      const GraphEntry& ge = r.ge;
      std::vector<uint32_t> roots;
      for (const Query& q : r.qs) roots.push_back(q.root);
      if (s->cfg.keep_results) r.out.resize((size_t)r.qs.size() * ge.nslots * elem_bytes(h.kind));
      la3_batch b;
      std::memset(&b, 0, sizeof(b));
      b.graph = ge.g;
      b.kind = h.kind;
      b.k = (uint32_t)r.qs.size();
      b.roots = roots.data();
      b.iters = h.iters;
      b.out = r.out.empty() ? nullptr : r.out.data();
      r.start = now;
      uint64_t id = 0;
      const la3_status st = la3_session_add(sess, &b, &id);
      if (st) failed.push_back({std::move(r), st});
      else running[id] = std::move(r);
    }
    uint32_t nrun = 0;
    la3_status sst = running.empty() ? LA3_OK : la3_session_step(sess, &nrun);
    uint64_t ids[64];
    uint32_t its[64];
    int32_t sts[64];
    uint32_t n = 0;
    std::vector<std::pair<uint64_t, la3_status>> fin;
    do {
      la3_session_poll(sess, ids, its, sts, 64, &n);
      for (uint32_t i = 0; i < n; i++) fin.push_back({ids[i], (la3_status)sts[i]});
    } while (n == 64);
    if (sst) {  // a failed step ends every running batch with its status
      for (auto& kv : running) fin.push_back({kv.first, sst});
    }
    const double done_t = s->now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (auto& f : failed) complete(s, f.first.qs, (uint32_t)f.first.qs.size(), f.second, f.first.start, done_t,
                                      nullptr, f.first.bytes);
      for (auto& f : fin) {
        auto it = running.find(f.first);
        if (it == running.end()) continue;
        Running& r = it->second;
        complete(s, r.qs, (uint32_t)r.qs.size(), f.second, r.start, done_t,
                 r.out.empty() ? nullptr : r.out.data(), r.bytes);
        running.erase(it);
      }
    }
    if (!fin.empty() || !failed.empty()) s->cv.notify_all();
    if (sst) {  // the session is unusable after a failed step: start a new one
      la3_session_destroy(sess);
      sess = nullptr;
      if (la3_session_create(ctx, s->cfg.ppr_alpha, &sess) != LA3_OK) sess = nullptr;
    }
  }
}

// la3_server_warm on one worker: each program on each graph, one query and
// max_batch queries (roots spread over the vertex ids), on the worker's
// context, results left on the device; statuses are not kept (a program a
// graph does not support fails at once).
void warm_worker(la3_server* s, la3_ctx* ctx, const std::vector<GraphEntry>& graphs) {
  const uint32_t kmax = s->cfg.max_batch;
  for (const GraphEntry& ge : graphs) {
    std::vector<uint32_t> roots(kmax);
    for (uint32_t i = 0; i < kmax; i++) roots[i] = (uint32_t)(((uint64_t)i * ge.nslots) / kmax);
    (void)la3_bfs_ctx(ctx, ge.g, roots[0], 0, nullptr, nullptr, nullptr);
    (void)la3_bfs_batch_ctx(ctx, ge.g, roots.data(), kmax, 0, nullptr, nullptr);
    if (ge.weighted && ge.directed) {
      (void)la3_sssp_ctx(ctx, ge.g, roots[0], 0, nullptr, nullptr);
      (void)la3_sssp_batch_ctx(ctx, ge.g, roots.data(), kmax, 0, nullptr, nullptr);
    }
    if (ge.directed) (void)la3_ppr_batch_ctx(ctx, ge.g, roots.data(), kmax, 1, s->cfg.ppr_alpha, nullptr, nullptr, nullptr);
  }
  (void)la3_ctx_synchronize(ctx);
}

void worker_main(la3_server* s, la3_ctx* ctx) {
  if (s->cfg.fused) return worker_fused(s, ctx);
  void* host = nullptr;
  uint64_t host_bytes = 0;
  uint64_t warmed = 0;
  for (;;) {
    std::vector<Query> batch;
    uint64_t bytes = 0;
    GraphEntry ge{};  // copied under the lock: add_graph may grow the table
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop) {
          lk.unlock();
          if (host) la3_host_free(host);
          la3_ctx_destroy(ctx);
          return;
        }
        if (warmed < s->warm_gen) {
          const std::vector<GraphEntry> graphs = s->graphs;
          warmed = s->warm_gen;
          lk.unlock();
          warm_worker(s, ctx, graphs);
          lk.lock();
          s->warm_done++;
          s->cv.notify_all();
          continue;
        }
        batch = pick_group(s, s->cfg.max_batch, bytes, true);
        if (!batch.empty()) break;
        s->cv.wait(lk);
      }
      ge = s->graphs[batch.front().graph];
    }
    const Query& h = batch.front();
    const uint32_t k = (uint32_t)batch.size();
    std::vector<uint32_t> roots(k);
    for (uint32_t i = 0; i < k; i++) roots[i] = batch[i].root;
    const double start = s->now();
    void* out = nullptr;
    const uint64_t need = (uint64_t)k * ge.nslots * elem_bytes(h.kind);
    if (s->cfg.keep_results) {
      if (host_bytes < need) {
        if (host) la3_host_free(host);
        host = nullptr;
        host_bytes = 0;
        if (la3_host_alloc(need, &host) == LA3_OK) host_bytes = need;
      }
      out = host;
    }
    // A lone BFS or SSSP query runs on the single-query program, which is
    // faster for one root; its levels / distances are exactly the batched
    // ones (integer results, tested equal for every root).
    la3_status st;
    if (s->cfg.keep_results && !out) st = LA3_ENOMEM;
    else if (h.kind == LA3_Q_BFS && k == 1)
      st = la3_bfs_ctx(ctx, ge.g, roots[0], 0, (uint8_t*)out, nullptr, nullptr);
    else if (h.kind == LA3_Q_SSSP && k == 1)
      st = la3_sssp_ctx(ctx, ge.g, roots[0], 0, (uint32_t*)out, nullptr);
    else if (h.kind == LA3_Q_BFS)
      st = la3_bfs_batch_ctx(ctx, ge.g, roots.data(), k, 0, (uint8_t*)out, nullptr);
    else if (h.kind == LA3_Q_SSSP)
      st = la3_sssp_batch_ctx(ctx, ge.g, roots.data(), k, 0, (uint32_t*)out, nullptr);
    else
      st = la3_ppr_batch_ctx(ctx, ge.g, roots.data(), k, h.iters, s->cfg.ppr_alpha, (float*)out,
                             nullptr, nullptr);
    if (st == LA3_OK && !out) st = la3_ctx_synchronize(ctx);
    const double fin = s->now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      complete(s, batch, k, st, start, fin, (const uint8_t*)out, bytes);
      s->running_groups.erase(group_of(h));
    }
    s->cv.notify_all();
  }
}

}  // namespace

extern "C" {

la3_status la3_server_create(int device, const la3_server_config* cfg, la3_server** out) {
  if (!cfg || !out || cfg->nworkers == 0 || cfg->nworkers > 64 || cfg->max_batch == 0 ||
      cfg->max_batch > 1024 || !(cfg->ppr_alpha > 0 && cfg->ppr_alpha < 1) || !(cfg->max_wait_s >= 0))
    return LA3_EINVAL;
  std::vector<la3_ctx*> ctxs;
  for (uint32_t i = 0; i < cfg->nworkers; i++) {
    la3_ctx* c = nullptr;
    la3_status st = la3_ctx_create(device, &c);
    if (st) {
      for (la3_ctx* x : ctxs) la3_ctx_destroy(x);
      return st;
    }
    ctxs.push_back(c);
  }
  auto* s = new la3_server();
  s->cfg = *cfg;
  s->device = device;
  for (la3_ctx* c : ctxs) s->workers.emplace_back(worker_main, s, c);
  *out = s;
  return LA3_OK;
}

la3_status la3_server_add_graph(la3_server* s, la3_graph* g, uint32_t* graph_id) {
  if (!s || !g || !graph_id) return LA3_EINVAL;
  la3_graph_info info;
  la3_status st = la3_graph_info_get(g, &info);
  if (st) return st;
  std::lock_guard<std::mutex> lk(s->mu);
  *graph_id = (uint32_t)s->graphs.size();
  s->graphs.push_back({g, info.nslots, info.weighted != 0, info.directed != 0, info.nnz});
  return LA3_OK;
}

la3_status la3_server_warm(la3_server* s) {
  if (!s) return LA3_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->cfg.fused) return LA3_OK;
  s->warm_gen++;
  const uint64_t target = s->warm_gen * s->workers.size();
  s->cv.notify_all();
  s->cv.wait(lk, [&] { return s->warm_done >= target || s->stop; });
  return LA3_OK;
}

uint64_t la3_server_query_bytes(const la3_server* s, la3_query_kind kind, uint32_t graph_id) {
  if (!s) return 0;
  std::lock_guard<std::mutex> lk(const_cast<la3_server*>(s)->mu);
  if (graph_id >= s->graphs.size()) return 0;
  return (uint64_t)(per_vertex_bytes(kind) * s->graphs[graph_id].nslots);
}

la3_status la3_server_submit(la3_server* s, la3_query_kind kind, uint32_t graph_id, uint32_t root,
                             uint32_t iters, uint64_t* qid) {
  if (!s || kind < LA3_Q_BFS || kind > LA3_Q_PPR) return LA3_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (graph_id >= s->graphs.size()) return LA3_EINVAL;
    if (root >= s->graphs[graph_id].nslots) return LA3_EINVAL;
    if (kind == LA3_Q_PPR && iters == 0) return LA3_EINVAL;
    Query q;
    q.qid = s->next_qid++;
    q.kind = kind;
    q.graph = graph_id;
    q.root = root;
    q.iters = kind == LA3_Q_PPR ? iters : 0;
    q.submit_s = s->now();
    s->pending.push_back(q);
    s->submitted++;
    if (qid) *qid = q.qid;
  }
  s->cv.notify_all();
  return LA3_OK;
}

la3_status la3_server_drain(la3_server* s) {
  if (!s) return LA3_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv.wait(lk, [&] { return s->completed == s->submitted; });
  return LA3_OK;
}

la3_status la3_server_completed(la3_server* s, la3_query_info* out, uint32_t cap, uint32_t* n) {
  if (!s || !n || (cap && !out)) return LA3_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  uint32_t i = 0;
  while (i < cap && !s->done.empty()) {
    out[i++] = s->done.front();
    s->done.pop_front();
  }
  *n = i;
  return LA3_OK;
}

la3_status la3_server_result(la3_server* s, uint64_t qid, void* dst, uint64_t bytes) {
  if (!s || !dst) return LA3_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->results.find(qid);
  if (it == s->results.end()) return LA3_EAGAIN;
  if (bytes < it->second.size()) return LA3_EINVAL;
  std::memcpy(dst, it->second.data(), it->second.size());
  s->results.erase(it);
  return LA3_OK;
}

double la3_server_clock(const la3_server* s) { return s ? s->now() : 0.0; }

void la3_server_destroy(la3_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (auto& t : s->workers) t.join();
  delete s;
}

}  // extern "C"
