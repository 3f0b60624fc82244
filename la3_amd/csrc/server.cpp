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
#include <thread>
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
  bool stop = false;
  std::vector<std::thread> workers;

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

// Mixed admission (cfg.fused): pending queries in arrival order join groups
// keyed by (kind, graph, iters), each up to max_batch, while the admitted
// state bytes fit the budget (a lone query is admitted when nothing runs)
// and the groups fit one fused run (at most 64 tasks of 256 queries).
// Called with the lock held; moves the picked queries into `groups`.
uint64_t pick_mixed(la3_server* s, std::vector<std::vector<Query>>& groups) {
  struct G { la3_query_kind kind; uint32_t graph, iters; std::vector<size_t> idx; };
  std::vector<G> gs;
  uint64_t bytes = 0;
  uint32_t tasks = 0;
  for (size_t i = 0; i < s->pending.size(); i++) {
    const Query& q = s->pending[i];
    G* gp = nullptr;
    for (G& g : gs)
      if (g.kind == q.kind && g.graph == q.graph && g.iters == q.iters) gp = &g;
    const uint32_t have = gp ? (uint32_t)gp->idx.size() : 0;
    if (have >= s->cfg.max_batch) continue;
    const uint64_t db = batch_bytes(s, q.kind, q.graph, have + 1) - (have ? batch_bytes(s, q.kind, q.graph, have) : 0);
    const uint32_t dt = (have % 256 == 0) ? 1 : 0;
    const bool first = gs.empty();
    if (!first && (s->in_flight_bytes + bytes + db > s->cfg.state_budget || tasks + dt > 64)) break;
    if (first && s->in_flight_bytes + db > s->cfg.state_budget && s->running_batches != 0) break;
    if (!gp) {
      gs.push_back({q.kind, q.graph, q.iters, {}});
      gp = &gs.back();
    }
    gp->idx.push_back(i);
    bytes += db;
    tasks += dt;
  }
  if (gs.empty()) return 0;
  std::vector<size_t> all;
  for (G& g : gs) {
    std::vector<Query> v;
    for (size_t i : g.idx) { v.push_back(s->pending[i]); all.push_back(i); }
    groups.push_back(std::move(v));
  }
  std::sort(all.begin(), all.end());
  for (size_t j = all.size(); j-- > 0;) s->pending.erase(s->pending.begin() + all[j]);
  return bytes;
}

// Runs the picked groups as one fused batch run and records completions.
void run_mixed(la3_server* s, la3_ctx* ctx, std::vector<std::vector<Query>>& groups, uint64_t bytes,
               void*& host, uint64_t& host_bytes) {
  const double start = s->now();
  std::vector<la3_batch> bs(groups.size());
  std::vector<std::vector<uint32_t>> roots(groups.size());
  std::vector<size_t> off(groups.size());
  uint64_t need = 0;
  for (size_t i = 0; i < groups.size(); i++) {
    const Query& h = groups[i].front();
    off[i] = need;
    need += (uint64_t)groups[i].size() * s->graphs[h.graph].nslots * elem_bytes(h.kind);
  }
  la3_status st = LA3_OK;
  if (s->cfg.keep_results && host_bytes < need) {
    if (host) la3_host_free(host);
    host = nullptr;
    host_bytes = 0;
    if (la3_host_alloc(need, &host) == LA3_OK) host_bytes = need;
    else st = LA3_ENOMEM;
  }
  for (size_t i = 0; i < groups.size(); i++) {
    const Query& h = groups[i].front();
    for (const Query& q : groups[i]) roots[i].push_back(q.root);
    la3_batch& b = bs[i];
    std::memset(&b, 0, sizeof(b));
    b.graph = s->graphs[h.graph].g;
    b.kind = h.kind;
    b.k = (uint32_t)groups[i].size();
    b.roots = roots[i].data();
    b.iters = h.iters;
    b.out = s->cfg.keep_results && host ? (uint8_t*)host + off[i] : nullptr;
  }
  if (st == LA3_OK) st = la3_batch_run(ctx, bs.data(), (uint32_t)bs.size(), s->cfg.ppr_alpha);
  const double fin = s->now();
  uint32_t total = 0;
  for (auto& g : groups) total += (uint32_t)g.size();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (size_t i = 0; i < groups.size(); i++) {
      const uint32_t k = (uint32_t)groups[i].size();
      for (uint32_t j = 0; j < k; j++) {
        const Query& q = groups[i][j];
        la3_query_info info{};
        info.qid = q.qid; info.kind = q.kind; info.graph = q.graph; info.root = q.root;
        info.batch_size = total; info.status = st; info.submit_s = q.submit_s;
        info.start_s = start; info.done_s = fin;
        s->done.push_back(info);
        if (s->cfg.keep_results && st == LA3_OK) {
          const size_t row = (size_t)s->graphs[q.graph].nslots * elem_bytes(q.kind);
          std::vector<uint8_t> v(row);
          std::memcpy(v.data(), (const uint8_t*)host + off[i] + (size_t)j * row, row);
          s->results[q.qid] = std::move(v);
        }
      }
    }
    s->completed += total;
    s->in_flight_bytes -= bytes;
    s->running_batches--;
  }
  s->cv.notify_all();
}

void worker_main(la3_server* s, la3_ctx* ctx) {
  void* host = nullptr;
  uint64_t host_bytes = 0;
  for (;;) {
    std::vector<Query> batch;
    std::vector<std::vector<Query>> groups;
    uint64_t bytes = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop) {
          lk.unlock();
          if (host) la3_host_free(host);
          la3_ctx_destroy(ctx);
          return;
        }
        if (!s->pending.empty() && s->cfg.fused) {
          bytes = pick_mixed(s, groups);
          if (!groups.empty()) {
            s->in_flight_bytes += bytes;
            s->running_batches++;
            break;
          }
        } else if (!s->pending.empty()) {
          const Query& head = s->pending.front();
          // Grow the batch with same-key queries in arrival order while the
          // budget admits it; a lone query is admitted when nothing runs.
          std::vector<size_t> pick;
          for (size_t i = 0; i < s->pending.size() && pick.size() < s->cfg.max_batch; i++) {
            const Query& q = s->pending[i];
            if (q.kind != head.kind || q.graph != head.graph || q.iters != head.iters) continue;
            const uint64_t b = batch_bytes(s, head.kind, head.graph, (uint32_t)pick.size() + 1);
            if (s->in_flight_bytes + b > s->cfg.state_budget && !(pick.empty() && s->running_batches == 0))
              break;
            pick.push_back(i);
          }
          if (!pick.empty()) {
            bytes = batch_bytes(s, head.kind, head.graph, (uint32_t)pick.size());
            for (size_t j = pick.size(); j-- > 0;) {
              batch.push_back(s->pending[pick[j]]);
              s->pending.erase(s->pending.begin() + pick[j]);
            }
            std::reverse(batch.begin(), batch.end());
            s->in_flight_bytes += bytes;
            s->running_batches++;
            break;
          }
        }
        s->cv.wait(lk);
      }
    }
    if (!groups.empty()) {
      run_mixed(s, ctx, groups, bytes, host, host_bytes);
      continue;
    }
    const Query& h = batch.front();
    const uint32_t k = (uint32_t)batch.size();
    const GraphEntry& ge = s->graphs[h.graph];
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
    la3_status st;
    if (s->cfg.keep_results && !out) st = LA3_ENOMEM;
    else if (h.kind == LA3_Q_BFS)
      st = la3_bfs_batch_ctx(ctx, ge.g, roots.data(), k, (uint8_t*)out, nullptr);
    else if (h.kind == LA3_Q_SSSP)
      st = la3_sssp_batch_ctx(ctx, ge.g, roots.data(), k, (uint32_t*)out, nullptr);
    else
      st = la3_ppr_batch_ctx(ctx, ge.g, roots.data(), k, h.iters, s->cfg.ppr_alpha, (float*)out,
                             nullptr, nullptr);
    if (st == LA3_OK && !out) st = la3_ctx_synchronize(ctx);
    const double fin = s->now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (uint32_t i = 0; i < k; i++) {
        const Query& q = batch[i];
        la3_query_info info{};
        info.qid = q.qid; info.kind = q.kind; info.graph = q.graph; info.root = q.root;
        info.batch_size = k; info.status = st; info.submit_s = q.submit_s;
        info.start_s = start; info.done_s = fin;
        s->done.push_back(info);
        if (s->cfg.keep_results && st == LA3_OK) {
          const size_t row = (size_t)ge.nslots * elem_bytes(h.kind);
          std::vector<uint8_t> v(row);
          std::memcpy(v.data(), (const uint8_t*)out + (size_t)i * row, row);
          s->results[q.qid] = std::move(v);
        }
      }
      s->completed += k;
      s->in_flight_bytes -= bytes;
      s->running_batches--;
    }
    s->cv.notify_all();
  }
}

}  // namespace

extern "C" {

la3_status la3_server_create(int device, const la3_server_config* cfg, la3_server** out) {
  if (!cfg || !out || cfg->nworkers == 0 || cfg->nworkers > 64 || cfg->max_batch == 0 ||
      cfg->max_batch > 1024 || !(cfg->ppr_alpha > 0 && cfg->ppr_alpha < 1))
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
  s->graphs.push_back({g, info.nslots});
  return LA3_OK;
}

uint64_t la3_server_query_bytes(const la3_server* s, la3_query_kind kind, uint32_t graph_id) {
  if (!s || graph_id >= s->graphs.size()) return 0;
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
