// C-ABI implementation (include/la3.h): contexts, device-resident graphs,
// per-launch HIP-event profiling and the host drivers of each program.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "../../include/la3.h"
#include "../../include/la3_vertex_program.hpp"
#include "internal.h"
#include "kernels_dev.h"
#include "kernels_batch.h"
#include "kernels_fused.h"
#include "kernels_topk.h"
#include "kernels_ingest.h"

#include <hipcub/hipcub.hpp>

using namespace la3;

// ------------------------------------------------------- device blocks ---
// Every device block the library allocates (context scratch, graph arrays,
// ingest temporaries) is registered, with its size and owner, from hipMalloc
// to hipFree. Every host <-> device copy and memset checks its device range
// against the registry first: a range that is not inside one live block (a
// stale pointer to a freed or regrown buffer) fails with LA3_EDEVICE and a
// message naming the range, instead of reaching the runtime, which treats an
// unknown pointer as host memory. Contexts and threads share one registry.
namespace {

struct BlockInfo {
  size_t bytes;
  const char* owner;  // static string
};
std::mutex g_blocks_mu;
std::map<uintptr_t, BlockInfo> g_blocks;
thread_local std::string g_copy_err;

hipError_t dev_alloc(void** p, size_t bytes, const char* owner) {
  *p = nullptr;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return e;
  }
  std::lock_guard<std::mutex> lk(g_blocks_mu);
  g_blocks[(uintptr_t)*p] = BlockInfo{bytes, owner};
  return hipSuccess;
}

void dev_free(void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_blocks_mu);
    g_blocks.erase((uintptr_t)p);
  }
  (void)hipFree(p);
}

// True iff [p, p + bytes) lies inside one live block.
bool dev_covers(const void* p, size_t bytes) {
  const uintptr_t a = (uintptr_t)p;
  std::lock_guard<std::mutex> lk(g_blocks_mu);
  auto it = g_blocks.upper_bound(a);
  if (it == g_blocks.begin()) return false;
  --it;
  return a >= it->first && a - it->first <= it->second.bytes && bytes <= it->second.bytes - (a - it->first);
}

hipError_t range_error(const void* p, size_t bytes, const char* what) {
  char buf[160];
  snprintf(buf, sizeof(buf), "%s: device range [%p, +%zu) is not inside a live device block", what, p, bytes);
  g_copy_err = buf;
  return hipErrorInvalidValue;
}

// hipMemcpy(Async) with the device side(s) checked against the registry.
// Zero-byte copies are no-ops.
hipError_t dcopy_async(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  if ((kind == hipMemcpyDeviceToHost || kind == hipMemcpyDeviceToDevice) && !dev_covers(src, bytes))
    return range_error(src, bytes, "copy source");
  if ((kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToDevice) && !dev_covers(dst, bytes))
    return range_error(dst, bytes, "copy destination");
  return hipMemcpyAsync(dst, src, bytes, kind, s);
}

hipError_t dcopy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  if (bytes == 0) return hipSuccess;
  if ((kind == hipMemcpyDeviceToHost || kind == hipMemcpyDeviceToDevice) && !dev_covers(src, bytes))
    return range_error(src, bytes, "copy source");
  if ((kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToDevice) && !dev_covers(dst, bytes))
    return range_error(dst, bytes, "copy destination");
  return hipMemcpy(dst, src, bytes, kind);
}

hipError_t dset_async(void* dst, int value, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  if (!dev_covers(dst, bytes)) return range_error(dst, bytes, "memset");
  return hipMemsetAsync(dst, value, bytes, s);
}

std::string hip_error_text(hipError_t e) {
  std::string m = hipGetErrorString(e);
  if (!g_copy_err.empty()) {
    m = g_copy_err;
    g_copy_err.clear();
  }
  return m;
}

}  // namespace

// ---------------------------------------------- allocation-free scratch ---
// Work buffers of la3_bfs_noalloc (levels, marks and the two frontier
// queues), for callers that cannot allocate device memory. The array exists
// once per device from the library's load to the end of the process; every
// call on a device uses the same copy, so calls that overlap on one device
// race. Host code uses the address hipGetSymbolAddress returns for the
// device (noalloc::base), never the array's name.
namespace noalloc {
constexpr size_t SCRATCH_BYTES = 64ull << 20;
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];
std::mutex mu;
void* dev_ptr[MAX_DEVICES] = {};
bool looked_up[MAX_DEVICES] = {};

// Called with `device` current (la3_ctx_create): looks the array's address
// up once per device, checks the result and registers the block so that
// counter copies out of it pass the range checks.
hipError_t setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (looked_up[device]) return dev_ptr[device] ? hipSuccess : hipErrorInvalidSymbol;
  looked_up[device] = true;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return e != hipSuccess ? e : hipErrorInvalidSymbol;
  }
  dev_ptr[device] = p;
  std::lock_guard<std::mutex> lb(g_blocks_mu);
  g_blocks[(uintptr_t)p] = BlockInfo{SCRATCH_BYTES, "allocation-free scratch"};
  return hipSuccess;
}

void* base(int device) {
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  return dev_ptr[device];
}
}  // namespace noalloc

// ------------------------------------------------------------------ ctx ---
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct la3_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool profiling = false;
  struct Ev { const char* name; hipEvent_t a, b; double bytes; uint64_t entries; };
  struct Stat { uint64_t launches = 0; double ms = 0, bytes = 0; uint64_t entries = 0; };
  std::map<std::string, Stat> stats;  // cumulative, while profiling
  std::vector<Ev> evs;
  std::vector<hipEvent_t> free_events;
  la3_timing last{};
  uint64_t edges_scanned = 0;
  std::map<std::string, DevBuf> scratch;
  void* host_ctr = nullptr;  // pinned, for frontier counters: [0, 4096) and PIPE_DEPTH ring slots after it
  hipEvent_t pipe_ev[4] = {};  // completion of each ring slot's copy (pipelined SSSP steps)
  int ncu = 256;
  int refs = 1;              // the caller's reference plus one per live graph
  const void* last_result = nullptr;  // device buffer of the last program's main output
  size_t last_result_bytes = 0;
  // Fused batch runs: task table and control block (fixed addresses for the
  // context's life) and the captured step graph that reads them.
  FTask* fz_tasks = nullptr;
  uint8_t* fz_ctl = nullptr;
  std::map<uint64_t, hipGraphExec_t> fz_exec;  // by (steps per replay, A grid, launch mask)
  bool fz_busy = false;             // a session holds the task table
  bool ingest_device = true;        // build loaded graphs on the device (else on the host)
  uint64_t scratch_epoch = 0;       // bumped when all scratch is released
};

namespace {
void ctx_release(la3_ctx* c);
}

struct la3_graph {
  la3_ctx* ctx = nullptr;
  uint32_t n = 0, nslots = 0, nreg = 0, nsink = 0, nsrc = 0, niso = 0, ncols = 0, nrows = 0;
  bool directed = true, weighted = false;
  int mode = 0;
  Hasher h;
  uint64_t nnz = 0, nnz_reg = 0, nnz_sink = 0, nrecords = 0, device_bytes = 0;
  uint64_t pnnz_reg = 0, nnz_long_reg = 0;
  uint32_t max_pseg = 0;  // most PSEG_ENTRIES-entry segments of one row
  // Near/far thresholds of the SSSP relaxation (0 = plain Bellman-Ford),
  // single-query and batched; set by la3_graph_set_tuning or la3_autotune.
  uint32_t sssp_delta = 0, mssp_delta = 0;
  // Fused runs: queries per task by kind (BFS, SSSP, PPR; 64, 128 or 256)
  // and blocks per CU of the A / pull launches; la3_autotune sets them.
  uint32_t task_k[3] = {256, 256, 256};
  uint32_t fused_blocks = 32;
  std::vector<uint32_t> hid_of, id_of, slot_of_id;
  std::vector<void*> allocs;
  DG dg{};
  std::unique_ptr<HostLayout> host;  // kept instead of device arrays on a host-only context
  // Fixed-iteration PR / PPR loop captured as a hipGraph, with the arguments
  // it was captured for (iterations, buffers, seed, alpha).
  hipGraphExec_t pr_exec = nullptr;
  std::vector<uint64_t> pr_key;
};

namespace {

// Frees the context once the caller has destroyed it and every graph on it
// has been freed.
void ctx_release(la3_ctx* c) {
  if (--c->refs > 0) return;
  if (c->device >= 0) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->scratch) dev_free(kv.second.p);
    for (auto& e : c->evs) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    for (auto e : c->free_events) (void)hipEventDestroy(e);
    for (auto& kv : c->fz_exec) (void)hipGraphExecDestroy(kv.second);
    if (c->host_ctr) (void)hipHostFree(c->host_ctr);
    for (auto e : c->pipe_ev)
      if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
  }
  delete c;
}

la3_status fail(la3_ctx* c, la3_status s, const std::string& msg) {
  if (c) c->err = msg;
  return s;
}

#define HIPC(ctx, call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail((ctx), LA3_EDEVICE, std::string(#call) + ": " + hip_error_text(e_));     \
  } while (0)

hipEvent_t get_event(la3_ctx* c) {
  if (!c->free_events.empty()) { hipEvent_t e = c->free_events.back(); c->free_events.pop_back(); return e; }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Launches a kernel on the context's stream; with profiling on, brackets it
// with an event pair. Returns the event slot (or -1) so that the calling
// program code can attach the launch's algorithmic bytes once its counters
// are known.
template <typename... KArgs, typename... Args>
int launch(la3_ctx* c, const char* name, void (*k)(KArgs...), uint32_t blocks, Args... args) {
  hipEvent_t a = nullptr, b = nullptr;
  if (blocks == 0) return -1;
  if (c->profiling) { a = get_event(c); b = get_event(c); (void)hipEventRecord(a, c->stream); }
  hipLaunchKernelGGL(k, dim3(blocks), dim3(BLOCK), 0, c->stream, args...);
  if (!c->profiling) return -1;
  (void)hipEventRecord(b, c->stream);
  c->evs.push_back({name, a, b, 0.0, 0});
  return (int)c->evs.size() - 1;
}

void annotate(la3_ctx* c, int ev, double bytes, uint64_t entries) {
  if (ev < 0 || ev >= (int)c->evs.size()) return;
  c->evs[ev].bytes += bytes;
  c->evs[ev].entries += entries;
}

uint32_t blocks_for(uint64_t threads, uint32_t cap = 8192) {
  uint64_t b = (threads + BLOCK - 1) / BLOCK;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, cap));
}

// Starts the timing window of one API call.
void timing_begin(la3_ctx* c) {
  for (auto& e : c->evs) {  // launches outside any timing window
    c->free_events.push_back(e.a);
    c->free_events.push_back(e.b);
  }
  c->evs.clear();
  c->edges_scanned = 0;
}

// Closes it (the stream must be synchronized): per-kernel totals.
void timing_end(la3_ctx* c) {
  la3_timing t{};
  std::map<std::string, std::pair<double, uint64_t>> per;
  for (auto& e : c->evs) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e.a, e.b);
    auto& p = per[e.name];
    p.first += ms;
    p.second++;
    t.total_ms += ms;
    auto& st = c->stats[e.name];
    st.launches++;
    st.ms += ms;
    st.bytes += e.bytes;
    st.entries += e.entries;
    c->free_events.push_back(e.a);
    c->free_events.push_back(e.b);
  }
  c->evs.clear();
  for (auto& kv : per)
    if (kv.second.first > t.dominant_ms) {
      t.dominant_ms = kv.second.first;
      t.dominant_launches = kv.second.second;
      snprintf(t.dominant_kernel, sizeof(t.dominant_kernel), "%s", kv.first.c_str());
    }
  t.edges_scanned = c->edges_scanned;
  c->last = t;
}

// A device-resident result (la3_result_fetch / la3_result_topk) stops being
// one when the block holding it is freed: forget it then.
void forget_result_in(la3_ctx* c, const void* p, size_t bytes) {
  const uintptr_t r = (uintptr_t)c->last_result, a = (uintptr_t)p;
  if (c->last_result && r >= a && r - a < bytes) {
    c->last_result = nullptr;
    c->last_result_bytes = 0;
  }
}

// Grow-only named device scratch.
void* scratch(la3_ctx* c, const char* name, size_t bytes) {
  DevBuf& b = c->scratch[name];
  if (b.bytes >= bytes) return b.p;
  if (b.p) {
    forget_result_in(c, b.p, b.bytes);
    dev_free(b.p);
  }
  b.p = nullptr;
  b.bytes = 0;
  size_t want = std::max<size_t>(bytes, 256);
  if (dev_alloc(&b.p, want, "context scratch") != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return nullptr;
  }
  b.bytes = want;
  return b.p;
}

// Frees the cached scratch whose names start with `prefix` (all with "").
void scratch_release(la3_ctx* c, const std::string& prefix) {
  (void)hipStreamSynchronize(c->stream);
  for (auto it = c->scratch.begin(); it != c->scratch.end();) {
    if (it->first.compare(0, prefix.size(), prefix) == 0) {
      if (it->second.p) {
        forget_result_in(c, it->second.p, it->second.bytes);
        dev_free(it->second.p);
      }
      it = c->scratch.erase(it);
    } else {
      ++it;
    }
  }
}

template <class T>
la3_status upload(la3_graph* g, const std::vector<T>& v, const T** out) {
  void* p = nullptr;
  size_t bytes = std::max<size_t>(v.size() * sizeof(T), 4);
  HIPC(g->ctx, dev_alloc(&p, bytes, "graph array"));
  g->allocs.push_back(p);
  g->device_bytes += bytes;
  if (!v.empty()) HIPC(g->ctx, dcopy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = (const T*)p;
  return LA3_OK;
}

la3_status graph_msg_order(la3_ctx* c, la3_graph* g);

la3_status graph_from_layout(la3_ctx* ctx, HostLayout& L, la3_graph** out) {
  auto g = std::make_unique<la3_graph>();
  g->ctx = ctx;
  g->n = L.n; g->nslots = L.nslots; g->nreg = L.nreg; g->nsink = L.nsink; g->nsrc = L.nsrc;
  g->niso = L.niso; g->ncols = L.ncols; g->nrows = L.nrows;
  g->directed = L.directed; g->weighted = L.weighted; g->mode = L.mode; g->h = L.h;
  g->nnz = L.nnz; g->nnz_reg = L.nnz_reg; g->nnz_sink = L.nnz_sink; g->nrecords = L.nrecords;
  g->pnnz_reg = L.pnnz_reg; g->nnz_long_reg = L.nnz_long_reg;
  if (ctx->device < 0) {
    g->hid_of = L.hid_of;
    g->id_of = L.id_of;
    g->slot_of_id = L.slot_of_id;
    g->host = std::make_unique<HostLayout>(std::move(L));
    ctx->refs++;
    *out = g.release();
    return LA3_OK;
  }
  HIPC(ctx, hipSetDevice(ctx->device));
  DG& d = g->dg;
  d.nslots = L.nslots; d.nreg = L.nreg; d.nsink = L.nsink; d.nsrc = L.nsrc; d.ncols = L.ncols;
  d.nrows = L.nrows; d.nnz = L.nnz; d.nlong = (uint32_t)L.long_rows.size();
  d.nplong = (uint32_t)L.plong.size(); d.nplong_reg = L.nplong_reg;
  d.npseg = (uint32_t)L.pseg_row.size(); d.npseg_reg = L.npseg_reg;
  for (size_t i = 0; i + 1 < L.pseg_first.size(); i++)
    g->max_pseg = std::max(g->max_pseg, L.pseg_first[i + 1] - L.pseg_first[i]);
  la3_status s;
  auto up = [&](const std::vector<uint32_t>& v, const uint32_t** o) { return upload(g.get(), v, o); };
  if ((s = up(L.col_ptr, &d.col_ptr)) || (s = up(L.col_split, &d.col_split)) ||
      (s = up(L.col_rows, &d.col_rows)) || (s = up(L.row_ptr, &d.row_ptr)) ||
      (s = up(L.row_split, &d.row_split)) || (s = up(L.row_cols, &d.row_cols)) ||
      (s = up(L.id_of, &d.id_of)) || (s = up(L.hid_of, &d.hid_of)) ||
      (s = up(L.long_rows, &d.long_rows)) || (s = up(L.slot_of_id, &d.slot_of_id)) ||
      (s = up(L.plong, &d.plong)) || (s = up(L.pseg_first, &d.pseg_first)) ||
      (s = up(L.pseg_row, &d.pseg_row)) || (s = up(L.pseg_beg, &d.pseg_beg))) {
    for (void* p : g->allocs) dev_free(p);
    return s;
  }
  if (L.weighted) {
    if ((s = up(L.col_w, &d.col_w)) || (s = up(L.row_w, &d.row_w))) {
      for (void* p : g->allocs) dev_free(p);
      return s;
    }
  }
  if ((s = graph_msg_order(ctx, g.get()))) {
    for (void* p : g->allocs) dev_free(p);
    return s;
  }
  g->hid_of.swap(L.hid_of);
  g->id_of.swap(L.id_of);
  g->slot_of_id.swap(L.slot_of_id);
  ctx->refs++;
  *out = g.release();
  return LA3_OK;
}

// ------------------------------------------------------ device ingest ---
// Device buffers of one ingest, freed together unless kept for the graph.
struct DevTemps {
  std::vector<void*> ptrs;
  ~DevTemps() { for (void* p : ptrs) dev_free(p); }
  template <class T>
  T* get(size_t count) {
    void* p = nullptr;
    if (dev_alloc(&p, std::max<size_t>(count * sizeof(T), 4), "graph build") != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    ptrs.push_back(p);
    return (T*)p;
  }
  // Hands a buffer over to the graph (not freed here).
  void keep(void* p) {
    for (auto& q : ptrs) if (q == p) q = nullptr;
  }
};

// The layout of kernels_ingest.h built on the device; fills g (counts, DG,
// host id maps). Equal, array for array, to build_layout + graph_from_layout.
la3_status ingest_device(la3_ctx* c, const uint32_t* rec, uint64_t nrec, uint32_t n, bool weighted, int mode,
                         bool bucket, la3_graph* g) {
  hipStream_t st = c->stream;
  const bool directed = mode != 2, reverse = mode == 1 || mode == 4, acyclic = mode == 3 || mode == 4;
  const uint32_t RW = weighted ? 3 : 2;
  const uint32_t S = n + 1;
  DevTemps T;
  g->n = n; g->nslots = S; g->directed = directed; g->weighted = weighted; g->mode = mode;
  g->nrecords = nrec;
  g->h = Hasher(n, bucket);
  IngestHash ih{bucket ? 1u : 0u, (uint64_t)g->h.nparts, (uint64_t)g->h.height, (uint64_t)g->h.max_range};
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  // hipcub temporary storage, grown as needed
  auto temp = [&](size_t need) -> void* {
    if (need <= tmp_bytes) return tmp;
    tmp = T.get<uint8_t>(need);
    tmp_bytes = tmp ? need : 0;
    return tmp;
  };
  auto exclusive_sum = [&](const uint32_t* in, uint32_t* out, uint64_t cnt) -> bool {
    if (cnt == 0) return true;
    size_t need = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, (int)cnt, st) != hipSuccess) return false;
    void* t = temp(need);
    return t && hipcub::DeviceScan::ExclusiveSum(t, need, in, out, (int)cnt, st) == hipSuccess;
  };
  auto sort_pairs = [&](hipcub::DoubleBuffer<uint64_t>& k, hipcub::DoubleBuffer<uint32_t>& v, uint64_t cnt,
                        int bits) -> bool {
    if (cnt == 0) return true;
    size_t need = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, need, k, v, (int)cnt, 0, bits, st) != hipSuccess) return false;
    void* t = temp(need);
    return t && hipcub::DeviceRadixSort::SortPairs(t, need, k, v, (int)cnt, 0, bits, st) == hipSuccess;
  };
  auto total = [&](const uint32_t* in, const uint32_t* ex, uint64_t cnt, uint64_t& out) -> la3_status {
    out = 0;
    if (cnt == 0) return LA3_OK;
    uint32_t a = 0, b = 0;
    HIPC(c, dcopy_async(&a, in + cnt - 1, 4, hipMemcpyDeviceToHost, st));
    HIPC(c, dcopy_async(&b, ex + cnt - 1, 4, hipMemcpyDeviceToHost, st));
    HIPC(c, hipStreamSynchronize(st));
    out = (uint64_t)a + b;
    return LA3_OK;
  };
  const char* oom = "device ingest: allocation failed";
  // 1. records: keep flags, positions, orient + hash
  uint32_t* drec = T.get<uint32_t>(nrec * RW);
  uint32_t* keep = T.get<uint32_t>(nrec);
  uint32_t* pos = T.get<uint32_t>(nrec);
  uint32_t* bad = T.get<uint32_t>(1);
  if (!drec || !keep || !pos || !bad) return fail(c, LA3_ENOMEM, oom);
  if (nrec) HIPC(c, dcopy_async(drec, rec, 4 * nrec * RW, hipMemcpyHostToDevice, st));
  HIPC(c, dset_async(bad, 0, 4, st));
  launch(c, "ingest", k_ing_keep, blocks_for(nrec), (const uint32_t*)drec, nrec, RW, n, directed ? 1u : 2u, keep, bad);
  if (!exclusive_sum(keep, pos, nrec)) return fail(c, LA3_EDEVICE, "device ingest: scan failed");
  uint64_t m = 0;
  la3_status e;
  if ((e = total(keep, pos, nrec, m))) return e;
  uint32_t hbad = 0;
  HIPC(c, dcopy(&hbad, bad, 4, hipMemcpyDeviceToHost));
  if (hbad) return fail(c, LA3_EINVAL, "vertex id out of range [0, nvertices]");
  if (m >= 0xffffffffull) return fail(c, LA3_ENOTSUP, "more than 2^32-1 stored entries");
  uint32_t* R = T.get<uint32_t>(m);
  uint32_t* C = T.get<uint32_t>(m);
  uint32_t* W = weighted ? T.get<uint32_t>(m) : nullptr;
  if (!R || !C || (weighted && !W)) return fail(c, LA3_ENOMEM, oom);
  launch(c, "ingest", k_ing_orient, blocks_for(nrec), (const uint32_t*)drec, nrec, RW, (const uint32_t*)pos,
         (int)directed, (int)reverse, (int)acyclic, ih, R, C, W);
  // 2. classes and slots
  uint8_t* rowf = T.get<uint8_t>(S);
  uint8_t* colf = T.get<uint8_t>(S);
  uint32_t* cls = T.get<uint32_t>(4ull * S);
  uint32_t* rank = T.get<uint32_t>(4ull * S);
  uint32_t* slot = T.get<uint32_t>(S);
  uint32_t* hid_of = T.get<uint32_t>(S);
  uint32_t* id_of = T.get<uint32_t>(S);
  uint32_t* slot_of_id = T.get<uint32_t>(S);
  if (!rowf || !colf || !cls || !rank || !slot || !hid_of || !id_of || !slot_of_id) return fail(c, LA3_ENOMEM, oom);
  HIPC(c, dset_async(rowf, 0, S, st));
  HIPC(c, dset_async(colf, 0, S, st));
  launch(c, "ingest", k_ing_occupy, blocks_for(m), (const uint32_t*)R, (const uint32_t*)C, m, rowf, colf);
  launch(c, "ingest", k_ing_class, blocks_for(S), (const uint8_t*)rowf, (const uint8_t*)colf, S, cls);
  uint64_t cnt[4];
  for (int q = 0; q < 4; q++) {
    if (!exclusive_sum(cls + (uint64_t)q * S, rank + (uint64_t)q * S, S))
      return fail(c, LA3_EDEVICE, "device ingest: scan failed");
    if ((e = total(cls + (uint64_t)q * S, rank + (uint64_t)q * S, S, cnt[q]))) return e;
  }
  g->nreg = (uint32_t)cnt[0]; g->nsink = (uint32_t)cnt[1]; g->nsrc = (uint32_t)cnt[2]; g->niso = (uint32_t)cnt[3];
  g->ncols = g->nreg + g->nsrc;
  g->nrows = g->nreg + g->nsink;
  const uint32_t nreg = g->nreg, nsink = g->nsink, nsrc = g->nsrc, NC = g->ncols, NR = g->nrows;
  launch(c, "ingest", k_ing_slots, blocks_for(S), (const uint8_t*)rowf, (const uint8_t*)colf, S, (const uint32_t*)rank,
         nreg, nreg + nsink, nreg + nsink + nsrc, ih, slot, hid_of, id_of, slot_of_id);
  // 3. entries by (column, row slot), first record of equal keys kept
  uint32_t b = 1;
  while ((1ull << b) < S) b++;
  uint64_t* k0 = T.get<uint64_t>(m);
  uint64_t* k1 = T.get<uint64_t>(m);
  uint32_t* v0 = T.get<uint32_t>(m);
  uint32_t* v1 = T.get<uint32_t>(m);
  uint32_t* first = T.get<uint32_t>(m);
  uint32_t* upos = T.get<uint32_t>(m);
  if (!k0 || !k1 || !v0 || !v1 || !first || !upos) return fail(c, LA3_ENOMEM, oom);
  launch(c, "ingest", k_ing_colkeys, blocks_for(m), (const uint32_t*)R, (const uint32_t*)C, m, (const uint32_t*)slot,
         nreg, nsink, b, k0, v0);
  hipcub::DoubleBuffer<uint64_t> kb(k0, k1);
  hipcub::DoubleBuffer<uint32_t> vb(v0, v1);
  if (!sort_pairs(kb, vb, m, 2 * b)) return fail(c, LA3_EDEVICE, "device ingest: sort failed");
  launch(c, "ingest", k_ing_firsts, blocks_for(m), (const uint64_t*)kb.Current(), m, first);
  if (!exclusive_sum(first, upos, m)) return fail(c, LA3_EDEVICE, "device ingest: scan failed");
  uint64_t nnz = 0;
  if ((e = total(first, upos, m, nnz))) return e;
  uint32_t* col_rows = T.get<uint32_t>(nnz);
  uint32_t* col_w = weighted ? T.get<uint32_t>(nnz) : nullptr;
  uint32_t* col_of = T.get<uint32_t>(nnz);
  uint32_t* col_ptr = T.get<uint32_t>((uint64_t)NC + 1);
  uint32_t* col_split = T.get<uint32_t>(NC);
  if (!col_rows || (weighted && !col_w) || !col_of || !col_ptr || !col_split) return fail(c, LA3_ENOMEM, oom);
  launch(c, "ingest", k_ing_compact, blocks_for(m), (const uint64_t*)kb.Current(), (const uint32_t*)vb.Current(),
         (const uint32_t*)first, (const uint32_t*)upos, m, b, (const uint32_t*)W, col_rows, col_w, col_of);
  HIPC(c, dset_async(col_ptr, 0, 4ull * (NC + 1), st));
  launch(c, "ingest", k_ing_ptr, blocks_for(nnz), (const uint32_t*)col_of, nnz, NC, col_ptr);
  launch(c, "ingest", k_ing_split, blocks_for(NC), (const uint32_t*)col_of, (const uint32_t*)col_rows,
         (const uint32_t*)col_ptr, nnz, NC, nreg, col_split);
  launch(c, "ingest", k_ing_split2, blocks_for(nnz), (const uint32_t*)col_of, (const uint32_t*)col_rows,
         (const uint32_t*)col_ptr, nnz, nreg, col_split);
  // 4. row-major companion by (row slot, message-order key)
  uint32_t* row_cols = T.get<uint32_t>(nnz);
  uint32_t* row_w = weighted ? T.get<uint32_t>(nnz) : nullptr;
  uint32_t* row_of = T.get<uint32_t>(nnz);
  uint32_t* mkey = T.get<uint32_t>(nnz);
  uint32_t* row_ptr = T.get<uint32_t>((uint64_t)NR + 1);
  uint32_t* row_split = T.get<uint32_t>(NR);
  if (!row_cols || (weighted && !row_w) || !row_of || !mkey || !row_ptr || !row_split) return fail(c, LA3_ENOMEM, oom);
  launch(c, "ingest", k_ing_rowkeys, blocks_for(nnz), (const uint32_t*)col_rows, (const uint32_t*)col_of, nnz, nreg,
         nsrc, b, k0, v0);
  hipcub::DoubleBuffer<uint64_t> kr(k0, k1);
  hipcub::DoubleBuffer<uint32_t> vr(v0, v1);
  if (!sort_pairs(kr, vr, nnz, 2 * b)) return fail(c, LA3_EDEVICE, "device ingest: sort failed");
  launch(c, "ingest", k_ing_rowfill, blocks_for(nnz), (const uint64_t*)kr.Current(), (const uint32_t*)vr.Current(), nnz,
         b, (const uint32_t*)col_of, (const uint32_t*)col_w, nreg, nsink, row_cols, row_w, row_of, mkey);
  HIPC(c, dset_async(row_ptr, 0, 4ull * (NR + 1), st));
  launch(c, "ingest", k_ing_ptr, blocks_for(nnz), (const uint32_t*)row_of, nnz, NR, row_ptr);
  launch(c, "ingest", k_ing_split, blocks_for(NR), (const uint32_t*)row_of, (const uint32_t*)mkey,
         (const uint32_t*)row_ptr, nnz, NR, nsrc, row_split);
  launch(c, "ingest", k_ing_split2, blocks_for(nnz), (const uint32_t*)row_of, (const uint32_t*)mkey,
         (const uint32_t*)row_ptr, nnz, nsrc, row_split);
  // 5. long rows and PPR segments
  uint32_t* lflag = T.get<uint32_t>(NR);
  uint32_t* lpos = T.get<uint32_t>(NR);
  if (!lflag || !lpos) return fail(c, LA3_ENOMEM, oom);
  auto compact_rows = [&](uint32_t lim, uint32_t*& list, uint64_t& count) -> la3_status {
    launch(c, "ingest", k_ing_longflag, blocks_for(NR), (const uint32_t*)row_ptr, NR, lim, lflag);
    if (!exclusive_sum(lflag, lpos, NR)) return fail(c, LA3_EDEVICE, "device ingest: scan failed");
    la3_status x;
    if ((x = total(lflag, lpos, NR, count))) return x;
    list = T.get<uint32_t>(count);
    if (!list) return fail(c, LA3_ENOMEM, oom);
    launch(c, "ingest", k_ing_compact_ids, blocks_for(NR), (const uint32_t*)lflag, (const uint32_t*)lpos, NR, list);
    return LA3_OK;
  };
  uint32_t* long_rows = nullptr;
  uint32_t* plong = nullptr;
  uint64_t nlong = 0, np = 0;
  if ((e = compact_rows(LONG_ROW, long_rows, nlong))) return e;
  if ((e = compact_rows(PSEG_ENTRIES, plong, np))) return e;
  uint32_t* nseg = T.get<uint32_t>(np);
  uint32_t* pseg_first = T.get<uint32_t>(np + 1);
  if (!nseg || !pseg_first) return fail(c, LA3_ENOMEM, oom);
  launch(c, "ingest", k_ing_segcount, blocks_for(np), (const uint32_t*)plong, (uint32_t)np, (const uint32_t*)row_ptr,
         PSEG_ENTRIES, nseg);
  if (!exclusive_sum(nseg, pseg_first, np)) return fail(c, LA3_EDEVICE, "device ingest: scan failed");
  uint64_t npseg = 0;
  if ((e = total(nseg, pseg_first, np, npseg))) return e;
  const uint32_t npseg32 = (uint32_t)npseg;
  HIPC(c, dcopy_async(pseg_first + np, &npseg32, 4, hipMemcpyHostToDevice, st));
  uint32_t* pseg_row = T.get<uint32_t>(npseg);
  uint32_t* pseg_beg = T.get<uint32_t>(npseg);
  unsigned long long* sums = T.get<unsigned long long>(2);
  if (!pseg_row || !pseg_beg || !sums) return fail(c, LA3_ENOMEM, oom);
  launch(c, "ingest", k_ing_segfill, blocks_for(np), (const uint32_t*)plong, (uint32_t)np, (const uint32_t*)row_ptr,
         (const uint32_t*)pseg_first, PSEG_ENTRIES, pseg_row, pseg_beg);
  HIPC(c, dset_async(sums, 0, 16, st));
  launch(c, "ingest", k_ing_lensum, blocks_for(nlong), (const uint32_t*)long_rows, (uint32_t)nlong,
         (const uint32_t*)row_ptr, nreg, sums);
  launch(c, "ingest", k_ing_lensum, blocks_for(np), (const uint32_t*)plong, (uint32_t)np, (const uint32_t*)row_ptr,
         nreg, sums + 1);
  // Host copies: counts, id maps, the regular part of the PPR lists.
  unsigned long long hs[2];
  std::vector<uint32_t> hplong(np), hpf(np + 1), hcp((size_t)NC + 1), hcs(NC);
  g->hid_of.resize(S); g->id_of.resize(S); g->slot_of_id.resize(S);
  HIPC(c, dcopy_async(hs, sums, 16, hipMemcpyDeviceToHost, st));
  if (np) HIPC(c, dcopy_async(hplong.data(), plong, 4 * np, hipMemcpyDeviceToHost, st));
  HIPC(c, dcopy_async(hpf.data(), pseg_first, 4 * (np + 1), hipMemcpyDeviceToHost, st));
  HIPC(c, dcopy_async(hcp.data(), col_ptr, 4ull * (NC + 1), hipMemcpyDeviceToHost, st));
  if (NC) HIPC(c, dcopy_async(hcs.data(), col_split, 4ull * NC, hipMemcpyDeviceToHost, st));
  HIPC(c, dcopy_async(g->hid_of.data(), hid_of, 4ull * S, hipMemcpyDeviceToHost, st));
  HIPC(c, dcopy_async(g->id_of.data(), id_of, 4ull * S, hipMemcpyDeviceToHost, st));
  HIPC(c, dcopy_async(g->slot_of_id.data(), slot_of_id, 4ull * S, hipMemcpyDeviceToHost, st));
  HIPC(c, hipStreamSynchronize(st));
  uint64_t nnz_reg = 0;
  for (uint32_t col = 0; col < NC; col++) nnz_reg += hcs[col] - hcp[col];
  uint32_t nplong_reg = 0;
  while (nplong_reg < np && hplong[nplong_reg] < nreg) nplong_reg++;
  for (uint64_t i = 0; i < np; i++) g->max_pseg = std::max(g->max_pseg, hpf[i + 1] - hpf[i]);
  g->nnz = nnz; g->nnz_reg = nnz_reg; g->nnz_sink = nnz - nnz_reg;
  g->nnz_long_reg = hs[0]; g->pnnz_reg = hs[1];
  DG& d = g->dg;
  d.nslots = S; d.nreg = nreg; d.nsink = nsink; d.nsrc = nsrc; d.ncols = NC; d.nrows = NR; d.nnz = nnz;
  d.col_ptr = col_ptr; d.col_split = col_split; d.col_rows = col_rows; d.col_w = col_w;
  d.row_ptr = row_ptr; d.row_split = row_split; d.row_cols = row_cols; d.row_w = row_w;
  d.id_of = id_of; d.hid_of = hid_of; d.long_rows = long_rows; d.nlong = (uint32_t)nlong;
  d.slot_of_id = slot_of_id; d.plong = plong; d.pseg_first = pseg_first; d.pseg_row = pseg_row;
  d.pseg_beg = pseg_beg; d.nplong = (uint32_t)np; d.nplong_reg = nplong_reg; d.npseg = npseg32;
  d.npseg_reg = hpf[nplong_reg];
  const std::pair<void*, uint64_t> kept[] = {
      {col_ptr, NC + 1ull}, {col_split, NC}, {col_rows, nnz}, {col_w, weighted ? nnz : 0}, {row_ptr, NR + 1ull},
      {row_split, NR}, {row_cols, nnz}, {row_w, weighted ? nnz : 0}, {id_of, S}, {hid_of, S}, {long_rows, nlong},
      {slot_of_id, S}, {plong, np}, {pseg_first, np + 1}, {pseg_row, npseg}, {pseg_beg, npseg}};
  for (const auto& kv : kept) {
    if (!kv.first) continue;
    T.keep(kv.first);
    g->allocs.push_back(kv.first);
    g->device_bytes += std::max<uint64_t>(4 * kv.second, 4);
  }
  return LA3_OK;
}

// Message rows of the narrow PPR kernels by decreasing out-degree
// (k_msg_keys): DG::rank_of and DG::row_rank, kept with the graph. An
// undirected load (PPR runs on directed loads only) and a graph of 2^31
// slots or more (beyond the sort's int count) keep slot order: rank_of is
// the identity and row_rank is row_cols.
la3_status graph_msg_order(la3_ctx* c, la3_graph* g) {
  DG& d = g->dg;
  const uint32_t S = d.nslots;
  hipStream_t st = c->stream;
  const char* oom = "message order: allocation failed";
  DevTemps T;
  uint32_t* rank_of = T.get<uint32_t>(S);
  uint32_t* row_rank = nullptr;
  if (!rank_of) return fail(c, LA3_ENOMEM, oom);
  if (S < (1u << 31) && g->directed) {
    // the key and value types of the ingest sorts (one instantiation)
    uint64_t* k0 = T.get<uint64_t>(S);
    uint64_t* k1 = T.get<uint64_t>(S);
    uint32_t* v0 = T.get<uint32_t>(S);
    uint32_t* v1 = T.get<uint32_t>(S);
    row_rank = T.get<uint32_t>(d.nnz);
    if (!k0 || !k1 || !v0 || !v1 || !row_rank) return fail(c, LA3_ENOMEM, oom);
    launch(c, "msg_order", k_msg_keys, blocks_for(S), d, k0, v0);
    hipcub::DoubleBuffer<uint64_t> kb(k0, k1);
    hipcub::DoubleBuffer<uint32_t> vb(v0, v1);
    size_t need = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, need, kb, vb, (int)S, 0, 32, st) != hipSuccess)
      return fail(c, LA3_EDEVICE, "message order: sort failed");
    void* tmp = T.get<uint8_t>(need);
    if (!tmp) return fail(c, LA3_ENOMEM, oom);
    if (hipcub::DeviceRadixSort::SortPairs(tmp, need, kb, vb, (int)S, 0, 32, st) != hipSuccess)
      return fail(c, LA3_EDEVICE, "message order: sort failed");
    launch(c, "msg_order", k_msg_rank, blocks_for(S), (const uint32_t*)vb.Current(), S, rank_of);
    if (d.nnz) launch(c, "msg_order", k_msg_entries, blocks_for(d.nnz), d.row_cols, d.nnz, (const uint32_t*)rank_of, row_rank);
  } else {
    launch(c, "msg_order", k_msg_identity, blocks_for(S), S, rank_of);
  }
  HIPC(c, hipStreamSynchronize(st));  // before the temporaries are freed
  d.rank_of = rank_of;
  d.row_rank = row_rank ? row_rank : d.row_cols;
  for (void* p : {(void*)rank_of, (void*)row_rank}) {
    if (!p) continue;
    T.keep(p);
    g->allocs.push_back(p);
  }
  g->device_bytes += 4ull * S + (row_rank ? 4 * std::max<uint64_t>(d.nnz, 1) : 0);
  return LA3_OK;
}

la3_status load_records(la3_ctx* ctx, const uint32_t* rec, uint64_t nrec, uint32_t n, int weighted,
                        la3_mode mode, la3_hashing hashing, la3_graph** out) {
  if (!ctx || !out || (nrec && !rec)) return fail(ctx, LA3_EINVAL, "null argument");
  if ((int)mode < 0 || (int)mode > 4) return fail(ctx, LA3_EINVAL, "bad load mode");
  if (n == 0xffffffffu) return fail(ctx, LA3_ENOTSUP, "nvertices must be < 2^32-1");
  // Device ingest sorts and scans with int counts: records and stored
  // entries (twice the records when undirected) below 2^31; larger graphs
  // take the host builder.
  const uint64_t stored = mode == LA3_UNDIRECTED ? 2 * nrec : nrec;
  if (ctx->device >= 0 && ctx->ingest_device && stored < (1ull << 31)) {
    HIPC(ctx, hipSetDevice(ctx->device));
    auto g = std::make_unique<la3_graph>();
    g->ctx = ctx;
    la3_status st = ingest_device(ctx, rec, nrec, n, weighted != 0, (int)mode, hashing != LA3_HASH_NONE, g.get());
    if (!st) st = graph_msg_order(ctx, g.get());
    if (st) {
      for (void* p : g->allocs) dev_free(p);
      return st;
    }
    ctx->refs++;
    *out = g.release();
    return LA3_OK;
  }
  HostLayout L;
  std::string e = build_layout(rec, nrec, n, weighted != 0, (int)mode, hashing != LA3_HASH_NONE, L);
  if (!e.empty()) return fail(ctx, e.find("2^32") != std::string::npos ? LA3_ENOTSUP : LA3_EINVAL, e);
  return graph_from_layout(ctx, L, out);
}

// Frontier queues: work items plus the counter block
// {u32 nitems, pad to 64 B, CtrSlot[NSLOT]}.
constexpr size_t CTR_BYTES = 64 + NSLOT * sizeof(CtrSlot);
static_assert(CTR_BYTES <= 4096, "one pinned counter slot");
constexpr uint64_t PIPE_MAX_NNZ = 1ull << 25;  // graphs whose single-query SSSP steps are pipelined

la3_status get_queue(la3_ctx* c, const std::string& name, uint64_t cap, Frontier& f) {
  f.items = (uint64_t*)scratch(c, (name + ".items").c_str(), cap * sizeof(uint64_t));
  uint8_t* ctr = (uint8_t*)scratch(c, (name + ".ctr").c_str(), CTR_BYTES);
  if (!f.items || !ctr) return fail(c, LA3_ENOMEM, "frontier allocation failed");
  f.nitems = (uint32_t*)ctr;
  f.slots = (CtrSlot*)(ctr + 64);
  return LA3_OK;
}

// Per-query frontier buffers (two, ping-ponged).
struct Fronts {
  Frontier f[2];
  uint64_t cap;
};

la3_status get_fronts(la3_ctx* c, la3_graph* g, Fronts& F, const char* tag) {
  F.cap = (uint64_t)g->nslots + g->nnz / CHUNK + 2;
  for (int i = 0; i < 2; i++) {
    la3_status st = get_queue(c, std::string(tag) + char('0' + i), F.cap, F.f[i]);
    if (st) return st;
  }
  return LA3_OK;
}

struct FrontCounts {
  uint32_t items = 0, verts = 0, cols = 0;
  unsigned long long edges = 0, scanned = 0;
  uint32_t nmin = 0xffffffffu;  // smallest distance put in the next frontier
};

// A frontier's counters as copied to the host (CTR_BYTES at h).
void parse_counts(const uint8_t* h, FrontCounts& out) {
  out = FrontCounts();
  std::memcpy(&out.items, h, 4);
  const CtrSlot* sl = (const CtrSlot*)(h + 64);
  uint32_t mx = 0;
  for (uint32_t i = 0; i < NSLOT; i++) {
    out.verts += sl[i].verts;
    out.cols += sl[i].cols;
    out.edges += sl[i].edges;
    out.scanned += sl[i].scanned;
    mx = std::max(mx, sl[i].nmin_inv);
  }
  out.nmin = ~mx;
}

la3_status read_counts(la3_ctx* c, const Frontier& f, FrontCounts& out) {
  HIPC(c, dcopy_async(c->host_ctr, f.nitems, CTR_BYTES, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  parse_counts((const uint8_t*)c->host_ctr, out);
  return LA3_OK;
}

void zero_counts(la3_ctx* c, const Frontier& f) {
  launch(c, "fill_u32", k_fill<uint32_t>, 3, f.nitems, (uint64_t)(CTR_BYTES / 4), 0u);
}

}  // namespace

// Next near/far threshold: unchanged while the next frontier holds pairs
// below it, else the frontier's smallest distance + delta.
uint32_t next_threshold(uint32_t T, uint32_t delta, uint32_t nmin) {
  if (delta == 0) return 0xffffffffu;
  if (nmin < T) return T;
  return (uint32_t)std::min<uint64_t>(0xffffffffull, (uint64_t)nmin + delta);
}

// Hands a device result to the caller: copies it into `dst` (host memory)
// when dst != NULL, and in every case records it as the context's last
// result so that la3_result_fetch can copy it later.
la3_status deliver(la3_ctx* c, void* dst, const void* dev, size_t bytes) {
  c->last_result = dev;
  c->last_result_bytes = bytes;
  if (dst) HIPC(c, dcopy_async(dst, dev, bytes, hipMemcpyDeviceToHost, c->stream));
  return LA3_OK;
}

// ================================================================== API ===
extern "C" {

la3_status la3_ctx_create(int device, la3_ctx** out) {
  if (!out) return LA3_EINVAL;
  auto c = std::make_unique<la3_ctx>();
  c->device = device;
  if (device == -1) { *out = c.release(); return LA3_OK; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return LA3_EDEVICE;
  if (device < 0 || device >= ndev) return LA3_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return LA3_EDEVICE;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return LA3_EDEVICE;
  if (hipHostMalloc(&c->host_ctr, 4096 * 4, hipHostMallocDefault) != hipSuccess) return LA3_EDEVICE;
  // The allocation-free entry point's array: its address on this device
  // (a failure only disables la3_bfs_noalloc, which then reports it).
  (void)noalloc::setup(device);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->ncu = prop.multiProcessorCount;
  *out = c.release();
  return LA3_OK;
}

void la3_ctx_destroy(la3_ctx* c) {
  if (c) ctx_release(c);
}

const char* la3_last_error(const la3_ctx* c) { return c ? c->err.c_str() : "null context"; }

la3_status la3_ctx_set_ingest(la3_ctx* c, int on_device) {
  if (!c) return LA3_EINVAL;
  c->ingest_device = on_device != 0;
  return LA3_OK;
}

la3_status la3_graph_array(const la3_graph* g, la3_array which, uint32_t* out, uint64_t cap, uint64_t* count) {
  if (!g || !count) return LA3_EINVAL;
  la3_ctx* c = g->ctx;
  const uint32_t* dev = nullptr;
  const std::vector<uint32_t>* hv = nullptr;
  uint64_t n = 0;
  const HostLayout* L = g->host.get();
  const DG& d = g->dg;
  switch (which) {
    case LA3_ARR_COL_PTR: n = g->ncols + 1ull; dev = d.col_ptr; if (L) hv = &L->col_ptr; break;
    case LA3_ARR_COL_SPLIT: n = g->ncols; dev = d.col_split; if (L) hv = &L->col_split; break;
    case LA3_ARR_COL_ROWS: n = g->nnz; dev = d.col_rows; if (L) hv = &L->col_rows; break;
    case LA3_ARR_COL_W: n = g->weighted ? g->nnz : 0; dev = d.col_w; if (L) hv = &L->col_w; break;
    case LA3_ARR_ROW_PTR: n = g->nrows + 1ull; dev = d.row_ptr; if (L) hv = &L->row_ptr; break;
    case LA3_ARR_ROW_SPLIT: n = g->nrows; dev = d.row_split; if (L) hv = &L->row_split; break;
    case LA3_ARR_ROW_COLS: n = g->nnz; dev = d.row_cols; if (L) hv = &L->row_cols; break;
    case LA3_ARR_ROW_W: n = g->weighted ? g->nnz : 0; dev = d.row_w; if (L) hv = &L->row_w; break;
    case LA3_ARR_ID_OF: n = g->nslots; hv = &g->id_of; break;
    case LA3_ARR_SLOT_OF_ID: n = g->nslots; hv = &g->slot_of_id; break;
    case LA3_ARR_LONG_ROWS: n = L ? L->long_rows.size() : d.nlong; dev = d.long_rows; if (L) hv = &L->long_rows; break;
    case LA3_ARR_PLONG: n = L ? L->plong.size() : d.nplong; dev = d.plong; if (L) hv = &L->plong; break;
    case LA3_ARR_PSEG_FIRST: n = L ? L->pseg_first.size() : d.nplong + 1ull; dev = d.pseg_first; if (L) hv = &L->pseg_first; break;
    case LA3_ARR_PSEG_ROW: n = L ? L->pseg_row.size() : d.npseg; dev = d.pseg_row; if (L) hv = &L->pseg_row; break;
    case LA3_ARR_PSEG_BEG: n = L ? L->pseg_beg.size() : d.npseg; dev = d.pseg_beg; if (L) hv = &L->pseg_beg; break;
    default: return fail(c, LA3_EINVAL, "unknown layout array");
  }
  *count = n;
  if (!out) return LA3_OK;
  if (cap < n) return fail(c, LA3_EINVAL, "output too small");
  if (n == 0) return LA3_OK;
  if (hv) {
    std::memcpy(out, hv->data(), 4 * n);
    return LA3_OK;
  }
  HIPC(c, hipSetDevice(c->device));
  HIPC(c, dcopy(out, dev, 4 * n, hipMemcpyDeviceToHost));
  return LA3_OK;
}

la3_status la3_ctx_release_scratch(la3_ctx* c) {
  if (!c) return LA3_EINVAL;
  if (c->device < 0) return LA3_OK;
  if (c->fz_busy) return fail(c, LA3_EAGAIN, "a session is running on the context");
  HIPC(c, hipSetDevice(c->device));
  HIPC(c, hipStreamSynchronize(c->stream));
  for (auto& kv : c->fz_exec) (void)hipGraphExecDestroy(kv.second);
  c->fz_exec.clear();
  c->fz_tasks = nullptr;
  c->fz_ctl = nullptr;
  c->last_result = nullptr;
  c->last_result_bytes = 0;
  scratch_release(c, "");
  c->scratch_epoch++;
  return LA3_OK;
}

la3_status la3_ctx_synchronize(la3_ctx* c) {
  if (!c) return LA3_EINVAL;
  if (c->device < 0) return LA3_OK;
  HIPC(c, hipStreamSynchronize(c->stream));
  return LA3_OK;
}

la3_status la3_set_profiling(la3_ctx* c, int enabled) {
  if (!c) return LA3_EINVAL;
  c->profiling = enabled != 0;
  return LA3_OK;
}

la3_status la3_last_timing(const la3_ctx* c, la3_timing* out) {
  if (!c || !out) return LA3_EINVAL;
  *out = c->last;
  return LA3_OK;
}

la3_status la3_kernel_stats(const la3_ctx* c, la3_kernel_stat* out, uint32_t cap, uint32_t* count) {
  if (!c || !count) return LA3_EINVAL;
  uint32_t i = 0;
  for (auto& kv : c->stats) {
    if (out && i < cap) {
      la3_kernel_stat& o = out[i];
      std::memset(&o, 0, sizeof(o));
      snprintf(o.name, sizeof(o.name), "%s", kv.first.c_str());
      o.launches = kv.second.launches;
      o.ms = kv.second.ms;
      o.alg_bytes = kv.second.bytes;
      o.entries = kv.second.entries;
    }
    i++;
  }
  *count = i;
  return LA3_OK;
}

la3_status la3_kernel_stats_reset(la3_ctx* c) {
  if (!c) return LA3_EINVAL;
  c->stats.clear();
  return LA3_OK;
}

la3_status la3_graph_load_records(la3_ctx* ctx, const uint32_t* records, uint64_t nrecords,
                                  uint32_t nvertices, int weighted, la3_mode mode,
                                  la3_hashing hashing, la3_graph** out) {
  return load_records(ctx, records, nrecords, nvertices, weighted, mode, hashing, out);
}

la3_status la3_graph_load_file(la3_ctx* ctx, const char* path, uint32_t nvertices, int weighted,
                               la3_mode mode, la3_hashing hashing, la3_graph** out) {
  if (!ctx || !path || !out) return fail(ctx, LA3_EINVAL, "null argument");
  int fd = open(path, O_RDONLY);
  if (fd < 0) return fail(ctx, LA3_EIO, std::string("Unable to open input file: ") + path);
  struct stat st;
  if (fstat(fd, &st) != 0) { close(fd); return fail(ctx, LA3_EIO, "stat() failure"); }
  const uint64_t size = (uint64_t)st.st_size;
  const uint8_t* base = nullptr;
  if (size > 0) {
    void* m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) { close(fd); return fail(ctx, LA3_EIO, "mmap() failure"); }
    base = (const uint8_t*)m;
  }
  close(fd);
  uint64_t off = 0;
  uint32_t n = nvertices;
  if (n == 0) {
    // 16-byte header {u32 nrows-1, u32 ncols-1, u64 nnz} (graph.hpp:89-117).
    if (size < 16) { if (base) munmap((void*)base, size); return fail(ctx, LA3_EIO, "fread() failure: short header"); }
    uint32_t hr, hc;
    std::memcpy(&hr, base, 4);
    std::memcpy(&hc, base + 4, 4);
    n = std::max(hr, hc) + 1;
    off = 16;
  }
  const uint64_t rsz = weighted ? 12 : 8;
  const uint64_t nrec = (size - off) / rsz;
  const uint32_t* rec = base ? (const uint32_t*)(base + off) : nullptr;
  la3_status s = load_records(ctx, rec, nrec, n, weighted, mode, hashing, out);
  if (base) munmap((void*)base, size);
  return s;
}

void la3_graph_free(la3_graph* g) {
  if (!g) return;
  la3_ctx* c = g->ctx;
  if (!g->host) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (g->pr_exec) (void)hipGraphExecDestroy(g->pr_exec);
    for (void* p : g->allocs) dev_free(p);
  }
  delete g;
  ctx_release(c);
}

la3_status la3_graph_info_get(const la3_graph* g, la3_graph_info* o) {
  if (!g || !o) return LA3_EINVAL;
  o->n = g->n; o->nslots = g->nslots; o->nregular = g->nreg; o->nsink = g->nsink;
  o->nsource = g->nsrc; o->nisolated = g->niso; o->ncols = g->ncols; o->directed = g->directed;
  o->weighted = g->weighted; o->mode = g->mode; o->nnz = g->nnz; o->nnz_regular = g->nnz_reg;
  o->nnz_sink = g->nnz_sink; o->nrecords = g->nrecords; o->device_bytes = g->device_bytes;
  return LA3_OK;
}

la3_status la3_graph_slots(const la3_graph* g, uint32_t* hid_of) {
  if (!g || !hid_of) return LA3_EINVAL;
  std::memcpy(hid_of, g->hid_of.data(), sizeof(uint32_t) * g->nslots);
  return LA3_OK;
}

uint32_t la3_graph_hash(const la3_graph* g, uint32_t v) { return g ? (uint32_t)g->h.hash(v) : v; }
uint32_t la3_graph_unhash(const la3_graph* g, uint32_t v) { return g ? (uint32_t)g->h.unhash(v) : v; }

la3_status la3_graph_csc(const la3_graph* g, int which, uint32_t* colptrs, uint32_t* colidxs,
                         uint32_t* global_idx, uint32_t* idx, uint32_t* weight) {
  if (!g || !colptrs || !colidxs || (which != 0 && which != 1)) return LA3_EINVAL;
  la3_ctx* c = g->ctx;
  const uint32_t NC = g->ncols;
  std::vector<uint32_t> cp, cs, rows, w;
  if (g->host) {
    cp = g->host->col_ptr; cs = g->host->col_split; rows = g->host->col_rows; w = g->host->col_w;
  } else {
    HIPC(c, hipSetDevice(c->device));
    cp.resize(NC + 1); cs.resize(NC); rows.resize(g->nnz);
    HIPC(c, dcopy(cp.data(), g->dg.col_ptr, sizeof(uint32_t) * (NC + 1), hipMemcpyDeviceToHost));
    HIPC(c, dcopy(cs.data(), g->dg.col_split, sizeof(uint32_t) * NC, hipMemcpyDeviceToHost));
    if (g->nnz) HIPC(c, dcopy(rows.data(), g->dg.col_rows, sizeof(uint32_t) * g->nnz, hipMemcpyDeviceToHost));
    if (g->weighted && g->nnz) {
      w.resize(g->nnz);
      HIPC(c, dcopy(w.data(), g->dg.col_w, sizeof(uint32_t) * g->nnz, hipMemcpyDeviceToHost));
    }
  }
  uint32_t k = 0;
  for (uint32_t col = 0; col < NC; col++) {
    const uint32_t b = which ? cs[col] : cp[col], e = which ? cp[col + 1] : cs[col];
    colptrs[col] = k;
    colidxs[col] = (e > b) ? g->hid_of[col < g->nreg ? col : col + g->nsink] : 0u;
    for (uint32_t j = b; j < e; j++, k++) {
      const uint32_t r = rows[j];
      if (global_idx) global_idx[k] = which ? r - g->nreg : r;
      if (idx) idx[k] = g->hid_of[r];
      if (weight) weight[k] = g->weighted ? w[j] : 0u;
    }
  }
  colptrs[NC] = k;
  return LA3_OK;
}

la3_status la3_graph_from_csc(la3_ctx* c, const la3_csc_arrays* a, la3_graph** out) {
  if (!c || !a || !out) return fail(c, LA3_EINVAL, "null argument");
  *out = nullptr;
  if ((int)a->mode < 0 || (int)a->mode > 4) return fail(c, LA3_EINVAL, "bad load mode");
  if (a->nvertices == 0xffffffffu) return fail(c, LA3_ENOTSUP, "nvertices must be < 2^32-1");
  const uint64_t S = (uint64_t)a->nvertices + 1;
  const uint32_t NC = a->ncols;
  if (NC > S) return fail(c, LA3_EINVAL, "more columns than vertices");
  const Hasher h(a->nvertices, a->hashing != LA3_HASH_NONE);
  // The stored matrix's entries as records (column vertex, row vertex[, w])
  // by original id: loaded directed and not reversed, they store exactly
  // these entries (graph.hpp:150-192), and the builder then derives the
  // classes, slots and both CSCs from them.
  const int wd = a->weighted ? 3 : 2;
  std::vector<uint32_t> rec;
  for (int which = 0; which < 2; which++) {
    const uint32_t* cp = a->colptrs[which];
    if (!cp || !a->colidxs[which]) return fail(c, LA3_EINVAL, "missing column arrays");
    if (cp[0] != 0) return fail(c, LA3_EINVAL, "colptrs must start at 0");
    for (uint32_t col = 0; col < NC; col++)
      if (cp[col + 1] < cp[col]) return fail(c, LA3_EINVAL, "colptrs must not decrease");
    const uint64_t nnz = cp[NC];
    if (nnz && (!a->idx[which] || (a->weighted && !a->weight[which])))
      return fail(c, LA3_EINVAL, "missing entry arrays");
    rec.reserve(rec.size() + wd * nnz);
    for (uint32_t col = 0; col < NC; col++) {
      if (cp[col + 1] == cp[col]) continue;
      const uint32_t hc = a->colidxs[which][col];
      if (hc >= S) return fail(c, LA3_EINVAL, "column id out of range");
      const uint32_t vc = (uint32_t)h.unhash(hc);
      for (uint32_t j = cp[col]; j < cp[col + 1]; j++) {
        const uint32_t hr = a->idx[which][j];
        if (hr >= S) return fail(c, LA3_EINVAL, "row id out of range");
        rec.push_back(vc);
        rec.push_back((uint32_t)h.unhash(hr));
        if (a->weighted) rec.push_back(a->weight[which][j]);
      }
    }
  }
  la3_graph* g = nullptr;
  la3_status st = load_records(c, rec.data(), rec.size() / wd, a->nvertices, a->weighted, LA3_DIRECTED,
                               a->hashing, &g);
  if (st) return st;
  rec = std::vector<uint32_t>();
  // The load mode the arrays came from: directedness and TC's orientation.
  g->mode = (int)a->mode;
  g->directed = a->mode != LA3_UNDIRECTED;
  if (g->host) {
    g->host->mode = (int)a->mode;
    g->host->directed = g->directed;
  }
  // Only a reference layout reproduces itself: compare both CSCs entry by entry.
  std::string bad;
  if (g->ncols != NC) bad = "column count " + std::to_string(g->ncols) + " != " + std::to_string(NC);
  for (int which = 0; which < 2 && bad.empty(); which++) {
    const uint32_t* cp = a->colptrs[which];
    const uint64_t nnz = cp[NC];
    std::vector<uint32_t> p(NC + 1), ci(NC + 1), gi(nnz + 1), ix(nnz + 1), w(nnz + 1);
    if ((st = la3_graph_csc(g, which, p.data(), ci.data(), gi.data(), ix.data(), w.data()))) break;
    const std::string name = which ? "sink CSC " : "regular CSC ";
    if (p[NC] != nnz) { bad = name + "entry count"; break; }
    for (uint32_t col = 0; col < NC && bad.empty(); col++) {
      if (p[col] != cp[col]) bad = name + "colptrs[" + std::to_string(col) + "]";
      else if (cp[col + 1] > cp[col] && ci[col] != a->colidxs[which][col]) bad = name + "colidxs[" + std::to_string(col) + "]";
    }
    for (uint64_t j = 0; j < nnz && bad.empty(); j++) {
      if (a->global_idx[which] && gi[j] != a->global_idx[which][j]) bad = name + "global_idx[" + std::to_string(j) + "]";
      else if (ix[j] != a->idx[which][j]) bad = name + "idx[" + std::to_string(j) + "]";
      else if (a->weighted && w[j] != a->weight[which][j]) bad = name + "weight[" + std::to_string(j) + "]";
    }
  }
  if (st || !bad.empty()) {
    la3_graph_free(g);
    if (st) return st;
    return fail(c, LA3_EINVAL, "not a CSC layout of the reference (csc.h:76-130): " + bad + " differs");
  }
  *out = g;
  return LA3_OK;
}

// ----------------------------------------------- vertex-program engine ---
}  // extern "C"
namespace {

la3_device_view device_view(const la3_graph* g) {
  const DG& d = g->dg;
  la3_device_view v;
  std::memset(&v, 0, sizeof(v));
  v.device = g->ctx->device;
  v.nslots = d.nslots; v.nreg = d.nreg; v.nsink = d.nsink; v.nsrc = d.nsrc; v.ncols = d.ncols; v.nrows = d.nrows;
  v.nnz = d.nnz;
  v.directed = g->directed; v.weighted = g->weighted;
  v.col_ptr = d.col_ptr; v.col_split = d.col_split; v.col_rows = d.col_rows; v.col_w = g->weighted ? d.col_w : nullptr;
  v.row_ptr = d.row_ptr; v.row_split = d.row_split; v.row_cols = d.row_cols; v.row_w = g->weighted ? d.row_w : nullptr;
  v.id_of = d.id_of; v.hid_of = d.hid_of; v.slot_of_id = d.slot_of_id;
  v.long_rows = d.long_rows; v.nlong = d.nlong;
  return v;
}
static_assert(LONG_ROW == vp_detail::VP_LONG, "the layout's long-row list must match the engine's lane-per-row limit");

// Engine buffers from the context's scratch, named under `prefix`; the
// scratch keeps them for the next call, so release does nothing. `sa` must
// outlive the engine.
struct ScratchAlloc {
  la3_ctx* c;
  std::string prefix;
};
vp_detail::Allocator scratch_allocator(ScratchAlloc& sa) {
  vp_detail::Allocator a;
  a.self = &sa;
  a.alloc = [](void* self, const char* name, size_t bytes) -> void* {
    ScratchAlloc* s = (ScratchAlloc*)self;
    return scratch(s->c, (s->prefix + name).c_str(), bytes);
  };
  a.release = [](void*, void*) {};
  return a;
}

// CcVertex's operations (src/apps/graph_analytics/cc.h:28-37).
struct CcOps {
  using W = Empty;
  using M = uint32_t;
  using A = uint32_t;
  struct S {
    uint32_t label = 0;
  };
  __device__ bool init(uint32_t vid, S& s) const {
    s.label = vid;
    return true;
  }
  __device__ M scatter(const S& s) const { return s.label; }
  __device__ A gather(const Edge<Empty>&, const M& m) const { return m; }
  __device__ void combine(const A& y1, A& y2) const { y2 = y2 == 0 ? y1 : min(y1, y2); }
  __device__ bool apply(const A& y, S& s) const {
    const uint32_t t = s.label;
    s.label = min(s.label, y);
    return t != s.label;
  }
  // Runs of a long row's messages folded separately and joined in order
  // (exactly the in-order fold): a 0 message empties the accumulator, so a
  // run holding one folds to its own value whatever came before it; a run
  // without one folds into the accumulator by combine.
  struct Seg {
    uint32_t v = 0;  // the run's fold from A() = 0
    uint32_t z = 0;  // the run holds a 0 message
  };
  __device__ void seg_add(const A& y, Seg& sg) const {
    combine(y, sg.v);
    sg.z |= y == 0;
  }
  __device__ void seg_join(const Seg& later, Seg& sg) const {
    if (later.z) sg.v = later.v;
    else if (later.v != 0) combine(later.v, sg.v);  // else: the run held nothing
    sg.z |= later.z;
  }
  __device__ A seg_value(const Seg& sg) const { return sg.v; }
};

}  // namespace
extern "C" {

la3_status la3_graph_device_view(const la3_graph* g, la3_device_view* out) {
  if (!g || !out) return LA3_EINVAL;
  if (g->host) return fail(g->ctx, LA3_EDEVICE, "graph lives on a host-only context");
  *out = device_view(g);
  return LA3_OK;
}

void* la3_ctx_stream(const la3_ctx* c) { return (c && c->device >= 0) ? (void*)c->stream : nullptr; }

// ------------------------------------------------------------- degree ---
la3_status la3_degree(la3_graph* g, uint32_t* degree) {
  if (!g) return LA3_EINVAL;
  la3_ctx* c = g->ctx;
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  HIPC(c, hipSetDevice(c->device));
  timing_begin(c);
  uint32_t* out = (uint32_t*)scratch(c, "out.u32", sizeof(uint32_t) * g->nslots);
  if (!out) return fail(c, LA3_ENOMEM, "scratch allocation failed");
  launch(c, "degree", k_degree, blocks_for(g->nslots), g->dg, out);
  la3_status st = deliver(c, degree, out, sizeof(uint32_t) * g->nslots);
  if (st) return st;
  HIPC(c, hipStreamSynchronize(c->stream));
  timing_end(c);
  return LA3_OK;
}

// ---------------------------------------------------------------- BFS ---
la3_status la3_bfs(la3_graph* g, uint32_t root, uint32_t max_iters, uint8_t* hops, uint32_t* parent,
                   uint32_t* iters_run) {
  if (!g) return LA3_EINVAL;
  return la3_bfs_ctx(g->ctx, g, root, max_iters, hops, parent, iters_run);
}

}  // extern "C"
namespace {

// The BFS traversal of la3_bfs_ctx on caller-chosen buffers: level[nslots],
// mark[nslots] and the two frontier queues F; *steps = traversal steps run.
la3_status bfs_core(la3_ctx* c, la3_graph* g, uint32_t root, uint32_t max_iters, uint8_t* level, uint32_t* mark,
                    const Fronts& F, uint32_t* steps) {
  const uint32_t S = g->nslots;
  la3_status st;
  const DG& d = g->dg;
  const uint32_t rs = g->slot_of_id[root];
  launch(c, "bfs_init", k_bfs_init, blocks_for(S), d, level, mark, rs, F.f[0], F.f[1].nitems);
  FrontCounts fc;
  if ((st = read_counts(c, F.f[0], fc))) return st;
  uint32_t iter = 0, cur = 0;
  bool bottom_up = false;
  uint64_t explored = fc.edges, visited = 1;
  const uint32_t wave_blocks = 4 * c->ncu * 8;
  // The reference stops after the first iteration that activates no regular
  // vertex (vertex_program_execute.hpp:211-212); a regular vertex is exactly one
  // that emits work items, so the loop runs while items exist. Step `iter`
  // gives level iter + 1, so max_iters = m stops after level m (a bounded
  // loop that settles earlier only repeats empty iterations).
  while (fc.items > 0 && (max_iters == 0 || iter < max_iters)) {
    const Frontier& nf = F.f[cur ^ 1];
    int ev = -1;
    if (!bottom_up) {
      const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((fc.items + 3) / 4, wave_blocks));
      ev = launch(c, "bfs_push", k_bfs_push, blocks, d, (const uint64_t*)F.f[cur].items, fc.items, level, mark, iter, nf, F.f[cur].nitems);
    } else {
      ev = launch(c, "bfs_pull", k_bfs_pull, blocks_for(g->nrows, 4 * c->ncu * 8), d, level, mark, iter, nf, F.f[cur].nitems);
      if (d.nlong) launch(c, "bfs_pull_long", k_bfs_pull_long, std::min<uint32_t>((d.nlong + 3) / 4, wave_blocks), d, level, mark, iter, nf);
    }
    FrontCounts nc;
    if ((st = read_counts(c, nf, nc))) return st;
    // Algorithmic bytes (m = a = 1/8 B frontier bit, s_r = s_w = 1 B level):
    // push: 4 B per entry, 8 B pointers + 1/8 B message per active column,
    //       (a + s_r + s_w) per activated row;
    // pull: 4 B + 1/8 B per entry read, 8 B pointers + 1 B state per row checked.
    if (!bottom_up) {
      annotate(c, ev, 4.0 * fc.edges + 8.125 * fc.cols + 2.25 * nc.verts, fc.edges);
      c->edges_scanned += fc.edges;
    } else {
      const uint64_t unvisited = g->nrows > visited ? g->nrows - visited : 0;
      annotate(c, ev, 4.125 * nc.scanned + 9.0 * unvisited + 2.25 * nc.verts, nc.scanned);
      c->edges_scanned += nc.scanned;
    }
    visited += nc.verts;
    iter++;
    explored += nc.edges;
    const uint64_t unexplored = g->nnz > explored ? g->nnz - explored : 0;
    if (!bottom_up && nc.edges > unexplored / 14 && nc.verts > 0) bottom_up = true;
    else if (bottom_up && (uint64_t)nc.verts * 24 < S) bottom_up = false;
    fc = nc;
    cur ^= 1;
  }
  *steps = iter;
  return LA3_OK;
}

}  // namespace
extern "C" {

la3_status la3_bfs_ctx(la3_ctx* c, la3_graph* g, uint32_t root, uint32_t max_iters, uint8_t* hops,
                       uint32_t* parent, uint32_t* iters_run) {
  if (!c || !g) return LA3_EINVAL;
  if (c->device != g->ctx->device) return fail(c, LA3_EINVAL, "graph lives on another device");
  if (root > g->n) return fail(c, LA3_EINVAL, "root id out of range");
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  HIPC(c, hipSetDevice(c->device));
  timing_begin(c);
  const uint32_t S = g->nslots;
  uint8_t* level = (uint8_t*)scratch(c, "bfs.level", S);
  uint32_t* mark = (uint32_t*)scratch(c, "bfs.mark", sizeof(uint32_t) * S);
  Fronts F;
  la3_status st = get_fronts(c, g, F, "bfs");
  if (st) return st;
  if (!level || !mark) return fail(c, LA3_ENOMEM, "scratch allocation failed");
  uint32_t iter = 0;
  if ((st = bfs_core(c, g, root, max_iters, level, mark, F, &iter))) return st;
  uint8_t* oh = (uint8_t*)scratch(c, "out.u8", S);
  uint32_t* op = parent ? (uint32_t*)scratch(c, "out.u32", sizeof(uint32_t) * S) : nullptr;
  if (!oh || (parent && !op)) return fail(c, LA3_ENOMEM, "scratch allocation failed");
  launch(c, "bfs_output", k_bfs_output, blocks_for(S), g->dg, (const uint8_t*)level, (const uint32_t*)mark, oh, op);
  if ((st = deliver(c, hops, oh, S))) return st;
  if (parent) HIPC(c, dcopy_async(parent, op, sizeof(uint32_t) * S, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  timing_end(c);
  if (iters_run) *iters_run = max_iters ? max_iters : std::max<uint32_t>(iter, 1);
  return LA3_OK;
}

la3_status la3_bfs_noalloc(la3_ctx* c, la3_graph* g, uint32_t root, uint32_t max_iters, uint8_t* hops_dev,
                           uint32_t* parent_dev, uint32_t* iters_run) {
  if (!c || !g || !hops_dev) return LA3_EINVAL;
  if (c->device != g->ctx->device) return fail(c, LA3_EINVAL, "graph lives on another device");
  if (root > g->n) return fail(c, LA3_EINVAL, "root id out of range");
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  HIPC(c, hipSetDevice(c->device));
  uint8_t* base = (uint8_t*)noalloc::base(c->device);
  if (!base) return fail(c, LA3_EDEVICE, "allocation-free scratch: hipGetSymbolAddress failed on this device");
  // Carve levels, marks and the two frontier queues out of the array.
  const uint64_t S = g->nslots;
  const uint64_t cap = S + g->nnz / CHUNK + 2;
  auto al = [](uint64_t b) { return (b + 255) & ~255ull; };
  const uint64_t need = al(S) + al(4 * S) + 2 * (al(8 * cap) + al(CTR_BYTES));
  if (need > noalloc::SCRATCH_BYTES)
    return fail(c, LA3_ENOTSUP, "la3_bfs_noalloc: the graph needs " + std::to_string(need) +
                                    " bytes of work buffers, more than the " +
                                    std::to_string(noalloc::SCRATCH_BYTES) + "-byte allocation-free scratch");
  uint64_t off = 0;
  auto take = [&](uint64_t b) { uint8_t* p = base + off; off += al(b); return p; };
  uint8_t* level = take(S);
  uint32_t* mark = (uint32_t*)take(4 * S);
  Fronts F;
  F.cap = cap;
  for (int i = 0; i < 2; i++) {
    F.f[i].items = (uint64_t*)take(8 * cap);
    uint8_t* ctr = take(CTR_BYTES);
    F.f[i].nitems = (uint32_t*)ctr;
    F.f[i].slots = (CtrSlot*)(ctr + 64);
  }
  // No per-launch profiling events here (event creation allocates).
  const bool prof = c->profiling;
  c->profiling = false;
  uint32_t iter = 0;
  la3_status st = bfs_core(c, g, root, max_iters, level, mark, F, &iter);
  if (!st) launch(c, "bfs_output", k_bfs_output, blocks_for(S), g->dg, (const uint8_t*)level, (const uint32_t*)mark,
                  hops_dev, parent_dev);
  c->profiling = prof;
  if (st) return st;
  HIPC(c, hipStreamSynchronize(c->stream));
  if (iters_run) *iters_run = max_iters ? max_iters : std::max<uint32_t>(iter, 1);
  return LA3_OK;
}

// --------------------------------------------------------------- SSSP ---
la3_status la3_sssp(la3_graph* g, uint32_t root, uint32_t max_iters, uint32_t* dist, uint32_t* iters_run) {
  if (!g) return LA3_EINVAL;
  return la3_sssp_ctx(g->ctx, g, root, max_iters, dist, iters_run);
}

la3_status la3_sssp_ctx(la3_ctx* c, la3_graph* g, uint32_t root, uint32_t max_iters, uint32_t* dist,
                        uint32_t* iters_run) {
  if (!c || !g) return LA3_EINVAL;
  if (c->device != g->ctx->device) return fail(c, LA3_EINVAL, "graph lives on another device");
  if (root > g->n) return fail(c, LA3_EINVAL, "root id out of range");
  if (!g->weighted) return fail(c, LA3_EINVAL, "SSSP needs a weighted graph");
  if (!g->directed) return fail(c, LA3_EINVAL, "SSSP runs on a directed load");
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  HIPC(c, hipSetDevice(c->device));
  timing_begin(c);
  const uint32_t S = g->nslots;
  uint32_t* dv = (uint32_t*)scratch(c, "sssp.dist", sizeof(uint32_t) * S);
  uint32_t* stamp = (uint32_t*)scratch(c, "sssp.stamp", sizeof(uint32_t) * S);
  // Bounded runs relax synchronously: messages are the distances at the end
  // of the previous step (sssp.msg), so m steps are the reference's m
  // iterations; unbounded runs relax asynchronously (same fixed point).
  const bool sync = max_iters > 0;
  uint32_t* msg = sync ? (uint32_t*)scratch(c, "sssp.msg", sizeof(uint32_t) * S) : nullptr;
  Fronts F;
  la3_status st = get_fronts(c, g, F, "bfs");
  if (st) return st;
  if (!dv || !stamp || (sync && !msg)) return fail(c, LA3_ENOMEM, "scratch allocation failed");
  const DG& d = g->dg;
  launch(c, "sssp_init", k_sssp_init, blocks_for(S), d, dv, stamp, g->slot_of_id[root], F.f[0], F.f[1].nitems);
  FrontCounts fc;
  if ((st = read_counts(c, F.f[0], fc))) return st;
  if (sync) {
    if (fc.items) launch(c, "sssp_snap", k_sssp_snap, blocks_for(fc.items), (const uint64_t*)F.f[0].items, fc.items,
                         (const uint32_t*)dv, msg);
    // execute(1) is execute_single (vertex_program_execute.hpp:20-32): the sink
    // rows take the root's initial message, so their step runs before the
    // regular rows' one.
    if (max_iters == 1 && g->nsink) launch(c, "sssp_sinks", k_sssp_sinks, blocks_for(g->nsink), d, dv);
  }
  uint32_t round = 0, cur = 0;
  const uint32_t wave_blocks = 4 * c->ncu * 8;
  const uint32_t delta = sync ? 0 : g->sssp_delta;
  uint32_t T = delta ? delta : 0xffffffffu;
  if (!sync && !delta && !c->profiling && fc.items > 0 && g->nnz <= PIPE_MAX_NNZ) {
    // Plain asynchronous Bellman-Ford on a small graph: no host decision
    // between steps but "stop when a step emits nothing", and steps short
    // enough that waiting for each step's counts costs as much as the step.
    // Up to PIPE_DEPTH steps are enqueued ahead, each reading its item count
    // on the device, clearing the counters it consumed and copying the ones
    // it produced into a pinned ring slot; the host reads a step's counts
    // once its copy has finished, while the next steps run. Steps enqueued
    // after the last non-empty one read a count of 0 and emit nothing. The
    // grid is fixed before the count is known: one residency of the device.
    // (R-MAT, one query: scale 18 1.19 -> 1.03 ms, scale 20 1.92 -> 1.58 ms;
    // at scale 22 and 24, with long steps and fixed grids, 3-20% slower,
    // so larger graphs keep the per-step host loop.)
    constexpr uint32_t PIPE_DEPTH = 3;
    const uint32_t pblocks = 4 * c->ncu * 2;
    for (uint32_t i = 0; i < PIPE_DEPTH; i++)
      if (!c->pipe_ev[i]) HIPC(c, hipEventCreateWithFlags(&c->pipe_ev[i], hipEventDisableTiming));
    uint32_t enq = 0;
    for (;;) {
      for (; enq < round + PIPE_DEPTH; enq++) {
        const uint32_t in = enq & 1;
        launch(c, "sssp_push", k_sssp_push<false>, pblocks, d, (const uint64_t*)F.f[in].items, 0u, dv, stamp,
               enq, T, F.f[in ^ 1], (uint32_t*)nullptr, (const uint32_t*)nullptr, (const uint32_t*)F.f[in].nitems);
        zero_counts(c, F.f[in]);
        uint8_t* slot = (uint8_t*)c->host_ctr + 4096 * (1 + enq % PIPE_DEPTH);
        HIPC(c, dcopy_async(slot, F.f[in ^ 1].nitems, CTR_BYTES, hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipEventRecord(c->pipe_ev[enq % PIPE_DEPTH], c->stream));
      }
      HIPC(c, hipEventSynchronize(c->pipe_ev[round % PIPE_DEPTH]));
      FrontCounts nc;
      parse_counts((const uint8_t*)c->host_ctr + 4096 * (1 + round % PIPE_DEPTH), nc);
      c->edges_scanned += fc.edges;
      fc = nc;
      round++;
      if (!nc.items) break;
    }
    HIPC(c, hipStreamSynchronize(c->stream));  // the steps enqueued past the last one
    cur = round & 1;
  }
  while (fc.items > 0 && (!sync || round < max_iters)) {
    const Frontier& nf = F.f[cur ^ 1];
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((fc.items + 3) / 4, wave_blocks));
    int ev;
    if (sync)
      ev = launch(c, "sssp_push", k_sssp_push<true>, blocks, d, (const uint64_t*)F.f[cur].items, fc.items, dv, stamp,
                  round, T, nf, F.f[cur].nitems, (const uint32_t*)msg, (const uint32_t*)nullptr);
    else
      ev = launch(c, "sssp_push", k_sssp_push<false>, blocks, d, (const uint64_t*)F.f[cur].items, fc.items, dv, stamp,
                  round, T, nf, F.f[cur].nitems, (const uint32_t*)nullptr, (const uint32_t*)nullptr);
    c->edges_scanned += fc.edges;
    FrontCounts nc;
    if ((st = read_counts(c, nf, nc))) return st;
    if (sync && nc.items)
      launch(c, "sssp_snap", k_sssp_snap, blocks_for(nc.items), (const uint64_t*)nf.items, nc.items,
             (const uint32_t*)dv, msg);
    // Algorithmic bytes (w = m = a = s_r = s_w = 4 B): (4 + w) per entry,
    // 8 B pointers + m per active column, (a + s_r + s_w) + m per activated row.
    annotate(c, ev, 8.0 * fc.edges + 12.0 * fc.cols + 16.0 * nc.verts, fc.edges);
    T = next_threshold(T, delta, nc.nmin);
    fc = nc;
    round++;
    cur ^= 1;
  }
  if (g->nsink && max_iters != 1) launch(c, "sssp_sinks", k_sssp_sinks, blocks_for(g->nsink), d, dv);
  uint32_t* out = (uint32_t*)scratch(c, "out.u32", sizeof(uint32_t) * S);
  if (!out) return fail(c, LA3_ENOMEM, "scratch allocation failed");
  launch(c, "gather_u32", k_gather_by_id<uint32_t>, blocks_for(S), d, (const uint32_t*)dv, out);
  if ((st = deliver(c, dist, out, sizeof(uint32_t) * S))) return st;
  HIPC(c, hipStreamSynchronize(c->stream));
  timing_end(c);
  if (iters_run) *iters_run = sync ? max_iters : round + 1;
  return LA3_OK;
}

// ----------------------------------------------------------- PageRank ---
}  // extern "C"

// One PR pass in MODE over the regular rows (sinks = false) or the sink rows.
// Rows longer than PSEG_ENTRIES: when no row has more than PR_BLOCK_SEGS
// segments, a block per row inside k_pr_step (one launch per pass); else
// their segments in k_pr_step, then k_pr_combine. Returns the event slot of
// the k_pr_step launch.
constexpr uint32_t PR_BLOCK_SEGS = 64;
template <int MODE>
static int pr_pass(la3_ctx* c, la3_graph* g, const PrArgs& a, bool sinks, double* part, const char* name) {
  const DG& d = g->dg;
  const uint32_t r0 = sinks ? g->nreg : 0, r1 = sinks ? g->nrows : g->nreg;
  uint32_t s0 = sinks ? d.npseg_reg : 0, s1 = sinks ? d.npseg : d.npseg_reg;
  const uint32_t p0 = sinks ? d.nplong_reg : 0, p1 = sinks ? d.nplong : d.nplong_reg;
  if (r1 <= r0) return -1;
  const bool block_rows = p1 > p0 && g->max_pseg <= PR_BLOCK_SEGS;
  if (block_rows) s1 = s0;  // no segments
  const uint32_t nbr = blocks_for(r1 - r0);
  const uint32_t nbw = std::min<uint32_t>(((uint64_t)d.nlong + (s1 - s0) + 3) / 4, 4096);
  const uint32_t nbb = block_rows ? std::min<uint32_t>(p1 - p0, 4096) : 0;
  const int ev = launch(c, name, k_pr_step<MODE>, nbr + nbw + nbb, d, a, r0, r1, nbr, s0, s1, part, nbw, p0,
                        block_rows ? p1 : p0);
  if (p1 > p0 && !block_rows)
    launch(c, "pr_combine", k_pr_combine<MODE>, blocks_for(p1 - p0), d, a, p0, p1, (const double*)part);
  return ev;
}

extern "C" {

static la3_status pr_common(la3_graph* g, uint32_t iters, double alpha, uint32_t seed_slot,
                            double* rank, uint32_t* degree, uint32_t* iters_run) {
  la3_ctx* c = g->ctx;
  if (!g->directed) return fail(c, LA3_EINVAL, "PageRank runs on a directed load");
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  HIPC(c, hipSetDevice(c->device));
  timing_begin(c);
  const uint32_t S = g->nslots;
  const DG& d = g->dg;
  double* x0 = (double*)scratch(c, "pr.x0", sizeof(double) * S);
  double* x1 = (double*)scratch(c, "pr.x1", sizeof(double) * S);
  double* rk = (double*)scratch(c, "pr.rank", sizeof(double) * S);
  double* ss = (double*)scratch(c, "pr.srcsum", sizeof(double) * (g->nreg + 1));
  uint32_t* deg = (uint32_t*)scratch(c, "pr.deg", sizeof(uint32_t) * S);
  uint32_t* act = (uint32_t*)scratch(c, "pr.active", 64);
  double* part = (double*)scratch(c, "pr.part", sizeof(double) * ((size_t)d.npseg + 1));
  if (!x0 || !x1 || !rk || !ss || !deg || !act || !part) return fail(c, LA3_ENOMEM, "scratch allocation failed");
  PrArgs a{};
  double* xa = x0;
  double* xb = x1;
  a.x = xa; a.xnext = xb; a.rank = rk; a.srcsum = ss; a.deg = deg; a.active = act;
  a.alpha = alpha; a.damp = 1.0 - alpha; a.tol = 1e-5; a.seed = seed_slot;
  launch(c, "outdeg", k_outdeg, blocks_for(S), d, deg);
  launch(c, "pr_init", k_pr_init, blocks_for(S), d, a);
  // Constant source contributions per regular row.
  pr_pass<0>(c, g, a, false, part, "pr_srcsum");
  const bool until = iters == 0;
  uint32_t it = 0;
  if (!until && !c->profiling) {
    // A fixed number of iterations needs no host decision between them: the
    // loop is captured once (per buffers and parameters) and replayed.
    uint64_t ab;
    std::memcpy(&ab, &alpha, 8);
    const std::vector<uint64_t> key = {(uint64_t)(uintptr_t)c, c->scratch_epoch, iters, (uint64_t)(uintptr_t)x0,
                                       (uint64_t)(uintptr_t)x1,
                                       (uint64_t)(uintptr_t)rk, (uint64_t)(uintptr_t)ss, (uint64_t)(uintptr_t)deg,
                                       (uint64_t)(uintptr_t)act, (uint64_t)(uintptr_t)part, seed_slot, ab};
    if (!g->pr_exec || g->pr_key != key) {
      if (g->pr_exec) (void)hipGraphExecDestroy(g->pr_exec);
      g->pr_exec = nullptr;
      hipGraph_t graph = nullptr;
      HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
      PrArgs b = a;
      for (uint32_t i = 0; i < iters; i++) {
        b.x = (i & 1) ? xb : xa;
        b.xnext = (i & 1) ? xa : xb;
        b.skip_rank = i + 1 < iters;  // only the last iteration's ranks are read
        pr_pass<1>(c, g, b, false, part, "pr_step");
      }
      hipError_t e = hipStreamEndCapture(c->stream, &graph);
      if (e != hipSuccess) return fail(c, LA3_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&g->pr_exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        g->pr_exec = nullptr;
        return fail(c, LA3_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
      g->pr_key = key;
    }
    HIPC(c, hipGraphLaunch(g->pr_exec, c->stream));
    it = iters;
    if (iters & 1) std::swap(xa, xb);  // xa: the final messages
  }
  const bool replayed = it > 0;
  for (; !replayed;) {
    if (!until && it >= iters) break;
    if (until) launch(c, "fill_u32", k_fill<uint32_t>, 1, act, (uint64_t)1, 0u);
    a.x = xa;
    a.xnext = xb;
    a.skip_rank = !until && it + 1 < iters;
    // Algorithmic bytes: 4 B index + 8 B message per regular-column entry,
    // 8 B source sum + 16 B rank read/written + 8 B next message per row.
    const int ev = pr_pass<1>(c, g, a, false, part, "pr_step");
    annotate(c, ev, 12.0 * g->nnz_reg + 32.0 * g->nreg, g->nnz_reg);
    std::swap(xa, xb);
    it++;
    if (until) {
      uint32_t h = 0;
      HIPC(c, dcopy_async(c->host_ctr, act, 4, hipMemcpyDeviceToHost, c->stream));
      HIPC(c, hipStreamSynchronize(c->stream));
      std::memcpy(&h, c->host_ctr, 4);
      if (!h) break;
    }
  }
  c->edges_scanned = (uint64_t)it * (g->nnz_reg) + g->nnz_sink;
  // Sink step: xa now holds the final regular messages. execute(1) is the
  // reference's execute_single (vertex_program_execute.hpp:20-32), whose sink
  // rows take the same initial messages as the regular rows: xb still holds
  // them after the one step.
  a.x = iters == 1 ? xb : xa;
  a.xnext = iters == 1 ? xa : xb;
  if (g->nsink) {
    pr_pass<2>(c, g, a, true, part, "pr_sinks");
  }
  double* outd = (double*)scratch(c, "out.f64", sizeof(double) * S);
  if (!outd) return fail(c, LA3_ENOMEM, "scratch allocation failed");
  launch(c, "gather_f64", k_gather_by_id<double>, blocks_for(S), d, (const double*)rk, outd);
  la3_status st = deliver(c, rank, outd, sizeof(double) * S);
  if (st) return st;
  if (degree) {
    uint32_t* outu = (uint32_t*)scratch(c, "out.u32", sizeof(uint32_t) * S);
    if (!outu) return fail(c, LA3_ENOMEM, "scratch allocation failed");
    launch(c, "gather_u32", k_gather_by_id<uint32_t>, blocks_for(S), d, (const uint32_t*)deg, outu);
    HIPC(c, dcopy_async(degree, outu, sizeof(uint32_t) * S, hipMemcpyDeviceToHost, c->stream));
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  timing_end(c);
  if (iters_run) *iters_run = it;
  return LA3_OK;
}

la3_status la3_pagerank(la3_graph* g, uint32_t iters, double* rank, uint32_t* degree,
                        uint32_t* iters_run) {
  if (!g) return LA3_EINVAL;
  return pr_common(g, iters, 0.15, 0xffffffffu, rank, degree, iters_run);
}

la3_status la3_ppr(la3_graph* g, uint32_t seed, uint32_t iters, double alpha, double* rank) {
  if (!g) return LA3_EINVAL;
  if (seed > g->n) return fail(g->ctx, LA3_EINVAL, "seed id out of range");
  if (!(alpha > 0.0 && alpha < 1.0)) return fail(g->ctx, LA3_EINVAL, "alpha must be in (0, 1)");
  if (iters == 0) return fail(g->ctx, LA3_EINVAL, "PPR needs a fixed iteration count");
  return pr_common(g, iters, alpha, g->slot_of_id[seed], rank, nullptr, nullptr);
}

// ----------------------------------------------------------------- CC ---
// Union-find components (no counterpart in the reference).
la3_status la3_components(la3_graph* g, uint32_t* label) {
  if (!g) return LA3_EINVAL;
  la3_ctx* c = g->ctx;
  if (g->directed) return fail(c, LA3_EINVAL, "CC runs on an undirected load");
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  HIPC(c, hipSetDevice(c->device));
  timing_begin(c);
  const uint32_t S = g->nslots;
  const DG& d = g->dg;
  uint32_t* p = (uint32_t*)scratch(c, "cc.parent", sizeof(uint32_t) * S);
  uint32_t* out = (uint32_t*)scratch(c, "out.u32", sizeof(uint32_t) * S);
  if (!p || !out) return fail(c, LA3_ENOMEM, "scratch allocation failed");
  launch(c, "cc_init", k_cc_init, blocks_for(S), d, p);
  if (g->nrows) {
    // Link every row to a smaller-id vertex among its first entries,
    // compress, and take the most frequent root of a sample of rows as the
    // largest tree; then only rows outside it union with all their entries.
    // Algorithmic bytes: row pointers, the scanned entries and their ids per
    // row (the finds' hops not counted); the finish pass's entries of rows
    // outside the largest tree are not counted either.
    constexpr uint32_t LINK_SCAN = 8, NSAMPLE = 512;
    const int ev = launch(c, "cc_link", k_cc_link, blocks_for(g->nrows), d, p, LINK_SCAN);
    annotate(c, ev, 8.0 * g->nrows + 8.0 * std::min<uint64_t>(g->nnz, (uint64_t)LINK_SCAN * g->nrows), g->nrows);
    const int evc = launch(c, "cc_compress", k_cc_compress, blocks_for(S), d, p);
    annotate(c, evc, 8.0 * S, 0);
    uint32_t* smp = (uint32_t*)scratch(c, "cc.sample", sizeof(uint32_t) * NSAMPLE);
    if (!smp) return fail(c, LA3_ENOMEM, "scratch allocation failed");
    launch(c, "cc_sample", k_cc_sample, 2, d, (const uint32_t*)p, NSAMPLE, 2654435761u, smp);
    HIPC(c, dcopy_async(c->host_ctr, smp, sizeof(uint32_t) * NSAMPLE, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    std::vector<uint32_t> roots((const uint32_t*)c->host_ctr, (const uint32_t*)c->host_ctr + NSAMPLE);
    std::sort(roots.begin(), roots.end());
    uint32_t big = roots[0], best = 0;
    for (size_t i = 0; i < roots.size();) {
      size_t j = i;
      while (j < roots.size() && roots[j] == roots[i]) j++;
      if (j - i > best) { best = (uint32_t)(j - i); big = roots[i]; }
      i = j;
    }
    const int evf = launch(c, "cc_finish", k_cc_finish, blocks_for(g->nrows), d, p, big, 0u);
    annotate(c, evf, 12.0 * g->nrows, g->nnz);
    if (d.nlong)
      launch(c, "cc_finish_long", k_cc_finish_long, std::min<uint32_t>((d.nlong + 3) / 4, 4096), d, p, big, 0u);
  }
  launch(c, "cc_output", k_cc_output, blocks_for(S), d, p, out);
  c->edges_scanned = g->nnz;
  la3_status st = deliver(c, label, out, sizeof(uint32_t) * S);
  if (st) return st;
  HIPC(c, hipStreamSynchronize(c->stream));
  timing_end(c);
  return LA3_OK;
}

// CcVertex (src/apps/graph_analytics/cc.h:22-38) on the vertex-program
// engine of include/la3_vertex_program.hpp, with the reference's operations.
la3_status la3_cc(la3_graph* g, uint32_t max_iters, uint32_t* label, uint32_t* iters_run) {
  if (!g) return LA3_EINVAL;
  la3_ctx* c = g->ctx;
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  HIPC(c, hipSetDevice(c->device));
  timing_begin(c);
  const uint32_t S = g->nslots;
  ScratchAlloc sa{c, "cc."};
  vp_detail::Engine<CcOps> e(device_view(g), c->stream, false, CcOps(), scratch_allocator(sa));
  hipError_t he = e.execute(max_iters);
  if (he != hipSuccess) return fail(c, LA3_EDEVICE, "CC: " + e.err);
  uint32_t* out = (uint32_t*)scratch(c, "out.u32", sizeof(uint32_t) * S);
  if (!out) return fail(c, LA3_ENOMEM, "scratch allocation failed");
  static_assert(sizeof(CcOps::S) == sizeof(uint32_t), "CC state is one label");
  launch(c, "vp_by_id", vp_detail::k_vp_by_id<CcOps::S>, blocks_for(S), device_view(g),
         (const CcOps::S*)e.bufs().v, (CcOps::S*)out);
  la3_status st = deliver(c, label, out, sizeof(uint32_t) * S);
  if (st) return st;
  HIPC(c, hipStreamSynchronize(c->stream));
  timing_end(c);
  if (iters_run) *iters_run = e.iters_run;
  return LA3_OK;
}

// ----------------------------------------------------------------- TC ---
la3_status la3_tc(la3_graph* g, uint64_t* total, uint32_t* per_vertex) {
  if (!g || !total) return LA3_EINVAL;
  la3_ctx* c = g->ctx;
  if (g->mode != LA3_ACYCLIC) return fail(c, LA3_EINVAL, "TC runs on an ACYCLIC load");
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  HIPC(c, hipSetDevice(c->device));
  timing_begin(c);
  const uint32_t S = g->nslots;
  const DG& d = g->dg;
  uint32_t* row_of = (uint32_t*)scratch(c, "tc.row_of", sizeof(uint32_t) * (g->nnz + 1));
  uint32_t* col_of = (uint32_t*)scratch(c, "tc.col_of", sizeof(uint32_t) * (g->nnz + 1));
  uint32_t* per = (uint32_t*)scratch(c, "tc.per", sizeof(uint32_t) * S);
  unsigned long long* tot = (unsigned long long*)scratch(c, "tc.total", 64);
  if (!row_of || !col_of || !per || !tot) return fail(c, LA3_ENOMEM, "scratch allocation failed");
  launch(c, "fill_u32", k_fill<uint32_t>, blocks_for(S), per, (uint64_t)S, 0u);
  launch(c, "fill_u64", k_fill<unsigned long long>, 1, tot, (uint64_t)6, 0ull);
  const uint32_t wmax = 4 * c->ncu * 8;
  // Entry -> row and entry -> column maps: each owner's first entry marked
  // with its index, then an inclusive max-scan (balanced whatever the list
  // lengths; a wavefront per 64 owners left hub columns on one wavefront).
  if (g->nnz > (uint64_t)INT32_MAX) return fail(c, LA3_EINVAL, "TC: more than 2^31 - 1 entries");
  uint32_t* mark = (uint32_t*)scratch(c, "tc.mark", sizeof(uint32_t) * (g->nnz + 1));
  size_t need = 0;
  if (hipcub::DeviceScan::InclusiveScan(nullptr, need, mark, row_of, hipcub::Max(), (int)g->nnz, c->stream) !=
      hipSuccess)
    return fail(c, LA3_EDEVICE, "TC: scan sizing failed");
  void* stmp = scratch(c, "tc.scan_tmp", std::max<size_t>(need, 16));
  if (!mark || !stmp) return fail(c, LA3_ENOMEM, "scratch allocation failed");
  const std::pair<const uint32_t*, uint32_t> owners[2] = {{d.row_ptr, g->nrows}, {d.col_ptr, d.ncols}};
  uint32_t* maps[2] = {row_of, col_of};
  for (int m = 0; m < 2; m++) {
    if (!g->nnz) break;
    launch(c, "fill_u32", k_fill<uint32_t>, blocks_for(g->nnz), mark, (uint64_t)g->nnz, 0u);
    launch(c, "tc_marks", k_tc_mark_owner, blocks_for(std::max<uint32_t>(owners[m].second, 1)), owners[m].first,
           owners[m].second, mark);
    if (hipcub::DeviceScan::InclusiveScan(stmp, need, mark, maps[m], hipcub::Max(), (int)g->nnz, c->stream) !=
        hipSuccess)
      return fail(c, LA3_EDEVICE, "TC: owner scan failed");
  }
  const int ev = launch(c, "tc_count", k_tc_count, blocks_for(g->nnz, wmax), d, (const uint32_t*)row_of, per, tot);
  const uint32_t hb = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)((g->nnz + BLOCK - 1) / BLOCK), 32 * c->ncu));
  // Owners of 33..1024 entries (8-KB tables: 8 blocks per CU), then
  // 1025..4096 (32 KB: 4 per CU). Splitting the upper band once more cost
  // more in its extra pass over the ranges than its occupancy gained.
  const int evh[4] = {
      launch(c, "tc_hash", k_tc_hash<0, TC_HASH_MIN, TC_HASH_SPLIT>, hb, d, (const uint32_t*)row_of, per, tot),
      launch(c, "tc_hash_big", k_tc_hash<0, TC_HASH_SPLIT, TC_HASH_MAX>, hb, d, (const uint32_t*)row_of, per, tot),
      launch(c, "tc_hash_cols", k_tc_hash<1, TC_HASH_MIN, TC_HASH_SPLIT>, hb, d, (const uint32_t*)col_of, per, tot),
      launch(c, "tc_hash_cols_big", k_tc_hash<1, TC_HASH_SPLIT, TC_HASH_MAX>, hb, d, (const uint32_t*)col_of, per,
             tot)};
  c->edges_scanned = g->nnz;
  unsigned long long h[6] = {0, 0, 0, 0, 0, 0};
  HIPC(c, dcopy_async(c->host_ctr, tot, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  if (per_vertex) {
    uint32_t* out = (uint32_t*)scratch(c, "out.u32", sizeof(uint32_t) * S);
    if (!out) return fail(c, LA3_ENOMEM, "scratch allocation failed");
    launch(c, "gather_u32", k_gather_by_id<uint32_t>, blocks_for(S), d, (const uint32_t*)per, out);
    la3_status st = deliver(c, per_vertex, out, sizeof(uint32_t) * S);
    if (st) return st;
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  std::memcpy(h, c->host_ctr, sizeof(h));
  *total = h[0];
  // Algorithmic bytes: 4 B per list entry a merge intersection of the pair
  // reads, by the kernel that counts the entry (h[1]: k_tc_count, h[2..5]:
  // the four hash launches), plus row id, column and both row pointers
  // (16 B) per entry for k_tc_count, which reads them for every entry.
  annotate(c, ev, 4.0 * h[1] + 16.0 * g->nnz, g->nnz);
  for (int k = 0; k < 4; k++) annotate(c, evh[k], 4.0 * h[2 + k], 0);
  timing_end(c);
  return LA3_OK;
}

// ------------------------------------------------------------ batched ---
}  // extern "C"

template <int W>
static la3_status msbfs_run(la3_ctx* c, la3_graph* g, const std::vector<uint32_t>& root_slots, uint32_t k,
                            uint32_t max_iters, uint8_t* hops, uint32_t* iters_run) {
  const uint32_t S = g->nslots;
  const uint32_t K = 64 * W;
  const DG& d = g->dg;
  uint64_t* vis = (uint64_t*)scratch(c, "msbfs.vis", sizeof(uint64_t) * S * W);
  uint64_t* fa = (uint64_t*)scratch(c, "msbfs.fa", sizeof(uint64_t) * S * W);
  uint64_t* fb = (uint64_t*)scratch(c, "msbfs.fb", sizeof(uint64_t) * S * W);
  uint8_t* lev = (uint8_t*)scratch(c, "msbfs.lev", (size_t)S * K);
  uint32_t* stamp = (uint32_t*)scratch(c, "msbfs.stamp", sizeof(uint32_t) * S);
  uint32_t* rs = (uint32_t*)scratch(c, "msbfs.roots", sizeof(uint32_t) * k);
  uint64_t* valid = (uint64_t*)scratch(c, "msbfs.valid", sizeof(uint64_t) * W);
  uint8_t* out = (uint8_t*)scratch(c, "msbfs.out", (size_t)S * k);
  Fronts F;
  la3_status st = get_fronts(c, g, F, "bfs");
  if (st) return st;
  Frontier T;
  if ((st = get_queue(c, "msbfs.touched", (uint64_t)S + 1, T))) return st;
  if (!vis || !fa || !fb || !lev || !stamp || !rs || !valid || !out)
    return fail(c, LA3_ENOMEM, "batched BFS: device allocation failed");
  uint64_t hv[W];
  for (int w = 0; w < W; w++) {
    const int bits = std::max(0, std::min(64, (int)k - 64 * w));
    hv[w] = bits == 64 ? ~0ull : ((1ull << bits) - 1);
  }
  HIPC(c, dcopy_async(valid, hv, sizeof(hv), hipMemcpyHostToDevice, c->stream));
  HIPC(c, dcopy_async(rs, root_slots.data(), sizeof(uint32_t) * k, hipMemcpyHostToDevice, c->stream));
  launch(c, "msbfs_init", k_msbfs_init<W>, blocks_for(S), d, vis, fa, fb, (uint64_t*)lev, stamp);
  launch(c, "msbfs_roots", k_msbfs_roots<W>, 1, d, (const uint32_t*)rs, k, vis, fa, lev);
  zero_counts(c, F.f[0]);
  launch(c, "msbfs_seed", k_msbfs_seed, blocks_for(k), d, (const uint32_t*)rs, k, stamp, F.f[0], 0);
  FrontCounts fc;
  if ((st = read_counts(c, F.f[0], fc))) return st;
  uint32_t iter = 0, cur = 0;
  bool bottom_up = false;
  const uint32_t wave_blocks = 4 * c->ncu * 8;
  while (fc.items > 0 && (max_iters == 0 || iter < max_iters)) {
    const Frontier& nf = F.f[cur ^ 1];
    zero_counts(c, nf);
    int ev = -1;
    if (!bottom_up) {
      zero_counts(c, T);
      const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((fc.items + 3) / 4, wave_blocks));
      ev = launch(c, "msbfs_push", k_msbfs_push<W>, blocks, d, (const uint64_t*)F.f[cur].items, fc.items,
                  (const uint64_t*)vis, (const uint64_t*)fa, fb, stamp, iter + 2, T);
      // The touched count bounds the apply grid.
      FrontCounts tc;
      if ((st = read_counts(c, T, tc))) return st;
      if (tc.items) launch(c, "msbfs_apply", k_msbfs_apply<W>, blocks_for(tc.items), d, (const uint64_t*)T.items,
                           tc.items, vis, fb, lev, iter, nf);
    } else {
      ev = launch(c, "msbfs_pull", k_msbfs_pull<W>, blocks_for(g->nrows, wave_blocks), d, vis,
                  (const uint64_t*)fa, fb, lev, iter, (const uint64_t*)valid, nf);
      if (d.nlong) launch(c, "msbfs_pull_long", k_msbfs_pull_long<W>, std::min<uint32_t>((d.nlong + 3) / 4, wave_blocks),
                          d, vis, (const uint64_t*)fa, fb, lev, iter, (const uint64_t*)valid, nf);
    }
    // The consumed frontier words are non-zero only at this step's work-list
    // vertices: clear those.
    launch(c, "msbfs_clear", k_clear_items<W>, blocks_for(fc.items), (const uint64_t*)F.f[cur].items, fc.items, fa);
    std::swap(fa, fb);
    FrontCounts nc;
    if ((st = read_counts(c, nf, nc))) return st;
    // Algorithmic bytes: k/8 B of frontier per entry and per active column,
    // plus k/8 B visited + k/8 B next frontier + the new level bytes per row.
    const double kb = k / 8.0;
    if (!bottom_up) {
      annotate(c, ev, (4.0 + kb) * fc.edges + (8.0 + kb) * fc.cols + 3 * kb * nc.verts, fc.edges);
      c->edges_scanned += fc.edges;
    } else {
      annotate(c, ev, (4.0 + kb) * nc.scanned + (8.0 + kb) * g->nrows + 2 * kb * nc.verts, nc.scanned);
      c->edges_scanned += nc.scanned;
    }
    iter++;
    if (!bottom_up && nc.edges * 20 > g->nnz) bottom_up = true;
    else if (bottom_up && (uint64_t)nc.verts * 24 < S) bottom_up = false;
    fc = nc;
    cur ^= 1;
  }
  launch(c, "msbfs_output", k_msbfs_output, blocks_for(S), d, (const uint8_t*)lev, K, k, out);
  if ((st = deliver(c, hops, out, (size_t)S * k))) return st;
  HIPC(c, hipStreamSynchronize(c->stream));
  if (iters_run) *iters_run = max_iters ? max_iters : std::max<uint32_t>(iter, 1);
  return LA3_OK;
}

extern "C" {

la3_status la3_bfs_batch(la3_graph* g, const uint32_t* roots, uint32_t k, uint32_t max_iters, uint8_t* hops,
                         uint32_t* iters_run) {
  if (!g) return LA3_EINVAL;
  return la3_bfs_batch_ctx(g->ctx, g, roots, k, max_iters, hops, iters_run);
}

la3_status la3_bfs_batch_ctx(la3_ctx* c, la3_graph* g, const uint32_t* roots, uint32_t k,
                             uint32_t max_iters, uint8_t* hops, uint32_t* iters_run) {
  if (!c || !g || !roots) return LA3_EINVAL;
  if (c->device != g->ctx->device) return fail(c, LA3_EINVAL, "graph lives on another device");
  if (k == 0 || k > 1024) return fail(c, LA3_EINVAL, "batch size must be in [1, 1024]");
  std::vector<uint32_t> rs(k);
  for (uint32_t q = 0; q < k; q++) {
    if (roots[q] > g->n) return fail(c, LA3_EINVAL, "root id out of range");
    rs[q] = g->slot_of_id[roots[q]];
  }
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  HIPC(c, hipSetDevice(c->device));
  timing_begin(c);
  la3_status st;
  const uint32_t W = (k + 63) / 64;
  if (W <= 1) st = msbfs_run<1>(c, g, rs, k, max_iters, hops, iters_run);
  else if (W <= 2) st = msbfs_run<2>(c, g, rs, k, max_iters, hops, iters_run);
  else if (W <= 4) st = msbfs_run<4>(c, g, rs, k, max_iters, hops, iters_run);
  else if (W <= 8) st = msbfs_run<8>(c, g, rs, k, max_iters, hops, iters_run);
  else st = msbfs_run<16>(c, g, rs, k, max_iters, hops, iters_run);
  timing_end(c);
  return st;
}

}  // extern "C"

// LPR < 64: batches of at most LPR queries on the narrow-row kernels
// (k_mssp_sub_*), K = LPR lanes per vertex; else K = 64 * QW.
template <int QW, int LPR = 64>
static la3_status mssp_run(la3_ctx* c, la3_graph* g, const std::vector<uint32_t>& root_slots, uint32_t k,
                           uint32_t* dist_out, uint32_t* iters_run) {
  const uint32_t S = g->nslots;
  constexpr bool SUB = LPR < 64;
  constexpr uint32_t K = SUB ? LPR : 64 * QW;
  const DG& d = g->dg;
  uint32_t* dist = (uint32_t*)scratch(c, "mssp.dist", sizeof(uint32_t) * (size_t)S * K);
  uint64_t* cm = (uint64_t*)scratch(c, "mssp.cm", sizeof(uint64_t) * (size_t)S * QW);
  uint64_t* nm = (uint64_t*)scratch(c, "mssp.nm", sizeof(uint64_t) * (size_t)S * QW);
  uint32_t* stamp = (uint32_t*)scratch(c, "mssp.stamp", sizeof(uint32_t) * S);
  uint32_t* rs = (uint32_t*)scratch(c, "mssp.roots", sizeof(uint32_t) * k);
  uint32_t* out = (uint32_t*)scratch(c, "mssp.out", sizeof(uint32_t) * (size_t)S * k);
  Fronts F;
  la3_status st = get_fronts(c, g, F, "bfs");
  if (st) return st;
  Frontier T;
  // One entry per improving scanned entry at most (with more than one mask
  // word per row, rows repeat until k_mssp_items deduplicates them).
  if ((st = get_queue(c, "mssp.touched", (uint64_t)S + g->nnz + 1, T))) return st;
  if (!dist || !cm || !nm || !stamp || !rs || !out)
    return fail(c, LA3_ENOMEM, "batched SSSP: device allocation failed");
  HIPC(c, dcopy_async(rs, root_slots.data(), sizeof(uint32_t) * k, hipMemcpyHostToDevice, c->stream));
  const uint64_t inf2 = ((uint64_t)SSSP_INF << 32) | SSSP_INF;
  if constexpr (K % 2)  // one lane per row: S may be odd
    launch(c, "fill_u32", k_fill<uint32_t>, blocks_for((uint64_t)S * K), dist, (uint64_t)S * K, SSSP_INF);
  else
    launch(c, "fill_u64", k_fill<unsigned long long>, blocks_for((uint64_t)S * K / 2), (unsigned long long*)dist,
           (uint64_t)S * K / 2, (unsigned long long)inf2);
  launch(c, "fill_u64", k_fill<unsigned long long>, blocks_for((uint64_t)S * QW), (unsigned long long*)cm,
         (uint64_t)S * QW, 0ull);
  launch(c, "fill_u64", k_fill<unsigned long long>, blocks_for((uint64_t)S * QW), (unsigned long long*)nm,
         (uint64_t)S * QW, 0ull);
  launch(c, "fill_u32", k_fill<uint32_t>, blocks_for(S), stamp, (uint64_t)S, 0u);
  launch(c, "mssp_roots", k_mssp_roots_k, 1, d, (const uint32_t*)rs, k, K, dist, cm);
  zero_counts(c, F.f[0]);
  zero_counts(c, F.f[1]);
  zero_counts(c, T);
  launch(c, "msbfs_seed", k_msbfs_seed, blocks_for(k), d, (const uint32_t*)rs, k, stamp, F.f[0], 1);
  FrontCounts fc;
  if ((st = read_counts(c, F.f[0], fc))) return st;
  uint32_t round = 0, cur = 0;
  const uint32_t wave_blocks = 4 * c->ncu * 8;
  const uint32_t delta = g->mssp_delta;
  uint32_t thr = delta ? delta : 0xffffffffu;
  while (fc.items > 0) {
    const Frontier& nf = F.f[cur ^ 1];
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((fc.items + 3) / 4, wave_blocks));
    if constexpr (SUB || QW == 1) {
      // One mask word per row: the push queues each improved row once and
      // emits its work items into nf itself (no touched list, no items pass,
      // one counter read per step).
      int ev;
      if constexpr (SUB)
        ev = launch(c, "mssp_push", k_mssp_sub_push<LPR>, blocks, d, (const uint64_t*)F.f[cur].items, fc.items,
                    dist, (const uint64_t*)cm, nm, thr, nf, F.f[cur].nitems);
      else
        ev = launch(c, "mssp_push", k_mssp_push<QW>, blocks, d, (const uint64_t*)F.f[cur].items, fc.items, dist,
                    (const uint64_t*)cm, nm, thr, nf, F.f[cur].nitems);
      launch(c, "mssp_clear", k_clear_items<QW>, blocks_for(fc.items), (const uint64_t*)F.f[cur].items, fc.items,
             cm);
      std::swap(cm, nm);
      FrontCounts nc;
      if ((st = read_counts(c, nf, nc))) return st;
      annotate(c, ev, (8.0 + 4.0 * K) * fc.edges + (8.0 + 4.0 * K) * fc.cols + (K / 8.0 + 4) * nc.verts, fc.edges);
      c->edges_scanned += fc.edges;
      thr = next_threshold(thr, delta, nc.nmin);
      round++;
      fc = nc;
      cur ^= 1;
      continue;
    }
    const int ev = launch(c, "mssp_push", k_mssp_push<QW>, blocks, d, (const uint64_t*)F.f[cur].items, fc.items,
                          dist, (const uint64_t*)cm, nm, thr, T, F.f[cur].nitems);
    FrontCounts tc;
    if ((st = read_counts(c, T, tc))) return st;
    // The touched list becomes the next frontier's work items; clearing T's
    // counters happens inside that launch.
    launch(c, "mssp_items", k_mssp_items, blocks_for(std::max<uint32_t>(tc.items, 1)), d, (const uint64_t*)T.items,
           tc.items, stamp, round + 2, (int)(QW > 1), nf, T.nitems);
    // cm of the finished step is cleared at the work-list vertices (the only
    // ones whose masks a push reads); nm becomes the next step's cm.
    launch(c, "mssp_clear", k_clear_items<QW>, blocks_for(fc.items), (const uint64_t*)F.f[cur].items, fc.items, cm);
    std::swap(cm, nm);
    FrontCounts nc;
    if ((st = read_counts(c, nf, nc))) return st;
    // Algorithmic bytes: (4 + w) per entry plus K x 4 B of distances read per
    // entry (the k-wide message row), K x 4 B per active column, and
    // K/8 B of masks per touched row.
    annotate(c, ev, (8.0 + 4.0 * K) * fc.edges + (8.0 + 4.0 * K) * fc.cols + (K / 8.0 + 4) * nc.verts, fc.edges);
    c->edges_scanned += fc.edges;
    thr = next_threshold(thr, delta, tc.nmin);
    round++;
    fc = nc;
    cur ^= 1;
  }
  if (g->nsink) {
    if constexpr (SUB)
      launch(c, "mssp_sinks", k_mssp_sub_sinks<LPR>,
             std::min<uint32_t>((g->nsink + 4 * (64 / LPR) - 1) / (4 * (64 / LPR)), wave_blocks), d, dist);
    else
      launch(c, "mssp_sinks", k_mssp_sinks<QW>, std::min<uint32_t>((g->nsink + 3) / 4, wave_blocks), d, dist);
  }
  launch(c, "mssp_output", k_batch_output<uint32_t>, blocks_for(S), d, (const uint32_t*)dist, K, k, out);
  if ((st = deliver(c, dist_out, out, sizeof(uint32_t) * (size_t)S * k))) return st;
  HIPC(c, hipStreamSynchronize(c->stream));
  if (iters_run) *iters_run = round + 1;
  return LA3_OK;
}

extern "C" {

la3_status la3_sssp_batch(la3_graph* g, const uint32_t* roots, uint32_t k, uint32_t max_iters, uint32_t* dist,
                          uint32_t* iters_run) {
  if (!g) return LA3_EINVAL;
  return la3_sssp_batch_ctx(g->ctx, g, roots, k, max_iters, dist, iters_run);
}

la3_status la3_sssp_batch_ctx(la3_ctx* c, la3_graph* g, const uint32_t* roots, uint32_t k,
                              uint32_t max_iters, uint32_t* dist, uint32_t* iters_run) {
  if (!c || !g || !roots) return LA3_EINVAL;
  if (c->device != g->ctx->device) return fail(c, LA3_EINVAL, "graph lives on another device");
  if (k == 0 || k > 1024) return fail(c, LA3_EINVAL, "batch size must be in [1, 1024]");
  if (!g->weighted) return fail(c, LA3_EINVAL, "SSSP needs a weighted graph");
  if (!g->directed) return fail(c, LA3_EINVAL, "SSSP runs on a directed load");
  std::vector<uint32_t> rs(k);
  for (uint32_t q = 0; q < k; q++) {
    if (roots[q] > g->n) return fail(c, LA3_EINVAL, "root id out of range");
    rs[q] = g->slot_of_id[roots[q]];
  }
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  HIPC(c, hipSetDevice(c->device));
  la3_status st;
  if (max_iters) {
    // Bounded: the queries one after another on the synchronous single-query
    // path (the k-wide push relaxes asynchronously), rows gathered on the
    // device in query order.
    const size_t S = g->nslots;
    for (uint32_t q = 0; q < k; q++) {
      if ((st = la3_sssp_ctx(c, g, roots[q], max_iters, nullptr, nullptr))) return st;
      uint32_t* rows = (uint32_t*)scratch(c, "msspb.out", sizeof(uint32_t) * S * k);
      if (!rows) return fail(c, LA3_ENOMEM, "batched SSSP: device allocation failed");
      HIPC(c, dcopy_async(rows + S * q, c->last_result, sizeof(uint32_t) * S, hipMemcpyDeviceToDevice, c->stream));
      if (q + 1 == k) {
        if ((st = deliver(c, dist, rows, sizeof(uint32_t) * S * k))) return st;
        HIPC(c, hipStreamSynchronize(c->stream));
      }
    }
    if (iters_run) *iters_run = max_iters;
    return LA3_OK;
  }
  timing_begin(c);
  const uint32_t QW = (k + 63) / 64;
  if (k == 1) st = mssp_run<1, 1>(c, g, rs, k, dist, iters_run);
  else if (k == 2) st = mssp_run<1, 2>(c, g, rs, k, dist, iters_run);
  else if (k <= 4) st = mssp_run<1, 4>(c, g, rs, k, dist, iters_run);
  else if (k <= 8) st = mssp_run<1, 8>(c, g, rs, k, dist, iters_run);
  else if (k <= 16) st = mssp_run<1, 16>(c, g, rs, k, dist, iters_run);
  else if (k <= 32) st = mssp_run<1, 32>(c, g, rs, k, dist, iters_run);
  else if (QW <= 1) st = mssp_run<1>(c, g, rs, k, dist, iters_run);
  else if (QW <= 2) st = mssp_run<2>(c, g, rs, k, dist, iters_run);
  else if (QW <= 4) st = mssp_run<4>(c, g, rs, k, dist, iters_run);
  else if (QW <= 8) st = mssp_run<8>(c, g, rs, k, dist, iters_run);
  else st = mssp_run<16>(c, g, rs, k, dist, iters_run);
  timing_end(c);
  return st;
}

}  // extern "C"

// LPR < 64: batches of at most LPR queries on the narrow-row kernels
// (k_mppr_sub_*), K = LPR lanes per vertex; else K = 64 * QW.
template <int QW, int LPR = 64>
static la3_status mppr_run(la3_ctx* c, la3_graph* g, const std::vector<uint32_t>& seed_slots, uint32_t k,
                           uint32_t iters, double alpha, float* rank, float* comp, double* residual) {
  const uint32_t S = g->nslots;
  constexpr bool SUB = LPR < 64;
  constexpr uint32_t K = SUB ? LPR : 64 * QW;
  constexpr uint32_t RPW = SUB ? 64 / LPR : 1;  // rows per wavefront
  const DG& d = g->dg;
  const size_t SK = (size_t)S * K;
  const uint32_t wave_blocks = 4 * c->ncu * 8;
  const uint32_t rblocks = std::max<uint32_t>(1, std::min<uint32_t>((g->nrows + 4 * RPW - 1) / (4 * RPW), wave_blocks));
  const uint32_t cblocks = std::max<uint32_t>(1, std::min<uint32_t>((d.nplong + 4 * RPW - 1) / (4 * RPW), wave_blocks));
  float* x0 = (float*)scratch(c, "mppr.x0", 4 * SK);
  float* x1 = (float*)scratch(c, "mppr.x1", 4 * SK);
  float* hi = (float*)scratch(c, "mppr.hi", 4 * SK);
  float* lo = (float*)scratch(c, "mppr.lo", 4 * SK);
  uint32_t* deg = (uint32_t*)scratch(c, "pr.deg", sizeof(uint32_t) * S);
  uint32_t* sd = (uint32_t*)scratch(c, "mppr.seeds", sizeof(uint32_t) * K);
  double* res_r = (double*)scratch(c, "mppr.res_rows", sizeof(double) * rblocks * K);
  double* res_c = (double*)scratch(c, "mppr.res_comb", sizeof(double) * cblocks * K);
  double* res_s = (double*)scratch(c, "mppr.res_sink", sizeof(double) * std::max(rblocks, cblocks) * K);
  float* part = (float*)scratch(c, "mppr.part", sizeof(float) * 2 * K * ((size_t)d.npseg + 1));
  float* out = (float*)scratch(c, "mppr.out", sizeof(float) * (size_t)S * k);
  float* outc = comp ? (float*)scratch(c, "mppr.outc", sizeof(float) * (size_t)S * k) : nullptr;
  if (!x0 || !x1 || !hi || !lo || !deg || !sd || !res_r || !res_c || !res_s || !part || !out || (comp && !outc))
    return fail(c, LA3_ENOMEM, "batched PPR: device allocation failed");
  std::vector<uint32_t> seeds(K, 0xffffffffu);
  std::copy(seed_slots.begin(), seed_slots.end(), seeds.begin());
  HIPC(c, dcopy_async(sd, seeds.data(), sizeof(uint32_t) * K, hipMemcpyHostToDevice, c->stream));
  PprBatchArgs a{};
  a.x = x0; a.xnext = x1; a.hi = hi; a.lo = lo; a.deg = deg; a.seeds = sd; a.k = k;
  a.alpha_hi = (float)alpha;
  a.alpha_lo = (float)(alpha - (double)a.alpha_hi);
  const double damp = 1.0 - alpha;
  a.damp_hi = (float)damp;
  a.damp_lo = (float)(damp - (double)a.damp_hi);
  a.part = part; a.seg_row = d.pseg_row; a.seg_beg = d.pseg_beg; a.long_ppr = d.plong;
  a.nlong_ppr = d.nplong; a.seg_first = d.pseg_first;
  launch(c, "outdeg", k_outdeg, blocks_for(S), d, deg);
  launch(c, "mppr_init", k_mppr_init_k<K>, blocks_for(SK), d, a, iters);
  float* xa = x0;
  float* xb = x1;
  int ev_rows = -1;
  // Rows and long-row segments in one launch (residual partials of its
  // rblocks blocks in rr), then the long rows' combine (partials in rc).
  auto step = [&](uint32_t r0, uint32_t r1, uint32_t s0, uint32_t s1, int loop, double* rr, double* rc) {
    a.x = xa; a.xnext = xb;
    a.resid = rr;
    if (s1 == s0) {
      if constexpr (SUB) ev_rows = launch(c, "mppr_rows", k_mppr_sub_rows<LPR>, rblocks, d, a, r0, r1, loop);
      else ev_rows = launch(c, "mppr_rows", k_mppr_rows<QW>, rblocks, d, a, r0, r1, loop);
    } else {
      PprBatchArgs as = a;
      as.seg_row = d.pseg_row + s0; as.seg_beg = d.pseg_beg + s0; as.part = part + (size_t)s0 * 2 * K;
      if constexpr (SUB)
        ev_rows = launch(c, "mppr_rows", k_mppr_sub_rowsegs<LPR>, rblocks, d, a, r0, r1, loop, as, s1 - s0);
      else
        ev_rows = launch(c, "mppr_rows", k_mppr_rowsegs<QW>, rblocks, d, a, r0, r1, loop, as, s1 - s0);
      a.resid = rc;
      if constexpr (SUB) launch(c, "mppr_combine", k_mppr_sub_combine<LPR>, cblocks, d, a, r0, r1, loop);
      else launch(c, "mppr_combine", k_mppr_combine<QW>, cblocks, d, a, r0, r1, loop);
    }
  };
  // Algorithmic bytes per iteration of the rows launch: 4 B index + 4k B
  // message row per entry of a regular row (live query lanes only), 20k B
  // of state and messages per regular row, 8k B per segment partial.
  const double kb = 4.0 * k;
  for (uint32_t it = 0; it < iters; it++) {
    a.skip_rank = it + 2 < iters;   // the last two iterations store their ranks,
    a.skip_resid = it + 1 < iters;  // the last one reads them back for its residual
    step(0, g->nreg, 0, d.npseg_reg, 1, res_r, res_c);
    annotate(c, ev_rows, (4.0 + kb) * g->nnz_reg + 5.0 * kb * g->nreg + 2.0 * kb * d.npseg_reg, g->nnz_reg);
    c->edges_scanned += g->nnz_reg;
    std::swap(xa, xb);
  }
  if (residual) {
    std::vector<double> hr((size_t)rblocks * K), hc((size_t)cblocks * K);
    HIPC(c, dcopy_async(hr.data(), res_r, sizeof(double) * hr.size(), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, dcopy_async(hc.data(), res_c, sizeof(double) * hc.size(), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    for (uint32_t q = 0; q < k; q++) {
      double t = 0;
      for (uint32_t b = 0; b < rblocks; b++) t += hr[(size_t)b * K + q];
      if (d.npseg_reg) for (uint32_t b = 0; b < cblocks; b++) t += hc[(size_t)b * K + q];
      residual[q] = iters ? t : 0.0;
    }
  }
  // Sink step from the final messages (vertex_program_execute.hpp:131-148);
  // with one iteration (execute_single, vertex_program_execute.hpp:20-32)
  // from the initial ones, which the other buffer still holds.
  if (g->nsink) {
    if (iters == 1) std::swap(xa, xb);
    a.skip_rank = 0;
    a.skip_resid = 1;  // sink ranks are stored; their residual is not reported
    step(g->nreg, g->nrows, d.npseg_reg, d.npseg, 0, res_s, res_s);
  }
  launch(c, "mppr_output", k_batch_output<float>, blocks_for(S), d, (const float*)hi, K, k, out);
  la3_status st = deliver(c, rank, out, sizeof(float) * (size_t)S * k);
  if (st) return st;
  if (comp) {
    launch(c, "mppr_output", k_batch_output<float>, blocks_for(S), d, (const float*)lo, K, k, outc);
    HIPC(c, dcopy_async(comp, outc, sizeof(float) * (size_t)S * k, hipMemcpyDeviceToHost, c->stream));
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  return LA3_OK;
}

extern "C" {

la3_status la3_ppr_batch(la3_graph* g, const uint32_t* seeds, uint32_t k, uint32_t iters,
                         double alpha, float* rank, float* rank_comp, double* residual) {
  if (!g) return LA3_EINVAL;
  return la3_ppr_batch_ctx(g->ctx, g, seeds, k, iters, alpha, rank, rank_comp, residual);
}

la3_status la3_ppr_batch_ctx(la3_ctx* c, la3_graph* g, const uint32_t* seeds, uint32_t k,
                             uint32_t iters, double alpha, float* rank, float* rank_comp,
                             double* residual) {
  if (!c || !g || !seeds) return LA3_EINVAL;
  if (c->device != g->ctx->device) return fail(c, LA3_EINVAL, "graph lives on another device");
  if (k == 0 || k > 1024) return fail(c, LA3_EINVAL, "batch size must be in [1, 1024]");
  if (!g->directed) return fail(c, LA3_EINVAL, "PPR runs on a directed load");
  if (!(alpha > 0.0 && alpha < 1.0)) return fail(c, LA3_EINVAL, "alpha must be in (0, 1)");
  if (iters == 0) return fail(c, LA3_EINVAL, "PPR needs a fixed iteration count");
  std::vector<uint32_t> ss(k);
  for (uint32_t q = 0; q < k; q++) {
    if (seeds[q] > g->n) return fail(c, LA3_EINVAL, "seed id out of range");
    ss[q] = g->slot_of_id[seeds[q]];
  }
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  HIPC(c, hipSetDevice(c->device));
  timing_begin(c);
  la3_status st;
  const uint32_t QW = (k + 63) / 64;
  if (k == 1) st = mppr_run<1, 1>(c, g, ss, k, iters, alpha, rank, rank_comp, residual);
  else if (k == 2) st = mppr_run<1, 2>(c, g, ss, k, iters, alpha, rank, rank_comp, residual);
  else if (k <= 4) st = mppr_run<1, 4>(c, g, ss, k, iters, alpha, rank, rank_comp, residual);
  else if (k <= 8) st = mppr_run<1, 8>(c, g, ss, k, iters, alpha, rank, rank_comp, residual);
  else if (k <= 16) st = mppr_run<1, 16>(c, g, ss, k, iters, alpha, rank, rank_comp, residual);
  else if (k <= 32) st = mppr_run<1, 32>(c, g, ss, k, iters, alpha, rank, rank_comp, residual);
  else if (QW <= 1) st = mppr_run<1>(c, g, ss, k, iters, alpha, rank, rank_comp, residual);
  else if (QW <= 2) st = mppr_run<2>(c, g, ss, k, iters, alpha, rank, rank_comp, residual);
  else if (QW <= 4) st = mppr_run<4>(c, g, ss, k, iters, alpha, rank, rank_comp, residual);
  else if (QW <= 8) st = mppr_run<8>(c, g, ss, k, iters, alpha, rank, rank_comp, residual);
  else st = mppr_run<16>(c, g, ss, k, iters, alpha, rank, rank_comp, residual);
  timing_end(c);
  return st;
}

la3_status la3_graph_set_tuning(la3_graph* g, la3_tunable what, uint32_t value) {
  if (!g) return LA3_EINVAL;
  if (what == LA3_TUNE_SSSP_DELTA) g->sssp_delta = value;
  else if (what == LA3_TUNE_BATCH_SSSP_DELTA) g->mssp_delta = value;
  else if (what >= LA3_TUNE_BFS_TASK_K && what <= LA3_TUNE_PPR_TASK_K) {
    if (value != 64 && value != 128 && value != 256) return fail(g->ctx, LA3_EINVAL, "task width must be 64, 128 or 256");
    g->task_k[what - LA3_TUNE_BFS_TASK_K] = value;
  } else if (what == LA3_TUNE_FUSED_BLOCKS) {
    if (value != 8 && value != 16 && value != 32) return fail(g->ctx, LA3_EINVAL, "blocks per CU must be 8, 16 or 32");
    g->fused_blocks = value;
  } else {
    return fail(g->ctx, LA3_EINVAL, "unknown tunable");
  }
  return LA3_OK;
}

la3_status la3_graph_get_tuning(const la3_graph* g, la3_tunable what, uint32_t* value) {
  if (!g || !value) return LA3_EINVAL;
  if (what == LA3_TUNE_SSSP_DELTA) *value = g->sssp_delta;
  else if (what == LA3_TUNE_BATCH_SSSP_DELTA) *value = g->mssp_delta;
  else if (what >= LA3_TUNE_BFS_TASK_K && what <= LA3_TUNE_PPR_TASK_K) *value = g->task_k[what - LA3_TUNE_BFS_TASK_K];
  else if (what == LA3_TUNE_FUSED_BLOCKS) *value = g->fused_blocks;
  else return LA3_EINVAL;
  return LA3_OK;
}

la3_status la3_autotune(la3_graph* g, const uint32_t* roots, uint32_t nroots, la3_autotune_report* rep) {
  if (!g || !roots || nroots == 0) return LA3_EINVAL;
  la3_ctx* c = g->ctx;
  if (!g->weighted || !g->directed) return fail(c, LA3_EINVAL, "autotune: SSSP needs a weighted directed load");
  static const uint32_t cand[] = {0, 32, 64, 128, 256, 512, 1024};
  const uint32_t nc = sizeof(cand) / sizeof(cand[0]);
  const uint32_t nsingle = std::min<uint32_t>(nroots, 4);
  const uint32_t kb = std::min<uint32_t>(nroots, 1024);
  auto clock = [] { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
  double best_s = 1e30, best_b = 1e30;
  uint32_t pick_s = 0, pick_b = 0;
  for (uint32_t i = 0; i < nc; i++) {
    g->sssp_delta = cand[i];
    la3_status st = la3_sssp(g, roots[0], 0, nullptr, nullptr);  // warm
    if (st) return st;
    double t0 = clock();
    for (uint32_t r = 0; r < nsingle; r++)
      if ((st = la3_sssp(g, roots[r], 0, nullptr, nullptr))) return st;
    const double ts = clock() - t0;
    if (rep) rep->single_ms[i] = ts * 1e3 / nsingle;
    if (ts < best_s) { best_s = ts; pick_s = cand[i]; }
    g->mssp_delta = cand[i];
    double tb = 1e30;  // best of three: candidates often differ by a few percent
    for (int r = 0; r < 3; r++) {
      t0 = clock();
      if ((st = la3_sssp_batch(g, roots, kb, 0, nullptr, nullptr))) return st;
      tb = std::min(tb, clock() - t0);
    }
    if (rep) rep->batch_ms[i] = tb * 1e3;
    if (tb < best_b) { best_b = tb; pick_b = cand[i]; }
  }
  g->sssp_delta = pick_s;
  g->mssp_delta = pick_b;
  if (rep) {
    rep->ncandidates = nc;
    for (uint32_t i = 0; i < nc; i++) rep->candidates[i] = cand[i];
    rep->sssp_delta = pick_s;
    rep->batch_sssp_delta = pick_b;
    rep->batch_k = kb;
    rep->fused_blocks = 0;
    for (int i = 0; i < 3; i++) rep->task_k[i] = 0;
  }
  if (nroots < 64) return LA3_OK;
  // Fused runs: task width per program, then blocks per CU. PPR state is 16
  // B per vertex and query lane, so its runs take at most 128 seeds.
  const uint32_t kf = std::min<uint32_t>(nroots, 256);
  auto run = [&](std::vector<la3_batch>& bs, double& ms) -> la3_status {
    la3_status st = la3_batch_run(c, bs.data(), (uint32_t)bs.size(), 0.15);  // warm (scratch, graph capture)
    if (st) return st;
    ms = 1e30;
    for (int r = 0; r < 3; r++) {  // best of three
      const double t0 = clock();
      if ((st = la3_batch_run(c, bs.data(), (uint32_t)bs.size(), 0.15))) return st;
      ms = std::min(ms, (clock() - t0) * 1e3);
    }
    return LA3_OK;
  };
  auto batch = [&](la3_query_kind kind) {
    la3_batch b;
    std::memset(&b, 0, sizeof(b));
    b.graph = g; b.kind = kind; b.roots = roots;
    b.k = kind == LA3_Q_PPR ? std::min<uint32_t>(kf, 128) : kf;
    b.iters = kind == LA3_Q_PPR ? 5 : 0;
    return b;
  };
  static const uint32_t widths[3] = {64, 128, 256};
  static const uint32_t blocks[3] = {8, 16, 32};
  for (int kind = 0; kind < 3; kind++) {
    double best = 1e30;
    uint32_t pick = 256;
    for (int wi = 0; wi < 3; wi++) {
      if (rep) rep->task_ms[kind][wi] = 0;
      if (wi > 0 && widths[wi - 1] >= batch((la3_query_kind)kind).k) continue;  // same split as the narrower width
      g->task_k[kind] = widths[wi];
      std::vector<la3_batch> bs = {batch((la3_query_kind)kind)};
      double ms = 0;
      la3_status st = run(bs, ms);
      if (st) return st;
      if (rep) rep->task_ms[kind][wi] = ms;
      // a wider task only when more than 1% faster alone: within that, the
      // narrower one shares fused launches better with the other programs
      // (scale 24: SSSP 128 and 256 within 0.03% alone, 913 vs 946 ms mixed)
      if (ms < best * 0.99) { best = ms; pick = widths[wi]; }
    }
    g->task_k[kind] = pick;
    if (rep) rep->task_k[kind] = pick;
  }
  double best = 1e30;
  uint32_t pick = 32;
  for (int bi = 0; bi < 3; bi++) {
    g->fused_blocks = blocks[bi];
    std::vector<la3_batch> bs = {batch(LA3_Q_BFS), batch(LA3_Q_SSSP)};
    double ms = 0;
    la3_status st = run(bs, ms);
    if (st) return st;
    if (rep) rep->fused_ms[bi] = ms;
    if (ms < best) { best = ms; pick = blocks[bi]; }
  }
  g->fused_blocks = pick;
  if (rep) rep->fused_blocks = pick;
  // The trial runs' task-slot buffers are not kept.
  scratch_release(c, "fz");
  return LA3_OK;
}

la3_status la3_copy_bandwidth(la3_ctx* c, uint64_t bytes, uint32_t reps, double* gbs) {
  if (!c || !gbs || bytes < 16 || reps == 0) return LA3_EINVAL;
  if (c->device < 0) return fail(c, LA3_EDEVICE, "host-only context");
  HIPC(c, hipSetDevice(c->device));
  const uint64_t n = bytes / 16;
  uint4* a = (uint4*)scratch(c, "copy.src", n * 16);
  uint4* b = (uint4*)scratch(c, "copy.dst", n * 16);
  if (!a || !b) return fail(c, LA3_ENOMEM, "copy buffers");
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + BLOCK - 1) / BLOCK, (uint64_t)c->ncu * 32);
  launch(c, "fill_u64", k_fill<unsigned long long>, blocks_for(n * 2), (unsigned long long*)a, n * 2, 1ull);
  hipLaunchKernelGGL(la3_copy, dim3(blocks), dim3(BLOCK), 0, c->stream, (const uint4*)a, b, n);
  hipEvent_t e0 = get_event(c), e1 = get_event(c);
  HIPC(c, hipEventRecord(e0, c->stream));
  for (uint32_t r = 0; r < reps; r++)
    hipLaunchKernelGGL(la3_copy, dim3(blocks), dim3(BLOCK), 0, c->stream, (const uint4*)a, b, n);
  HIPC(c, hipEventRecord(e1, c->stream));
  HIPC(c, hipEventSynchronize(e1));
  float ms = 0.f;
  HIPC(c, hipEventElapsedTime(&ms, e0, e1));
  c->free_events.push_back(e0);
  c->free_events.push_back(e1);
  *gbs = 2.0 * n * 16 * reps / (ms * 1e-3) / 1e9;
  return LA3_OK;
}

la3_status la3_result_fetch(la3_ctx* c, void* dst, uint64_t bytes) {
  if (!c || !dst) return LA3_EINVAL;
  if (!c->last_result) return fail(c, LA3_EINVAL, "no result on the device");
  if (bytes > c->last_result_bytes) return fail(c, LA3_EINVAL, "fetch larger than the last result");
  if (!dev_covers(c->last_result, bytes)) return fail(c, LA3_EINVAL, "the last result is no longer in device memory");
  HIPC(c, dcopy_async(dst, c->last_result, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LA3_OK;
}

la3_status la3_result_fetch_slots(la3_ctx* c, const la3_graph* g, uint32_t elem_bytes, uint32_t rows, void* dst) {
  if (!c || !g || !dst || rows == 0 || (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8))
    return LA3_EINVAL;
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  if (g->ctx->device != c->device) return fail(c, LA3_EINVAL, "graph lives on another device");
  if (!c->last_result) return fail(c, LA3_EINVAL, "no result on the device");
  const uint64_t S = g->nslots, bytes = (uint64_t)elem_bytes * S * rows;
  if (bytes > c->last_result_bytes) return fail(c, LA3_EINVAL, "fetch larger than the last result");
  if (!dev_covers(c->last_result, bytes)) return fail(c, LA3_EINVAL, "the last result is no longer in device memory");
  const void* src = c->last_result;  // the scratch call below must not move it
  void* out = scratch(c, "out.slots", bytes);
  if (!out) return fail(c, LA3_ENOMEM, "scratch allocation failed");
  if (!dev_covers(src, bytes)) return fail(c, LA3_EINVAL, "the last result is no longer in device memory");
  launch(c, "by_slot", k_rows_by_slot, blocks_for(S * rows), g->dg.id_of, (const uint8_t*)src, (uint8_t*)out, S,
         (uint32_t)rows, elem_bytes);
  HIPC(c, dcopy_async(dst, out, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LA3_OK;
}

la3_status la3_host_alloc(uint64_t bytes, void** out) {
  if (!out) return LA3_EINVAL;
  *out = nullptr;
  if (hipHostMalloc(out, std::max<uint64_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) return LA3_ENOMEM;
  return LA3_OK;
}

void la3_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

// -------------------------------------------------------------- tools ---
la3_status la3_rmat_generate(uint32_t scale, uint32_t edgefactor, double a, double b, double c,
                             uint64_t seed, int weighted, uint32_t* out, uint64_t* nrecords) {
  if (!nrecords || scale == 0 || scale > 31 || edgefactor == 0) return LA3_EINVAL;
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) return LA3_EINVAL;
  // The generator is deterministic; the count query runs it once and caches
  // the records for the immediately following fill call with the same args.
  static std::vector<uint32_t> cache;
  static std::string key;
  char kb[160];
  snprintf(kb, sizeof(kb), "%u/%u/%.17g/%.17g/%.17g/%llu/%d", scale, edgefactor, a, b, c,
           (unsigned long long)seed, weighted);
  if (key != kb) {
    cache = rmat_generate(scale, edgefactor, a, b, c, seed, weighted != 0);
    key = kb;
  }
  const uint64_t W = weighted ? 3 : 2;
  *nrecords = cache.size() / W;
  if (out) {
    std::memcpy(out, cache.data(), cache.size() * sizeof(uint32_t));
    std::vector<uint32_t>().swap(cache);
    key.clear();
  }
  return LA3_OK;
}

}  // extern "C"

// ====================================================== fused batch runs ===
// la3_batch_run: every batch is split into tasks of at most 256 queries;
// all tasks advance together through k_fused_a / k_fused_b / k_fused_c
// (kernels_fused.h), several steps per replay of one captured hipGraph:
// FZ_STEPS for la3_batch_run (whose batches all start together), one for a
// session, so that its batches can join and leave at every step.
namespace {

constexpr uint32_t FZ_STEPS = 4;
// Session sub-steps: tasks on graphs with at most 1/SUB_RATIO of the entries
// of the largest running task's graph, up to SUB_MAX sub-steps per step.
constexpr uint64_t SUB_RATIO = 8;
constexpr uint32_t SUB_MAX = 3;
constexpr uint32_t FZ_BLOCKS_PER_CU = 32;
constexpr size_t FZ_SLOT_KEEP_BYTES = 256ull << 20;  // a finished session slot keeps up to this much

struct FzTask {
  uint32_t batch, q0, k;
};

// Blocks of the A / pull launches (split among the tasks by k_fused_c):
// blocks_per_cu per CU (LA3_TUNE_FUSED_BLOCKS, default 32); B launches 8 per CU.
uint32_t fz_ga(const la3_ctx* c, uint32_t blocks_per_cu) { return blocks_per_cu * c->ncu; }
uint32_t fz_gb(const la3_ctx* c) { return 4u * c->ncu * 2; }

std::string fz_name(uint32_t t, const char* what) { return "fz" + std::to_string(t) + "." + what; }

// Allocates and initialises one task's state on the context's stream and
// fills its FTask entry (the single-family drivers' initialisation).
la3_status fz_setup(la3_ctx* c, const la3_batch& b, const FzTask& ft, uint32_t t, double alpha, uint32_t ga_blocks,
                    FTask& T) {
  la3_graph* g = b.graph;
  const DG& d = g->dg;
  const uint32_t S = g->nslots;
  // SSSP and PPR tasks of at most 32 queries take LPR lanes per vertex row
  // (the batched calls' narrow rows); the others 64, 128 or 256.
  const uint32_t k = ft.k;
  const bool narrow = b.kind != LA3_Q_BFS && k <= 32;
  const uint32_t LPR = !narrow ? 0u : k <= 1 ? 1u : k <= 2 ? 2u : k <= 4 ? 4u : k <= 8 ? 8u : k <= 16 ? 16u : 32u;
  const uint32_t W = narrow ? FW_NARROW : (k + 63) / 64 == 3 ? 4 : (k + 63) / 64;
  const uint32_t K = narrow ? LPR : 64 * W;
  std::memset(&T, 0, sizeof(T));
  T.g = d;
  T.kind = (uint32_t)b.kind;
  T.width = W;
  T.lpr = LPR;
  T.cls = narrow ? narrow_cls(T.kind, LPR) : wide_cls(W);
  T.k = k;
  T.iters = b.iters;
  std::vector<uint32_t> slots(k);
  for (uint32_t q = 0; q < k; q++) slots[q] = g->slot_of_id[b.roots[ft.q0 + q]];
  la3_status st;
  auto buf = [&](const char* what, size_t bytes) { return scratch(c, fz_name(t, what).c_str(), bytes); };
  const size_t SW = (size_t)S * (narrow ? 1 : W), SK = (size_t)S * K;  // mask words, lanes
  if (b.kind != LA3_Q_PPR) {
    const bool bfs = b.kind == LA3_Q_BFS;
    uint32_t* rs = (uint32_t*)buf("roots", sizeof(uint32_t) * k);
    T.stamp = (uint32_t*)buf("stamp", sizeof(uint32_t) * S);
    if (!rs || !T.stamp) return fail(c, LA3_ENOMEM, "fused run: device allocation failed");
    // Measurement runs (profiling on) count the rows that receive a message
    // in each step, for the algorithmic-byte accounting.
    T.tstamp = nullptr;
    if (c->profiling) {
      T.tstamp = (uint32_t*)buf("tstamp", sizeof(uint32_t) * S);
      if (!T.tstamp) return fail(c, LA3_ENOMEM, "fused run: device allocation failed");
      launch(c, "fill_u32", k_fill<uint32_t>, blocks_for(S), T.tstamp, (uint64_t)S, 0u);
    }
    const uint64_t qcap = (uint64_t)S + g->nnz / CHUNK + 2;
    // (a narrow SSSP push queues rows straight into the next list: its
    // touched list only holds counters)
    const uint64_t tcap = narrow ? 64 : bfs ? (uint64_t)S + 1 : (uint64_t)S + g->nnz + 1;
    if ((st = get_queue(c, fz_name(t, "q0"), qcap, T.q[0])) || (st = get_queue(c, fz_name(t, "q1"), qcap, T.q[1])) ||
        (st = get_queue(c, fz_name(t, "touched"), tcap, T.touched)))
      return st;
    HIPC(c, dcopy_async(rs, slots.data(), sizeof(uint32_t) * k, hipMemcpyHostToDevice, c->stream));
    if (bfs) {
      T.vis = (uint64_t*)buf("vis", 8 * SW);
      T.fr[0] = (uint64_t*)buf("fr0", 8 * SW);
      T.fr[1] = (uint64_t*)buf("fr1", 8 * SW);
      T.lev = (uint8_t*)buf("lev", SK);
      uint64_t* valid = (uint64_t*)buf("valid", 8 * W);
      if (!T.vis || !T.fr[0] || !T.fr[1] || !T.lev || !valid) return fail(c, LA3_ENOMEM, "fused run: device allocation failed");
      uint64_t hv[FMAX_WIDTH] = {0, 0, 0, 0};
      for (uint32_t w = 0; w < W; w++) {
        const int bits = std::max(0, std::min(64, (int)k - 64 * (int)w));
        hv[w] = bits == 64 ? ~0ull : ((1ull << bits) - 1);
      }
      HIPC(c, dcopy_async(valid, hv, 8 * W, hipMemcpyHostToDevice, c->stream));
      T.valid = valid;
      if (W == 1) {
        launch(c, "msbfs_init", k_msbfs_init<1>, blocks_for(S), d, T.vis, T.fr[0], T.fr[1], (uint64_t*)T.lev, T.stamp);
        launch(c, "msbfs_roots", k_msbfs_roots<1>, 1, d, (const uint32_t*)rs, k, T.vis, T.fr[0], T.lev);
      } else if (W == 2) {
        launch(c, "msbfs_init", k_msbfs_init<2>, blocks_for(S), d, T.vis, T.fr[0], T.fr[1], (uint64_t*)T.lev, T.stamp);
        launch(c, "msbfs_roots", k_msbfs_roots<2>, 1, d, (const uint32_t*)rs, k, T.vis, T.fr[0], T.lev);
      } else {
        launch(c, "msbfs_init", k_msbfs_init<4>, blocks_for(S), d, T.vis, T.fr[0], T.fr[1], (uint64_t*)T.lev, T.stamp);
        launch(c, "msbfs_roots", k_msbfs_roots<4>, 1, d, (const uint32_t*)rs, k, T.vis, T.fr[0], T.lev);
      }
    } else {
      T.delta = g->mssp_delta;
      T.dist = (uint32_t*)buf("dist", 4 * SK);
      T.m[0] = (uint64_t*)buf("m0", 8 * SW);
      T.m[1] = (uint64_t*)buf("m1", 8 * SW);
      if (!T.dist || !T.m[0] || !T.m[1]) return fail(c, LA3_ENOMEM, "fused run: device allocation failed");
      const uint64_t inf2 = ((uint64_t)SSSP_INF << 32) | SSSP_INF;
      if (SK % 2)  // one lane per row: S may be odd
        launch(c, "fill_u32", k_fill<uint32_t>, blocks_for(SK), T.dist, (uint64_t)SK, SSSP_INF);
      else
        launch(c, "fill_u64", k_fill<unsigned long long>, blocks_for(SK / 2), (unsigned long long*)T.dist,
               (uint64_t)(SK / 2), (unsigned long long)inf2);
      launch(c, "fill_u64", k_fill<unsigned long long>, blocks_for(SW), (unsigned long long*)T.m[0], (uint64_t)SW, 0ull);
      launch(c, "fill_u64", k_fill<unsigned long long>, blocks_for(SW), (unsigned long long*)T.m[1], (uint64_t)SW, 0ull);
      launch(c, "fill_u32", k_fill<uint32_t>, blocks_for(S), T.stamp, (uint64_t)S, 0u);
      if (narrow) launch(c, "mssp_roots", k_mssp_roots_k, 1, d, (const uint32_t*)rs, k, K, T.dist, T.m[0]);
      else if (W == 1) launch(c, "mssp_roots", k_mssp_roots<1>, 1, d, (const uint32_t*)rs, k, T.dist, T.m[0]);
      else if (W == 2) launch(c, "mssp_roots", k_mssp_roots<2>, 1, d, (const uint32_t*)rs, k, T.dist, T.m[0]);
      else launch(c, "mssp_roots", k_mssp_roots<4>, 1, d, (const uint32_t*)rs, k, T.dist, T.m[0]);
    }
    zero_counts(c, T.q[0]);
    zero_counts(c, T.q[1]);
    zero_counts(c, T.touched);
    launch(c, "msbfs_seed", k_msbfs_seed, blocks_for(k), d, (const uint32_t*)rs, k, T.stamp, T.q[0], bfs ? 0 : 1);
    return LA3_OK;
  }
  // PPR
  const uint32_t ga = ga_blocks + FMAX_TASKS, gb = fz_gb(c) + FMAX_TASKS;
  T.xs[0] = (float*)buf("x0", 4 * SK);
  T.xs[1] = (float*)buf("x1", 4 * SK);
  float* hi = (float*)buf("hi", 4 * SK);
  float* lo = (float*)buf("lo", 4 * SK);
  uint32_t* deg = (uint32_t*)buf("deg", sizeof(uint32_t) * S);
  uint32_t* sd = (uint32_t*)buf("seeds", sizeof(uint32_t) * K);
  float* part = (float*)buf("part", sizeof(float) * 2 * K * ((size_t)d.npseg + 1));
  T.res_rows = (double*)buf("res_rows", sizeof(double) * ga * K);
  T.res_comb = (double*)buf("res_comb", sizeof(double) * gb * K);
  T.resid = (double*)buf("resid", sizeof(double) * K);
  if (!T.xs[0] || !T.xs[1] || !hi || !lo || !deg || !sd || !part || !T.res_rows || !T.res_comb || !T.resid)
    return fail(c, LA3_ENOMEM, "fused run: device allocation failed");
  std::vector<uint32_t> seeds(K, 0xffffffffu);
  std::copy(slots.begin(), slots.end(), seeds.begin());
  HIPC(c, dcopy_async(sd, seeds.data(), sizeof(uint32_t) * K, hipMemcpyHostToDevice, c->stream));
  PprBatchArgs& a = T.pa;
  a.x = T.xs[0]; a.xnext = T.xs[1]; a.hi = hi; a.lo = lo; a.deg = deg; a.seeds = sd; a.k = k;
  a.alpha_hi = (float)alpha;
  a.alpha_lo = (float)(alpha - (double)a.alpha_hi);
  const double damp = 1.0 - alpha;
  a.damp_hi = (float)damp;
  a.damp_lo = (float)(damp - (double)a.damp_hi);
  a.part = part; a.seg_row = d.pseg_row; a.seg_beg = d.pseg_beg; a.long_ppr = d.plong;
  a.nlong_ppr = d.nplong; a.seg_first = d.pseg_first;
  // Algorithmic bytes per iteration ((4 + w) E + (8 + m) C + (a + s_r + s_w)
  // R_touch + m R_act): every regular-row entry, every column and every
  // regular row, per live query an f32 message (m = 4k), a compensated sum
  // (a = 8k) and the compensated rank read and written (s_r = s_w = 8k).
  const double kb = 4.0 * k;
  T.step_bytes = 4.0 * g->nnz_reg + (8.0 + kb) * g->ncols + 6.0 * kb * g->nreg + kb * g->nreg;
  T.step_edges = g->nnz_reg;
  launch(c, "outdeg", k_outdeg, blocks_for(S), d, deg);
  switch (narrow ? LPR : 64 * W) {
    case 1: launch(c, "mppr_init", k_mppr_init_k<1>, blocks_for(SK), d, a, T.iters); break;
    case 2: launch(c, "mppr_init", k_mppr_init_k<2>, blocks_for(SK), d, a, T.iters); break;
    case 4: launch(c, "mppr_init", k_mppr_init_k<4>, blocks_for(SK), d, a, T.iters); break;
    case 8: launch(c, "mppr_init", k_mppr_init_k<8>, blocks_for(SK), d, a, T.iters); break;
    case 16: launch(c, "mppr_init", k_mppr_init_k<16>, blocks_for(SK), d, a, T.iters); break;
    case 32: launch(c, "mppr_init", k_mppr_init_k<32>, blocks_for(SK), d, a, T.iters); break;
    case 64: launch(c, "mppr_init", k_mppr_init<1>, blocks_for(SK), d, a, T.iters); break;
    case 128: launch(c, "mppr_init", k_mppr_init<2>, blocks_for(SK), d, a, T.iters); break;
    default: launch(c, "mppr_init", k_mppr_init<4>, blocks_for(SK), d, a, T.iters); break;
  }
  return LA3_OK;
}

// Post-loop work of one task and delivery of its results (LPR < 64: a
// narrow task, LPR lanes per vertex row).
template <int QW, int LPR = 64>
la3_status fz_finish(la3_ctx* c, const la3_batch& b, const FzTask& ft, uint32_t t, const FTask& T, const FCtl& C) {
  la3_graph* g = b.graph;
  const DG& d = g->dg;
  constexpr bool SUB = LPR < 64;
  constexpr uint32_t RPW = SUB ? 64 / LPR : 1;  // rows per wavefront
  const uint32_t S = g->nslots, K = SUB ? LPR : 64 * QW, k = ft.k;
  const uint32_t wave_blocks = 4 * c->ncu * 8;
  const size_t esz = b.kind == LA3_Q_BFS ? 1 : 4;
  void* out = scratch(c, fz_name(t, "out").c_str(), esz * (size_t)S * k);
  if (!out) return fail(c, LA3_ENOMEM, "fused run: device allocation failed");
  void* host = b.out ? (uint8_t*)b.out + esz * (size_t)S * ft.q0 : nullptr;
  if (b.kind == LA3_Q_BFS) {
    launch(c, "msbfs_output", k_msbfs_output, blocks_for(S), d, (const uint8_t*)T.lev, K, k, (uint8_t*)out);
  } else if (b.kind == LA3_Q_SSSP) {
    if (g->nsink) {
      if constexpr (SUB)
        launch(c, "mssp_sinks", k_mssp_sub_sinks<LPR>,
               std::min<uint32_t>((g->nsink + 4 * RPW - 1) / (4 * RPW), wave_blocks), d, T.dist);
      else
        launch(c, "mssp_sinks", k_mssp_sinks<QW>, std::min<uint32_t>((g->nsink + 3) / 4, wave_blocks), d, T.dist);
    }
    launch(c, "mssp_output", k_batch_output<uint32_t>, blocks_for(S), d, (const uint32_t*)T.dist, K, k, (uint32_t*)out);
  } else {
    // Sink step from the final messages (vertex_program_execute.hpp:131-148);
    // with one iteration (execute_single) from the initial ones.
    if (g->nsink) {
      PprBatchArgs a = T.pa;
      a.skip_resid = 1;  // sink ranks are stored; their residual is not reported
      const uint32_t xi = T.iters == 1 ? C.cur ^ 1 : C.cur;
      a.x = T.xs[xi];
      a.xnext = T.xs[xi ^ 1];
      a.resid = (double*)scratch(c, fz_name(t, "res_sink").c_str(), sizeof(double) * wave_blocks * K);
      if (!a.resid) return fail(c, LA3_ENOMEM, "fused run: device allocation failed");
      const uint32_t rblocks = std::max<uint32_t>(1, std::min<uint32_t>((g->nrows + 4 * RPW - 1) / (4 * RPW), wave_blocks));
      const uint32_t cblocks = std::max<uint32_t>(1, std::min<uint32_t>((d.nplong + 4 * RPW - 1) / (4 * RPW), wave_blocks));
      if constexpr (SUB) launch(c, "mppr_rows", k_mppr_sub_rows<LPR>, rblocks, d, a, g->nreg, g->nrows, 0);
      else launch(c, "mppr_rows", k_mppr_rows<QW>, rblocks, d, a, g->nreg, g->nrows, 0);
      if (d.npseg > d.npseg_reg) {
        PprBatchArgs as = a;
        const uint32_t s0 = d.npseg_reg;
        as.seg_row = d.pseg_row + s0; as.seg_beg = d.pseg_beg + s0; as.part = a.part + (size_t)s0 * 2 * K;
        const uint32_t sblocks =
            std::max<uint32_t>(1, std::min<uint32_t>((d.npseg - s0 + 4 * RPW - 1) / (4 * RPW), wave_blocks));
        if constexpr (SUB) {
          launch(c, "mppr_segs", k_mppr_sub_segs<LPR>, sblocks, d, as, d.npseg - s0);
          launch(c, "mppr_combine", k_mppr_sub_combine<LPR>, cblocks, d, a, g->nreg, g->nrows, 0);
        } else {
          launch(c, "mppr_segs", k_mppr_segs<QW>, sblocks, d, as, d.npseg - s0);
          launch(c, "mppr_combine", k_mppr_combine<QW>, cblocks, d, a, g->nreg, g->nrows, 0);
        }
      }
    }
    launch(c, "mppr_output", k_batch_output<float>, blocks_for(S), d, (const float*)T.pa.hi, K, k, (float*)out);
    if (b.out_comp) {
      float* oc = (float*)scratch(c, fz_name(t, "outc").c_str(), sizeof(float) * (size_t)S * k);
      if (!oc) return fail(c, LA3_ENOMEM, "fused run: device allocation failed");
      launch(c, "mppr_output", k_batch_output<float>, blocks_for(S), d, (const float*)T.pa.lo, K, k, oc);
      HIPC(c, dcopy_async(b.out_comp + (size_t)S * ft.q0, oc, sizeof(float) * (size_t)S * k, hipMemcpyDeviceToHost,
                             c->stream));
    }
    if (b.residual)
      HIPC(c, dcopy_async(b.residual + ft.q0, T.resid, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream));
  }
  return deliver(c, host, out, esz * (size_t)S * k);
}

// Kernels of the narrow launch classes 3..14 (narrow_cls): SSSP, then PPR,
// LPR = 1, 2, 4, 8, 16, 32.
using FzKernel = void (*)(const FTask*, const FHdr*, const FCtl*);
const FzKernel fz_narrow_a[12] = {
    k_fused_an<FK_SSSP, 1>, k_fused_an<FK_SSSP, 2>, k_fused_an<FK_SSSP, 4>,  k_fused_an<FK_SSSP, 8>,
    k_fused_an<FK_SSSP, 16>, k_fused_an<FK_SSSP, 32>, k_fused_an<FK_PPR, 1>, k_fused_an<FK_PPR, 2>,
    k_fused_an<FK_PPR, 4>,  k_fused_an<FK_PPR, 8>,  k_fused_an<FK_PPR, 16>, k_fused_an<FK_PPR, 32>};
const FzKernel fz_narrow_b[12] = {
    k_fused_bn<FK_SSSP, 1>, k_fused_bn<FK_SSSP, 2>, k_fused_bn<FK_SSSP, 4>,  k_fused_bn<FK_SSSP, 8>,
    k_fused_bn<FK_SSSP, 16>, k_fused_bn<FK_SSSP, 32>, k_fused_bn<FK_PPR, 1>, k_fused_bn<FK_PPR, 2>,
    k_fused_bn<FK_PPR, 4>,  k_fused_bn<FK_PPR, 8>,  k_fused_bn<FK_PPR, 16>, k_fused_bn<FK_PPR, 32>};
constexpr uint32_t FZ_PULL_BIT = NCLS;  // bits NCLS + 0..2: bottom-up BFS launches of the wide classes

// One step: the A and B launches of each launch class in `mask` (bit c: class
// c, wide classes 0-2 = 64, 128, 256 queries, narrow classes 3-14), the
// bottom-up BFS launch of each wide class with BFS tasks (bits
// FZ_PULL_BIT + 0..2), then the control launch. With profiling each launch is
// bracketed by events (last_a: the event slot of the last A).
void fz_step(la3_ctx* c, uint32_t mask, uint32_t ga_blocks, bool events, int* last_a, int* last_p,
             uint32_t cmode = FC_ALL) {
  const uint32_t ga = ga_blocks + FMAX_TASKS, gb = fz_gb(c) + FMAX_TASKS;
  const FTask* tk = (const FTask*)c->fz_tasks;
  FHdr* hdr = (FHdr*)c->fz_ctl;
  FCtl* ctl = (FCtl*)(c->fz_ctl + sizeof(FHdr));
  auto go = [&](const char* name, auto kern, uint32_t grid) {
    if (events) {
      const int ev = launch(c, name, kern, grid, tk, (const FHdr*)hdr, (const FCtl*)ctl);
      if (name[6] == 'a') *last_a = ev;
      if (name[6] == 'p') *last_p = ev;
    } else {
      hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, c->stream, tk, (const FHdr*)hdr, (const FCtl*)ctl);
    }
  };
  if (mask & 1) go("fused_a", k_fused_a<1>, ga);
  if (mask & 2) go("fused_a", k_fused_a<2>, ga);
  if (mask & 4) go("fused_a", k_fused_a<4>, ga);
  for (uint32_t n = 0; n < 12; n++)
    if (mask >> (3 + n) & 1) go("fused_a", fz_narrow_a[n], ga);
  if (mask >> FZ_PULL_BIT & 1) go("fused_pull", k_fused_pull<1>, ga);
  if (mask >> (FZ_PULL_BIT + 1) & 1) go("fused_pull", k_fused_pull<2>, ga);
  if (mask >> (FZ_PULL_BIT + 2) & 1) go("fused_pull", k_fused_pull<4>, ga);
  if (mask & 1) go("fused_b", k_fused_b<1>, gb);
  if (mask & 2) go("fused_b", k_fused_b<2>, gb);
  if (mask & 4) go("fused_b", k_fused_b<4>, gb);
  for (uint32_t n = 0; n < 12; n++)
    if (mask >> (3 + n) & 1) go("fused_b", fz_narrow_b[n], gb);
  if (events) launch(c, "fused_c", k_fused_c, 1, tk, hdr, ctl, cmode);
  else hipLaunchKernelGGL(k_fused_c, dim3(1), dim3(BLOCK), 0, c->stream, tk, hdr, ctl, cmode);
}

// steps full steps, each followed by `sub` sub-steps in which only the tasks
// of the header's sub_mask run and advance; the control pass before each
// launch sequence splits the blocks among the tasks that sequence runs.
// Captured once per key.
la3_status fz_capture(la3_ctx* c, uint32_t mask, uint32_t steps, uint32_t ga_blocks, uint32_t sub,
                      hipGraphExec_t* out) {
  hipGraphExec_t& ex =
      c->fz_exec[((uint64_t)steps << 52) | ((uint64_t)sub << 48) | ((uint64_t)ga_blocks << 24) | mask];
  *out = ex;
  if (ex) return LA3_OK;
  hipGraph_t graph = nullptr;
  HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  for (uint32_t i = 0; i < steps; i++) {
    fz_step(c, mask, ga_blocks, false, nullptr, nullptr, sub ? FC_SPLIT_SUB : FC_ALL);
    for (uint32_t j = 0; j < sub; j++)
      fz_step(c, mask, ga_blocks, false, nullptr, nullptr, FC_ADV_SUB | (j + 1 < sub ? FC_SPLIT_SUB : 0u));
  }
  hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (e != hipSuccess) return fail(c, LA3_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    ex = nullptr;
    return fail(c, LA3_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  *out = ex;
  return LA3_OK;
}

la3_status fz_check(la3_ctx* c, const la3_batch& b, double alpha) {
  la3_graph* g = b.graph;
  if (!g || !b.roots) return fail(c, LA3_EINVAL, "batch without graph or roots");
  if (g->host) return fail(c, LA3_EDEVICE, "graph lives on a host-only context");
  if (g->ctx->device != c->device) return fail(c, LA3_EINVAL, "graph lives on another device");
  if (b.k == 0 || b.k > 1024) return fail(c, LA3_EINVAL, "batch size must be in [1, 1024]");
  if (b.kind != LA3_Q_BFS && b.kind != LA3_Q_SSSP && b.kind != LA3_Q_PPR)
    return fail(c, LA3_EINVAL, "unknown query kind");
  if (b.kind == LA3_Q_SSSP && !g->weighted) return fail(c, LA3_EINVAL, "SSSP needs a weighted graph");
  if (b.kind != LA3_Q_BFS && !g->directed) return fail(c, LA3_EINVAL, "SSSP and PPR run on a directed load");
  if (b.kind == LA3_Q_PPR && b.iters == 0) return fail(c, LA3_EINVAL, "PPR needs a fixed iteration count");
  if (b.kind == LA3_Q_PPR && !(alpha > 0.0 && alpha < 1.0)) return fail(c, LA3_EINVAL, "alpha must be in (0, 1)");
  for (uint32_t q = 0; q < b.k; q++)
    if (b.roots[q] > g->n) return fail(c, LA3_EINVAL, "root id out of range");
  return LA3_OK;
}

struct FzBatch {
  la3_batch b;
  std::vector<uint32_t> roots;
  uint32_t tasks_left = 0, iters_run = 0;
  la3_status st = LA3_OK;
};

}  // namespace

// A fused run that batches join and leave between steps (include/la3.h).
struct la3_session {
  la3_ctx* c = nullptr;
  double alpha = 0.15;
  struct Slot {
    bool used = false;
    uint64_t batch = 0;
    FzTask ft{};
    uint64_t max_steps = 0;
  };
  Slot slot[FMAX_TASKS];
  FTask ht[FMAX_TASKS];     // host copies of the task table (staging for uploads)
  FCtl hc[FMAX_TASKS];      // control blocks read back after each replay
  FCtl fresh[FMAX_TASKS];   // staging for control blocks written by the host
  FHdr hdr{};               // staging for the header
  uint32_t ntasks = 0;      // slots [0, ntasks) are known to the device
  uint32_t steps = 1;       // steps per replay
  bool keep_slots = false;  // keep every slot's buffers (la3_batch_run: the caller releases)
  size_t keep_bytes = FZ_SLOT_KEEP_BYTES;  // a finished slot keeps its buffers up to this much
  bool pending_init = false;
  bool sub_steps = false;   // small tasks take sub-steps (sessions; not la3_batch_run)
  uint64_t sub_mask = 0;    // the device header's sub_mask (host copy, source of its uploads)
  std::map<uint64_t, FzBatch> batches;
  std::deque<uint64_t> done;
  uint64_t next_id = 1;
};

namespace {

uint32_t sess_free_slots(const la3_session* s) {
  uint32_t n = 0;
  for (const auto& sl : s->slot) n += !sl.used;
  return n;
}

la3_status sess_add(la3_session* s, const la3_batch& in, uint64_t* id) {
  la3_ctx* c = s->c;
  la3_status st;
  if ((st = fz_check(c, in, s->alpha))) return st;
  const uint32_t per = in.graph->task_k[in.kind];  // queries per task (LA3_TUNE_*_TASK_K)
  const uint32_t need = (in.k + per - 1) / per;
  if (need > sess_free_slots(s)) return fail(c, LA3_EAGAIN, "session: no free task slots");
  const uint64_t bid = s->next_id++;
  FzBatch& fb = s->batches[bid];
  fb.b = in;
  fb.roots.assign(in.roots, in.roots + in.k);
  fb.b.roots = fb.roots.data();
  fb.tasks_left = need;
  FHdr* dh = (FHdr*)c->fz_ctl;
  FCtl* dc = (FCtl*)(c->fz_ctl + sizeof(FHdr));
  uint32_t t = 0;
  std::vector<uint32_t> taken;
  // A failed set-up (device memory) gives the batch's slots back and frees
  // what they allocated.
  auto undo = [&](la3_status e) {
    for (uint32_t u : taken) {
      s->slot[u].used = false;
      s->fresh[u] = FCtl{};
      s->fresh[u].state = FS_FREE;
      s->fresh[u].done = 1;
      (void)dcopy_async(dc + u, &s->fresh[u], sizeof(FCtl), hipMemcpyHostToDevice, c->stream);
      scratch_release(c, fz_name(u, ""));
    }
    s->batches.erase(bid);
    return e;
  };
  for (uint32_t q0 = 0; q0 < in.k; q0 += per) {
    while (s->slot[t].used) t++;
    la3_session::Slot& sl = s->slot[t];
    sl.used = true;
    sl.batch = bid;
    taken.push_back(t);
    sl.ft = FzTask{0, q0, std::min(per, in.k - q0)};
    if ((st = fz_setup(c, fb.b, sl.ft, t, s->alpha, s->hdr.ga, s->ht[t]))) return undo(st);
    const FTask& T = s->ht[t];
    // Safety net against a run that does not terminate: far more steps than
    // any traversal here takes (BFS <= nslots + 1; SSSP rounds are bounded
    // by the near/far phases; PPR = iters).
    sl.max_steps = T.kind == FK_PPR ? (uint64_t)T.iters + 8 : 8ull * T.g.nslots + 4096;
    s->fresh[t] = FCtl{};
    s->fresh[t].state = FS_NEW;
    if (dcopy_async(c->fz_tasks + t, &s->ht[t], sizeof(FTask), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        dcopy_async(dc + t, &s->fresh[t], sizeof(FCtl), hipMemcpyHostToDevice, c->stream) != hipSuccess)
      return undo(fail(c, LA3_EDEVICE, "session: task upload failed"));
    s->ntasks = std::max(s->ntasks, t + 1);
  }
  s->hdr.ntasks = s->ntasks;
  HIPC(c, dcopy_async(&dh->ntasks, &s->hdr.ntasks, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  s->pending_init = true;
  if (id) *id = bid;
  return LA3_OK;
}

// One replay (s->steps steps) of the running tasks, then the post-loop work
// of every task that finished; a batch completes when its last task does.
la3_status sess_step(la3_session* s) {
  la3_ctx* c = s->c;
  FHdr* dh = (FHdr*)c->fz_ctl;
  FCtl* dc = (FCtl*)(c->fz_ctl + sizeof(FHdr));
  la3_status st;
  if (s->pending_init) {
    launch(c, "fused_c", k_fused_c, 1, (const FTask*)c->fz_tasks, dh, dc, (uint32_t)FC_INIT);
    s->pending_init = false;
  }
  uint32_t mask = 0;
  uint64_t big = 0;  // entries of the largest running task's graph
  for (uint32_t t = 0; t < s->ntasks; t++) {
    if (!s->slot[t].used) continue;
    const uint32_t cls = s->ht[t].cls;
    mask |= 1u << cls;
    if (s->ht[t].kind == FK_BFS) mask |= 1u << (FZ_PULL_BIT + cls);  // BFS tasks are wide (cls 0-2)
    big = std::max<uint64_t>(big, s->batches[s->slot[t].batch].b.graph->nnz);
  }
  // Sub-steps: a step lasts as long as its largest task's, so tasks on
  // graphs of at most 1/SUB_RATIO of the largest running graph's entries
  // also run `sub` sub-steps per step (about log2 of the size ratio, at
  // most SUB_MAX): they advance several iterations while the large tasks
  // advance one, instead of waiting on them every iteration.
  uint64_t sub_mask = 0, small = 0;
  for (uint32_t t = 0; t < s->ntasks; t++) {
    if (!s->slot[t].used) continue;
    const uint64_t e = s->batches[s->slot[t].batch].b.graph->nnz;
    if (e * SUB_RATIO <= big) {
      sub_mask |= 1ull << t;
      small = std::max(small, e);
    }
  }
  uint32_t sub = 0;
  if (sub_mask && s->sub_steps) {
    const double r = (double)big / (double)std::max<uint64_t>(small, 1);
    sub = std::min<uint32_t>(SUB_MAX, std::max<uint32_t>(1, (uint32_t)std::log2(r)));
  }
  if (sub && sub_mask != s->sub_mask) {
    s->sub_mask = sub_mask;
    HIPC(c, dcopy_async(&dh->sub_mask, &s->sub_mask, sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
  }
  int last_a = -1, last_pull = -1;
  if (mask) {
    if (!c->profiling) {
      hipGraphExec_t ex = nullptr;
      if ((st = fz_capture(c, mask, s->steps, s->hdr.ga, sub, &ex))) return st;
      HIPC(c, hipGraphLaunch(ex, c->stream));
    } else {
      // Profiling: the same steps as separate launches, each bracketed by
      // events (per-kernel device time; the graph has none).
      for (uint32_t i = 0; i < s->steps; i++) {
        fz_step(c, mask, s->hdr.ga, true, &last_a, &last_pull, sub ? FC_SPLIT_SUB : FC_ALL);
        for (uint32_t j = 0; j < sub; j++)
          fz_step(c, mask, s->hdr.ga, true, &last_a, &last_pull, FC_ADV_SUB | (j + 1 < sub ? FC_SPLIT_SUB : 0u));
      }
    }
  }
  HIPC(c, dcopy_async(s->hc, dc, sizeof(FCtl) * s->ntasks, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  double bytes = 0, pbytes = 0;
  uint64_t edges = 0;
  bool finished = false;
  std::vector<uint32_t> freed;
  for (uint32_t t = 0; t < s->ntasks; t++) {
    la3_session::Slot& sl = s->slot[t];
    if (!sl.used) continue;
    const FCtl& C = s->hc[t];
    FzBatch& fb = s->batches[sl.batch];
    if (C.state == FS_RUN && !C.done && C.iter > sl.max_steps)
      return fail(c, LA3_EDEVICE, "fused run: step limit reached");
    if (C.state != FS_RUN || !C.done) continue;
    bytes += C.bytes - C.pbytes;
    pbytes += C.pbytes;
    edges += C.edges;
    const la3_query_kind kind = fb.b.kind;
    const uint32_t it = kind == LA3_Q_BFS ? std::max<uint32_t>(C.iter, 1) : kind == LA3_Q_SSSP ? C.iter + 1 : C.iter;
    fb.iters_run = std::max(fb.iters_run, it);
    const uint32_t W = s->ht[t].width;
    if (W == FW_NARROW) {
      switch (s->ht[t].lpr) {
        case 1: st = fz_finish<1, 1>(c, fb.b, sl.ft, t, s->ht[t], C); break;
        case 2: st = fz_finish<1, 2>(c, fb.b, sl.ft, t, s->ht[t], C); break;
        case 4: st = fz_finish<1, 4>(c, fb.b, sl.ft, t, s->ht[t], C); break;
        case 8: st = fz_finish<1, 8>(c, fb.b, sl.ft, t, s->ht[t], C); break;
        case 16: st = fz_finish<1, 16>(c, fb.b, sl.ft, t, s->ht[t], C); break;
        default: st = fz_finish<1, 32>(c, fb.b, sl.ft, t, s->ht[t], C); break;
      }
    } else {
      st = W == 1 ? fz_finish<1>(c, fb.b, sl.ft, t, s->ht[t], C)
           : W == 2 ? fz_finish<2>(c, fb.b, sl.ft, t, s->ht[t], C)
                    : fz_finish<4>(c, fb.b, sl.ft, t, s->ht[t], C);
    }
    if (st) fb.st = st;
    // The slot is free again: the device skips it from now on.
    s->fresh[t] = FCtl{};
    s->fresh[t].state = FS_FREE;
    s->fresh[t].done = 1;
    HIPC(c, dcopy_async(dc + t, &s->fresh[t], sizeof(FCtl), hipMemcpyHostToDevice, c->stream));
    sl.used = false;
    if (--fb.tasks_left == 0) s->done.push_back(sl.batch);
    finished = true;
    freed.push_back(t);
  }
  // Bytes of the finished tasks: bottom-up BFS steps to k_fused_pull, the
  // rest to k_fused_a (the B and control launches move no counted bytes).
  annotate(c, last_a, bytes, edges);
  annotate(c, last_pull, pbytes, 0);
  c->edges_scanned += edges;
  if (finished) HIPC(c, hipStreamSynchronize(c->stream));
  // A slot keeps its buffers for the next task unless they are large: each
  // of the 64 slots growing to the largest task it ever ran would hold far
  // more memory than the running tasks need.
  for (uint32_t t : freed) {
    if (s->keep_slots) break;
    const std::string pre = fz_name(t, "");
    size_t held = 0;
    for (const auto& kv : c->scratch)
      if (kv.first.compare(0, pre.size(), pre) == 0) held += kv.second.bytes;
    if (held > s->keep_bytes) scratch_release(c, pre);
  }
  return LA3_OK;
}

la3_status sess_init(la3_ctx* c, double alpha, uint32_t blocks_per_cu, la3_session* s) {
  if (c->device < 0) return fail(c, LA3_EDEVICE, "host-only context");
  if (c->fz_busy) return fail(c, LA3_EAGAIN, "the context already runs a session");
  HIPC(c, hipSetDevice(c->device));
  if (!c->fz_tasks) {
    c->fz_tasks = (FTask*)scratch(c, "fused.tasks", sizeof(FTask) * FMAX_TASKS);
    c->fz_ctl = (uint8_t*)scratch(c, "fused.ctl", sizeof(FHdr) + sizeof(FCtl) * FMAX_TASKS);
    if (!c->fz_tasks || !c->fz_ctl) {
      c->fz_tasks = nullptr;
      return fail(c, LA3_ENOMEM, "fused run: device allocation failed");
    }
  }
  s->c = c;
  s->alpha = alpha;
  // LA3_SESSION_KEEP_BYTES overrides how much a finished slot keeps.
  if (const char* kb = getenv("LA3_SESSION_KEEP_BYTES")) s->keep_bytes = (size_t)strtoull(kb, nullptr, 10);
  s->hdr = FHdr{};
  s->hdr.ga = fz_ga(c, blocks_per_cu);
  s->hdr.gb = fz_gb(c);
  HIPC(c, dcopy_async(c->fz_ctl, &s->hdr, sizeof(FHdr), hipMemcpyHostToDevice, c->stream));
  c->fz_busy = true;
  return LA3_OK;
}

void sess_release(la3_session* s) {
  if (!s->c) return;
  (void)hipStreamSynchronize(s->c->stream);
  s->c->fz_busy = false;
  s->c = nullptr;
}

}  // namespace

extern "C" {

la3_status la3_batch_run(la3_ctx* c, la3_batch* batches, uint32_t nbatches, double ppr_alpha) {
  if (!c || !batches || nbatches == 0) return LA3_EINVAL;
  uint32_t need = 0, bpc = 0;
  for (uint32_t i = 0; i < nbatches; i++) {
    la3_status st = fz_check(c, batches[i], ppr_alpha);
    if (st) return st;
    const uint32_t per = batches[i].graph->task_k[batches[i].kind];
    need += (batches[i].k + per - 1) / per;
    bpc = std::max(bpc, batches[i].graph->fused_blocks);
  }
  if (need > FMAX_TASKS) return fail(c, LA3_EINVAL, "fused run: more than 64 tasks");
  auto s = std::make_unique<la3_session>();
  la3_status st = sess_init(c, ppr_alpha, bpc, s.get());
  if (st) return st;
  s->steps = FZ_STEPS;
  s->keep_slots = true;
  timing_begin(c);
  std::vector<uint64_t> ids(nbatches);
  for (uint32_t i = 0; i < nbatches && !st; i++) st = sess_add(s.get(), batches[i], &ids[i]);
  while (!st && s->done.size() < nbatches) st = sess_step(s.get());
  timing_end(c);
  for (uint32_t i = 0; i < nbatches && !st; i++) {
    const FzBatch& fb = s->batches[ids[i]];
    batches[i].iters_run = fb.iters_run;
    if (fb.st) st = fb.st;
  }
  sess_release(s.get());
  return st;
}

la3_status la3_session_create(la3_ctx* c, double ppr_alpha, la3_session** out) {
  if (!c || !out) return LA3_EINVAL;
  if (!(ppr_alpha > 0.0 && ppr_alpha < 1.0)) return fail(c, LA3_EINVAL, "alpha must be in (0, 1)");
  auto s = std::make_unique<la3_session>();
  la3_status st = sess_init(c, ppr_alpha, FZ_BLOCKS_PER_CU, s.get());
  if (st) return st;
  s->sub_steps = true;
  *out = s.release();
  return LA3_OK;
}

la3_status la3_session_add(la3_session* s, const la3_batch* batch, uint64_t* id) {
  if (!s || !s->c || !batch) return LA3_EINVAL;
  HIPC(s->c, hipSetDevice(s->c->device));
  return sess_add(s, *batch, id);
}

la3_status la3_session_step(la3_session* s, uint32_t* running) {
  if (!s || !s->c) return LA3_EINVAL;
  la3_ctx* c = s->c;
  HIPC(c, hipSetDevice(c->device));
  timing_begin(c);
  la3_status st = LA3_OK;
  if (s->batches.size() > s->done.size()) st = sess_step(s);
  timing_end(c);
  if (running) {
    uint32_t n = 0;
    for (const auto& kv : s->batches) n += kv.second.tasks_left > 0;
    *running = n;
  }
  return st;
}

la3_status la3_session_poll(la3_session* s, uint64_t* ids, uint32_t* iters_run, int32_t* status, uint32_t cap,
                            uint32_t* n) {
  if (!s || !n || (cap && !ids)) return LA3_EINVAL;
  uint32_t i = 0;
  while (i < cap && !s->done.empty()) {
    const uint64_t id = s->done.front();
    s->done.pop_front();
    auto it = s->batches.find(id);
    ids[i] = id;
    if (iters_run) iters_run[i] = it->second.iters_run;
    if (status) status[i] = it->second.st;
    s->batches.erase(it);
    i++;
  }
  *n = i;
  return LA3_OK;
}

uint32_t la3_session_free_slots(const la3_session* s) { return s ? sess_free_slots(s) : 0; }

void la3_session_destroy(la3_session* s) {
  if (!s) return;
  sess_release(s);
  delete s;
}

}  // extern "C"

// ============================================================== top-k ===
extern "C" {

}  // extern "C"

namespace {
template <class T, class Acc>
la3_status row_reduce(la3_ctx* c, uint32_t rows, uint64_t n, la3_reduce_op op, T ne, Acc* out) {
  const uint32_t nb = (uint32_t)((n + RD_CHUNK - 1) / RD_CHUNK);
  Acc* part = (Acc*)scratch(c, "reduce.part", sizeof(Acc) * (size_t)rows * nb);
  Acc* dout = (Acc*)scratch(c, "reduce.out", sizeof(Acc) * (size_t)rows);
  if (!part || !dout) return fail(c, LA3_ENOMEM, "reduce: device allocation failed");
  launch(c, "row_reduce", k_row_reduce<T, Acc>, rows * nb, (const T*)c->last_result, n, nb, (uint32_t)op, ne, part);
  launch(c, "row_reduce_fin", k_row_reduce_fin<Acc>, (rows + BLOCK - 1) / BLOCK, (const Acc*)part, rows, nb, dout);
  HIPC(c, dcopy_async(out, dout, sizeof(Acc) * (size_t)rows, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return LA3_OK;
}
}  // namespace

extern "C" {

la3_status la3_result_reduce(la3_ctx* c, la3_dtype dtype, uint32_t rows, uint64_t n, la3_reduce_op op, double value,
                             void* out) {
  if (!c || !out || rows == 0 || n == 0 || rows > (1u << 20)) return LA3_EINVAL;
  if (dtype < LA3_DT_U8 || dtype > LA3_DT_F64) return fail(c, LA3_EINVAL, "unknown value type");
  if (op != LA3_REDUCE_SUM && op != LA3_REDUCE_COUNT_NE) return fail(c, LA3_EINVAL, "unknown reduction");
  if (c->device < 0) return fail(c, LA3_EDEVICE, "host-only context");
  const uint32_t esz = dtype == LA3_DT_U8 ? 1 : dtype == LA3_DT_F64 ? 8 : 4;
  if (!c->last_result || (uint64_t)rows * n * esz > c->last_result_bytes)
    return fail(c, LA3_EINVAL, "no device-resident result of that size (call a program with a NULL output first)");
  if (!dev_covers(c->last_result, (size_t)rows * n * esz))
    return fail(c, LA3_EINVAL, "the last result is no longer in device memory");
  if ((n + RD_CHUNK - 1) / RD_CHUNK * (uint64_t)rows > 0x7fffffffull) return fail(c, LA3_EINVAL, "result too large");
  HIPC(c, hipSetDevice(c->device));
  timing_begin(c);
  la3_status st;
  const bool count = op == LA3_REDUCE_COUNT_NE;
  switch (dtype) {
    case LA3_DT_U8:
      st = row_reduce<uint8_t, uint64_t>(c, rows, n, op, (uint8_t)value, (uint64_t*)out);
      break;
    case LA3_DT_U32:
      st = row_reduce<uint32_t, uint64_t>(c, rows, n, op, (uint32_t)value, (uint64_t*)out);
      break;
    case LA3_DT_F32:
      st = count ? row_reduce<float, uint64_t>(c, rows, n, op, (float)value, (uint64_t*)out)
                 : row_reduce<float, double>(c, rows, n, op, (float)value, (double*)out);
      break;
    default:
      st = count ? row_reduce<double, uint64_t>(c, rows, n, op, value, (uint64_t*)out)
                 : row_reduce<double, double>(c, rows, n, op, value, (double*)out);
      break;
  }
  if (st) return st;
  timing_end(c);
  return LA3_OK;
}

la3_status la3_result_topk(la3_ctx* c, la3_dtype dtype, uint32_t rows, uint64_t n, uint32_t k, int largest,
                           uint32_t* ids, void* values) {
  if (!c || !ids || rows == 0 || n == 0 || k == 0 || n > 0xffffffffull) return LA3_EINVAL;
  if (dtype < LA3_DT_U8 || dtype > LA3_DT_F64) return fail(c, LA3_EINVAL, "unknown value type");
  if (c->device < 0) return fail(c, LA3_EDEVICE, "host-only context");
  const uint32_t esz = dtype == LA3_DT_U8 ? 1 : dtype == LA3_DT_F64 ? 8 : 4;
  if (!c->last_result || (uint64_t)rows * n * esz > c->last_result_bytes)
    return fail(c, LA3_EINVAL, "no device-resident result of that size (call a program with a NULL output first)");
  if (!dev_covers(c->last_result, (size_t)rows * n * esz))
    return fail(c, LA3_EINVAL, "the last result is no longer in device memory");
  HIPC(c, hipSetDevice(c->device));
  timing_begin(c);
  const void* v = c->last_result;
  const uint32_t bits = dtype == LA3_DT_U8 ? 8 : dtype == LA3_DT_F64 ? 64 : 32;
  const uint32_t kk = (uint32_t)std::min<uint64_t>(k, n);
  const uint32_t chunks = (uint32_t)((n + TK_CHUNK - 1) / TK_CHUNK);
  const uint32_t nb = rows * chunks;
  uint64_t* dprefix = (uint64_t*)scratch(c, "topk.prefix", 8ull * rows);
  uint32_t* dhist = (uint32_t*)scratch(c, "topk.hist", 4ull * 256 * rows);
  uint32_t* dceq = (uint32_t*)scratch(c, "topk.ceq", 4ull * nb);
  uint32_t* dsmall = (uint32_t*)scratch(c, "topk.small", 4ull * 4 * rows);  // need, gtc, gt_next, cnt
  uint32_t* dbase = (uint32_t*)scratch(c, "topk.base", 4ull * nb);
  uint32_t* dids = (uint32_t*)scratch(c, "topk.ids", 4ull * rows * k);
  void* dvals = scratch(c, "topk.vals", (size_t)esz * rows * k);
  if (!dprefix || !dhist || !dceq || !dsmall || !dbase || !dids || !dvals)
    return fail(c, LA3_ENOMEM, "top-k: device allocation failed");
  // 1. radix select of the kk-th key per row, 8 bits per pass from the top
  std::vector<uint64_t> prefix(rows, 0);
  std::vector<uint32_t> above(rows, 0), hist(256ull * rows);
  uint64_t hi_mask = 0;
  for (int shift = (int)bits - 8; shift >= 0; shift -= 8) {
    HIPC(c, dcopy_async(dprefix, prefix.data(), 8ull * rows, hipMemcpyHostToDevice, c->stream));
    launch(c, "fill_u32", k_fill<uint32_t>, blocks_for(256ull * rows), dhist, 256ull * rows, 0u);
    launch(c, "topk_hist", k_topk_hist, nb, v, (uint32_t)dtype, largest, n, chunks, (const uint64_t*)dprefix, hi_mask,
           (uint32_t)shift, dhist);
    HIPC(c, dcopy_async(hist.data(), dhist, 4ull * 256 * rows, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    for (uint32_t r = 0; r < rows; r++) {
      uint32_t acc = above[r];
      int d = 255;
      for (; d > 0; d--) {
        if (acc + hist[256ull * r + d] >= kk) break;
        acc += hist[256ull * r + d];
      }
      prefix[r] |= (uint64_t)d << shift;
      above[r] = acc;
    }
    hi_mask |= 0xffull << shift;
  }
  // 2. ordered compaction: every key above the threshold, equal keys by id
  std::vector<uint32_t> small(4ull * rows, 0), ceq(nb), base(nb);
  HIPC(c, dcopy_async(dprefix, prefix.data(), 8ull * rows, hipMemcpyHostToDevice, c->stream));
  launch(c, "topk_count", k_topk_count_eq, nb, v, (uint32_t)dtype, largest, n, chunks, (const uint64_t*)dprefix, dceq);
  HIPC(c, dcopy_async(ceq.data(), dceq, 4ull * nb, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  for (uint32_t r = 0; r < rows; r++) {
    uint32_t b = 0;
    for (uint32_t ch = 0; ch < chunks; ch++) { base[(size_t)r * chunks + ch] = b; b += ceq[(size_t)r * chunks + ch]; }
    small[r] = kk - above[r];           // need
    small[rows + r] = above[r];         // gtc
    small[3ull * rows + r] = kk;        // cnt
  }
  HIPC(c, dcopy_async(dsmall, small.data(), 16ull * rows, hipMemcpyHostToDevice, c->stream));
  HIPC(c, dcopy_async(dbase, base.data(), 4ull * nb, hipMemcpyHostToDevice, c->stream));
  launch(c, "topk_write", k_topk_write, nb, v, (uint32_t)dtype, largest, n, chunks, (const uint64_t*)dprefix,
         (const uint32_t*)dsmall, (const uint32_t*)(dsmall + rows), (const uint32_t*)dbase, dsmall + 2 * rows, k, dids);
  launch(c, "topk_gather", k_topk_gather, blocks_for((uint64_t)rows * k), v, esz, n, rows, k,
         (const uint32_t*)(dsmall + 3 * rows), (const uint32_t*)dids, dvals);
  std::vector<uint32_t> hid((size_t)rows * k);
  std::vector<uint8_t> hv((size_t)rows * k * esz);
  HIPC(c, dcopy_async(hid.data(), dids, 4ull * rows * k, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, dcopy_async(hv.data(), dvals, (size_t)esz * rows * k, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  timing_end(c);
  // 3. order each row's kk entries: value (largest or smallest first), then id
  for (uint32_t r = 0; r < rows; r++) {
    std::vector<uint32_t> idx(kk);
    for (uint32_t j = 0; j < kk; j++) idx[j] = j;
    auto val = [&](uint32_t j) -> double {
      const size_t x = (size_t)r * k + j;
      if (esz == 1) return hv[x];
      if (dtype == LA3_DT_U32) { uint32_t u; std::memcpy(&u, &hv[x * 4], 4); return u; }
      if (dtype == LA3_DT_F32) { float f; std::memcpy(&f, &hv[x * 4], 4); return f; }
      double d; std::memcpy(&d, &hv[x * 8], 8); return d;
    };
    std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
      const double va = val(a), vb = val(b);
      if (va != vb) return largest ? va > vb : va < vb;
      return hid[(size_t)r * k + a] < hid[(size_t)r * k + b];
    });
    for (uint32_t j = 0; j < k; j++) {
      const size_t o = (size_t)r * k + j;
      if (j >= kk) {
        ids[o] = 0xffffffffu;
        if (values) std::memset((uint8_t*)values + o * esz, 0, esz);
        continue;
      }
      const size_t x = (size_t)r * k + idx[j];
      ids[o] = hid[x];
      if (values) std::memcpy((uint8_t*)values + o * esz, &hv[x * esz], esz);
    }
  }
  return LA3_OK;
}

}  // extern "C"
