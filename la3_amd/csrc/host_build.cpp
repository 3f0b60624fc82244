// Host-side ingest: LA3 binary records -> vertex classes, slot order, DCSC and
// its row-major companion, ready for upload. Multi-threaded with std::thread;
// the output is identical for any thread count.
//
// Semantics follow the reference at one rank:
//   record filtering / orientation / hashing   src/matrix/graph.hpp:151-191
//   vertex classes                            src/matrix/processed_matrix2d.hpp:182-229
//   slot order (dashboard locator)            src/utils/locator.h:84-120
//   dedup, first record wins                  src/matrix/csc_matrix2d.hpp:60-72 with
//                                             unordered_set::insert (src/matrix/csc_matrix2d.hpp:62-72)
//   DCSC columns in col-group locator order,  src/utils/csc.h:94-125
//   entries by row position
#include <algorithm>
#include <atomic>
#include <cstring>
#include <numeric>
#include <thread>

#include "internal.h"

namespace la3 {

int host_threads() {
  static int t = [] {
    const char* e = getenv("LA3_THREADS");
    int v = e ? atoi(e) : 0;
    if (v <= 0) v = (int)std::thread::hardware_concurrency();
    if (v <= 0) v = 1;
    return std::min(v, 64);
  }();
  return t;
}

namespace {

// Static chunking: f(tid, begin, end).
template <class F>
void parallel_chunks(uint64_t n, F f) {
  int T = host_threads();
  if (n < 1u << 16 || T == 1) { f(0, 0, n); return; }
  std::vector<std::thread> th;
  for (int t = 0; t < T; t++) {
    uint64_t b = n * t / T, e = n * (t + 1) / T;
    th.emplace_back([=, &f] { f(t, b, e); });
  }
  for (auto& x : th) x.join();
}

// Dynamic scheduling over [0, n) in blocks of `grain`: f(i).
template <class F>
void parallel_dynamic(uint64_t n, uint64_t grain, F f) {
  int T = host_threads();
  if (n < 2 * grain || T == 1) { for (uint64_t i = 0; i < n; i++) f(i); return; }
  std::atomic<uint64_t> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < T; t++)
    th.emplace_back([&] {
      for (;;) {
        uint64_t b = next.fetch_add(grain);
        if (b >= n) break;
        uint64_t e = std::min(n, b + grain);
        for (uint64_t i = b; i < e; i++) f(i);
      }
    });
  for (auto& x : th) x.join();
}

inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Stable LSD radix sort of 64-bit keys on their low `bits` bits.
void radix_sort_u64(std::vector<uint64_t>& keys, int bits) {
  const int D = 11, B = 1 << D;
  std::vector<uint64_t> tmp(keys.size());
  const int T = host_threads();
  const uint64_t n = keys.size();
  for (int shift = 0; shift < bits; shift += D) {
    std::vector<uint64_t> hist((size_t)T * B, 0);
    auto chunk = [&](int t, uint64_t& b, uint64_t& e) { b = n * t / T; e = n * (t + 1) / T; };
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
      th.emplace_back([&, t] {
        uint64_t b, e; chunk(t, b, e);
        uint64_t* h = &hist[(size_t)t * B];
        for (uint64_t i = b; i < e; i++) h[(keys[i] >> shift) & (B - 1)]++;
      });
    for (auto& x : th) x.join();
    th.clear();
    uint64_t sum = 0;
    for (int d = 0; d < B; d++)
      for (int t = 0; t < T; t++) { uint64_t c = hist[(size_t)t * B + d]; hist[(size_t)t * B + d] = sum; sum += c; }
    for (int t = 0; t < T; t++)
      th.emplace_back([&, t] {
        uint64_t b, e; chunk(t, b, e);
        uint64_t* h = &hist[(size_t)t * B];
        for (uint64_t i = b; i < e; i++) tmp[h[(keys[i] >> shift) & (B - 1)]++] = keys[i];
      });
    for (auto& x : th) x.join();
    keys.swap(tmp);
  }
}

}  // namespace

std::vector<uint32_t> rmat_generate(uint32_t scale, uint32_t edgefactor, double a, double b,
                                    double c, uint64_t seed, bool weighted) {
  const uint64_t nsamples = (uint64_t)edgefactor << scale;
  std::vector<uint64_t> keys(nsamples);
  const double ab = a + b, abc = a + b + c;
  const double inv = 1.0 / 9007199254740992.0;  // 2^-53
  parallel_chunks(nsamples, [&](int, uint64_t beg, uint64_t end) {
    for (uint64_t i = beg; i < end; i++) {
      uint64_t s = seed * 0xD1B54A32D192ED03ull + i * 0x9E3779B97F4A7C15ull;
      uint64_t src = 0, dst = 0;
      for (uint32_t l = 0; l < scale; l++) {
        double u = (double)(splitmix64(s) >> 11) * inv;
        uint64_t sb = u >= ab, db = (u >= a && u < ab) || u >= abc;
        src = (src << 1) | sb;
        dst = (dst << 1) | db;
      }
      keys[i] = (src << scale) | dst;
    }
  });
  radix_sort_u64(keys, 2 * scale);
  // Unique, dropping self-loops.
  const uint64_t mask = (1ull << scale) - 1;
  uint64_t m = 0;
  for (uint64_t i = 0; i < nsamples; i++) {
    uint64_t k = keys[i];
    if ((k >> scale) == (k & mask)) continue;
    if (m > 0 && keys[m - 1] == k) continue;
    keys[m++] = k;
  }
  const uint32_t W = weighted ? 3 : 2;
  std::vector<uint32_t> out(m * W);
  parallel_chunks(m, [&](int, uint64_t beg, uint64_t end) {
    for (uint64_t i = beg; i < end; i++) {
      out[i * W] = (uint32_t)(keys[i] >> scale) + 1;
      out[i * W + 1] = (uint32_t)(keys[i] & mask) + 1;
      if (weighted) {
        uint64_t s = (seed ^ 0x5851F42D4C957F2Dull) + i * 0x9E3779B97F4A7C15ull;
        out[i * W + 2] = 1 + (uint32_t)(splitmix64(s) % 128);
      }
    }
  });
  return out;
}

std::string build_layout(const uint32_t* rec, uint64_t nrec, uint32_t n, bool weighted, int mode,
                         bool bucket, HostLayout& L) {
  const bool directed = mode != 2;
  const bool reverse = mode == 1 || mode == 4;
  const bool acyclic = mode == 3 || mode == 4;
  L = HostLayout();
  L.n = n;
  L.nslots = n + 1;
  L.directed = directed;
  L.weighted = weighted;
  L.mode = mode;
  L.nrecords = nrec;
  L.h = Hasher(n, bucket);
  const uint32_t RW = weighted ? 3 : 2;
  const uint32_t S = L.nslots;
  const Hasher H = L.h;

  // 1. Validate ids and count kept records per chunk.
  const int T = host_threads();
  std::vector<uint64_t> kept(T + 1, 0);
  std::atomic<bool> bad{false};
  parallel_chunks(nrec, [&](int t, uint64_t b, uint64_t e) {
    uint64_t k = 0;
    for (uint64_t i = b; i < e; i++) {
      uint32_t u = rec[i * RW], v = rec[i * RW + 1];
      if (u > n || v > n) { bad = true; break; }
      k += (u != v);
    }
    kept[t + 1] = k;
  });
  if (bad) return "vertex id out of range [0, nvertices]";
  for (int t = 0; t < T; t++) kept[t + 1] += kept[t];
  const uint64_t per = directed ? 1 : 2;
  const uint64_t m = kept[T] * per;
  if (m >= 0xffffffffull) return "more than 2^32-1 stored entries";

  // 2. Orient and hash (graph.hpp:165-191), preserving record order.
  std::vector<uint32_t> R(m), C(m), Wt(weighted ? m : 0);
  parallel_chunks(nrec, [&](int t, uint64_t b, uint64_t e) {
    uint64_t o = kept[t] * per;
    for (uint64_t i = b; i < e; i++) {
      uint32_t row = rec[i * RW], col = rec[i * RW + 1];
      if (row == col) continue;
      if (directed && !reverse) std::swap(row, col);
      if (acyclic && ((!reverse && col > row) || (reverse && col < row))) std::swap(row, col);
      row = (uint32_t)H.hash(row);
      col = (uint32_t)H.hash(col);
      uint32_t w = weighted ? rec[i * RW + 2] : 0;
      R[o] = row; C[o] = col; if (weighted) Wt[o] = w; o++;
      if (!directed) { R[o] = col; C[o] = row; if (weighted) Wt[o] = w; o++; }
    }
  });

  // 3. Row / column occupancy -> classes.
  std::vector<uint8_t> rowf(S, 0), colf(S, 0);
  parallel_chunks(m, [&](int, uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; i++) {
      __atomic_store_n(&rowf[R[i]], 1, __ATOMIC_RELAXED);
      __atomic_store_n(&colf[C[i]], 1, __ATOMIC_RELAXED);
    }
  });
  std::vector<uint32_t> slot(S);
  L.hid_of.resize(S);
  uint32_t pos = 0;
  for (uint32_t i = 0; i < S; i++) if (rowf[i] && colf[i]) { slot[i] = pos; L.hid_of[pos++] = i; }
  L.nreg = pos;
  for (uint32_t i = 0; i < S; i++) if (rowf[i] && !colf[i]) { slot[i] = pos; L.hid_of[pos++] = i; }
  L.nsink = pos - L.nreg;
  for (uint32_t i = 0; i < S; i++) if (!rowf[i] && colf[i]) { slot[i] = pos; L.hid_of[pos++] = i; }
  L.nsrc = pos - L.nreg - L.nsink;
  for (uint32_t i = 0; i < S; i++) if (!rowf[i] && !colf[i]) { slot[i] = pos; L.hid_of[pos++] = i; }
  L.niso = S - L.nreg - L.nsink - L.nsrc;
  L.ncols = L.nreg + L.nsrc;
  L.nrows = L.nreg + L.nsink;
  const uint32_t nreg = L.nreg, nsink = L.nsink;
  auto col_of_slot = [=](uint32_t s) { return s < nreg ? s : s - nsink; };
  L.id_of.resize(S);
  L.slot_of_id.resize(S);
  parallel_chunks(S, [&](int, uint64_t b, uint64_t e) {
    for (uint64_t p = b; p < e; p++) {
      uint32_t id = (uint32_t)H.unhash(L.hid_of[p]);
      L.id_of[p] = id;
      L.slot_of_id[id] = (uint32_t)p;
    }
  });

  // 4. Stable bucket of entries by column (keeps record order per column).
  const uint32_t NC = L.ncols;
  std::vector<uint32_t> cstart(NC + 1, 0);
  for (uint64_t i = 0; i < m; i++) cstart[col_of_slot(slot[C[i]]) + 1]++;
  for (uint32_t c = 0; c < NC; c++) cstart[c + 1] += cstart[c];
  std::vector<uint32_t> brow(m), bw(weighted ? m : 0);
  {
    std::vector<uint32_t> cur(cstart.begin(), cstart.end() - 1);
    for (uint64_t i = 0; i < m; i++) {
      uint32_t p = cur[col_of_slot(slot[C[i]])]++;
      brow[p] = slot[R[i]];
      if (weighted) bw[p] = Wt[i];
    }
  }
  { std::vector<uint32_t>().swap(R); std::vector<uint32_t>().swap(C); std::vector<uint32_t>().swap(Wt); }

  // 5. Per column: stable sort by row slot, keep the first of equal rows.
  std::vector<uint32_t> ucount(NC + 1, 0);
  parallel_dynamic(NC, 256, [&](uint64_t c) {
    uint32_t b = cstart[c], e = cstart[c + 1];
    uint32_t len = e - b;
    if (len > 1) {
      bool sorted = true;
      for (uint32_t j = b + 1; j < e; j++) if (brow[j] <= brow[j - 1]) { sorted = false; break; }
      if (!sorted) {
        std::vector<uint64_t> tmp(len);
        for (uint32_t j = 0; j < len; j++)
          tmp[j] = ((uint64_t)brow[b + j] << 32) | (weighted ? bw[b + j] : 0);
        std::stable_sort(tmp.begin(), tmp.end(),
                         [](uint64_t x, uint64_t y) { return (x >> 32) < (y >> 32); });
        uint32_t k = 0;
        for (uint32_t j = 0; j < len; j++) {
          if (k > 0 && (tmp[j] >> 32) == brow[b + k - 1]) continue;
          brow[b + k] = (uint32_t)(tmp[j] >> 32);
          if (weighted) bw[b + k] = (uint32_t)tmp[j];
          k++;
        }
        len = k;
      }
    }
    ucount[c + 1] = len;
  });
  L.col_ptr.assign(NC + 1, 0);
  for (uint32_t c = 0; c < NC; c++) L.col_ptr[c + 1] = L.col_ptr[c] + ucount[c + 1];
  const uint64_t nnz = L.col_ptr[NC];
  L.nnz = nnz;
  L.col_rows.resize(nnz);
  if (weighted) L.col_w.resize(nnz);
  L.col_split.resize(NC);
  parallel_dynamic(NC, 1024, [&](uint64_t c) {
    uint32_t b = cstart[c], o = L.col_ptr[c], len = L.col_ptr[c + 1] - o;
    std::memcpy(&L.col_rows[o], &brow[b], sizeof(uint32_t) * len);
    if (weighted) std::memcpy(&L.col_w[o], &bw[b], sizeof(uint32_t) * len);
    L.col_split[c] = (uint32_t)(std::lower_bound(L.col_rows.begin() + o, L.col_rows.begin() + o + len,
                                                 nreg) - L.col_rows.begin());
  });
  { std::vector<uint32_t>().swap(brow); std::vector<uint32_t>().swap(bw); }
  uint64_t nnz_reg = 0;
  for (uint32_t c = 0; c < NC; c++) nnz_reg += L.col_split[c] - L.col_ptr[c];
  L.nnz_reg = nnz_reg;
  L.nnz_sink = nnz - nnz_reg;

  // 6. Row-major companion in the reference's message order: all source
  //    columns (ascending), then all regular columns (ascending)
  //    (vertex_program_execute.hpp:477-482).
  const uint32_t NR = L.nrows;
  L.row_ptr.assign(NR + 1, 0);
  for (uint64_t j = 0; j < nnz; j++) L.row_ptr[L.col_rows[j] + 1]++;
  for (uint32_t r = 0; r < NR; r++) L.row_ptr[r + 1] += L.row_ptr[r];
  L.row_cols.resize(nnz);
  if (weighted) L.row_w.resize(nnz);
  L.row_split.resize(NR);
  std::vector<uint32_t> cur(L.row_ptr.begin(), L.row_ptr.end() - 1);
  auto fill_col = [&](uint32_t c) {
    uint32_t s = c < nreg ? c : c + nsink;
    for (uint32_t j = L.col_ptr[c]; j < L.col_ptr[c + 1]; j++) {
      uint32_t p = cur[L.col_rows[j]]++;
      L.row_cols[p] = s;
      if (weighted) L.row_w[p] = L.col_w[j];
    }
  };
  for (uint32_t c = nreg; c < NC; c++) fill_col(c);
  for (uint32_t r = 0; r < NR; r++) L.row_split[r] = cur[r];
  for (uint32_t c = 0; c < nreg; c++) fill_col(c);
  for (uint32_t r = 0; r < NR; r++)
    if (L.row_ptr[r + 1] - L.row_ptr[r] > LONG_ROW) {
      L.long_rows.push_back(r);
      if (r < nreg) L.nnz_long_reg += L.row_ptr[r + 1] - L.row_ptr[r];
    }
  for (uint32_t r = 0; r < NR; r++) {
    const uint32_t b = L.row_ptr[r], e = L.row_ptr[r + 1];
    if (e - b <= PSEG_ENTRIES) continue;
    if (r < nreg) { L.nplong_reg++; L.pnnz_reg += e - b; }
    L.plong.push_back(r);
    L.pseg_first.push_back((uint32_t)L.pseg_row.size());
    for (uint32_t j = b; j < e; j += PSEG_ENTRIES) {
      L.pseg_row.push_back(r);
      L.pseg_beg.push_back(j);
    }
    if (r < nreg) L.npseg_reg = (uint32_t)L.pseg_row.size();
  }
  L.pseg_first.push_back((uint32_t)L.pseg_row.size());
  return "";
}

}  // namespace la3
