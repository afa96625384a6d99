// Host-side preprocessing for the HIP runtime: a locality-preserving node
// order inside each level (see include/grg_host.h).

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/grg_host.h"

namespace {

thread_local std::string g_last_error;

// Runs f(lo, hi) over [0, n) cut into contiguous ranges on up to 16 host threads (one
// thread below min_chunk items per thread); an exception in any range is rethrown.
template <typename F>
void parallel_ranges(int64_t n, F f, int64_t min_chunk = 1 << 15) {
  const int64_t hw = std::max<int64_t>(1, (int64_t)std::thread::hardware_concurrency());
  const int64_t T = std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(hw, 16), n / std::max<int64_t>(1, min_chunk)));
  if (T <= 1) {
    if (n > 0) f((int64_t)0, n);
    return;
  }
  std::vector<std::thread> pool;
  std::vector<std::string> errs((size_t)T);
  for (int64_t t = 0; t < T; ++t) {
    pool.emplace_back([&, t]() {
      try {
        f(n * t / T, n * (t + 1) / T);
      } catch (const std::exception& e) {
        errs[(size_t)t] = e.what();
      }
    });
  }
  for (auto& th : pool) th.join();
  for (const auto& e : errs)
    if (!e.empty()) throw std::invalid_argument(e);
}

// ids[0, n) ordered by key (ties by id, which is the order the ids come in when they
// start ascending): contiguous ranges sorted on threads, then merged pairwise.
void sort_ids_by_key(int64_t* ids, int64_t n, const std::vector<double>& key) {
  if (n < 2) return;
  std::vector<std::pair<double, int64_t>> kv((size_t)n);
  for (int64_t i = 0; i < n; ++i) kv[(size_t)i] = {key[ids[i]], ids[i]};
  const int64_t hw = std::max<int64_t>(1, (int64_t)std::thread::hardware_concurrency());
  int64_t T = std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(hw, 16), n / (1 << 16)));
  std::vector<int64_t> cut((size_t)T + 1);
  for (int64_t t = 0; t <= T; ++t) cut[(size_t)t] = n * t / T;
  parallel_ranges(T, [&](int64_t t0, int64_t t1) {
    for (int64_t t = t0; t < t1; ++t) std::sort(kv.begin() + cut[(size_t)t], kv.begin() + cut[(size_t)t + 1]);
  }, 1);
  for (int64_t w = 1; w < T; w *= 2) {  // merge runs of w ranges pairwise
    const int64_t pairs = (T + 2 * w - 1) / (2 * w);
    parallel_ranges(pairs, [&](int64_t p0, int64_t p1) {
      for (int64_t p = p0; p < p1; ++p) {
        const int64_t a = 2 * w * p, m = std::min(T, a + w), e = std::min(T, a + 2 * w);
        if (m < e)
          std::inplace_merge(kv.begin() + cut[(size_t)a], kv.begin() + cut[(size_t)m], kv.begin() + cut[(size_t)e]);
      }
    }, 1);
  }
  for (int64_t i = 0; i < n; ++i) ids[i] = kv[(size_t)i].second;
}

void locality_order(int64_t N, int64_t S, int32_t L, const int64_t* off, const GrgHostBlock* blocks, int32_t nb,
                    int64_t degree_window, double degree_min_mean, int64_t* new_pos) {
  if (N < 0 || S < 0 || S > N || L < 0) throw std::invalid_argument("bad sizes");
  if (degree_window < 0) throw std::invalid_argument("degree_window < 0");
  if (off[0] != 0 || off[L] != N) throw std::invalid_argument("level offsets do not cover the nodes");
  // Global children CSR (internal ids), children of a node ordered by source level then column:
  // the blocks of each destination level in source-level order, rows filled on threads.
  std::vector<std::vector<int32_t>> by_dst((size_t)std::max<int32_t>(L, 0));
  for (int32_t b = 0; b < nb; ++b) {
    const GrgHostBlock& B = blocks[b];
    if (B.dst_level <= B.src_level || B.dst_level >= L) throw std::invalid_argument("block levels out of order");
    by_dst[(size_t)B.dst_level].push_back(b);
  }
  for (auto& v : by_dst)
    std::stable_sort(v.begin(), v.end(), [&](int32_t a, int32_t b) { return blocks[a].src_level < blocks[b].src_level; });
  std::vector<int64_t> deg(N + 1, 0);
  for (int32_t d = 0; d < L; ++d) {
    const int64_t lo = off[d];
    parallel_ranges(off[d + 1] - lo, [&](int64_t r0, int64_t r1) {
      for (int32_t bi : by_dst[(size_t)d]) {
        const GrgHostBlock& B = blocks[bi];
        for (int64_t r = r0; r < r1; ++r) deg[lo + r + 1] += B.indptr[r + 1] - B.indptr[r];
      }
    });
  }
  for (int64_t v = 0; v < N; ++v) deg[v + 1] += deg[v];
  std::vector<int32_t> child((size_t)deg[N]);
  for (int32_t d = 0; d < L; ++d) {
    const int64_t lo = off[d];
    parallel_ranges(off[d + 1] - lo, [&](int64_t r0, int64_t r1) {
      for (int64_t r = r0; r < r1; ++r) {
        int64_t pos = deg[lo + r];
        for (int32_t bi : by_dst[(size_t)d]) {
          const GrgHostBlock& B = blocks[bi];
          const int64_t src0 = off[B.src_level], src1 = off[B.src_level + 1];
          for (int64_t j = B.indptr[r]; j < B.indptr[r + 1]; ++j) {
            const int64_t c = src0 + B.indices[j];
            if (c < src0 || c >= src1) throw std::invalid_argument("block column out of range");
            child[(size_t)pos++] = (int32_t)c;
          }
        }
      }
    });
  }
  // Sample keys: first-visit order of a depth-first walk from the top levels down.
  std::vector<double> key(N, -1.0);
  std::vector<uint8_t> seen(N, 0);
  std::vector<int64_t> stack;
  int64_t next_leaf = 0;
  // A node is queued once (by the first parent that reaches it), so the stack stays below N.
  for (int64_t root = N - 1; root >= 0; --root) {
    if (seen[root]) continue;
    seen[root] = 1;
    stack.push_back(root);
    while (!stack.empty()) {
      int64_t v = stack.back();
      stack.pop_back();
      if (v < S) key[v] = (double)next_leaf++;
      for (int64_t j = deg[v + 1] - 1; j >= deg[v]; --j) {  // reversed push: first child visited first
        int64_t c = child[j];
        if (!seen[c]) {
          seen[c] = 1;
          stack.push_back(c);
        }
      }
    }
  }
  // Internal keys: mean of children keys, bottom up (children are in lower levels), each
  // level's nodes on threads.
  const double last = (double)N * 4.0;
  for (int32_t l = 0; l < L; ++l) {
    const int64_t lo = off[l];
    parallel_ranges(off[l + 1] - lo, [&](int64_t r0, int64_t r1) {
      for (int64_t v = lo + r0; v < lo + r1; ++v) {
        if (v < S) continue;
        const int64_t n = deg[v + 1] - deg[v];
        if (n == 0) {
          key[v] = last + (double)v;
          continue;
        }
        double acc = 0.0;
        for (int64_t j = deg[v]; j < deg[v + 1]; ++j) acc += key[child[j]];
        key[v] = acc / (double)n;
      }
    });
  }
  std::vector<int64_t> ids;
  for (int32_t l = 0; l < L; ++l) {
    int64_t lo = off[l], hi = off[l + 1];
    ids.resize(hi - lo);
    std::iota(ids.begin(), ids.end(), lo);
    if (l == 0) {
      // samples keep the prefix [0, S); other childless nodes follow
      sort_ids_by_key(ids.data(), S - lo, key);
      sort_ids_by_key(ids.data() + (S - lo), hi - S, key);
    } else {
      sort_ids_by_key(ids.data(), hi - lo, key);
      // Degree windows: inside each run of degree_window consecutive nodes, order by
      // child count (ties keep key order). A wave sums several consecutive rows in
      // lockstep rounds until its longest row is done, so rows of similar length
      // side by side waste fewer rounds; the window keeps the locality of the keys.
      const int64_t n = hi - lo;
      if (degree_window > 1 && n > 0 && (double)(deg[hi] - deg[lo]) >= degree_min_mean * (double)n) {
        const int64_t nw = (n + degree_window - 1) / degree_window;
        parallel_ranges(nw, [&](int64_t w0, int64_t w1) {
          for (int64_t w = w0 * degree_window; w < std::min(n, w1 * degree_window); w += degree_window) {
            const int64_t we = std::min(n, w + degree_window);
            std::stable_sort(ids.begin() + w, ids.begin() + we,
                             [&](int64_t a, int64_t b) { return deg[a + 1] - deg[a] < deg[b + 1] - deg[b]; });
          }
        }, 8);
      }
    }
    for (int64_t i = 0; i < hi - lo; ++i) new_pos[ids[i]] = lo + i;
  }
}

// Min-child order inside windows of a level (see grg_host_cluster_order).
void cluster_order(int64_t N, int32_t L, const int64_t* off, const GrgHostBlock* blocks, int32_t nb, int64_t window,
                   int64_t degree_window, double min_mean, int64_t min_rows, int64_t* new_pos) {
  if (N < 0 || L < 0 || off[0] != 0 || off[L] != N) throw std::invalid_argument("level offsets do not cover the nodes");
  if (window < 2) throw std::invalid_argument("window must be >= 2");
  if (degree_window < 0) throw std::invalid_argument("degree_window < 0");
  for (int64_t v = 0; v < N; ++v)
    if (new_pos[v] < 0 || new_pos[v] >= N) throw std::invalid_argument("new_pos out of range");
  for (int32_t b = 0; b < nb; ++b)
    if (blocks[b].dst_level <= blocks[b].src_level || blocks[b].dst_level >= L) throw std::invalid_argument("block levels out of order");
  std::vector<int32_t> order(nb);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return blocks[a].dst_level < blocks[b].dst_level; });
  std::vector<int64_t> mn, dg, ids;
  size_t bi = 0;
  for (int32_t l = 1; l < L; ++l) {
    const int64_t lo = off[l], hi = off[l + 1], n = hi - lo;
    // the level's blocks (children at lower levels, whose positions are final by now)
    mn.assign(n, INT64_MAX);
    dg.assign(n, 0);
    for (; bi < order.size() && blocks[order[bi]].dst_level == l; ++bi) {
      const GrgHostBlock& B = blocks[order[bi]];
      const int64_t src0 = off[B.src_level], src1 = off[B.src_level + 1];
      for (int64_t r = 0; r < n; ++r) {
        for (int64_t j = B.indptr[r]; j < B.indptr[r + 1]; ++j) {
          const int64_t c = src0 + B.indices[j];
          if (c < src0 || c >= src1) throw std::invalid_argument("block column out of range");
          mn[r] = std::min(mn[r], new_pos[c]);
        }
        dg[r] += B.indptr[r + 1] - B.indptr[r];
      }
    }
    int64_t e = 0;
    for (int64_t r = 0; r < n; ++r) e += dg[r];
    if (n < std::max<int64_t>(2, min_rows) || (double)e < min_mean * (double)n) continue;
    // rows (level-local) in their current order
    ids.assign(n, -1);
    for (int64_t r = 0; r < n; ++r) {
      const int64_t q = new_pos[lo + r] - lo;
      if (q < 0 || q >= n || ids[q] >= 0) throw std::invalid_argument("new_pos is not a permutation inside each level");
      ids[q] = r;
    }
    for (int64_t w = 0; w < n; w += window) {
      const int64_t we = std::min(n, w + window);
      // rows whose smallest child comes first lead; rows sharing their smallest child end up side by side
      std::stable_sort(ids.begin() + w, ids.begin() + we, [&](int64_t a, int64_t b) { return mn[a] < mn[b]; });
      if (degree_window > 1) {
        for (int64_t u = w; u < we; u += degree_window) {
          const int64_t ue = std::min(we, u + degree_window);
          std::stable_sort(ids.begin() + u, ids.begin() + ue, [&](int64_t a, int64_t b) { return dg[a] < dg[b]; });
        }
      }
    }
    for (int64_t i = 0; i < n; ++i) new_pos[lo + ids[i]] = lo + i;
  }
}

void build_tiles(int64_t R, int32_t L, const int64_t* off, int32_t nseg, const int32_t* const* sp,
                 const int32_t* const* si, const int64_t* src_rows, const int32_t* eptr, const int64_t* heavy,
                 int32_t max_staged, int32_t max_rows, int32_t max_entries, int32_t* tile_row, int32_t* tile_src,
                 uint32_t* src_list, uint16_t* lidx, int64_t* level_tile_off, int32_t* level_max) {
  if (R < 0 || L < 0 || nseg < 1 || nseg > 3) throw std::invalid_argument("bad sizes");
  if (max_staged < 1 || max_staged > 65535 || max_rows < 1 || max_entries < 1) throw std::invalid_argument("bad tile limits");
  if (off[0] != 0 || off[L] != R) throw std::invalid_argument("level offsets do not cover the rows");
  if (eptr[0] != 0) throw std::invalid_argument("eptr must start at 0");
  // stamp[s][src] = id of the last tile that staged src, pos[s][src] = its staged position there
  std::vector<std::vector<int32_t>> stamp(nseg);
  std::vector<std::vector<uint16_t>> pos(nseg);
  for (int32_t s = 0; s < nseg; ++s) {
    if (src_rows[s] < 0 || src_rows[s] >= (int64_t(1) << 30)) throw std::invalid_argument("source rows must be below 2^30");
    stamp[s].assign((size_t)src_rows[s], -1);
    pos[s].assign((size_t)src_rows[s], 0);
  }
  int64_t nt = 0;  // tiles so far (the open tile, if any, is tile nt)
  int64_t ns = 0;  // staged sources so far
  int64_t cur_rows = 0, cur_entries = 0;
  int32_t cur_staged = 0;
  bool open = false;
  int32_t* lvl_staged = level_max;
  int32_t* lvl_rows = level_max + L;
  int32_t* lvl_entries = level_max + 2 * L;
  int32_t lvl = 0;
  auto close = [&]() {
    if (!open) return;
    if (cur_staged > 0) {  // staged tile: record the level's LDS needs
      lvl_staged[lvl] = std::max<int32_t>(lvl_staged[lvl], cur_staged);
      lvl_rows[lvl] = std::max<int32_t>(lvl_rows[lvl], (int32_t)cur_rows);
      lvl_entries[lvl] = std::max<int32_t>(lvl_entries[lvl], (int32_t)cur_entries);
    }
    ++nt;
    open = false;
  };
  auto start = [&](int64_t r) {
    tile_row[nt] = (int32_t)r;
    tile_src[nt] = (int32_t)ns;
    cur_rows = cur_entries = 0;
    cur_staged = 0;
    open = true;
  };
  for (int32_t l = 0; l < L; ++l) {
    lvl = l;
    level_tile_off[l] = nt;
    lvl_staged[l] = lvl_rows[l] = lvl_entries[l] = 0;
    const int64_t lo = off[l], hi = off[l + 1];
    for (int64_t r = lo; r < hi; ++r) {
      const int64_t deg = (int64_t)eptr[r + 1] - eptr[r];
      int64_t check = 0;
      for (int32_t s = 0; s < nseg; ++s) check += (int64_t)sp[s][r + 1] - sp[s][r];
      if (check != deg) throw std::invalid_argument("eptr does not match the segment degrees");
      const bool is_heavy = heavy[l] >= 0 && deg > heavy[l];
      if (is_heavy || deg > max_staged || deg > max_entries) {
        // a direct tile of its own (a heavy row in it is produced by its heavy workgroup)
        close();
        start(r);
        cur_rows = 1;
        close();
        continue;
      }
      if (open) {
        int64_t fresh = 0;
        for (int32_t s = 0; s < nseg; ++s)
          for (int32_t j = sp[s][r]; j < sp[s][r + 1]; ++j) fresh += stamp[s][(size_t)si[s][j]] != (int32_t)nt;
        if (cur_staged + fresh > max_staged || cur_rows + 1 > max_rows || cur_entries + deg > max_entries) close();
      }
      if (!open) start(r);
      int64_t e = eptr[r];
      for (int32_t s = 0; s < nseg; ++s) {
        for (int32_t j = sp[s][r]; j < sp[s][r + 1]; ++j, ++e) {
          const int32_t c = si[s][j];
          if (c < 0 || c >= src_rows[s]) throw std::invalid_argument("segment source out of range");
          if (stamp[s][(size_t)c] != (int32_t)nt) {
            stamp[s][(size_t)c] = (int32_t)nt;
            pos[s][(size_t)c] = (uint16_t)cur_staged++;
            src_list[ns++] = ((uint32_t)s << 30) | (uint32_t)c;
          }
          lidx[e] = pos[s][(size_t)c];
        }
      }
      ++cur_rows;
      cur_entries += deg;
    }
    close();
  }
  level_tile_off[L] = nt;
  tile_row[nt] = (int32_t)R;
  tile_src[nt] = (int32_t)ns;
  if (ns > INT32_MAX) throw std::invalid_argument("more than 2^31 - 1 staged sources");
}


// Panels of the panel launch: runs of panel_rows consecutive rows whose
// distinct sources (the union of their entries' keys, ascending) are staged
// max_staged at a time ("passes"). Per pass, every row's entries that fall in
// the staged window are listed as 8-bit window positions, ascending.
constexpr int32_t kPanelPad = 255;  // entry position of the padding entries (a zero row in the kernel's LDS)

struct PanelSet {
  int32_t panel_rows = 0;
  int32_t max_staged = 0;
  int32_t max_pass_entries = 0;
  int32_t max_passes = 0;
  std::vector<int32_t> pan_pass{0};
  std::vector<int32_t> pass_src{0};
  std::vector<int32_t> pass_ent{0};
  std::vector<uint16_t> pass_rowoff;
  std::vector<uint32_t> src_list;
  std::vector<uint8_t> lidx;
};

struct PanelPart {  // panels [p0, p1) built by one thread, offsets local to the part
  std::vector<int32_t> n_pass;  // passes per panel
  std::vector<int32_t> pass_src, pass_ent;  // per pass: sources, entries
  std::vector<uint16_t> pass_rowoff;
  std::vector<uint32_t> src_list;
  std::vector<uint8_t> lidx;
  int32_t max_pass_entries = 0;
};

void build_panel_part(int64_t row_begin, int64_t row_end, const int32_t* eptr, const uint32_t* keys, int32_t R, int32_t P,
                      int32_t E, int64_t heavy, int64_t p0, int64_t p1, PanelPart& out) {
  std::vector<std::pair<uint32_t, int32_t>> ent;  // (key, row in panel)
  std::vector<uint32_t> uni;
  std::vector<int32_t> uses;   // entries per distinct source
  std::vector<int32_t> pass_of, first_of;  // pass of each distinct source; first source of each pass
  std::vector<int32_t> cnt;
  std::vector<int32_t> rowcnt, touched;  // entries per row of the pass being built; rows with entries
  for (int64_t p = p0; p < p1; ++p) {
    const int64_t r0 = row_begin + p * R, r1 = std::min<int64_t>(row_end, r0 + R);
    ent.clear();
    for (int64_t r = r0; r < r1; ++r) {
      const int64_t deg = (int64_t)eptr[r + 1] - eptr[r];
      if (deg < 0) throw std::invalid_argument("eptr is not non-decreasing");
      if (heavy >= 0 && deg > heavy) continue;  // produced by its heavy workgroup
      for (int64_t j = eptr[r]; j < eptr[r + 1]; ++j) ent.emplace_back(keys[j], (int32_t)(r - r0));
    }
    std::sort(ent.begin(), ent.end());
    uni.clear();
    uses.clear();
    for (size_t i = 0; i < ent.size(); ++i) {
      if (i > 0 && ent[i].first == ent[i - 1].first && ent[i].second == ent[i - 1].second)
        throw std::invalid_argument("a row lists the same source twice");
      if (uni.empty() || uni.back() != ent[i].first) {
        uni.push_back(ent[i].first);
        uses.push_back(0);
      }
      ++uses.back();
    }
    // passes: runs of consecutive distinct sources, closed before the (P + 1)-th source or
    // before the padded entries would exceed E. A row's entries in a pass are padded to
    // a multiple of 4 (with position kPanelPad), so adding a source costs 4 padded
    // entries for each of its rows whose count in the pass is a multiple of 4 (a
    // source's padded entries, at most 4 R <= E, always fit an empty pass).
    pass_of.assign(uni.size(), 0);
    first_of.clear();
    {
      rowcnt.assign((size_t)R, 0);
      touched.clear();
      int32_t q = -1, n_src = P;
      int64_t padded = 0;
      size_t i0 = 0;
      for (size_t u = 0; u < uni.size(); ++u) {
        const size_t i1 = i0 + (size_t)uses[u];  // entries of source u are ent[i0, i1)
        int64_t delta = 0;
        for (size_t i = i0; i < i1; ++i) delta += rowcnt[(size_t)ent[i].second] % 4 == 0 ? 4 : 0;
        if (n_src == P || padded + delta > E) {
          ++q;
          first_of.push_back((int32_t)u);
          n_src = 0;
          padded = 0;
          for (int32_t r : touched) rowcnt[(size_t)r] = 0;
          touched.clear();
          delta = 4 * (int64_t)uses[u];
        }
        pass_of[u] = q;
        ++n_src;
        padded += delta;
        for (size_t i = i0; i < i1; ++i) {
          const int32_t r = ent[i].second;
          if (rowcnt[(size_t)r]++ == 0) touched.push_back(r);
        }
        i0 = i1;
      }
    }
    const int32_t np = (int32_t)first_of.size();
    first_of.push_back((int32_t)uni.size());
    out.n_pass.push_back(np);
    out.src_list.insert(out.src_list.end(), uni.begin(), uni.end());
    // entries bucketed by (pass, row); within a bucket in ascending position, padded to a
    // multiple of 4 with kPanelPad
    cnt.assign((size_t)np * R, 0);
    size_t u = 0;
    std::vector<int32_t> pos(ent.size());
    for (size_t i = 0; i < ent.size(); ++i) {
      while (uni[u] != ent[i].first) ++u;
      pos[i] = (int32_t)u;
      ++cnt[(size_t)pass_of[u] * R + ent[i].second];
    }
    const size_t base = out.lidx.size();
    std::vector<int64_t> fill((size_t)np * R);
    int64_t acc = 0;
    for (int32_t q = 0; q < np; ++q) {
      out.pass_src.push_back(first_of[q + 1] - first_of[q]);
      const int64_t q0 = acc;
      for (int32_t r = 0; r < R; ++r) {
        out.pass_rowoff.push_back((uint16_t)(acc - q0));
        fill[(size_t)q * R + r] = acc;
        acc += (cnt[(size_t)q * R + r] + 3) / 4 * 4;
      }
      out.pass_rowoff.push_back((uint16_t)(acc - q0));
      if (acc - q0 > E) throw std::logic_error("a pass holds more entries than its limit");
      out.pass_ent.push_back((int32_t)(acc - q0));
      out.max_pass_entries = std::max<int32_t>(out.max_pass_entries, (int32_t)(acc - q0));
    }
    out.lidx.resize(base + (size_t)acc, (uint8_t)kPanelPad);
    for (size_t i = 0; i < ent.size(); ++i) {
      const int32_t q = pass_of[(size_t)pos[i]];
      out.lidx[base + (size_t)fill[(size_t)q * R + ent[i].second]++] = (uint8_t)(pos[i] - first_of[q]);
    }
  }
}

PanelSet* build_panels(int64_t row_begin, int64_t row_end, const int32_t* eptr, const uint32_t* keys, int32_t R, int32_t P,
                       int32_t E, int64_t heavy, int32_t threads) {
  if (row_end < row_begin || row_begin < 0) throw std::invalid_argument("bad row range");
  if (R < 1 || R > 1024 || P < 1 || P > kPanelPad)
    throw std::invalid_argument("panel_rows must be in [1, 1024] and max_staged in [1, 255]");
  if (E < 4 * R || E > 65535) throw std::invalid_argument("max_pass_entries must be in [4 * panel_rows, 65535]");
  const int64_t n_panels = (row_end - row_begin + R - 1) / R;
  const int64_t T = std::max<int64_t>(1, std::min<int64_t>(threads > 0 ? threads : 1, n_panels));
  std::vector<PanelPart> parts((size_t)T);
  std::vector<std::string> errs((size_t)T);
  std::vector<std::thread> pool;
  for (int64_t t = 0; t < T; ++t) {
    pool.emplace_back([&, t]() {
      try {
        build_panel_part(row_begin, row_end, eptr, keys, R, P, E, heavy, n_panels * t / T, n_panels * (t + 1) / T,
                         parts[(size_t)t]);
      } catch (const std::exception& e) {
        errs[(size_t)t] = e.what();
      }
    });
  }
  for (auto& th : pool) th.join();
  for (const auto& e : errs)
    if (!e.empty()) throw std::invalid_argument(e);
  auto* ps = new PanelSet();
  ps->panel_rows = R;
  ps->max_staged = P;
  for (auto& part : parts) {
    int64_t q = 0;
    for (int32_t np : part.n_pass) {
      ps->max_passes = std::max(ps->max_passes, np);
      ps->pan_pass.push_back(ps->pan_pass.back() + np);
      for (int32_t i = 0; i < np; ++i, ++q) {
        const int64_t s = (int64_t)ps->pass_src.back() + part.pass_src[(size_t)q];
        const int64_t e = (int64_t)ps->pass_ent.back() + part.pass_ent[(size_t)q];
        if (s > INT32_MAX || e > INT32_MAX) throw std::invalid_argument("more than 2^31 - 1 panel sources or entries");
        ps->pass_src.push_back((int32_t)s);
        ps->pass_ent.push_back((int32_t)e);
      }
    }
    ps->pass_rowoff.insert(ps->pass_rowoff.end(), part.pass_rowoff.begin(), part.pass_rowoff.end());
    ps->src_list.insert(ps->src_list.end(), part.src_list.begin(), part.src_list.end());
    ps->lidx.insert(ps->lidx.end(), part.lidx.begin(), part.lidx.end());
    ps->max_pass_entries = std::max(ps->max_pass_entries, part.max_pass_entries);
  }
  return ps;
}

}  // namespace

extern "C" {

int grg_host_locality_order(int64_t num_nodes, int64_t num_samples, int32_t num_levels, const int64_t* level_offsets,
                            const GrgHostBlock* blocks, int32_t num_blocks, int64_t degree_window, double degree_min_mean,
                            int64_t* new_pos) {
  if (!level_offsets || !new_pos || (num_blocks > 0 && !blocks)) {
    g_last_error = "null argument";
    return 1;
  }
  try {
    locality_order(num_nodes, num_samples, num_levels, level_offsets, blocks, num_blocks, degree_window, degree_min_mean,
                   new_pos);
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 2;
  }
}

int grg_host_cluster_order(int64_t num_nodes, int32_t num_levels, const int64_t* level_offsets, const GrgHostBlock* blocks,
                           int32_t num_blocks, int64_t window, int64_t degree_window, double min_mean, int64_t min_rows,
                           int64_t* new_pos) {
  if (!level_offsets || !new_pos || (num_blocks > 0 && !blocks)) {
    g_last_error = "null argument";
    return 1;
  }
  try {
    cluster_order(num_nodes, num_levels, level_offsets, blocks, num_blocks, window, degree_window, min_mean, min_rows, new_pos);
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 2;
  }
}

int grg_host_build_tiles(int64_t num_rows, int32_t num_levels, const int64_t* level_offsets, int32_t nseg,
                         const int32_t* const* seg_ptr, const int32_t* const* seg_idx, const int64_t* seg_src_rows,
                         const int32_t* eptr, const int64_t* heavy_threshold, int32_t max_staged, int32_t max_rows,
                         int32_t max_entries, int32_t* tile_row, int32_t* tile_src, uint32_t* src_list, uint16_t* lidx,
                         int64_t* level_tile_off, int32_t* level_max) {
  if (!level_offsets || !seg_ptr || !seg_idx || !seg_src_rows || !eptr || !heavy_threshold || !tile_row || !tile_src ||
      !src_list || !lidx || !level_tile_off || !level_max) {
    g_last_error = "null argument";
    return 1;
  }
  try {
    build_tiles(num_rows, num_levels, level_offsets, nseg, seg_ptr, seg_idx, seg_src_rows, eptr, heavy_threshold,
                max_staged, max_rows, max_entries, tile_row, tile_src, src_list, lidx, level_tile_off, level_max);
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 2;
  }
}

int grg_host_merged_keys(int64_t num_rows, int32_t nseg, const int32_t* const* seg_ptr, const int32_t* const* seg_idx,
                         const int32_t* eptr, uint32_t* ekey) {
  if (!seg_ptr || !seg_idx || !eptr || !ekey || num_rows < 0 || nseg < 1 || nseg > 3) {
    g_last_error = "bad arguments";
    return 1;
  }
  if (eptr[0] != 0) {
    g_last_error = "eptr must start at 0";
    return 2;
  }
  for (int64_t r = 0; r < num_rows; ++r) {
    int64_t e = eptr[r];
    for (int32_t s = 0; s < nseg; ++s) {
      for (int32_t j = seg_ptr[s][r]; j < seg_ptr[s][r + 1]; ++j, ++e) {
        const int32_t c = seg_idx[s][j];
        if (c < 0 || c >= (1 << 30)) {
          g_last_error = "source row outside [0, 2^30)";
          return 2;
        }
        ekey[e] = ((uint32_t)s << 30) | (uint32_t)c;
      }
    }
    if (e != eptr[r + 1]) {
      g_last_error = "eptr does not match the segment degrees";
      return 2;
    }
  }
  return 0;
}

const char* grg_host_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

extern "C" {

void* grg_host_build_panels(int64_t row_begin, int64_t row_end, const int32_t* eptr, const uint32_t* keys, int32_t panel_rows,
                            int32_t max_staged, int32_t max_pass_entries, int64_t heavy_threshold, int32_t threads) {
  if (!eptr || (!keys && row_end > row_begin)) {
    g_last_error = "null argument";
    return nullptr;
  }
  try {
    return build_panels(row_begin, row_end, eptr, keys, panel_rows, max_staged, max_pass_entries, heavy_threshold, threads);
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return nullptr;
  }
}

int grg_host_panels_sizes(const void* panels, int64_t* sizes) {
  if (!panels || !sizes) {
    g_last_error = "null argument";
    return 1;
  }
  const PanelSet* ps = static_cast<const PanelSet*>(panels);
  sizes[0] = (int64_t)ps->pan_pass.size() - 1;
  sizes[1] = (int64_t)ps->pass_src.size() - 1;
  sizes[2] = (int64_t)ps->src_list.size();
  sizes[3] = (int64_t)ps->lidx.size();
  sizes[4] = ps->max_pass_entries;
  sizes[5] = ps->max_passes;
  return 0;
}

int grg_host_panels_copy(const void* panels, int32_t* pan_pass, int32_t* pass_src, int32_t* pass_ent, uint16_t* pass_rowoff,
                         uint32_t* src_list, uint8_t* lidx) {
  if (!panels || !pan_pass || !pass_src || !pass_ent || !pass_rowoff || !src_list || !lidx) {
    g_last_error = "null argument";
    return 1;
  }
  const PanelSet* ps = static_cast<const PanelSet*>(panels);
  std::copy(ps->pan_pass.begin(), ps->pan_pass.end(), pan_pass);
  std::copy(ps->pass_src.begin(), ps->pass_src.end(), pass_src);
  std::copy(ps->pass_ent.begin(), ps->pass_ent.end(), pass_ent);
  std::copy(ps->pass_rowoff.begin(), ps->pass_rowoff.end(), pass_rowoff);
  std::copy(ps->src_list.begin(), ps->src_list.end(), src_list);
  std::copy(ps->lidx.begin(), ps->lidx.end(), lidx);
  return 0;
}

void grg_host_panels_free(void* panels) { delete static_cast<PanelSet*>(panels); }

}  // extern "C"

// ---- row groups of the MFMA row-group launch ---------------------------------

namespace {

constexpr int kGroupRows = 16;
constexpr uint32_t kAbsentKey = 0xffffffffu;

struct GroupPart {
  std::vector<int64_t> count;   // padded union entries per group
  std::vector<uint16_t> store;  // rows each group produces
  std::vector<uint32_t> key;
  std::vector<uint16_t> mask;
};

struct GroupSet {
  std::vector<int64_t> off{0};
  std::vector<uint16_t> store;
  std::vector<uint32_t> key;
  std::vector<uint16_t> mask;
};

// Groups [g0, g1) of rows [row_begin, row_end): each group's light rows' entries
// as (key << 16 | row bit), sorted, then runs of one key merged into one entry
// whose mask holds the bits of the rows that list it.
void build_group_part(int64_t row_begin, int64_t row_end, const int32_t* eptr, const uint32_t* keys, int64_t heavy,
                      int64_t g0, int64_t g1, GroupPart& out) {
  std::vector<uint64_t> items;
  for (int64_t g = g0; g < g1; ++g) {
    const int64_t r0 = row_begin + g * kGroupRows;
    const int64_t r1 = std::min<int64_t>(row_end, r0 + kGroupRows);
    items.clear();
    uint16_t store = 0;
    for (int64_t r = r0; r < r1; ++r) {
      const int64_t deg = (int64_t)eptr[r + 1] - eptr[r];
      if (deg < 0) throw std::invalid_argument("eptr is not non-decreasing");
      if (heavy >= 0 && deg > heavy) continue;  // produced by its heavy workgroup
      const uint64_t bit = (uint64_t)1 << (r - r0);
      store |= (uint16_t)bit;
      for (int32_t j = eptr[r]; j < eptr[r + 1]; ++j) {
        if (keys[j] == kAbsentKey) throw std::invalid_argument("entry key 0xffffffff is reserved for padding");
        items.push_back(((uint64_t)keys[j] << 16) | bit);
      }
    }
    std::sort(items.begin(), items.end());
    int64_t n = 0;
    for (size_t i = 0; i < items.size();) {
      const uint32_t k = (uint32_t)(items[i] >> 16);
      uint16_t m = 0;
      for (; i < items.size() && (uint32_t)(items[i] >> 16) == k; ++i) {
        const uint16_t bit = (uint16_t)(items[i] & 0xffffu);
        if (m & bit) throw std::invalid_argument("a row lists the same entry key twice");
        m |= bit;
      }
      out.key.push_back(k);
      out.mask.push_back(m);
      ++n;
    }
    for (; n % 4; ++n) {
      out.key.push_back(kAbsentKey);
      out.mask.push_back(0);
    }
    out.count.push_back(n);
    out.store.push_back(store);
  }
}

GroupSet* build_groups(int64_t row_begin, int64_t row_end, const int32_t* eptr, const uint32_t* keys, int64_t heavy,
                       int32_t threads) {
  if (row_end < row_begin || row_begin < 0) throw std::invalid_argument("bad row range");
  const int64_t n_groups = (row_end - row_begin + kGroupRows - 1) / kGroupRows;
  const int64_t T = std::max<int64_t>(1, std::min<int64_t>(threads > 0 ? threads : 1, n_groups));
  std::vector<GroupPart> parts((size_t)T);
  std::vector<std::string> errs((size_t)T);
  std::vector<std::thread> pool;
  for (int64_t t = 0; t < T; ++t) {
    pool.emplace_back([&, t]() {
      try {
        build_group_part(row_begin, row_end, eptr, keys, heavy, n_groups * t / T, n_groups * (t + 1) / T, parts[(size_t)t]);
      } catch (const std::exception& e) {
        errs[(size_t)t] = e.what();
      }
    });
  }
  for (auto& th : pool) th.join();
  for (const auto& e : errs)
    if (!e.empty()) throw std::invalid_argument(e);
  auto* gs = new GroupSet();
  for (auto& part : parts) {
    for (int64_t c : part.count) gs->off.push_back(gs->off.back() + c);
    gs->store.insert(gs->store.end(), part.store.begin(), part.store.end());
    gs->key.insert(gs->key.end(), part.key.begin(), part.key.end());
    gs->mask.insert(gs->mask.end(), part.mask.begin(), part.mask.end());
  }
  return gs;
}

}  // namespace

extern "C" {

void* grg_host_build_groups(int64_t row_begin, int64_t row_end, const int32_t* eptr, const uint32_t* keys,
                            int64_t heavy_threshold, int32_t threads) {
  if (!eptr || (!keys && row_end > row_begin)) {
    g_last_error = "null argument";
    return nullptr;
  }
  try {
    return build_groups(row_begin, row_end, eptr, keys, heavy_threshold, threads);
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return nullptr;
  }
}

int grg_host_groups_sizes(const void* groups, int64_t* sizes) {
  if (!groups || !sizes) {
    g_last_error = "null argument";
    return 1;
  }
  const GroupSet* gs = static_cast<const GroupSet*>(groups);
  sizes[0] = (int64_t)gs->store.size();
  sizes[1] = (int64_t)gs->key.size();
  return 0;
}

int grg_host_groups_copy(const void* groups, int64_t* off, uint16_t* store, uint32_t* key, uint16_t* mask) {
  if (!groups || !off || !store || !key || !mask) {
    g_last_error = "null argument";
    return 1;
  }
  const GroupSet* gs = static_cast<const GroupSet*>(groups);
  std::copy(gs->off.begin(), gs->off.end(), off);
  std::copy(gs->store.begin(), gs->store.end(), store);
  std::copy(gs->key.begin(), gs->key.end(), key);
  std::copy(gs->mask.begin(), gs->mask.end(), mask);
  return 0;
}

void grg_host_groups_free(void* groups) { delete static_cast<GroupSet*>(groups); }

}  // extern "C"
