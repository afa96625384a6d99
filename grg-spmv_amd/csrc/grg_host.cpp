// Host-side preprocessing for the HIP runtime: a locality-preserving node
// order inside each level (see include/grg_host.h).

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/grg_host.h"

namespace {

thread_local std::string g_last_error;

void locality_order(int64_t N, int64_t S, int32_t L, const int64_t* off, const GrgHostBlock* blocks, int32_t nb,
                    int64_t* new_pos) {
  if (N < 0 || S < 0 || S > N || L < 0) throw std::invalid_argument("bad sizes");
  if (off[0] != 0 || off[L] != N) throw std::invalid_argument("level offsets do not cover the nodes");
  // Global children CSR (internal ids), children of a node ordered by source level then column.
  std::vector<int64_t> deg(N + 1, 0);
  for (int32_t b = 0; b < nb; ++b) {
    const GrgHostBlock& B = blocks[b];
    if (B.dst_level <= B.src_level || B.dst_level >= L) throw std::invalid_argument("block levels out of order");
    const int64_t rows = off[B.dst_level + 1] - off[B.dst_level];
    for (int64_t r = 0; r < rows; ++r) deg[off[B.dst_level] + r + 1] += B.indptr[r + 1] - B.indptr[r];
  }
  for (int64_t v = 0; v < N; ++v) deg[v + 1] += deg[v];
  std::vector<int32_t> child((size_t)deg[N]);
  std::vector<int64_t> fill(deg.begin(), deg.end() - 1);
  std::vector<int32_t> order(nb);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return blocks[a].src_level < blocks[b].src_level; });
  for (int32_t bi : order) {
    const GrgHostBlock& B = blocks[bi];
    const int64_t rows = off[B.dst_level + 1] - off[B.dst_level];
    const int64_t src0 = off[B.src_level], src1 = off[B.src_level + 1];
    for (int64_t r = 0; r < rows; ++r) {
      int64_t v = off[B.dst_level] + r;
      for (int64_t j = B.indptr[r]; j < B.indptr[r + 1]; ++j) {
        int64_t c = src0 + B.indices[j];
        if (c < src0 || c >= src1) throw std::invalid_argument("block column out of range");
        child[fill[v]++] = (int32_t)c;
      }
    }
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
  // Internal keys: mean of children keys, bottom up (children are in lower levels).
  const double last = (double)N * 4.0;
  for (int32_t l = 0; l < L; ++l) {
    for (int64_t v = off[l]; v < off[l + 1]; ++v) {
      if (v < S) continue;
      const int64_t n = deg[v + 1] - deg[v];
      if (n == 0) { key[v] = last + (double)v; continue; }
      double acc = 0.0;
      for (int64_t j = deg[v]; j < deg[v + 1]; ++j) acc += key[child[j]];
      key[v] = acc / (double)n;
    }
  }
  std::vector<int64_t> ids;
  for (int32_t l = 0; l < L; ++l) {
    int64_t lo = off[l], hi = off[l + 1];
    ids.resize(hi - lo);
    std::iota(ids.begin(), ids.end(), lo);
    if (l == 0) {
      // samples keep the prefix [0, S); other childless nodes follow
      std::stable_sort(ids.begin(), ids.begin() + (S - lo), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
      std::stable_sort(ids.begin() + (S - lo), ids.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
    } else {
      std::stable_sort(ids.begin(), ids.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
    }
    for (int64_t i = 0; i < hi - lo; ++i) new_pos[ids[i]] = lo + i;
  }
}

}  // namespace

extern "C" {

int grg_host_locality_order(int64_t num_nodes, int64_t num_samples, int32_t num_levels, const int64_t* level_offsets,
                            const GrgHostBlock* blocks, int32_t num_blocks, int64_t* new_pos) {
  if (!level_offsets || !new_pos || (num_blocks > 0 && !blocks)) {
    g_last_error = "null argument";
    return 1;
  }
  try {
    locality_order(num_nodes, num_samples, num_levels, level_offsets, blocks, num_blocks, new_pos);
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 2;
  }
}

const char* grg_host_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
