// Seeded synthetic GRG generator, compiled straight into the level-blocked
// layout of a .grg_spmv artifact.
//
// Model (one chromosome):
//  * Samples sit at fixed positions of a random leaf order. A local tree is the
//    Cartesian tree (max at the root) of the S-1 "gap times" between adjacent
//    leaves; gap times are Kingman coalescent times assigned to gaps in random
//    order, so tree shapes follow the coalescent's ranked shapes.
//  * Moving along the genome, each next local tree re-shuffles the gap times
//    inside one random window. Clades (leaf intervals) untouched by the window
//    persist, so adjacent trees share most of their nodes, as in a GRG.
//  * Mutations land on branches with probability proportional to branch
//    length (the neutral 1/f site-frequency spectrum); clades above S/2 are
//    redrawn, giving minor-allele frequencies on [1/S, 1/2].
//  * GRG nodes: samples, every mutated clade, and a hash-chosen fraction of
//    the remaining clades (hierarchy). A node's children are the nearest
//    retained clades / samples below it in the tree where it first appears,
//    which partition its sample set, so every sample is reached along exactly
//    one path: X[s, m] = 1 iff sample s lies below mutation m's node.
//  * Optional missingness: a pool of nodes, each over a random fraction of the
//    samples; every mutation points at one pool node.
//  * Optional diploid coalescence counts (ploidy 2, individual i = samples 2i,
//    2i+1): coal(v) = both(v) - sum over children of both(c), where both(v)
//    counts individuals with both haplotypes below v.
// The genome is cut into independent segments generated in parallel threads.
// The output is already in internal (stable height) order with per
// (parent level, child level) binary CSR blocks and level-local columns.

#include <algorithm>
#include <climits>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/grg_synth.h"

namespace {

thread_local std::string g_last_error;

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct Rng {  // xoshiro256**
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) { z = splitmix64(z); s[i] = z; }
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
  double exponential(double rate) { return -std::log1p(-uniform()) / rate; }
};

// A child reference inside one local tree: >= 0 is an internal node (gap
// index), < 0 encodes leaf position p as -(p + 1).
struct LocalTree {
  int64_t n_leaves = 0;
  std::vector<int32_t> left, right, lo, hi;  // per gap
  std::vector<int32_t> parent;               // per gap, -1 at root
  int32_t root = -1;

  void build(const std::vector<double>& gap_time) {
    const int32_t G = (int32_t)gap_time.size();
    const int32_t kUnset = INT32_MIN;
    left.assign(G, kUnset); right.assign(G, kUnset); lo.assign(G, 0); hi.assign(G, 0); parent.assign(G, -1);
    std::vector<int32_t> stack;
    stack.reserve(64);
    for (int32_t i = 0; i < G; ++i) {
      int32_t last = -1;
      while (!stack.empty() && gap_time[stack.back()] < gap_time[i]) { last = stack.back(); stack.pop_back(); }
      left[i] = (last >= 0) ? last : kUnset;
      if (last >= 0) parent[last] = i;
      if (!stack.empty()) { right[stack.back()] = i; parent[i] = stack.back(); }
      stack.push_back(i);
    }
    root = stack.empty() ? -1 : stack.front();
    // Children the stack pass left unset are the adjacent leaves.
    for (int32_t i = 0; i < G; ++i) {
      if (left[i] == kUnset) left[i] = -(i + 1);    // leaf at position i
      if (right[i] == kUnset) right[i] = -(i + 2);  // leaf at position i + 1
    }
    // Leaf intervals, children before parents (post-order via explicit stack).
    if (root < 0) return;
    std::vector<int32_t> order;
    order.reserve(G);
    std::vector<int32_t> st{root};
    while (!st.empty()) {
      int32_t v = st.back(); st.pop_back();
      order.push_back(v);
      if (left[v] >= 0) st.push_back(left[v]);
      if (right[v] >= 0) st.push_back(right[v]);
    }
    for (auto it = order.rbegin(); it != order.rend(); ++it) {
      int32_t v = *it;
      lo[v] = left[v] >= 0 ? lo[left[v]] : -left[v] - 1;
      hi[v] = right[v] >= 0 ? hi[right[v]] : -right[v] - 1;
    }
  }
};

struct SegmentOut {
  // Nodes created in this segment, local ids 0..n-1 (global = base + local).
  std::vector<int64_t> child_ptr{0};
  std::vector<int64_t> child_ref;  // >= 0: local node id; < 0: sample -(s + 1)
  std::vector<int32_t> node_lo, node_hi;  // leaf-position interval
  // Mutations of this segment in genome order.
  std::vector<int64_t> mut_ref;  // local node id, or sample -(s + 1)
  std::vector<double> mut_pos, mut_time;
};

inline uint64_t clade_key(int64_t lo, int64_t hi, int64_t S) { return (uint64_t)lo * (uint64_t)S + (uint64_t)hi; }

void run_segment(const GrgSynthParams& P, int64_t seg, int64_t n_seg, const std::vector<int64_t>& leaf_sample,
                 int64_t tree_begin, int64_t tree_end, int64_t mut_begin, int64_t mut_end, SegmentOut& out) {
  const int64_t S = P.num_samples;
  const int64_t G = S - 1;
  const uint64_t keep_threshold = (uint64_t)(std::min(std::max(P.keep_fraction, 0.0), 1.0) * 18446744073709551615.0);
  Rng rng(splitmix64(P.seed * 1000003ull + (uint64_t)seg));
  // Kingman merge times assigned to gaps in random order (segment's base tree).
  std::vector<double> gap_time(G);
  {
    double t = 0.0;
    for (int64_t k = S; k >= 2; --k) { t += rng.exponential(0.5 * (double)k * (double)(k - 1)); gap_time[S - k] = t; }
    for (int64_t i = G - 1; i > 0; --i) std::swap(gap_time[i], gap_time[rng.below(i + 1)]);
  }
  const int64_t n_trees = tree_end - tree_begin;
  const int64_t window = std::max<int64_t>(2, (int64_t)std::llround(P.window_fraction * (double)G));
  // Window starts and shuffle seeds, so pass 2 replays the same tree sequence.
  std::vector<int64_t> win_start(n_trees);
  std::vector<uint64_t> win_seed(n_trees);
  for (int64_t b = 0; b < n_trees; ++b) {
    win_start[b] = (G > window) ? (int64_t)rng.below(G - window + 1) : 0;
    win_seed[b] = rng.next();
  }
  auto apply_window = [&](std::vector<double>& gt, int64_t b) {
    if (b == 0) return;
    Rng wr(win_seed[b]);
    int64_t a = win_start[b], w = std::min<int64_t>(window, G - a);
    for (int64_t i = w - 1; i > 0; --i) std::swap(gt[a + i], gt[a + (int64_t)wr.below(i + 1)]);
  };
  // Mutations per tree.
  const int64_t n_mut = mut_end - mut_begin;
  std::vector<int64_t> mut_per_tree(n_trees, n_mut / std::max<int64_t>(n_trees, 1));
  for (int64_t b = 0; b < n_mut % std::max<int64_t>(n_trees, 1); ++b) mut_per_tree[b] += 1;

  // ---- pass 1: place mutations, record mutated clades ----
  std::unordered_set<uint64_t> mutated;
  mutated.reserve((size_t)(n_mut * 1.3) + 16);
  struct MutRec { uint64_t key; double pos, time; };
  std::vector<MutRec> muts;
  muts.reserve((size_t)n_mut);
  LocalTree tree;
  std::vector<double> gt = gap_time;
  std::vector<double> cum;
  std::vector<int32_t> bref;   // branch owner: >= 0 gap, < 0 leaf
  for (int64_t b = 0; b < n_trees; ++b) {
    apply_window(gt, b);
    tree.build(gt);
    // Branches: every non-root internal node and every leaf.
    cum.clear(); bref.clear();
    double total = 0.0;
    for (int32_t v = 0; v < (int32_t)G; ++v) {
      if (tree.parent[v] < 0) continue;
      total += gt[tree.parent[v]] - gt[v];
      cum.push_back(total); bref.push_back(v);
    }
    for (int32_t v = 0; v < (int32_t)G; ++v) {
      if (tree.left[v] < 0) { total += gt[v]; cum.push_back(total); bref.push_back(tree.left[v]); }
      if (tree.right[v] < 0) { total += gt[v]; cum.push_back(total); bref.push_back(tree.right[v]); }
    }
    std::vector<MutRec> here;
    here.reserve(mut_per_tree[b]);
    for (int64_t j = 0; j < mut_per_tree[b]; ++j) {
      for (int attempt = 0;; ++attempt) {
        double u = rng.uniform() * total;
        size_t at = (size_t)(std::upper_bound(cum.begin(), cum.end(), u) - cum.begin());
        if (at >= cum.size()) at = cum.size() - 1;
        int32_t br = bref[at];
        int64_t lo, hi;
        double t_lo, t_hi;
        if (br >= 0) { lo = tree.lo[br]; hi = tree.hi[br]; t_lo = gt[br]; t_hi = gt[tree.parent[br]]; }
        else { lo = hi = -br - 1; t_lo = 0.0; t_hi = 0.0; }
        if (br < 0) {
          // parent of a leaf at position p is gap p-1 or p (the larger-time-lower one that owns it)
          int64_t p = lo;
          double a = (p > 0) ? gt[p - 1] : 1e300, c = (p < G) ? gt[p] : 1e300;
          t_hi = std::min(a, c);
        }
        if ((hi - lo + 1) * 2 > S && attempt < 64) continue;  // minor allele only
        MutRec r;
        r.key = clade_key(lo, hi, S);
        r.pos = (double)(tree_begin + b) + rng.uniform();
        r.time = t_lo + (t_hi - t_lo) * rng.uniform();
        here.push_back(r);
        if (lo != hi) mutated.insert(r.key);
        break;
      }
    }
    std::sort(here.begin(), here.end(), [](const MutRec& a, const MutRec& b) { return a.pos < b.pos; });
    muts.insert(muts.end(), here.begin(), here.end());
  }
  auto kept = [&](uint64_t key) -> bool {
    if (mutated.count(key)) return true;
    return splitmix64(key ^ (P.seed * 0x9E3779B97F4A7C15ull)) < keep_threshold;
  };

  // ---- pass 2: emit retained clades with their frontier children ----
  std::unordered_map<uint64_t, int64_t> node_of;
  node_of.reserve(mutated.size() * 2 + 16);
  gt = gap_time;
  std::vector<int32_t> post, st, dfs;
  for (int64_t b = 0; b < n_trees; ++b) {
    apply_window(gt, b);
    tree.build(gt);
    post.clear();
    st.assign(1, tree.root);
    while (!st.empty()) {
      int32_t v = st.back(); st.pop_back();
      post.push_back(v);
      if (tree.left[v] >= 0) st.push_back(tree.left[v]);
      if (tree.right[v] >= 0) st.push_back(tree.right[v]);
    }
    for (auto it = post.rbegin(); it != post.rend(); ++it) {
      int32_t v = *it;
      uint64_t key = clade_key(tree.lo[v], tree.hi[v], S);
      if (!kept(key) || node_of.count(key)) continue;
      // Frontier of the left and right subtrees.
      dfs.clear();
      dfs.push_back(tree.right[v]);
      dfs.push_back(tree.left[v]);
      while (!dfs.empty()) {
        int32_t c = dfs.back(); dfs.pop_back();
        if (c < 0) { out.child_ref.push_back(-(int64_t)leaf_sample[-c - 1] - 1); continue; }
        uint64_t ck = clade_key(tree.lo[c], tree.hi[c], S);
        auto f = node_of.find(ck);
        if (f != node_of.end()) { out.child_ref.push_back(f->second); continue; }
        dfs.push_back(tree.right[c]);
        dfs.push_back(tree.left[c]);
      }
      int64_t id = (int64_t)out.node_lo.size();
      node_of.emplace(key, id);
      out.node_lo.push_back(tree.lo[v]);
      out.node_hi.push_back(tree.hi[v]);
      out.child_ptr.push_back((int64_t)out.child_ref.size());
    }
  }
  out.mut_ref.resize(muts.size());
  out.mut_pos.resize(muts.size());
  out.mut_time.resize(muts.size());
  for (size_t j = 0; j < muts.size(); ++j) {
    uint64_t key = muts[j].key;
    int64_t lo = (int64_t)(key / (uint64_t)S), hi = (int64_t)(key % (uint64_t)S);
    if (lo == hi) out.mut_ref[j] = -leaf_sample[lo] - 1;
    else {
      auto f = node_of.find(key);
      if (f == node_of.end()) throw std::runtime_error("internal error: mutated clade was not emitted");
      out.mut_ref[j] = f->second;
    }
    out.mut_pos[j] = muts[j].pos;
    out.mut_time[j] = muts[j].time;
  }
  (void)n_seg;
}

template <class T>
T* dup_array(const std::vector<T>& v) {
  T* p = (T*)std::malloc(std::max<size_t>(1, v.size()) * sizeof(T));
  if (!p) throw std::bad_alloc();
  if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

void generate(const GrgSynthParams& P, GrgSynthResult& R) {
  const int64_t S = P.num_samples, M = P.num_mutations;
  if (S < 2) throw std::invalid_argument("num_samples must be >= 2");
  if (M < 1) throw std::invalid_argument("num_mutations must be >= 1");
  if (P.num_trees < 1) throw std::invalid_argument("num_trees must be >= 1");
  if (P.ploidy != 1 && P.ploidy != 2) throw std::invalid_argument("ploidy must be 1 or 2");
  if (P.ploidy == 2 && (S % 2)) throw std::invalid_argument("ploidy 2 needs an even number of samples");
  Rng rng(splitmix64(P.seed ^ 0x5EEDull));
  std::vector<int64_t> leaf_sample(S);  // leaf position -> sample id
  std::iota(leaf_sample.begin(), leaf_sample.end(), 0);
  for (int64_t i = S - 1; i > 0; --i) std::swap(leaf_sample[i], leaf_sample[rng.below(i + 1)]);
  std::vector<int64_t> pos_of(S);
  for (int64_t p = 0; p < S; ++p) pos_of[leaf_sample[p]] = p;

  const int64_t n_seg = std::max<int64_t>(1, std::min<int64_t>(P.num_segments > 0 ? P.num_segments : 1, P.num_trees));
  std::vector<SegmentOut> segs(n_seg);
  std::vector<std::string> errors(n_seg);
  int n_threads = P.num_threads > 0 ? P.num_threads : (int)std::max(1u, std::thread::hardware_concurrency());
  std::atomic<int64_t> next_seg{0};
  auto worker = [&]() {
    for (;;) {
      int64_t s = next_seg.fetch_add(1);
      if (s >= n_seg) return;
      int64_t tb = P.num_trees * s / n_seg, te = P.num_trees * (s + 1) / n_seg;
      int64_t mb = M * s / n_seg, me = M * (s + 1) / n_seg;
      try { run_segment(P, s, n_seg, leaf_sample, tb, te, mb, me, segs[s]); }
      catch (const std::exception& e) { errors[s] = e.what(); }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < std::min<int64_t>(n_threads, n_seg); ++t) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  for (auto& e : errors) if (!e.empty()) throw std::runtime_error(e);

  // ---- missingness pool ----
  const int64_t pool_n = (P.missing_fraction > 0.0) ? std::max<int64_t>(1, P.missing_pool) : 0;
  const int64_t miss_k = std::max<int64_t>(1, (int64_t)std::llround(P.missing_fraction * (double)S));

  // ---- global original ids: samples, then segment nodes, then missing pool ----
  std::vector<int64_t> seg_base(n_seg + 1, S);
  for (int64_t s = 0; s < n_seg; ++s) seg_base[s + 1] = seg_base[s] + (int64_t)segs[s].node_lo.size();
  const int64_t pool_base = seg_base[n_seg];
  const int64_t N = pool_base + pool_n;
  std::vector<int64_t> cptr(N + 1, 0);
  for (int64_t s = 0; s < n_seg; ++s)
    for (size_t v = 0; v + 1 < segs[s].child_ptr.size(); ++v)
      cptr[seg_base[s] + (int64_t)v + 1] = segs[s].child_ptr[v + 1] - segs[s].child_ptr[v];
  for (int64_t p = 0; p < pool_n; ++p) cptr[pool_base + p + 1] = miss_k;
  for (int64_t v = 0; v < N; ++v) cptr[v + 1] += cptr[v];
  const int64_t E = cptr[N];
  std::vector<int32_t> cidx((size_t)E);
  for (int64_t s = 0; s < n_seg; ++s) {
    const auto& sg = segs[s];
    for (size_t v = 0; v + 1 < sg.child_ptr.size(); ++v) {
      int64_t dst = cptr[seg_base[s] + (int64_t)v];
      for (int64_t j = sg.child_ptr[v]; j < sg.child_ptr[v + 1]; ++j) {
        int64_t r = sg.child_ref[j];
        cidx[dst++] = (int32_t)(r < 0 ? -r - 1 : seg_base[s] + r);
      }
    }
  }
  {
    Rng mr(splitmix64(P.seed ^ 0xABCDEFull));
    std::vector<int64_t> perm(S);
    std::iota(perm.begin(), perm.end(), 0);
    for (int64_t p = 0; p < pool_n; ++p) {
      for (int64_t i = 0; i < miss_k; ++i) std::swap(perm[i], perm[i + mr.below(S - i)]);
      std::vector<int32_t> pick(perm.begin(), perm.begin() + miss_k);
      std::sort(pick.begin(), pick.end());
      std::copy(pick.begin(), pick.end(), cidx.begin() + cptr[pool_base + p]);
    }
  }
  if (N > (int64_t)INT32_MAX) throw std::runtime_error("too many nodes for int32 node ids");

  // ---- levels, stable height order ----
  std::vector<int32_t> level(N, 0);
  for (int64_t v = S; v < N; ++v) {
    int32_t m = -1;
    for (int64_t j = cptr[v]; j < cptr[v + 1]; ++j) m = std::max(m, level[cidx[j]]);
    level[v] = m + 1;
  }
  int32_t L = 0;
  for (int64_t v = 0; v < N; ++v) L = std::max(L, level[v] + 1);
  std::vector<int64_t> off(L + 1, 0);
  for (int64_t v = 0; v < N; ++v) off[level[v] + 1]++;
  for (int32_t l = 0; l < L; ++l) off[l + 1] += off[l];
  std::vector<int64_t> perm(N), inv(N);
  {
    std::vector<int64_t> cur(off.begin(), off.end() - 1);
    for (int64_t v = 0; v < N; ++v) { int64_t p = cur[level[v]]++; perm[p] = v; inv[v] = p; }
  }
  // ---- blocks: per (d, s) CSR, rows in internal order, columns sorted ----
  std::vector<int64_t> bnnz((size_t)L * L, 0);
  for (int64_t v = S; v < N; ++v)
    for (int64_t j = cptr[v]; j < cptr[v + 1]; ++j) bnnz[(size_t)level[v] * L + level[cidx[j]]]++;
  std::vector<int64_t> bstart((size_t)L * L + 1, 0);   // edge offset of block (d, s), row-major d*L+s
  std::vector<int64_t> pstart((size_t)L * L + 1, 0);   // indptr offset of block (d, s)
  for (int32_t d = 0; d < L; ++d)
    for (int32_t s = 0; s < L; ++s) {
      size_t b = (size_t)d * L + s;
      bstart[b + 1] = bstart[b] + bnnz[b];
      pstart[b + 1] = pstart[b] + ((s < d) ? (off[d + 1] - off[d] + 1) : 0);
    }
  std::vector<int32_t> bidx((size_t)E);
  std::vector<int32_t> bptr((size_t)pstart[(size_t)L * L]);
  {
    std::vector<int64_t> fill(bstart.begin(), bstart.end() - 1);
    std::vector<std::pair<int32_t, int32_t>> kids;
    for (int32_t d = 1; d < L; ++d) {
      for (int64_t r = 0; r < off[d + 1] - off[d]; ++r) {
        int64_t v = perm[off[d] + r];
        kids.clear();
        for (int64_t j = cptr[v]; j < cptr[v + 1]; ++j) {
          int32_t c = cidx[j];
          kids.emplace_back(level[c], (int32_t)(inv[c] - off[level[c]]));
        }
        std::sort(kids.begin(), kids.end());
        for (int32_t s = 0; s < d; ++s) {
          size_t b = (size_t)d * L + s;
          bptr[pstart[b] + r] = (int32_t)(fill[b] - bstart[b]);
        }
        for (auto& kc : kids) { size_t b = (size_t)d * L + kc.first; bidx[fill[b]++] = kc.second; }
      }
      for (int32_t s = 0; s < d; ++s) {
        size_t b = (size_t)d * L + s;
        bptr[pstart[b] + (off[d + 1] - off[d])] = (int32_t)(fill[b] - bstart[b]);
        if (bnnz[b] > (int64_t)INT32_MAX) throw std::runtime_error("block too large for int32 indptr");
      }
    }
  }
  // ---- mutations (genome order: segments in order) ----
  std::vector<int64_t> mut_node, miss_node;
  std::vector<double> mpos, mtime;
  mut_node.reserve(M); mpos.reserve(M); mtime.reserve(M);
  for (int64_t s = 0; s < n_seg; ++s) {
    const auto& sg = segs[s];
    for (size_t j = 0; j < sg.mut_ref.size(); ++j) {
      int64_t r = sg.mut_ref[j];
      int64_t orig = r < 0 ? -r - 1 : seg_base[s] + r;
      mut_node.push_back(inv[orig]);
      mpos.push_back(sg.mut_pos[j]);
      mtime.push_back(sg.mut_time[j]);
    }
  }
  if ((int64_t)mut_node.size() != M) throw std::runtime_error("internal error: mutation count mismatch");
  miss_node.assign(M, -1);
  if (pool_n > 0)
    for (int64_t m = 0; m < M; ++m)
      miss_node[m] = inv[pool_base + (int64_t)(splitmix64((uint64_t)m ^ P.seed) % (uint64_t)pool_n)];
  // ---- diploid coalescence counts ----
  std::vector<int64_t> coal;
  if (P.ploidy == 2) {
    // both(v) for retained clades by an offline 2D dominance count over
    // individuals' (min, max) haplotype positions; pool nodes use their
    // explicit sample lists.
    const int64_t I = S / 2;
    std::vector<int64_t> both(N, 0);
    std::vector<std::pair<int64_t, int64_t>> pts(I);
    for (int64_t i = 0; i < I; ++i) {
      int64_t a = pos_of[2 * i], b = pos_of[2 * i + 1];
      pts[i] = {std::max(a, b), std::min(a, b)};
    }
    std::sort(pts.begin(), pts.end());
    struct Q { int64_t hi, lo, node; };
    std::vector<Q> qs;
    qs.reserve(pool_base - S);
    for (int64_t s = 0; s < n_seg; ++s)
      for (size_t v = 0; v < segs[s].node_lo.size(); ++v)
        qs.push_back({segs[s].node_hi[v], segs[s].node_lo[v], seg_base[s] + (int64_t)v});
    std::sort(qs.begin(), qs.end(), [](const Q& a, const Q& b) { return a.hi < b.hi; });
    std::vector<int64_t> bit(S + 1, 0);
    auto bit_add = [&](int64_t i) { for (++i; i <= S; i += i & -i) bit[i]++; };
    auto bit_sum = [&](int64_t i) { int64_t r = 0; for (++i; i > 0; i -= i & -i) r += bit[i]; return r; };  // count <= i
    size_t pi = 0;
    int64_t inserted = 0;
    for (const Q& q : qs) {
      while (pi < pts.size() && pts[pi].first <= q.hi) { bit_add(pts[pi].second); ++inserted; ++pi; }
      both[q.node] = inserted - (q.lo > 0 ? bit_sum(q.lo - 1) : 0);
    }
    std::vector<uint8_t> seen(S, 0);
    for (int64_t p = 0; p < pool_n; ++p) {
      int64_t v = pool_base + p, cnt = 0;
      for (int64_t j = cptr[v]; j < cptr[v + 1]; ++j) seen[cidx[j]] = 1;
      for (int64_t j = cptr[v]; j < cptr[v + 1]; ++j) {
        int64_t smp = cidx[j];
        if ((smp & 1) == 0 && seen[smp + 1]) ++cnt;
      }
      for (int64_t j = cptr[v]; j < cptr[v + 1]; ++j) seen[cidx[j]] = 0;
      both[v] = cnt;
    }
    coal.assign(N, 0);
    for (int64_t v = S; v < N; ++v) {
      int64_t c = both[v];
      for (int64_t j = cptr[v]; j < cptr[v + 1]; ++j) c -= both[cidx[j]];
      if (c < 0) throw std::runtime_error("internal error: negative coalescence count");
      coal[inv[v]] = c;
    }
  }

  R.num_samples = S;
  R.num_mutations = M;
  R.num_nodes = N;
  R.num_edges = E;
  R.num_levels = L;
  R.ploidy = P.ploidy;
  R.level_offsets = dup_array(off);
  R.node_perm = dup_array(perm);
  R.block_edge_offset = dup_array(bstart);
  R.block_indptr_offset = dup_array(pstart);
  R.block_indices = dup_array(bidx);
  R.block_indptr = dup_array(bptr);
  R.mutation_node = dup_array(mut_node);
  R.miss_node = dup_array(miss_node);
  R.coalescence_counts = coal.empty() ? nullptr : dup_array(coal);
  R.mutation_positions = dup_array(mpos);
  R.mutation_times = dup_array(mtime);
}

}  // namespace

extern "C" {

int grg_synth_generate(const GrgSynthParams* params, GrgSynthResult* out) {
  if (!params || !out) { g_last_error = "null argument"; return 1; }
  std::memset(out, 0, sizeof(*out));
  try {
    generate(*params, *out);
    return 0;
  } catch (const std::exception& e) {
    grg_synth_free(out);
    g_last_error = e.what();
    return 2;
  }
}

void grg_synth_free(GrgSynthResult* r) {
  if (!r) return;
  void* ptrs[] = {r->level_offsets, r->node_perm, r->block_edge_offset, r->block_indptr_offset, r->block_indices,
                  r->block_indptr, r->mutation_node, r->miss_node, r->coalescence_counts, r->mutation_positions,
                  r->mutation_times};
  for (void* p : ptrs) std::free(p);
  std::memset(r, 0, sizeof(*r));
}

const char* grg_synth_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
