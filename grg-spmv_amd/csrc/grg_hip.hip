// GRG matmul kernels for MI355X (gfx950, CDNA4, wave64) and their C ABI.
//
// One kernel, "row gather", runs every part of a sweep: each level of the up
// sweep (parents pull their children), each level of the down sweep (nodes
// pull their parents plus their selector inputs), and the write-out passes.
// A row's value is a plain sum of gathered k-wide rows (the GRG matrix is
// 0/1, so there are no multiplies), seeded by fused terms, and is written
// exactly once: no atomics anywhere.
//
// The sweep is bound by the latency of dependent gathers (row pointer ->
// column index -> node row), not by bytes, so the kernel is built to keep
// many independent gathers in flight with few registers:
//
// Light path. A wave of 64 lanes is split into lane groups of G = kl * es
// lanes, R = 1, 2 or 4 consecutive rows per group (the R rows' gathers share
// each round, so short rows overlap their latency); kl lanes cover a chunk of kl * cpl columns with
// cpl-wide vector loads, es lanes interleave over the row's entries (so the
// group's lanes read adjacent entries, which the locality order of the node
// ids keeps close in memory). Each lane works in rounds of U (4, 8 or 16)
// predicated gathers: all index loads of the round, then all row loads, then
// the adds, so a row of up to U * es entries costs one round trip; loads in
// flight stay in the storage type (U * cpl values per lane).
// The es partial sums are combined with cross-lane xor shuffles.
//
// Heavy path (same launch): rows of high degree get one 256-thread
// workgroup each (256 / kl edge lanes), reduced with shuffles plus one LDS
// pass.
//
// Split rows: a row too long for one workgroup (for example a missingness
// pool node referenced by every mutation) is cut into fixed-size chunks of
// its concatenated segment entries; a partials launch (flag
// GRG_HIP_FLAG_PARTIALS) gives each chunk a workgroup that writes one FP64
// partial row, and the level launch's heavy path then sums the row's
// partials instead of its entries. Both launches are ordinary stream-ordered
// kernels, so the pair stays deterministic and graph-capturable.
//
// Optional XCD remap: workgroups are dispatched round-robin over the 8 XCDs,
// so consecutive row blocks would land on different L2s; the remap gives each
// XCD one contiguous eighth of the rows, keeping neighbouring rows (which
// share children in the locality order) on one L2.
//
// FP64 values are summed in FP64. FP32 values are summed as two-float
// compensated pairs (TwoSum), with loads in flight held as FP32 so the FP32
// form keeps half the registers per gathered row; split-row partials are
// kept in FP64 for both.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/grg_hip.h"

namespace {

thread_local std::string g_last_error;

int fail(const std::string& msg, int code = 1) {
  g_last_error = msg;
  return code;
}

constexpr int kBlock = 256;  // threads per workgroup
constexpr int kWavesPerBlock = kBlock / 64;

template <typename T, int N>
struct VecT { typedef T type __attribute__((ext_vector_type(N))); };

// Load cpl consecutive values (one vector load), kept in their storage type
// so that loads in flight hold as few registers as possible.
template <typename T, int CPL>
__device__ __forceinline__ void load_raw(const T* p, T (&v)[CPL]) {
  if constexpr (CPL == 1) {
    v[0] = *p;
  } else {
    typename VecT<T, CPL>::type w = *reinterpret_cast<const typename VecT<T, CPL>::type*>(p);
#pragma unroll
    for (int c = 0; c < CPL; ++c) v[c] = w[c];
  }
}

template <typename T, int CPL>
__device__ __forceinline__ void store_raw(T* p, const T (&v)[CPL]) {
  if constexpr (CPL == 1) {
    *p = v[0];
  } else {
    typename VecT<T, CPL>::type w;
#pragma unroll
    for (int c = 0; c < CPL; ++c) w[c] = v[c];
    *reinterpret_cast<typename VecT<T, CPL>::type*>(p) = w;
  }
}

// A device pointer to global memory, marked as such: loads through it become
// global (not flat) instructions, which count against the vector-memory wait
// counter only. Flat loads also count against the LDS counter, so a kernel that
// mixes LDS reads with flat loads waits for every load in flight at each LDS
// wait. Only for pointers into device (global) memory, never LDS.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* global_ptr(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}

template <typename T, int CPL>
__device__ __forceinline__ void load_raw_global(const T* p, T (&v)[CPL]) {
  if constexpr (CPL == 1) {
    v[0] = *global_ptr(p);
  } else {
    typedef typename VecT<T, CPL>::type V;
    const V x = *reinterpret_cast<const __attribute__((address_space(1))) V*>(global_ptr(p));
#pragma unroll
    for (int c = 0; c < CPL; ++c) v[c] = x[c];
  }
}

// FP64 values: plain FP64 sum.
struct SumF64 {
  double s;
  __device__ __forceinline__ void zero() { s = 0.0; }
  __device__ __forceinline__ void add(double v) { s += v; }
  __device__ __forceinline__ void add_wide(double v) { s += v; }
  __device__ __forceinline__ void merge(const SumF64& o) { s += o.s; }
  __device__ __forceinline__ void merge_xor(int off) { s += __shfl_xor(s, off, 64); }
  __device__ __forceinline__ double wide() const { return s; }
  __device__ __forceinline__ double value() const { return s; }
};

// FP32 values: two-float compensated sum. hi carries the running FP32 sum,
// lo the exact rounding error of every addition (Knuth's TwoSum), so a row
// sum is accurate to about FP32 epsilon squared before its single rounding
// to FP32 at the store.
struct SumF32C {
  float hi, lo;
  __device__ __forceinline__ void zero() { hi = 0.f; lo = 0.f; }
  __device__ __forceinline__ void add2(float h2, float l2) {
    const float s = hi + h2;
    const float bp = s - hi;
    const float err = (hi - (s - bp)) + (h2 - bp);
    hi = s;
    lo += l2 + err;
  }
  __device__ __forceinline__ void add(float v) { add2(v, 0.f); }
  __device__ __forceinline__ void add_wide(double v) {  // an FP64 partial as an unevaluated float pair
    const float h = (float)v;
    add2(h, (float)(v - (double)h));
  }
  __device__ __forceinline__ void merge(const SumF32C& o) { add2(o.hi, o.lo); }
  __device__ __forceinline__ void merge_xor(int off) { add2(__shfl_xor(hi, off, 64), __shfl_xor(lo, off, 64)); }
  __device__ __forceinline__ double wide() const { return (double)hi + (double)lo; }
  __device__ __forceinline__ float value() const { return hi + lo; }
};

template <typename T> struct SumOf;
template <> struct SumOf<double> { using type = SumF64; };
template <> struct SumOf<float> { using type = SumF32C; };

// Sum src[idx[j]] for j = first, first + stride, ... < end into acc, in rounds
// of U predicated gathers (absent entries add an exact zero).
template <typename T, int CPL, int U, typename Acc>
__device__ __forceinline__ void gather_entries(const int32_t* __restrict__ idx, const T* __restrict__ src, int64_t ld,
                                               int32_t first, int32_t end, int32_t stride, Acc (&acc)[CPL]) {
  for (int32_t j = first; j < end; j += U * stride) {
    int32_t ii[U];
#pragma unroll
    for (int t = 0; t < U; ++t) {
      const int32_t jt = j + t * stride;
      ii[t] = jt < end ? global_ptr(idx)[jt] : -1;
    }
    T v[U][CPL];
#pragma unroll
    for (int t = 0; t < U; ++t) {
      if (ii[t] >= 0) {
        load_raw_global<T, CPL>(src + (int64_t)ii[t] * ld, v[t]);
      } else {
#pragma unroll
        for (int c = 0; c < CPL; ++c) v[t][c] = T(0);
      }
    }
#pragma unroll
    for (int t = 0; t < U; ++t)
#pragma unroll
      for (int c = 0; c < CPL; ++c) acc[c].add(v[t][c]);
  }
}

// gather_entries for long entry ranges (heavy rows, split-row chunks): the indices
// of round r + 1 are loaded while the gathers of round r are in flight, so a round
// costs one memory round trip instead of two. Slots past the end reload the range's
// first index and gather its row without adding it, so no load depends on a
// select of a loaded value. Entries are added in the same order as gather_entries.
template <typename T, int CPL, int U, typename Acc>
__device__ __forceinline__ void gather_entries_pipelined(const int32_t* __restrict__ idx, const T* __restrict__ src,
                                                         int64_t ld, int32_t first, int32_t end, int32_t stride,
                                                         Acc (&acc)[CPL]) {
  if (first >= end) return;
  const auto gi = global_ptr(idx);
  int32_t ii[U];
#pragma unroll
  for (int t = 0; t < U; ++t) {
    const int32_t jt = first + t * stride;
    ii[t] = gi[jt < end ? jt : first];
  }
  for (int32_t j = first; j < end; j += U * stride) {
    T v[U][CPL];
#pragma unroll
    for (int t = 0; t < U; ++t) load_raw_global<T, CPL>(src + (int64_t)ii[t] * ld, v[t]);
    const int32_t jn = j + U * stride;
#pragma unroll
    for (int t = 0; t < U; ++t) {
      const int32_t jt = jn + t * stride;
      ii[t] = gi[jt < end ? jt : first];
    }
#pragma unroll
    for (int t = 0; t < U; ++t) {
      if (j + t * stride < end) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[c].add(v[t][c]);
      }
    }
  }
}

// R rows at once: the same rounds of U predicated gathers, for R independent
// entry ranges [j[t], end[t]) side by side, so one round trip serves R short
// rows. Rounds continue while any of the R ranges has entries left.
template <typename T, int CPL, int U, int R, typename Acc>
__device__ __forceinline__ void gather_rows(const int32_t* __restrict__ idx, const T* __restrict__ src, int64_t ld,
                                            int32_t (&j)[R], const int32_t (&end)[R], int32_t stride,
                                            Acc (&acc)[R][CPL]) {
  while (true) {
    bool any = false;
#pragma unroll
    for (int t = 0; t < R; ++t) any |= j[t] < end[t];
    if (!any) break;
    int32_t ii[R][U];
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t jt = j[t] + u * stride;
        ii[t][u] = jt < end[t] ? global_ptr(idx)[jt] : -1;
      }
    T v[R][U][CPL];
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (ii[t][u] >= 0) {
          load_raw_global<T, CPL>(src + (int64_t)ii[t][u] * ld, v[t][u]);
        } else {
#pragma unroll
          for (int c = 0; c < CPL; ++c) v[t][u][c] = T(0);
        }
      }
#pragma unroll
    for (int t = 0; t < R; ++t) {
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[t][c].add(v[t][u][c]);
      j[t] += U * stride;
    }
  }
}

// gather_rows over merged entry keys: each entry names its segment in the top
// two bits (3 = absent), so one round serves all segments of the R rows.
template <typename T, int CPL, int U, int R, typename Acc>
__device__ __forceinline__ void gather_rows_keyed(const uint32_t* __restrict__ ekey, const T* src0, const T* src1,
                                                  const T* src2, int64_t ld0, int64_t ld1, int64_t ld2, int32_t (&j)[R],
                                                  const int32_t (&end)[R], int32_t stride, Acc (&acc)[R][CPL]) {
  while (true) {
    bool any = false;
#pragma unroll
    for (int t = 0; t < R; ++t) any |= j[t] < end[t];
    if (!any) break;
    uint32_t kk[R][U];
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t jt = j[t] + u * stride;
        kk[t][u] = jt < end[t] ? global_ptr(ekey)[jt] : 0xffffffffu;
      }
    T v[R][U][CPL];
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t sg = kk[t][u] >> 30;
        if (sg != 3u) {
          const int64_t r = (int64_t)(kk[t][u] & 0x3fffffffu);
          const T* base = sg == 0 ? src0 : (sg == 1 ? src1 : src2);
          const int64_t ld = sg == 0 ? ld0 : (sg == 1 ? ld1 : ld2);
          load_raw_global<T, CPL>(base + r * ld, v[t][u]);
        } else {
#pragma unroll
          for (int c = 0; c < CPL; ++c) v[t][u][c] = T(0);
        }
      }
#pragma unroll
    for (int t = 0; t < R; ++t) {
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[t][c].add(v[t][u][c]);
      j[t] += U * stride;
    }
  }
}

// One row per lane group (G lanes, R = 1): the group loads each round's U * es
// consecutive entry keys (merged entries) or source indices (one segment)
// together, one dword per lane per G entries, and hands lane (e, c) its entries
// with lane shuffles, in place of U loads per lane. Lane e still sums entries
// e, e + es, e + 2 es, ... in the same rounds, so the sums are bit for bit those
// of gather_rows_keyed / gather_rows. The loop bounds are the group's, so every
// shuffle reads an active lane of the same group; lanes past the last column
// chunk (col_ok false) load entries but gather nothing.
template <int U>
__device__ __forceinline__ void group_entries(const uint32_t* __restrict__ ent, int32_t b, int32_t end, int e, int es,
                                              int kl_shift, int lane, int g_shift, uint32_t (&kk)[U]) {
  const int G = 1 << g_shift;
  const int lg = lane & (G - 1);
  const int group0 = lane - lg;
  const int kl_mask = (1 << kl_shift) - 1;
  const int nslot = (U >> kl_shift) > 0 ? (U >> kl_shift) : 1;  // G-entry loads per round
  uint32_t ks[U];
#pragma unroll
  for (int q = 0; q < U; ++q) {
    const int32_t i = b + q * G + lg;
    ks[q] = (q < nslot && i < end) ? global_ptr(ent)[i] : 0xffffffffu;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    // entry b + e + u * es was loaded by group lane e + (u mod kl) * es into slot u / kl
    const int slot = u >> kl_shift;
    uint32_t v = ks[0];
#pragma unroll
    for (int q = 1; q < U; ++q)
      if (q == slot) v = ks[q];
    kk[u] = __shfl(v, group0 + e + (u & kl_mask) * es, 64);
  }
}

template <typename T, int CPL, int U, typename Acc>
__device__ __forceinline__ void gather_row_keyed_group(const uint32_t* __restrict__ ekey, const T* src0, const T* src1,
                                                       const T* src2, int64_t ld0, int64_t ld1, int64_t ld2, int32_t begin,
                                                       int32_t end, int e, int es, int kl_shift, int lane, int g_shift,
                                                       bool col_ok, Acc (&acc)[CPL]) {
  for (int32_t b = begin; b < end; b += U * es) {
    uint32_t kk[U];
    group_entries<U>(ekey, b, end, e, es, kl_shift, lane, g_shift, kk);
    T v[U][CPL];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t sg = kk[u] >> 30;
      if (col_ok && sg != 3u) {
        const int64_t r = (int64_t)(kk[u] & 0x3fffffffu);
        const T* base = sg == 0 ? src0 : (sg == 1 ? src1 : src2);
        const int64_t ld = sg == 0 ? ld0 : (sg == 1 ? ld1 : ld2);
        load_raw_global<T, CPL>(base + r * ld, v[u]);
      } else {
#pragma unroll
        for (int c = 0; c < CPL; ++c) v[u][c] = T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int c = 0; c < CPL; ++c) acc[c].add(v[u][c]);
  }
}

template <typename T, int CPL, int U, typename Acc>
__device__ __forceinline__ void gather_row_indexed_group(const int32_t* __restrict__ idx, const T* src, int64_t ld,
                                                         int32_t begin, int32_t end, int e, int es, int kl_shift, int lane,
                                                         int g_shift, bool col_ok, Acc (&acc)[CPL]) {
  for (int32_t b = begin; b < end; b += U * es) {
    uint32_t kk[U];
    group_entries<U>(reinterpret_cast<const uint32_t*>(idx), b, end, e, es, kl_shift, lane, g_shift, kk);
    T v[U][CPL];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (col_ok && kk[u] != 0xffffffffu) {
        load_raw_global<T, CPL>(src + (int64_t)kk[u] * ld, v[u]);
      } else {
#pragma unroll
        for (int c = 0; c < CPL; ++c) v[u][c] = T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int c = 0; c < CPL; ++c) acc[c].add(v[u][c]);
  }
}

// Row-wise seed / epilogue terms (added once per row by one lane).
template <typename T, int CPL, typename Acc>
__device__ __forceinline__ void add_row_terms(const GrgHipGatherArgs& a, int64_t row, int col, Acc (&acc)[CPL]) {
  for (int m = 0; m < a.nmap; ++m) {
    const GrgHipMap& mp = a.map[m];
    if (row >= mp.row_lo && row < mp.row_hi) {
      const int64_t r = mp.map ? (int64_t)mp.map[row - mp.row_lo] : (row - mp.row_lo);
      T v[CPL];
      load_raw_global<T, CPL>(static_cast<const T*>(mp.src) + r * mp.ld + col, v);
#pragma unroll
      for (int c = 0; c < CPL; ++c) acc[c].add(v[c]);
    }
  }
  if (a.row_scale || a.col_vec) {
    const T rs = a.row_scale ? static_cast<const T*>(a.row_scale)[row - a.scale_base] : T(1);
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const T cv = a.col_vec ? static_cast<const T*>(a.col_vec)[col + c] : T(1);
      acc[c].add_wide((double)rs * (double)cv);
    }
  }
  if (a.accumulate) {
    T v[CPL];
    load_raw_global<T, CPL>(static_cast<const T*>(a.out) + (row - a.out_base) * a.ld_out + col, v);
#pragma unroll
    for (int c = 0; c < CPL; ++c) acc[c].add(v[c]);
  }
}

// Workgroup reduction of the heavy / partials paths: lanes with the same
// column lane (threadIdx.x % KL) are summed; the result is valid in wave 0,
// lanes 0..KL-1.
template <int CPL, typename Acc>
__device__ __forceinline__ void block_reduce(Acc (&acc)[CPL], int KL, Acc (&partial)[kWavesPerBlock][64]) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  for (int off = KL; off < 64; off <<= 1) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) acc[c].merge_xor(off);
  }
  if (lane < KL) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) partial[wave][lane * CPL + c] = acc[c];
  }
  __syncthreads();
  if (wave == 0 && lane < KL) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      Acc t = partial[0][lane * CPL + c];
      for (int w = 1; w < kWavesPerBlock; ++w) t.merge(partial[w][lane * CPL + c]);
      acc[c] = t;
    }
  }
}

template <typename T, int CPL, typename Acc>
__device__ __forceinline__ void store_row(const GrgHipGatherArgs& a, int64_t row, int col, const Acc (&acc)[CPL]) {
  T v[CPL];
  if (a.out_scale) {
    const double sc = (double)static_cast<const T*>(a.out_scale)[row - a.out_scale_base];
#pragma unroll
    for (int c = 0; c < CPL; ++c) v[c] = (T)(acc[c].wide() * sc);
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) v[c] = (T)acc[c].value();
  }
  store_raw<T, CPL>(static_cast<T*>(a.out) + (row - a.out_base) * a.ld_out + col, v);
}

// One row by the whole workgroup (heavy rows; rows of direct tiles): 256 / kl
// edge lanes over the row's segments (or, for a split heavy row h, over its
// FP64 partial rows), a shuffle + LDS reduction, one store. Uses partial[] and
// one barrier: callers looping over rows add a barrier between rows. PIPE: the
// entry indices of the next round are loaded during the current round's gathers
// (gather_entries_pipelined); it costs registers, so the kernels whose light
// work is register-bound (tiles, row groups, streamed) keep the plain rounds.
template <typename T, int CPL, int U, bool PIPE = false, typename Acc>
__device__ __forceinline__ void workgroup_row(const GrgHipGatherArgs& a, int64_t row, int64_t h, int by, int kl_shift,
                                              Acc (&partial)[kWavesPerBlock][64]) {
  const int KL = 1 << kl_shift;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int cl = threadIdx.x & (KL - 1);
  const int e = threadIdx.x >> kl_shift;
  const int ES = kBlock >> kl_shift;
  const int col = (by * KL + cl) * CPL;
  const bool live = col < a.k;
  Acc acc[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) acc[c].zero();
  if (live) {
    const int32_t it0 = (h >= 0 && a.split_ptr) ? a.split_ptr[h] : 0;
    const int32_t it1 = (h >= 0 && a.split_ptr) ? a.split_ptr[h + 1] : 0;
    if (it1 > it0) {
      // split row: sum the FP64 partial rows written by the partials launch
      const double* part = static_cast<const double*>(a.partials) + col;
      for (int32_t it = it0 + e; it < it1; it += ES) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[c].add_wide(part[(int64_t)it * a.k + c]);
      }
    } else {
      for (int s = 0; s < a.nseg; ++s) {
        const int32_t* ptr = a.seg[s].ptr + (row - a.seg[s].ptr_base);
        if constexpr (PIPE)
          gather_entries_pipelined<T, CPL, U>(a.seg[s].idx, static_cast<const T*>(a.seg[s].src) + col, a.seg[s].ld,
                                              ptr[0] + e, ptr[1], ES, acc);
        else
          gather_entries<T, CPL, U>(a.seg[s].idx, static_cast<const T*>(a.seg[s].src) + col, a.seg[s].ld, ptr[0] + e, ptr[1],
                                    ES, acc);
      }
    }
    if (e == 0) add_row_terms<T, CPL>(a, row, col, acc);
  }
  block_reduce<CPL>(acc, KL, partial);
  if (wave == 0 && lane < KL && live) store_row<T, CPL>(a, row, col, acc);
}

// Block order over the XCDs (block b runs on XCD b % 8). GRG_HIP_FLAG_XCD_REMAP
// gives XCD x the x-th contiguous eighth of [0, n); GRG_HIP_FLAG_XCD_RUNS gives it
// runs of 2^s consecutive blocks, the eight XCDs' runs side by side in one window
// (blocks past the last full window keep their order). Both are permutations of
// [0, n).
__device__ __forceinline__ int64_t xcd_order(int64_t b, int64_t n, int32_t flags) {
  if (flags & GRG_HIP_FLAG_XCD_REMAP) {
    const int64_t q = n >> 3, r = n & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
  }
  if (flags & GRG_HIP_FLAG_XCD_RUNS) {
    const int s = (flags >> GRG_HIP_XCD_RUN_SHIFT) & 15;
    const int64_t window = (int64_t)8 << s;
    if (b < (n / window) * window) {
      const int64_t x = b & 7, j = b >> 3;
      return ((j >> s) << (s + 3)) + (x << s) + (j & (((int64_t)1 << s) - 1));
    }
  }
  return b;
}

// Block (bx, by) of a row-gather launch: a partials item, R-row lane groups of
// the light path, or a heavy row. Shared by the one-launch kernel and the
// job-table kernel.
template <typename T, int CPL, int U, int R, typename Acc>
__device__ __forceinline__ void gather_block(const GrgHipGatherArgs& a, int64_t bx, int by, int64_t n_light_blocks,
                                             int g_shift, int kl_shift, Acc (&partial)[kWavesPerBlock][64]) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int KL = 1 << kl_shift;
  Acc acc[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) acc[c].zero();

  if (a.flags & GRG_HIP_FLAG_PARTIALS) {
    // ---- partials launch: workgroup i sums chunk j of split heavy row h ----
    const int64_t item = bx;
    const int32_t h = a.split_row[item];
    const int64_t row = a.heavy_rows[h];
    const int64_t lo = (item - a.split_ptr[h]) * a.split_chunk;
    const int64_t hi = lo + a.split_chunk;
    const int cl = threadIdx.x & (KL - 1);
    const int e = threadIdx.x >> kl_shift;
    const int ES = kBlock >> kl_shift;
    const int col = (by * KL + cl) * CPL;
    if (col < a.k) {
      int64_t g = 0;  // offset of segment s in the concatenated entry list
      for (int s = 0; s < a.nseg; ++s) {
        const int32_t* ptr = a.seg[s].ptr + (row - a.seg[s].ptr_base);
        const int32_t p0 = ptr[0], p1 = ptr[1];
        const int64_t n = p1 - p0;
        const int64_t b = lo > g ? lo - g : 0;
        const int64_t t = hi - g < n ? hi - g : n;
        if (b < t)
          gather_entries_pipelined<T, CPL, U>(a.seg[s].idx, static_cast<const T*>(a.seg[s].src) + col, a.seg[s].ld,
                                              p0 + (int32_t)b + e, p0 + (int32_t)t, ES, acc);
        g += n;
      }
    }
    block_reduce<CPL>(acc, KL, partial);
    if (wave == 0 && lane < KL && col < a.k) {
      double* dst = static_cast<double*>(a.partials) + item * a.k + col;
#pragma unroll
      for (int c = 0; c < CPL; ++c) dst[c] = acc[c].wide();
    }
    return;
  }

  if (bx < n_light_blocks) {
    // ---- light path: R consecutive rows per G-lane group ----
    const int G = 1 << g_shift;
    const int lg = lane & (G - 1);
    const int e = lg >> kl_shift;
    const int ES = G >> kl_shift;
    const int col = (by * KL + (lg & (KL - 1))) * CPL;
    int64_t blk = bx;
    blk = xcd_order(blk, n_light_blocks, a.flags);
    const int64_t row0 = a.row_begin + (int64_t)R * (((blk * kWavesPerBlock + wave) << (6 - g_shift)) + (lane >> g_shift));
    const bool any_live = row0 < a.row_end && col < a.k;
    bool live[R];
    Acc racc[R][CPL];
#pragma unroll
    for (int t = 0; t < R; ++t) {
      live[t] = false;
#pragma unroll
      for (int c = 0; c < CPL; ++c) racc[t][c].zero();
    }
    if constexpr (R == 1) {
      if (a.flags & GRG_HIP_FLAG_GROUP_KEYS) {
        // one row per lane group (row0 is the group's): entries loaded by the whole group
        if (row0 < a.row_end) {
          const bool merged = (a.flags & GRG_HIP_FLAG_MERGED) != 0;
          const int32_t* ptr = merged ? a.eptr : a.seg[0].ptr - a.seg[0].ptr_base;
          const int32_t p0 = ptr[row0], p1 = ptr[row0 + 1];
          const bool light = !(a.heavy_threshold >= 0 && p1 - p0 > a.heavy_threshold);
          const bool col_ok = col < a.k;
          live[0] = light && col_ok;
          const T* src0 = static_cast<const T*>(a.seg[0].src) + col;
          if (merged) {
            const T* src1 = (a.nseg > 1 ? static_cast<const T*>(a.seg[1].src) : static_cast<const T*>(a.seg[0].src)) + col;
            const T* src2 = (a.nseg > 2 ? static_cast<const T*>(a.seg[2].src) : static_cast<const T*>(a.seg[0].src)) + col;
            const int64_t ld0 = a.seg[0].ld;
            const int64_t ld1 = a.nseg > 1 ? a.seg[1].ld : ld0;
            const int64_t ld2 = a.nseg > 2 ? a.seg[2].ld : ld0;
            gather_row_keyed_group<T, CPL, U>(a.ekey, src0, src1, src2, ld0, ld1, ld2, p0, light ? p1 : p0, e, ES, kl_shift,
                                              lane, g_shift, col_ok, racc[0]);
          } else {
            gather_row_indexed_group<T, CPL, U>(a.seg[0].idx, src0, a.seg[0].ld, p0, light ? p1 : p0, e, ES, kl_shift, lane,
                                                g_shift, col_ok, racc[0]);
          }
          if (e == 0 && live[0]) add_row_terms<T, CPL>(a, row0, col, racc[0]);
        }
        goto reduce_and_store;
      }
    }
    if (any_live && (a.flags & GRG_HIP_FLAG_MERGED)) {
      // one pointer pair per row over the merged entries of all segments
      int32_t pb[R + 1];
#pragma unroll
      for (int t = 0; t <= R; ++t) pb[t] = a.eptr[row0 + t < a.row_end ? row0 + t : a.row_end];
      int32_t j[R], end[R];
#pragma unroll
      for (int t = 0; t < R; ++t) {
        live[t] = row0 + t < a.row_end && !(a.heavy_threshold >= 0 && pb[t + 1] - pb[t] > a.heavy_threshold);
        j[t] = pb[t] + e;
        end[t] = live[t] ? pb[t + 1] : 0;
      }
      const T* src0 = static_cast<const T*>(a.seg[0].src) + col;
      const T* src1 = (a.nseg > 1 ? static_cast<const T*>(a.seg[1].src) : static_cast<const T*>(a.seg[0].src)) + col;
      const T* src2 = (a.nseg > 2 ? static_cast<const T*>(a.seg[2].src) : static_cast<const T*>(a.seg[0].src)) + col;
      const int64_t ld0 = a.seg[0].ld;
      const int64_t ld1 = a.nseg > 1 ? a.seg[1].ld : ld0;
      const int64_t ld2 = a.nseg > 2 ? a.seg[2].ld : ld0;
      gather_rows_keyed<T, CPL, U, R>(a.ekey, src0, src1, src2, ld0, ld1, ld2, j, end, ES, racc);
      if (e == 0) {
#pragma unroll
        for (int t = 0; t < R; ++t)
          if (live[t]) add_row_terms<T, CPL>(a, row0 + t, col, racc[t]);
      }
    } else if (any_live) {
      // row pointers of the R rows in every segment (rows past row_end read the level's end)
      int32_t pb[GRG_HIP_MAX_SEGS][R + 1];
      int64_t deg[R];
#pragma unroll
      for (int t = 0; t < R; ++t) deg[t] = 0;
#pragma unroll
      for (int s = 0; s < GRG_HIP_MAX_SEGS; ++s) {
        if (s < a.nseg) {
          const int32_t* ptr = a.seg[s].ptr - a.seg[s].ptr_base;
#pragma unroll
          for (int t = 0; t <= R; ++t) pb[s][t] = ptr[row0 + t < a.row_end ? row0 + t : a.row_end];
#pragma unroll
          for (int t = 0; t < R; ++t) deg[t] += pb[s][t + 1] - pb[s][t];
        }
      }
#pragma unroll
      for (int t = 0; t < R; ++t)
        live[t] = row0 + t < a.row_end && !(a.heavy_threshold >= 0 && deg[t] > a.heavy_threshold);  // heavy: heavy path
#pragma unroll
      for (int s = 0; s < GRG_HIP_MAX_SEGS; ++s) {
        if (s < a.nseg) {
          int32_t j[R], end[R];
#pragma unroll
          for (int t = 0; t < R; ++t) {
            j[t] = pb[s][t] + e;
            end[t] = live[t] ? pb[s][t + 1] : 0;
          }
          gather_rows<T, CPL, U, R>(a.seg[s].idx, static_cast<const T*>(a.seg[s].src) + col, a.seg[s].ld, j, end, ES, racc);
        }
      }
      if (e == 0) {
#pragma unroll
        for (int t = 0; t < R; ++t)
          if (live[t]) add_row_terms<T, CPL>(a, row0 + t, col, racc[t]);
      }
    }
  reduce_and_store:
    for (int off = KL; off < G; off <<= 1) {
#pragma unroll
      for (int t = 0; t < R; ++t)
#pragma unroll
        for (int c = 0; c < CPL; ++c) racc[t][c].merge_xor(off);
    }
    if (any_live && e == 0) {
#pragma unroll
      for (int t = 0; t < R; ++t)
        if (live[t]) store_row<T, CPL>(a, row0 + t, col, racc[t]);
    }
    return;
  }

  // ---- heavy path: the whole workgroup on one row ----
  const int64_t h = bx - n_light_blocks;
  workgroup_row<T, CPL, U, true>(a, a.heavy_rows[h], h, by, kl_shift, partial);
}

template <typename T, int CPL, int U, int R>
__global__ __launch_bounds__(kBlock) void row_gather_kernel(const GrgHipGatherArgs a, int64_t n_light_blocks,
                                                             int g_shift, int kl_shift) {
  using Acc = typename SumOf<T>::type;
  __shared__ Acc partial[kWavesPerBlock][64];
  gather_block<T, CPL, U, R>(a, (int64_t)blockIdx.x, (int)blockIdx.y, n_light_blocks, g_shift, kl_shift, partial);
}

// Streamed light path (GRG_HIP_FLAG_STREAM): each wave loops over batches of
// 64 / G rows (one per lane group), batch b, b + W, ... for the W waves of the
// grid. A row costs three dependent loads (row pointers -> entry keys -> node
// rows); the loop runs them as a three-stage pipeline, issuing in one
// iteration the pointers of batch i + 2, the keys of batch i + 1 and the first
// U gathers of batch i, so a row waits for one round trip instead of three.
// Entries beyond the first U per lane (long rows) are gathered in place. Keys
// are merged entry keys (GRG_HIP_FLAG_MERGED) or, with one segment, its
// indices. Heavy rows follow, one workgroup each, over the same grid.
// Keys of the first U entries (per lane) of a row [p0, p1) of a streamed
// launch; absent entries (and rows that are not live) read as 0xffffffff.
template <int U>
__device__ __forceinline__ void stream_keys(const uint32_t* __restrict__ keys, int32_t p0, int32_t p1, bool live, int e,
                                            int ES, uint32_t (&kk)[U]) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int32_t j = p0 + e + u * ES;
    kk[u] = live && j < p1 ? keys[j] : 0xffffffffu;
  }
}

// Gathers of U keyed entries (segment in the top two bits, 3 = absent).
template <typename T, int CPL, int U>
__device__ __forceinline__ void stream_gather(const uint32_t (&kk)[U], const T* src0, const T* src1, const T* src2,
                                              int64_t ld0, int64_t ld1, int64_t ld2, T (&v)[U][CPL]) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint32_t sg = kk[u] >> 30;
    if (sg != 3u) {
      const int64_t r = (int64_t)(kk[u] & 0x3fffffffu);
      const T* base = sg == 0 ? src0 : (sg == 1 ? src1 : src2);
      const int64_t ld = sg == 0 ? ld0 : (sg == 1 ? ld1 : ld2);
      load_raw_global<T, CPL>(base + r * ld, v[u]);
    } else {
#pragma unroll
      for (int c = 0; c < CPL; ++c) v[u][c] = T(0);
    }
  }
}

template <typename T, int CPL, int U>
__global__ __launch_bounds__(kBlock) void row_stream_kernel(const GrgHipGatherArgs a, int g_shift, int kl_shift) {
  using Acc = typename SumOf<T>::type;
  __shared__ Acc partial[kWavesPerBlock][64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int KL = 1 << kl_shift;
  const int G = 1 << g_shift;
  const int lg = lane & (G - 1);
  const int e = lg >> kl_shift;
  const int ES = G >> kl_shift;
  const int col = ((int)blockIdx.y * KL + (lg & (KL - 1))) * CPL;
  const bool merged = (a.flags & GRG_HIP_FLAG_MERGED) != 0;
  const int32_t* ptr = merged ? a.eptr : a.seg[0].ptr - a.seg[0].ptr_base;
  const uint32_t* keys = merged ? a.ekey : reinterpret_cast<const uint32_t*>(a.seg[0].idx);
  const T* src0 = static_cast<const T*>(a.seg[0].src) + col;
  const T* src1 = (a.nseg > 1 ? static_cast<const T*>(a.seg[1].src) : static_cast<const T*>(a.seg[0].src)) + col;
  const T* src2 = (a.nseg > 2 ? static_cast<const T*>(a.seg[2].src) : static_cast<const T*>(a.seg[0].src)) + col;
  const int64_t ld0 = a.seg[0].ld;
  const int64_t ld1 = a.nseg > 1 ? a.seg[1].ld : ld0;
  const int64_t ld2 = a.nseg > 2 ? a.seg[2].ld : ld0;
  const int64_t rows_per_wave = 64 >> g_shift;
  const int64_t n_batches = (a.row_end - a.row_begin + rows_per_wave - 1) / rows_per_wave;
  const int64_t n_waves = (int64_t)gridDim.x * kWavesPerBlock;
  const int64_t my_row = lane >> g_shift;
  const bool col_live = col < a.k;
  const int64_t thr = a.heavy_threshold;

  // batch b's row for this lane group: r, its pointers [p0, p1) (p1 < p0: no row)
  int64_t b = (int64_t)blockIdx.x * kWavesPerBlock + wave;
  int64_t rc = a.row_begin + b * rows_per_wave + my_row;
  int64_t rn = rc + n_waves * rows_per_wave;
  int32_t c0 = 0, c1 = -1, n0 = 0, n1 = -1;
  if (b < n_batches && rc < a.row_end) {
    c0 = ptr[rc];
    c1 = ptr[rc + 1];
  }
  if (b + n_waves < n_batches && rn < a.row_end) {
    n0 = ptr[rn];
    n1 = ptr[rn + 1];
  }
  bool live_c = col_live && c1 >= c0 && !(thr >= 0 && c1 - c0 > thr);
  uint32_t kc[U];
  stream_keys<U>(keys, c0, c1, live_c, e, ES, kc);
  // The previous batch's row is stored only after this batch's loads are issued:
  // stores and loads share one in-order counter (vmcnt), so a store issued ahead
  // of the gathers would make the wait for the gathers wait for the store too.
  Acc pend[CPL];
  int64_t pend_row = 0;
  bool pend_live = false;
  while (b < n_batches) {
    // pointers of batch b + 2W, keys of batch b + W, gathers of batch b: independent loads
    const int64_t rf = rn + n_waves * rows_per_wave;
    int32_t f0 = 0, f1 = -1;
    if (b + 2 * n_waves < n_batches && rf < a.row_end) {
      f0 = ptr[rf];
      f1 = ptr[rf + 1];
    }
    const bool live_n = col_live && n1 >= n0 && !(thr >= 0 && n1 - n0 > thr);
    uint32_t kn[U];
    stream_keys<U>(keys, n0, n1, live_n, e, ES, kn);
    T v[U][CPL];
    stream_gather<T, CPL, U>(kc, src0, src1, src2, ld0, ld1, ld2, v);
    if (pend_live) store_row<T, CPL>(a, pend_row, col, pend);
    Acc acc[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) acc[c].zero();
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int c = 0; c < CPL; ++c) acc[c].add(v[u][c]);
    if (live_c) {
      // entries past the first U per lane: gathered in place
      for (int32_t j = c0 + e + U * ES; j < c1; j += U * ES) {
        uint32_t kx[U];
        stream_keys<U>(keys, j - e, c1, true, e, ES, kx);
        T vx[U][CPL];
        stream_gather<T, CPL, U>(kx, src0, src1, src2, ld0, ld1, ld2, vx);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int c = 0; c < CPL; ++c) acc[c].add(vx[u][c]);
      }
      if (e == 0) add_row_terms<T, CPL>(a, rc, col, acc);
    }
    for (int off = KL; off < G; off <<= 1) {
#pragma unroll
      for (int c = 0; c < CPL; ++c) acc[c].merge_xor(off);
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) pend[c] = acc[c];
    pend_row = rc;
    pend_live = live_c && e == 0;
    b += n_waves;
    rc = rn;
    rn = rf;
    c0 = n0;
    c1 = n1;
    n0 = f0;
    n1 = f1;
    live_c = live_n;
#pragma unroll
    for (int u = 0; u < U; ++u) kc[u] = kn[u];
  }
  if (pend_live) store_row<T, CPL>(a, pend_row, col, pend);
  // heavy rows: one workgroup each, strided over the grid
  for (int64_t h = blockIdx.x; h < a.num_heavy; h += gridDim.x) {
    workgroup_row<T, CPL, U>(a, a.heavy_rows[h], h, (int)blockIdx.y, kl_shift, partial);
    __syncthreads();
  }
}

// Padding (elements) after each staged row of a tile: at least 16 bytes and one
// lane vector, so rows start on a vector boundary and consecutive rows start in
// different LDS banks.
template <typename T, int CPL>
constexpr int tile_pad() { return CPL * (int)sizeof(T) >= 16 ? CPL : 16 / (int)sizeof(T); }

template <typename T>
constexpr int tile_static_lds() { return (int)sizeof(typename SumOf<T>::type) * kWavesPerBlock * 64; }

// LDS of a tiled workgroup, in bytes: the staged rows, then the tile's row
// pointers, then its entry indices (each part a multiple of 16 bytes).
struct TileLds {
  int64_t staged, rowptr, entries;
  __host__ __device__ int64_t total() const { return staged + rowptr + entries; }
};

__host__ __device__ inline int64_t round16(int64_t b) { return (b + 15) & ~int64_t(15); }

__host__ __device__ inline TileLds tile_lds_layout(int64_t staged_rows, int64_t stride_bytes, int64_t max_rows,
                                                   int64_t max_entries) {
  TileLds t;
  t.staged = round16(staged_rows * stride_bytes);
  t.rowptr = round16((max_rows + 1) * 4);
  t.entries = round16((max_entries + 16) * 2);
  return t;
}

// Tiled launch: workgroup t copies into LDS, with coalesced loads, the
// distinct source rows of tile t (one column chunk of kl * cpl columns each),
// the tile's row pointers and its 16-bit entry indices; its lane groups then
// sum the tile's rows from LDS alone and write them. Rows of one tile share
// many sources, so each source row crosses L2 once per tile instead of once
// per entry, and no global load sits between a row and its sum. Heavy rows
// sit in direct tiles (skipped here) and follow as ordinary workgroup rows.
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void tile_gather_kernel(const GrgHipGatherArgs a, int g_shift, int kl_shift) {
  using Acc = typename SumOf<T>::type;
  extern __shared__ __align__(16) unsigned char tile_lds[];
  __shared__ Acc partial[kWavesPerBlock][64];
  const int lane = threadIdx.x & 63;
  const int KL = 1 << kl_shift;
  const int64_t nt = a.num_tiles;
  int64_t t = blockIdx.x;
  if (t >= nt) {
    workgroup_row<T, CPL, 8>(a, a.heavy_rows[t - nt], t - nt, (int)blockIdx.y, kl_shift, partial);
    return;
  }
  t = xcd_order(t, nt, a.flags);
  const int32_t r0 = a.tile_row[t], r1 = a.tile_row[t + 1];
  const int32_t s0 = a.tile_src[t], s1 = a.tile_src[t + 1];
  if (s1 == s0) {
    // direct tile (its rows stage nothing): one workgroup per row, from the segments
    for (int32_t r = r0; r < r1; ++r) {
      const int32_t deg = a.eptr[r + 1] - a.eptr[r];
      if (a.heavy_threshold >= 0 && deg > a.heavy_threshold) continue;
      workgroup_row<T, CPL, 8>(a, r, -1, (int)blockIdx.y, kl_shift, partial);
      __syncthreads();  // partial[] is reused by the next row
    }
    return;
  }
  const int chunk = KL * CPL;
  const int stride = chunk + tile_pad<T, CPL>();
  const int col0 = (int)blockIdx.y * chunk;
  const TileLds lay = tile_lds_layout(a.tile_lds_rows, (int64_t)stride * (int64_t)sizeof(T), a.tile_max_rows, a.tile_max_entries);
  T* buf = reinterpret_cast<T*>(tile_lds);
  int32_t* rowptr = reinterpret_cast<int32_t*>(tile_lds + lay.staged);
  uint16_t* ent = reinterpret_cast<uint16_t*>(tile_lds + lay.staged + lay.rowptr);
  const int32_t nrows = r1 - r0;
  const int32_t e0 = a.eptr[r0];
  const int32_t e1 = a.eptr[r1];
  const int32_t eo = e0 & 7;  // entry e0 sits at ent[eo]
  {
    // ---- row pointers and entry indices (16-byte words of 8 indices) ----
    for (int i = threadIdx.x; i <= nrows; i += kBlock) rowptr[i] = a.eptr[r0 + i] - e0;
    typedef uint32_t U4 __attribute__((ext_vector_type(4)));
    const U4* src_w = reinterpret_cast<const U4*>(a.lidx) + (e0 >> 3);
    U4* dst_w = reinterpret_cast<U4*>(ent);
    const int n_words = ((e1 + 7) >> 3) - (e0 >> 3);
    for (int i = threadIdx.x; i < n_words; i += kBlock) dst_w[i] = src_w[i];
  }
  {
    // ---- staged source rows: D rows x KL lane vectors, 4 loads in flight per thread ----
    const T* src0 = static_cast<const T*>(a.seg[0].src);
    const T* src1 = a.nseg > 1 ? static_cast<const T*>(a.seg[1].src) : src0;
    const T* src2 = a.nseg > 2 ? static_cast<const T*>(a.seg[2].src) : src0;
    const int64_t ld0 = a.seg[0].ld;
    const int64_t ld1 = a.nseg > 1 ? a.seg[1].ld : ld0;
    const int64_t ld2 = a.nseg > 2 ? a.seg[2].ld : ld0;
    const int n_pieces = (s1 - s0) << kl_shift;
    for (int i = threadIdx.x; i < n_pieces; i += 4 * kBlock) {
      T v[4][CPL];
      int dst[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int ii = i + u * kBlock;
        const int d = ii >> kl_shift, p = ii & (KL - 1);
        const int col = col0 + p * CPL;
        dst[u] = -1;
        if (ii < n_pieces && col < a.k) {
          const uint32_t key = a.src_list[s0 + d];
          const uint32_t sg = key >> 30;
          const int64_t srow = (int64_t)(key & 0x3fffffffu);
          const T* base = sg == 0 ? src0 : (sg == 1 ? src1 : src2);
          const int64_t ld = sg == 0 ? ld0 : (sg == 1 ? ld1 : ld2);
          load_raw_global<T, CPL>(base + srow * ld + col, v[u]);
          dst[u] = d * stride + p * CPL;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (dst[u] >= 0) store_raw<T, CPL>(buf + dst[u], v[u]);
    }
  }
  __syncthreads();
  // ---- sum: lane groups of G = kl * es lanes, one row each per pass, from LDS ----
  const int G = 1 << g_shift;
  const int lg = lane & (G - 1);
  const int cl = lg & (KL - 1);
  const int e = lg >> kl_shift;
  const int ES = G >> kl_shift;
  const int groups = kBlock >> g_shift;
  const int gid = (int)threadIdx.x >> g_shift;
  const int col = col0 + cl * CPL;
  const T* lds_col = buf + cl * CPL;
  const uint16_t* ent0 = ent + eo;
  for (int32_t b = 0; b < nrows; b += groups) {  // same trip count for every lane (shuffles below)
    const int32_t i = b + gid;
    bool live = i < nrows && col < a.k;
    Acc acc[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) acc[c].zero();
    int32_t p0 = 0, p1 = 0;
    if (live) {
      p0 = rowptr[i];
      p1 = rowptr[i + 1];
      if (a.heavy_threshold >= 0 && p1 - p0 > a.heavy_threshold) live = false;
    }
    if (live) {
      for (int32_t j = p0 + e; j < p1; j += 4 * ES) {
        int li[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int32_t jj = j + u * ES;
          li[u] = jj < p1 ? (int)ent0[jj] : -1;
        }
        T v[4][CPL];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (li[u] >= 0) {
            load_raw<T, CPL>(lds_col + li[u] * stride, v[u]);
          } else {
#pragma unroll
            for (int c = 0; c < CPL; ++c) v[u][c] = T(0);
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int c = 0; c < CPL; ++c) acc[c].add(v[u][c]);
      }
    }
    for (int off = KL; off < G; off <<= 1) {
#pragma unroll
      for (int c = 0; c < CPL; ++c) acc[c].merge_xor(off);
    }
    if (live && e == 0) {
      const int32_t r = r0 + i;
      add_row_terms<T, CPL>(a, r, col, acc);
      store_row<T, CPL>(a, r, col, acc);
    }
  }
}

// ---- panel launch -----------------------------------------------------------
//
// A panel is R = panel_rows consecutive rows of one launch (256, 512 or 1024).
// Its distinct sources, in key order, are cut into passes of at most
// panel_staged (<= 256) sources and panel_max_entries entries. One workgroup of
// 1024 threads runs one panel: lane groups of kl lanes own rows_per_group
// consecutive rows each and keep their sums in registers over all passes; a
// pass stages its sources in LDS (one kl-lane vector of one column chunk per
// thread and piece) and every row then sums its entries of the pass from LDS.
// A source shared by several rows of the panel is read from memory once per
// panel instead of once per entry: on the sample level of a down sweep, where
// each sample sums hundreds of parents, 1024-row panels read each parent row
// about 1.1 times instead of about 9 times.

constexpr int kPanelBlock = 1024;                   // threads of a panel workgroup
constexpr int kPanelEntryBytes = 8 * kPanelBlock;  // one 8-byte word of entry positions per thread and pass

// LDS of a panel workgroup, in bytes: the panel's pass source and entry offsets
// (int32 each), then two pass buffers, each one pass's staged rows, the end
// offsets of the panel's rows within the pass (uint16) and its entry positions
// (uint8, one 8-byte word per thread).
struct PanelLds {
  int64_t meta, staged, rowoff, entries;
  __host__ __device__ int64_t buffer() const { return staged + rowoff + entries; }
  __host__ __device__ int64_t total() const { return meta + 2 * buffer(); }
};

constexpr int kPanelPad = 255;  // entry position of padding entries: staged row 255, kept zero

__host__ __device__ inline PanelLds panel_lds_layout(int64_t max_passes, int64_t stride_bytes, int64_t panel_rows) {
  PanelLds t;
  t.meta = round16(2 * (max_passes + 1) * 4);
  t.staged = round16((kPanelPad + 1) * stride_bytes);
  t.rowoff = round16(panel_rows * 2);
  t.entries = kPanelEntryBytes;
  return t;
}

typedef uint32_t PanelWord __attribute__((ext_vector_type(2)));

// Keys of one pass's staged pieces (kl-lane vectors): the pass stages sources
// pan_src[s0 .. s0 + n) (n >= 1), at most panel_staged of them. Every thread
// loads a key for each of its PPT pieces (absent pieces reload the pass's last
// key and are marked 0xffffffff), so the load count never depends on the pass.
template <int PPT>
__device__ __forceinline__ void panel_keys(const GrgHipGatherArgs& a, int32_t s0, int32_t n, int kl_shift, uint32_t (&key)[PPT]) {
  const int n_pieces = (n < a.panel_staged ? n : a.panel_staged) << kl_shift;
  const auto src = global_ptr(a.pan_src);
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int i = threadIdx.x + u * kPanelBlock;
    const int r = i < n_pieces ? (i >> kl_shift) : (n > 0 ? n - 1 : 0);
    const uint32_t k = src[s0 + r];
    key[u] = i < n_pieces ? k : 0xffffffffu;
  }
}

// Loads of one pass into registers: the staged pieces by key, the pass's entry
// positions [e0, e1) as 8-byte words (one per thread), and the end offset of row
// threadIdx.x of the panel within the pass. Absent pieces and columns past k load
// row 0, column 0 of the first segment, words past the pass reload its first
// word, and threads past the panel's rows reload its last offset: none of them is
// stored or read, so the loaded registers are never touched before their store
// and the compiler keeps every pass in flight until then (a select on a loaded
// value would make it wait for all loads in flight).
template <typename T, int CPL, int PPT>
__device__ __forceinline__ void panel_issue(const GrgHipGatherArgs& a, const uint32_t (&key)[PPT], int col0, int kl_shift,
                                            int32_t e0, int32_t e1, int64_t rowoff_at, T (&v)[PPT][CPL], PanelWord& w,
                                            uint32_t& roff) {
  const int KL = 1 << kl_shift;
  const T* src0 = static_cast<const T*>(a.seg[0].src);
  const T* src1 = a.nseg > 1 ? static_cast<const T*>(a.seg[1].src) : src0;
  const T* src2 = a.nseg > 2 ? static_cast<const T*>(a.seg[2].src) : src0;
  const int64_t ld0 = a.seg[0].ld;
  const int64_t ld1 = a.nseg > 1 ? a.seg[1].ld : ld0;
  const int64_t ld2 = a.nseg > 2 ? a.seg[2].ld : ld0;
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int i = threadIdx.x + u * kPanelBlock;
    const int c = col0 + (i & (KL - 1)) * CPL;
    const bool ok = key[u] != 0xffffffffu && c < a.k;
    const uint32_t kk = ok ? key[u] : 0u;
    const uint32_t sg = kk >> 30;
    const int64_t srow = (int64_t)(kk & 0x3fffffffu);
    const T* base = sg == 0 ? src0 : (sg == 1 ? src1 : src2);
    const int64_t ld = sg == 0 ? ld0 : (sg == 1 ? ld1 : ld2);
    load_raw_global<T, CPL>(base + srow * ld + (ok ? c : 0), v[u]);
  }
  const int32_t w0 = e0 >> 3;
  int32_t nw = ((e1 + 7) >> 3) - w0;
  if (nw > kPanelBlock) nw = kPanelBlock;  // never more than the LDS holds
  const auto ent = global_ptr(reinterpret_cast<const PanelWord*>(a.pan_lidx));
  const int i = threadIdx.x;
  w = ent[w0 + (i < nw ? i : 0)];  // word w0 is readable (padding after the entries)
  // pass_rowoff holds R + 1 offsets per pass, the first 0: thread i reads row i's end
  const int R = a.panel_rows;
  roff = (uint32_t)global_ptr(a.pass_rowoff)[rowoff_at + (i < R ? i + 1 : R)];
}

// One pass from registers into an LDS pass buffer.
template <typename T, int CPL, int PPT>
__device__ __forceinline__ void panel_store(int n_pieces, int kl_shift, int stride, int R, const T (&v)[PPT][CPL],
                                            const PanelWord& w, uint32_t roff, T* buf, uint16_t* ends, PanelWord* ent) {
  const int KL = 1 << kl_shift;
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int i = threadIdx.x + u * kPanelBlock;
    if (i < n_pieces) store_raw<T, CPL>(buf + (i >> kl_shift) * stride + (i & (KL - 1)) * CPL, v[u]);
  }
  ent[threadIdx.x] = w;
  if ((int)threadIdx.x < R) ends[threadIdx.x] = (uint16_t)roff;
}

// A lane group's rows r0 .. r0 + RPG - 1 of the panel: their entries of one
// stored pass, summed from LDS in ascending entry order. Each row's entries in a
// pass are padded to a multiple of four with position kPanelPad (a staged row of
// zeros), so one aligned 4-byte read gives four positions and the four staged-row
// reads that follow carry no predicates.
template <typename T, int CPL, int RPG, typename Acc>
__device__ __forceinline__ void panel_sum(const T* lds_col, const uint16_t* ends, const uint32_t* ent4, int r0, int stride,
                                          Acc (&acc)[RPG][CPL]) {
  int p0 = r0 > 0 ? (int)ends[r0 - 1] : 0;
#pragma unroll
  for (int t = 0; t < RPG; ++t) {
    const int p1 = ends[r0 + t];
    for (int j = p0 >> 2; j < (p1 >> 2); ++j) {
      const uint32_t w = ent4[j];
      T x[4][CPL];
#pragma unroll
      for (int u = 0; u < 4; ++u) load_raw<T, CPL>(lds_col + (int)((w >> (8 * u)) & 255u) * stride, x[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[t][c].add(x[u][c]);
    }
    p0 = p1;
  }
}

// Panel launch: workgroup t (of 1024 threads) owns panel t (panel_rows
// consecutive rows; t in the launch's XCD block order). Lane group g (kl lanes,
// one kl * cpl column chunk) owns rows g * RPG .. g * RPG + RPG - 1 of the panel.
// The panel's pass offsets are read into LDS once. Its passes then stream
// through a pipeline: D passes are in flight in registers (each thread's pieces
// of the pass's staged source rows, its word of the pass's entry positions and
// one row's end offset) while the pass before them is summed from one of two
// LDS buffers; the keys of the next pass to issue are loaded one pass ahead. Per
// pass: store the oldest pass in flight to the buffer the pass before last
// used, one barrier, issue the pass D ahead into the freed registers (past the
// panel's last pass, the last pass again: the load count stays the same and
// those loads are never stored), sum the stored pass. The barrier also orders
// the buffer's reuse: every thread finished summing from it (two passes
// earlier) before any thread reaches the barrier of the pass in between. Each
// lane group keeps its rows' sums in registers over all passes and writes each
// row once. Heavy rows (no entries in the panels) are produced by a separate
// launch of the per-row kernel.
template <typename T, int CPL, int RPG, int PPT, int D>
__global__ __launch_bounds__(kPanelBlock) void panel_gather_kernel(const GrgHipGatherArgs a, int kl_shift) {
  using Acc = typename SumOf<T>::type;
  extern __shared__ __align__(16) unsigned char panel_lds[];
  const int KL = 1 << kl_shift;
  const int64_t np = a.num_panels;
  const int64_t t = xcd_order((int64_t)blockIdx.x, np, a.flags);
  const int R = a.panel_rows;
  const int chunk = KL * CPL;
  const int stride = chunk + tile_pad<T, CPL>();
  const int col0 = (int)blockIdx.y * chunk;
  const PanelLds lay = panel_lds_layout(a.panel_max_passes, (int64_t)stride * (int64_t)sizeof(T), R);
  int32_t* msrc = reinterpret_cast<int32_t*>(panel_lds);
  int32_t* ment = msrc + (a.panel_max_passes + 1);
  unsigned char* buf0 = panel_lds + lay.meta;
  unsigned char* buf1 = buf0 + lay.buffer();

  const int cl = threadIdx.x & (KL - 1);
  const int r0 = ((int)threadIdx.x >> kl_shift) * RPG;  // the lane group's first row in the panel
  const int col = col0 + cl * CPL;
  const bool col_ok = col < a.k;
  Acc acc[RPG][CPL];
#pragma unroll
  for (int r = 0; r < RPG; ++r)
#pragma unroll
    for (int c = 0; c < CPL; ++c) acc[r][c].zero();

  const int32_t g0 = a.pan_pass[t];
  int npass = a.pan_pass[t + 1] - g0;
  if (npass > a.panel_max_passes) npass = a.panel_max_passes;  // never more than the LDS holds
  for (int i = threadIdx.x; i <= npass; i += kPanelBlock) {
    msrc[i] = a.pass_src[g0 + i];
    ment[i] = a.pass_ent[g0 + i];
  }
  if ((int)threadIdx.x < 2 * stride) {  // staged row kPanelPad of both buffers: zeros
    T* z = reinterpret_cast<T*>((threadIdx.x < (unsigned)stride ? buf0 : buf1)) + kPanelPad * stride;
    z[threadIdx.x % stride] = T(0);
  }
  __syncthreads();

  if (npass > 0) {  // the same for every thread of the workgroup
    // pass p's registers: v / w / roff[p % D]. Keys run D passes ahead of the rows:
    // while pass p's rows are in flight, slot p % D of key holds the keys of pass p + D,
    // loaded right after pass p's rows were issued. Loads complete in issue order
    // (one vector-memory counter), so each wait is for loads issued D passes earlier
    // and never for a younger pass.
    T v[D][PPT][CPL];
    PanelWord w[D];
    uint32_t roff[D];
    uint32_t key[D][PPT];
    const int last = npass - 1;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int p = j < last ? j : last;
      panel_keys<PPT>(a, msrc[p], msrc[p + 1] - msrc[p], kl_shift, key[j]);
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int p = j < last ? j : last;
      panel_issue<T, CPL, PPT>(a, key[j], col0, kl_shift, ment[p], ment[p + 1], (int64_t)(g0 + p) * (R + 1), v[j], w[j], roff[j]);
      const int pk = j + D < last ? j + D : last;
      panel_keys<PPT>(a, msrc[pk], msrc[pk + 1] - msrc[pk], kl_shift, key[j]);
    }
    const T* lds_col0 = reinterpret_cast<const T*>(buf0) + cl * CPL;
    const T* lds_col1 = reinterpret_cast<const T*>(buf1) + cl * CPL;
    int q0 = 0;
    for (; q0 + D <= npass; q0 += D) {
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const int q = q0 + j;
        unsigned char* b = (q & 1) ? buf1 : buf0;
        const int n_q = msrc[q + 1] - msrc[q];
        panel_store<T, CPL, PPT>((n_q < a.panel_staged ? n_q : a.panel_staged) << kl_shift, kl_shift, stride, R, v[j], w[j],
                                 roff[j], reinterpret_cast<T*>(b), reinterpret_cast<uint16_t*>(b + lay.staged),
                                 reinterpret_cast<PanelWord*>(b + lay.staged + lay.rowoff));
        __syncthreads();
        // the rows of pass q + D (their keys sit in slot j), then the keys of pass q + 2D
        const int p = q + D < last ? q + D : last;
        panel_issue<T, CPL, PPT>(a, key[j], col0, kl_shift, ment[p], ment[p + 1], (int64_t)(g0 + p) * (R + 1), v[j], w[j],
                                 roff[j]);
        const int pk = q + 2 * D < last ? q + 2 * D : last;
        panel_keys<PPT>(a, msrc[pk], msrc[pk + 1] - msrc[pk], kl_shift, key[j]);
        if (col_ok)
          panel_sum<T, CPL, RPG, Acc>((q & 1) ? lds_col1 : lds_col0, reinterpret_cast<const uint16_t*>(b + lay.staged),
                                      reinterpret_cast<const uint32_t*>(b + lay.staged + lay.rowoff) + ((ment[q] & 7) >> 2), r0,
                                      stride, acc);
      }
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int q = q0 + j;
      if (q < npass) {
        unsigned char* b = (q & 1) ? buf1 : buf0;
        const int n_q = msrc[q + 1] - msrc[q];
        panel_store<T, CPL, PPT>((n_q < a.panel_staged ? n_q : a.panel_staged) << kl_shift, kl_shift, stride, R, v[j], w[j],
                                 roff[j], reinterpret_cast<T*>(b), reinterpret_cast<uint16_t*>(b + lay.staged),
                                 reinterpret_cast<PanelWord*>(b + lay.staged + lay.rowoff));
        __syncthreads();
        if (col_ok)
          panel_sum<T, CPL, RPG, Acc>((q & 1) ? lds_col1 : lds_col0, reinterpret_cast<const uint16_t*>(b + lay.staged),
                                      reinterpret_cast<const uint32_t*>(b + lay.staged + lay.rowoff) + ((ment[q] & 7) >> 2), r0,
                                      stride, acc);
      }
    }
  }
  if (!col_ok) return;
#pragma unroll
  for (int r = 0; r < RPG; ++r) {
    const int64_t row = a.row_begin + t * R + r0 + r;
    if (row >= a.row_end) break;
    if (a.heavy_threshold >= 0) {
      int64_t deg = 0;
      for (int s = 0; s < a.nseg; ++s) {
        const int32_t* ptr = a.seg[s].ptr + (row - a.seg[s].ptr_base);
        deg += ptr[1] - ptr[0];
      }
      if (deg > a.heavy_threshold) continue;  // a heavy row is produced by its heavy workgroup
    }
    add_row_terms<T, CPL>(a, row, col, acc[r]);
    store_row<T, CPL>(a, row, col, acc[r]);
  }
}

// ---- row-group launch (FP64, MFMA) ------------------------------------------

// The 16 x 16 FP64 tile of one MFMA (v_mfma_f64_16x16x4f64): lane j + 16 (i % 4),
// register i / 4 holds tile[i][j]. Its operands: lane i + 16 q holds A[i][q]
// and B[q][i] (as measured on gfx950 with integer-valued products).
typedef double GroupTile __attribute__((ext_vector_type(4)));

// Rounds r = r0, r0 + 4, r0 + 8, ... < q1 of four group entries each, summed into
// one wave's 16 x 16 tile (rows of the group x the 16 columns of the chunk whose
// column col this lane holds). Per round, lane i + 16 q holds entry q's
// membership bit of row i (A, 16 x 4) and column col of entry q's source row
// (B, 4 x 16); one MFMA adds A x B, i.e. every source row read once is added to
// each row of its mask. The rounds go in batches of U: a batch's keys and masks
// (4 U entries) are loaded by the whole wave, one or two per lane, into the
// wave's LDS slice (keys[q * U + u], masks likewise), while the previous batch's
// rows are still in flight; each lane then reads its U keys back, gathers the U
// rows, and reads each mask just before its MFMA. Registers thus hold little
// besides the gathered values (U doubles per lane). Two tiles take alternate
// rounds so that consecutive MFMAs do not wait on one another.
template <int U>
__device__ __forceinline__ GroupTile group_tile(const GrgHipGatherArgs& a, int64_t r0, int64_t q1, int col, bool col_ok, int i,
                                                int q, int lane, uint32_t* skey, uint16_t* smask) {
  static_assert(U % 16 == 0, "a batch is a multiple of 16 rounds (64 entries, one per lane)");
  constexpr int kPer = 4 * U / 64;  // staged entries per lane and batch
  const double* src0 = static_cast<const double*>(a.seg[0].src);
  const double* src1 = a.nseg > 1 ? static_cast<const double*>(a.seg[1].src) : src0;
  const double* src2 = a.nseg > 2 ? static_cast<const double*>(a.seg[2].src) : src0;
  const int64_t ld0 = a.seg[0].ld;
  const int64_t ld1 = a.nseg > 1 ? a.seg[1].ld : ld0;
  const int64_t ld2 = a.nseg > 2 ? a.seg[2].ld : ld0;
  GroupTile t0 = {0.0, 0.0, 0.0, 0.0};
  GroupTile t1 = {0.0, 0.0, 0.0, 0.0};
  // entry slot t = lane + 64 p of a batch: round u = t >> 2, entry qq = t & 3
  uint32_t nk[kPer];
  uint32_t nm[kPer];
  auto fetch = [&](int64_t r) {
#pragma unroll
    for (int p = 0; p < kPer; ++p) {
      const int t = lane + 64 * p;
      const int64_t rr = r + 4 * (int64_t)(t >> 2);
      if (rr < q1) {
        nk[p] = a.grp_key[rr * 4 + (t & 3)];
        nm[p] = (uint32_t)a.grp_mask[rr * 4 + (t & 3)];
      } else {
        nk[p] = 0xffffffffu;
        nm[p] = 0u;
      }
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int p = 0; p < kPer; ++p) {
      const int t = lane + 64 * p;
      skey[(t & 3) * U + (t >> 2)] = nk[p];
      smask[(t & 3) * U + (t >> 2)] = (uint16_t)nm[p];
    }
    // the slice is this wave's alone: order its writes before the lanes' reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  fetch(r0);
  for (int64_t r = r0; r < q1; r += 4 * U) {
    stage();
    double b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t key = skey[q * U + u];
      const uint32_t sg = key >> 30;
      b[u] = 0.0;
      if (sg != 3u && col_ok) {
        const int64_t srow = (int64_t)(key & 0x3fffffffu);
        const double* base = sg == 0 ? src0 : (sg == 1 ? src1 : src2);
        const int64_t ld = sg == 0 ? ld0 : (sg == 1 ? ld1 : ld2);
        b[u] = base[srow * ld + col];
      }
    }
    fetch(r + 4 * U);  // the next batch's keys fly while these rows arrive
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double av = ((smask[q * U + u] >> i) & 1u) ? 1.0 : 0.0;
      if (u & 1) t1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b[u], t1, 0, 0, 0);
      else t0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b[u], t0, 0, 0, 0);
    }
    // every lane has read this batch's keys and masks before the next stage() overwrites them
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  return t0 + t1;
}

// Row-group launch: workgroup b sums group b' (b in the launch's XCD block
// order) per 16-column chunk. Its four waves take interleaved rounds of the
// group's entries (wave w: rounds w, w + 4, ...), each into a tile of its own;
// waves 1-3 hand their tiles to wave 0 through LDS, wave 0 adds them in wave
// order and stores the rows of the group's store mask with their row terms.
// Workgroups past the groups' produce the heavy rows. Every call of the MFMA is
// made by all 64 lanes of a wave: the round loop's bounds are the wave's.
template <int CPL, int U>
__global__ __launch_bounds__(kBlock) void group_gather_kernel(const GrgHipGatherArgs a, int kl_shift) {
  __shared__ SumF64 partial[kWavesPerBlock][64];
  __shared__ double handoff[kWavesPerBlock - 1][4][64];
  __shared__ uint32_t skey[kWavesPerBlock][4 * U];
  __shared__ uint16_t smask[kWavesPerBlock][4 * U];
  const int64_t bx = blockIdx.x;
  const int64_t ng = a.num_groups;
  if (bx >= ng) {
    const int64_t h = bx - ng;
    workgroup_row<double, CPL, 8>(a, a.heavy_rows[h], h, (int)blockIdx.y, kl_shift, partial);
    return;
  }
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t g = xcd_order(bx, ng, a.flags);
  const int i = lane & 15;
  const int q = lane >> 4;
  const int col = (int)blockIdx.y * 16 + i;
  const bool col_ok = col < a.k;
  GroupTile t = group_tile<U>(a, (a.grp_off[g] >> 2) + w, a.grp_off[g + 1] >> 2, col, col_ok, i, q, lane, skey[w], smask[w]);
  if (w > 0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) handoff[w - 1][r][lane] = t[r];
  }
  __syncthreads();
  if (w != 0) return;
#pragma unroll
  for (int v = 0; v < kWavesPerBlock - 1; ++v) {
#pragma unroll
    for (int r = 0; r < 4; ++r) t[r] += handoff[v][r][lane];
  }
  const uint32_t store = a.grp_store[g];
  const int64_t row0 = a.row_begin + g * 16;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int il = q + 4 * r;  // the tile row in register r of this lane
    if (col_ok && ((store >> il) & 1u)) {
      SumF64 acc[1];
      acc[0].s = t[r];
      add_row_terms<double, 1>(a, row0 + il, col, acc);
      store_row<double, 1>(a, row0 + il, col, acc);
    }
  }
}


// ---- row-weighted sums ------------------------------------------------------

// out[j] = sum_m x[j + m * ld] * w[m] (j < rows, m < cols), in two fixed-order
// passes: workgroup b sums the flattened (m, j) elements b*256*iters + ... of a
// grid-stride walk whose stride is a multiple of rows, so each thread keeps one
// j; the workgroup's threads of equal j are summed in thread order into
// partials[b * rows + j]; the second pass sums the partials of each j in
// workgroup order. FP64 throughout the sums, so the result is deterministic.
template <typename T>
__global__ __launch_bounds__(kBlock) void row_weighted_partials_kernel(const T* __restrict__ x, const T* __restrict__ w,
                                                                      int64_t rows, int64_t cols, int64_t ld,
                                                                      double* __restrict__ partials) {
  __shared__ double part[kBlock];
  const int64_t per = (kBlock / rows) * rows;  // threads in use: a multiple of rows
  const int t = threadIdx.x;
  double acc = 0.0;
  if (t < per) {
    const int64_t j = t % rows;
    const int64_t stride = (int64_t)gridDim.x * per;
    const int64_t total = rows * cols;
    for (int64_t f = (int64_t)blockIdx.x * per + t; f < total; f += stride) {
      const int64_t m = f / rows;
      acc += (double)x[j + m * ld] * (double)w[m];
    }
  }
  part[t] = acc;
  __syncthreads();
  if (t < rows) {
    double s = 0.0;
    for (int64_t i = t; i < per; i += rows) s += part[i];
    partials[(int64_t)blockIdx.x * rows + t] = s;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void row_weighted_finish_kernel(const double* __restrict__ partials, int64_t rows,
                                                                    int32_t n_blocks, T* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= rows) return;
  double s = 0.0;
  for (int32_t b = 0; b < n_blocks; ++b) s += partials[(int64_t)b * rows + j];
  out[j] = (T)s;
}

// ---- job tables -------------------------------------------------------------

constexpr uint32_t kSpinLimit = 1u << 24;  // barrier polls before a wait is declared failed

// Grid barrier of the persistent job kernel, shaped by the caches and by the
// cost of device-scope atomics. The CUs of one XCD share its L2 (vector L1 is
// write-through); only agent-scope fences move data between the XCDs' L2s; and
// device-scope atomics on one word serialize, so no word takes more than a few
// dozen of them per barrier. Every counter sits in its own slot of kBarStride
// words (its own 256-byte line) of the barrier buffer, zeroed before each launch:
constexpr int kBarStride = 64;
constexpr int kMaxXcd = 16;   // XCC_ID is 4 bits
constexpr int kMaxSub = 16;   // arrival sub-groups per XCD: workgroup w joins sub-group (w / 8) % 16
constexpr int kJoinFan = 32;  // registration sub-counters: workgroup w registers on w % 32
constexpr int kSlotGrid = 0;                              // XCDs done with a barrier (monotonic)
constexpr int kSlotJoin = 1;                              // registration sub-counters that filled up
constexpr int kSlotGroups = 2;                            // XCDs holding at least one workgroup
constexpr int kSlotJoinSub = 3;                           // [kJoinFan]: workgroups registered
constexpr int kSlotSubs = kSlotJoinSub + kJoinFan;        // [XCD]: its non-empty sub-groups
constexpr int kSlotXcd = kSlotSubs + kMaxXcd;             // [XCD]: its sub-groups done (monotonic)
constexpr int kSlotRelease = kSlotXcd + kMaxXcd;          // [XCD]: barriers it is released from
constexpr int kSlotCount = kSlotRelease + kMaxXcd;        // [XCD][sub-group]: workgroups in it
constexpr int kSlotSub = kSlotCount + kMaxXcd * kMaxSub;  // [XCD][sub-group]: arrivals (monotonic)
constexpr int kBarSlots = kSlotSub + kMaxXcd * kMaxSub;
static_assert(kBarSlots * kBarStride <= GRG_HIP_BARRIER_WORDS, "barrier buffer too small for its slots");
// s_getreg operand for HW_REG_XCC_ID bits [3:0]: register 20, offset 0, 4 bits.
constexpr int kHwRegXccId = 20 | (0 << 6) | ((4 - 1) << 11);

// One workgroup's place in the barrier (kept in LDS, used by thread 0).
struct GridSync {
  uint32_t* bar;
  int32_t* status;
  uint32_t xcd;       // this workgroup's XCD
  uint32_t sub;       // its arrival sub-group on that XCD
  uint32_t sub_size;  // workgroups in that sub-group
  uint32_t subs;      // non-empty sub-groups on that XCD
  uint32_t groups;    // XCDs holding at least one workgroup of the grid
};

__device__ __forceinline__ uint32_t* bar_slot(uint32_t* bar, int slot) { return bar + (size_t)slot * kBarStride; }

__device__ __forceinline__ uint32_t atomic_get(const uint32_t* w) {
  return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t atomic_bump(uint32_t* w) {
  return __hip_atomic_fetch_add(w, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
}

// Poll until *w >= target; past kSpinLimit polls set *status and give up.
__device__ __forceinline__ void wait_at_least(const uint32_t* w, uint32_t target, int32_t* status) {
  uint32_t spins = 0;
  while (atomic_get(w) < target) {
    __builtin_amdgcn_s_sleep(1);
    if (++spins > kSpinLimit) {
      __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
  }
}

// Kernel start (thread 0): count this workgroup in its sub-group (the first
// of a sub-group counts it on its XCD, the first of an XCD counts the XCD),
// register through a two-level counter and wait until the whole grid has
// registered, so that every count is final. Only atomics are exchanged: no
// cache maintenance is needed.
__device__ __forceinline__ GridSync grid_join(uint32_t* bar, int32_t* status) {
  GridSync g{bar, status, 0u, 0u, 0u, 0u, 0u};
  g.xcd = (uint32_t)__builtin_amdgcn_s_getreg(kHwRegXccId) & (kMaxXcd - 1);
  g.sub = (blockIdx.x / 8u) % (uint32_t)kMaxSub;
  const int cell = (int)(g.xcd * kMaxSub + g.sub);
  if (atomic_bump(bar_slot(bar, kSlotCount + cell)) == 1u && atomic_bump(bar_slot(bar, kSlotSubs + (int)g.xcd)) == 1u)
    atomic_bump(bar_slot(bar, kSlotGroups));
  __builtin_amdgcn_s_waitcnt(0);  // counted before registered
  const uint32_t j = blockIdx.x % (uint32_t)kJoinFan;
  const uint32_t in_j = (gridDim.x - j + kJoinFan - 1) / kJoinFan;  // workgroups w < gridDim.x with w % kJoinFan == j
  if (atomic_bump(bar_slot(bar, kSlotJoinSub + (int)j)) == in_j) atomic_bump(bar_slot(bar, kSlotJoin));
  wait_at_least(bar_slot(bar, kSlotJoin), min(gridDim.x, (uint32_t)kJoinFan), status);
  g.sub_size = atomic_get(bar_slot(bar, kSlotCount + cell));
  g.subs = atomic_get(bar_slot(bar, kSlotSubs + (int)g.xcd));
  g.groups = atomic_get(bar_slot(bar, kSlotGroups));
  return g;
}

// Barrier number n (1, 2, ... within a launch). Every wave first waits for its
// own stores to reach the XCD's L2. Thread 0 arrives on its sub-group's
// counter; the last of a sub-group arrives on the XCD's counter; the last of an
// XCD writes that L2 back once (agent-scope release) and arrives on the grid
// counter; the last XCD to arrive raises every XCD's release word. Each
// workgroup waits on its XCD's word and then drops stale L1/L2 lines with an
// agent-scope acquire (an invalidate at workgroup scope would leave its CU's
// L1 alone). Every wait is bounded (wait_at_least), so the kernel always ends;
// a failed wait sets *status.
__device__ __forceinline__ void grid_barrier(const GridSync& g, uint32_t n) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    bool last = atomic_bump(bar_slot(g.bar, kSlotSub + (int)(g.xcd * kMaxSub + g.sub))) == g.sub_size * n;
    if (last) last = atomic_bump(bar_slot(g.bar, kSlotXcd + (int)g.xcd)) == g.subs * n;
    if (last) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __builtin_amdgcn_s_waitcnt(0);  // the writeback completes before the grid arrival
      if (atomic_bump(bar_slot(g.bar, kSlotGrid)) == g.groups * n) {
        for (int x = 0; x < kMaxXcd; ++x)
          __hip_atomic_store(bar_slot(g.bar, kSlotRelease + x), n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    wait_at_least(bar_slot(g.bar, kSlotRelease + (int)g.xcd), n, g.status);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __builtin_amdgcn_s_waitcnt(0);
  }
  __syncthreads();
}

template <typename T, typename Acc>
__device__ __forceinline__ void run_job_block(const GrgHipJob& J, int64_t bx, int by, Acc (&partial)[kWavesPerBlock][64]) {
  const GrgHipGatherArgs& a = J.args;
  if constexpr (sizeof(T) == 8) {
    switch (a.cpl) {
      case 1: gather_block<T, 1, 8, 1>(a, bx, by, J.n_light, J.g_shift, J.kl_shift, partial); break;
      case 2: gather_block<T, 2, 8, 1>(a, bx, by, J.n_light, J.g_shift, J.kl_shift, partial); break;
      default: gather_block<T, 4, 8, 1>(a, bx, by, J.n_light, J.g_shift, J.kl_shift, partial); break;
    }
  } else {
    switch (a.cpl) {
      case 1: gather_block<T, 1, 8, 1>(a, bx, by, J.n_light, J.g_shift, J.kl_shift, partial); break;
      case 2: gather_block<T, 2, 8, 1>(a, bx, by, J.n_light, J.g_shift, J.kl_shift, partial); break;
      case 8: gather_block<T, 8, 8, 1>(a, bx, by, J.n_light, J.g_shift, J.kl_shift, partial); break;
      default: gather_block<T, 4, 8, 1>(a, bx, by, J.n_light, J.g_shift, J.kl_shift, partial); break;
    }
  }
}

// Phases [phase_begin, phase_end) of a job table. Each phase's virtual
// workgroups (jobs laid end to end, each job gx x gy) are taken in strides of
// the grid; with a barrier counter (persistent launch) the grid meets between
// phases, without one the launch covers a single phase.
template <typename T>
__global__ __launch_bounds__(kBlock, 4) void jobs_kernel(const GrgHipJob* __restrict__ jobs,
                                                      const int32_t* __restrict__ phase_job_off,
                                                      const int64_t* __restrict__ phase_blocks, int32_t phase_begin,
                                                      int32_t phase_end, uint32_t* barrier, int32_t* status) {
  using Acc = typename SumOf<T>::type;
  __shared__ Acc partial[kWavesPerBlock][64];
  __shared__ GridSync sync;  // in LDS: nothing of the barrier stays live in registers across the gathers
  if (barrier && threadIdx.x == 0) sync = grid_join(barrier, status);
  for (int32_t p = phase_begin; p < phase_end; ++p) {
    const int32_t j0 = phase_job_off[p], j1 = phase_job_off[p + 1];
    const int64_t nb = phase_blocks[p];
    for (int64_t vb = blockIdx.x; vb < nb; vb += gridDim.x) {
      int32_t lo = j0, hi = j1 - 1;  // the last job of the phase starting at or before vb
      while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].block_begin <= vb) lo = mid;
        else hi = mid - 1;
      }
      const GrgHipJob& J = jobs[lo];
      const int64_t local = vb - J.block_begin;
      run_job_block<T>(J, local % J.gx, (int)(local / J.gx), partial);
      __syncthreads();  // partial[] is reused by the next block
    }
    if (barrier && p + 1 < phase_end) grid_barrier(sync, (uint32_t)(p - phase_begin + 1));
  }
}

__global__ void reset_counter_kernel(uint32_t* bar) {
  for (int slot = (int)threadIdx.x; slot < kBarSlots; slot += (int)blockDim.x) *bar_slot(bar, slot) = 0u;
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
int ilog2(int v) {
  int s = 0;
  while ((1 << s) < v) ++s;
  return s;
}

bool misaligned(const void* p, int64_t ld, int cpl, size_t elem) {
  if (!p) return false;
  return (reinterpret_cast<uintptr_t>(p) % (cpl * elem)) != 0 || (ld % cpl) != 0;
}

constexpr int64_t kTileLdsLimit = 160 * 1024;  // LDS per CU on gfx950, including the kernel's static LDS

// Dynamic LDS of a tiled launch plus the kernel's static LDS.
int64_t tile_lds_bytes(const GrgHipGatherArgs& a) {
  const bool f64 = a.dtype == GRG_HIP_F64;
  const int64_t elem = f64 ? 8 : 4;
  const int64_t pad = a.cpl * elem >= 16 ? a.cpl : 16 / elem;
  const int64_t stride = (int64_t)a.kl * a.cpl + pad;
  const int64_t stat = f64 ? tile_static_lds<double>() : tile_static_lds<float>();
  return tile_lds_layout(a.tile_lds_rows, stride * elem, a.tile_max_rows, a.tile_max_entries).total() + stat;
}

// LDS of a panel launch (all of it dynamic).
int64_t panel_lds_bytes(const GrgHipGatherArgs& a) {
  const int64_t elem = a.dtype == GRG_HIP_F64 ? 8 : 4;
  const int64_t pad = a.cpl * elem >= 16 ? a.cpl : 16 / elem;
  const int64_t stride = (int64_t)a.kl * a.cpl + pad;
  return panel_lds_layout(a.panel_max_passes, stride * elem, a.panel_rows).total();
}

int validate(const GrgHipGatherArgs& a) {
  if (a.dtype != GRG_HIP_F64 && a.dtype != GRG_HIP_F32C) return fail("unknown dtype code " + std::to_string(a.dtype));
  if (a.row_end < a.row_begin) return fail("row_end < row_begin");
  if (a.k < 1) return fail("k must be >= 1");
  if (a.cpl != 1 && a.cpl != 2 && a.cpl != 4 && !(a.cpl == 8 && a.dtype == GRG_HIP_F32C))
    return fail("cpl must be 1, 2 or 4 (or 8 for FP32 values: 32 bytes per lane)");
  if (!is_pow2(a.kl) || !is_pow2(a.es) || a.kl * a.es > 64) return fail("kl and es must be powers of two with kl * es <= 64");
  if (a.kl * a.cpl > 64) return fail("kl * cpl must be <= 64");
  if (a.flags & GRG_HIP_FLAG_PANEL) {
    const int u = a.unroll ? a.unroll : 8;
    const int r = a.rows_per_group ? a.rows_per_group : 1;
    if ((u != 4 && u != 8) || (r != 1 && r != 2 && r != 4))
      return fail("panel launch: unroll (passes in flight) must be 4 or 8 and rows_per_group 1, 2 or 4");
  } else {
    const int u = a.unroll ? a.unroll : 8;
    const int r = a.rows_per_group ? a.rows_per_group : 1;
    const bool ok = (r == 1 && (u == 4 || u == 8 || u == 16)) || (r == 2 && (u == 4 || u == 8)) || (r == 4 && u == 4);
    if (!ok) return fail("(unroll, rows_per_group) must be one of (4|8|16, 1), (4|8, 2), (4, 4)");
  }
  if (a.flags & ~(GRG_HIP_FLAG_XCD_REMAP | GRG_HIP_FLAG_PARTIALS | GRG_HIP_FLAG_TILED | GRG_HIP_FLAG_MERGED |
                  GRG_HIP_FLAG_STREAM | GRG_HIP_FLAG_PANEL | GRG_HIP_FLAG_XCD_RUNS | (15 << GRG_HIP_XCD_RUN_SHIFT) |
                  GRG_HIP_FLAG_GROUP_KEYS | GRG_HIP_FLAG_GROUPS))
    return fail("unknown flags");
  if (a.flags & GRG_HIP_FLAG_GROUPS) {
    if (a.dtype != GRG_HIP_F64) return fail("the row-group launch sums FP64 values");
    if (a.flags & (GRG_HIP_FLAG_TILED | GRG_HIP_FLAG_PANEL | GRG_HIP_FLAG_STREAM | GRG_HIP_FLAG_MERGED | GRG_HIP_FLAG_GROUP_KEYS |
                   GRG_HIP_FLAG_PARTIALS))
      return fail("a row-group launch is not tiled, panel, streamed, merged, group-keyed or a partials launch");
    if (a.num_groups < 0) return fail("num_groups < 0");
    if (a.num_groups > 0 && (!a.grp_off || !a.grp_store || !a.grp_key || !a.grp_mask))
      return fail("row-group launch needs grp_off, grp_store, grp_key and grp_mask");
    if (a.num_groups * 16 < a.row_end - a.row_begin) return fail("row-group launch: the groups do not cover the rows");
    if (a.kl * a.cpl != 16) return fail("row-group launch: kl * cpl must be 16 (one 16-column chunk)");
    if (a.nseg < 1) return fail("row-group launch needs at least one segment");
    if (a.unroll != 0 && a.unroll != 4 && a.unroll != 8 && a.unroll != 16) return fail("row-group launch: unroll must be 4, 8 or 16");
  }
  if (a.flags & GRG_HIP_FLAG_GROUP_KEYS) {
    if (a.flags & (GRG_HIP_FLAG_TILED | GRG_HIP_FLAG_PANEL | GRG_HIP_FLAG_STREAM | GRG_HIP_FLAG_PARTIALS))
      return fail("group-loaded entries apply to the per-row light path (not tiled, panel, streamed or partials)");
    if (!(a.flags & GRG_HIP_FLAG_MERGED) && a.nseg != 1)
      return fail("group-loaded entries need merged entries or one segment");
    if (a.rows_per_group != 0 && a.rows_per_group != 1) return fail("group-loaded entries need one row per lane group");
  }
  if ((a.flags & (15 << GRG_HIP_XCD_RUN_SHIFT)) && !(a.flags & GRG_HIP_FLAG_XCD_RUNS))
    return fail("a run length without GRG_HIP_FLAG_XCD_RUNS");
  if ((a.flags & GRG_HIP_FLAG_XCD_RUNS) && (a.flags & (GRG_HIP_FLAG_XCD_REMAP | GRG_HIP_FLAG_STREAM)))
    return fail("XCD runs apply to the per-row light path, tiles and panels (not with XCD_REMAP or STREAM)");
  if (a.flags & GRG_HIP_FLAG_PANEL) {
    if (a.flags & (GRG_HIP_FLAG_TILED | GRG_HIP_FLAG_PARTIALS | GRG_HIP_FLAG_STREAM | GRG_HIP_FLAG_MERGED))
      return fail("a panel launch is not tiled, merged, streamed or a partials launch");
    if (!a.pan_pass || !a.pass_src || !a.pass_ent || !a.pass_rowoff || !a.pan_src || !a.pan_lidx)
      return fail("panel launch needs pan_pass, pass_src, pass_ent, pass_rowoff, pan_src and pan_lidx");
    if (a.nseg < 1) return fail("panel launch needs at least one segment");
    if (a.num_panels < 0) return fail("num_panels < 0");
    const int rpg = a.rows_per_group ? a.rows_per_group : 1;
    if (a.es != 1) return fail("panel launch: es must be 1 (a lane group of kl lanes owns whole rows)");
    if (rpg > a.kl || (int64_t)a.panel_rows * a.kl != (int64_t)kPanelBlock * rpg)
      return fail("panel launch: panel_rows must be 1024 / kl * rows_per_group, with rows_per_group <= kl");
    if (a.num_panels * (int64_t)a.panel_rows < a.row_end - a.row_begin) return fail("panel launch: the panels do not cover the rows");
    if (a.panel_staged < 1 || a.panel_staged > kPanelPad || (int64_t)a.panel_staged * a.kl > 2 * (int64_t)kPanelBlock)
      return fail("panel launch: panel_staged must be in [1, 255] with panel_staged * kl <= 2048");
    if (a.panel_max_passes < 0) return fail("panel launch: panel_max_passes < 0");
    if (a.panel_max_entries < 0 || (int64_t)a.panel_max_entries + 32 > (int64_t)kPanelEntryBytes)
      return fail("panel launch: panel_max_entries out of range");
    if (reinterpret_cast<uintptr_t>(a.pan_lidx) % 16) return fail("pan_lidx must be 16-byte aligned");
    if (panel_lds_bytes(a) > kTileLdsLimit) return fail("panel launch: one pass exceeds the LDS");
  }
  if (a.flags & GRG_HIP_FLAG_STREAM) {
    if (a.flags & (GRG_HIP_FLAG_TILED | GRG_HIP_FLAG_PARTIALS)) return fail("a streamed launch is neither tiled nor a partials launch");
    if (!(a.flags & GRG_HIP_FLAG_MERGED) && a.nseg != 1) return fail("a streamed launch needs merged entries or one segment");
    if (a.nseg < 1) return fail("a streamed launch needs at least one segment");
    if (a.unroll != 0 && a.unroll != 4 && a.unroll != 8) return fail("a streamed launch runs 4 or 8 gathers in flight");
    if (a.stream_blocks < 0) return fail("stream_blocks < 0");
  }
  if ((a.flags & GRG_HIP_FLAG_MERGED) && (!a.eptr || !a.ekey || a.nseg < 1))
    return fail("merged-entry launch needs eptr, ekey and at least one segment");
  if ((a.flags & GRG_HIP_FLAG_TILED) && (a.flags & GRG_HIP_FLAG_PARTIALS)) return fail("a launch is either tiled or a partials launch");
  if (a.k % a.cpl) return fail("k must be a multiple of cpl");
  if (!a.out) return fail("out is null");
  if (a.nseg < 0 || a.nseg > GRG_HIP_MAX_SEGS) return fail("nseg out of range");
  if (a.nmap < 0 || a.nmap > GRG_HIP_MAX_MAPS) return fail("nmap out of range");
  const size_t elem = a.dtype == GRG_HIP_F64 ? 8 : 4;
  if (misaligned(a.out, a.ld_out, a.cpl, elem)) return fail("out pointer / ld_out not aligned to cpl columns");
  for (int s = 0; s < a.nseg; ++s) {
    if (!a.seg[s].ptr || !a.seg[s].idx || !a.seg[s].src) return fail("segment " + std::to_string(s) + " has a null pointer");
    if (misaligned(a.seg[s].src, a.seg[s].ld, a.cpl, elem)) return fail("segment src / ld not aligned to cpl columns");
  }
  for (int m = 0; m < a.nmap; ++m) {
    if (!a.map[m].src) return fail("map " + std::to_string(m) + " has a null src");
    if (misaligned(a.map[m].src, a.map[m].ld, a.cpl, elem)) return fail("map src / ld not aligned to cpl columns");
  }
  if (a.num_heavy < 0 || (a.num_heavy > 0 && !a.heavy_rows)) return fail("heavy rows missing");
  if (a.num_heavy > 0 && a.heavy_threshold < 0) return fail("heavy rows need heavy_threshold >= 0");
  if (a.split_ptr && (!a.partials || a.num_heavy < 1)) return fail("split rows need partials and heavy rows");
  if (a.flags & GRG_HIP_FLAG_PARTIALS) {
    if (!a.split_ptr || !a.split_row || !a.partials) return fail("partials launch needs split_ptr, split_row and partials");
    if (!a.heavy_rows || a.num_heavy < 1) return fail("partials launch needs the heavy rows");
    if (a.split_items < 0) return fail("split_items < 0");
    if (a.split_chunk < 1 || a.split_chunk > 0x40000000LL) return fail("split_chunk out of range");
    if (reinterpret_cast<uintptr_t>(a.partials) % (a.cpl * 8)) return fail("partials not aligned to cpl doubles");
  }
  if (a.flags & GRG_HIP_FLAG_TILED) {
    if (!a.tile_row || !a.tile_src || !a.src_list || !a.eptr || !a.lidx) return fail("tiled launch needs tile_row, tile_src, src_list, eptr and lidx");
    if (a.nseg < 1) return fail("tiled launch needs at least one segment");
    if (a.num_tiles < 0) return fail("num_tiles < 0");
    if (a.tile_lds_rows < 0 || a.tile_lds_rows > GRG_HIP_TILE_MAX_STAGED) return fail("tile_lds_rows out of range");
    if (a.tile_max_rows < 0 || a.tile_max_entries < 0) return fail("tile_max_rows / tile_max_entries < 0");
    if (reinterpret_cast<uintptr_t>(a.lidx) % 16) return fail("lidx must be 16-byte aligned");
    if ((int64_t)a.kl * a.cpl * (int64_t)elem > GRG_HIP_TILE_CHUNK_BYTES) return fail("tiled launch: kl * cpl columns exceed the staged row bytes");
    if (tile_lds_bytes(a) > kTileLdsLimit) return fail("tiled launch: the staged rows exceed the LDS");
  }
  return 0;
}

template <typename T, int CPL, int U, int R>
void launch_typed(const GrgHipGatherArgs& a, dim3 grid, int64_t n_light, int g_shift, int kl_shift, hipStream_t s) {
  hipLaunchKernelGGL((row_gather_kernel<T, CPL, U, R>), grid, dim3(kBlock), 0, s, a, n_light, g_shift, kl_shift);
}

// (unroll, rows per group) pairs with at most 16 gathers in flight per lane
template <typename T, int CPL>
void launch_unrolled(const GrgHipGatherArgs& a, dim3 grid, int64_t n_light, int g_shift, int kl_shift, hipStream_t s) {
  const int u = a.unroll ? a.unroll : 8;
  const int r = a.rows_per_group ? a.rows_per_group : 1;
  if (u == 4 && r == 1) launch_typed<T, CPL, 4, 1>(a, grid, n_light, g_shift, kl_shift, s);
  else if (u == 16 && r == 1) launch_typed<T, CPL, 16, 1>(a, grid, n_light, g_shift, kl_shift, s);
  else if (u == 4 && r == 2) launch_typed<T, CPL, 4, 2>(a, grid, n_light, g_shift, kl_shift, s);
  else if (u == 8 && r == 2) launch_typed<T, CPL, 8, 2>(a, grid, n_light, g_shift, kl_shift, s);
  else if (u == 4 && r == 4) launch_typed<T, CPL, 4, 4>(a, grid, n_light, g_shift, kl_shift, s);
  else launch_typed<T, CPL, 8, 1>(a, grid, n_light, g_shift, kl_shift, s);
}

template <typename T, int CPL>
int launch_tiled_typed(const GrgHipGatherArgs& a, dim3 grid, int g_shift, int kl_shift, hipStream_t s) {
  const int64_t stat = tile_static_lds<T>();
  const size_t dyn = (size_t)(tile_lds_bytes(a) - stat);
  if (dyn > 64 * 1024) {
    hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void*>(&tile_gather_kernel<T, CPL>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    if (err != hipSuccess) return fail(std::string("hipFuncSetAttribute: ") + hipGetErrorString(err), 2);
  }
  hipLaunchKernelGGL((tile_gather_kernel<T, CPL>), grid, dim3(kBlock), dyn, s, a, g_shift, kl_shift);
  return 0;
}

template <typename T, int CPL, int RPG, int PPT, int D>
int launch_panel_depth(const GrgHipGatherArgs& a, dim3 grid, int kl_shift, hipStream_t s) {
  const size_t dyn = (size_t)panel_lds_bytes(a);
  if (dyn > 64 * 1024) {
    hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void*>(&panel_gather_kernel<T, CPL, RPG, PPT, D>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    if (err != hipSuccess) return fail(std::string("hipFuncSetAttribute: ") + hipGetErrorString(err), 2);
  }
  hipLaunchKernelGGL((panel_gather_kernel<T, CPL, RPG, PPT, D>), grid, dim3(kPanelBlock), dyn, s, a, kl_shift);
  return 0;
}

// passes in flight: 2 for unroll 4, else 3
template <typename T, int CPL, int RPG, int PPT>
int launch_panel_pieces(const GrgHipGatherArgs& a, dim3 grid, int kl_shift, hipStream_t s) {
  if (a.unroll == 4) return launch_panel_depth<T, CPL, RPG, PPT, 2>(a, grid, kl_shift, s);
  return launch_panel_depth<T, CPL, RPG, PPT, 3>(a, grid, kl_shift, s);
}

template <typename T, int CPL, int RPG>
int launch_panel_rows(const GrgHipGatherArgs& a, dim3 grid, int kl_shift, hipStream_t s) {
  const int64_t pieces = ((int64_t)a.panel_staged * a.kl + kPanelBlock - 1) / kPanelBlock;
  if (pieces <= 1) return launch_panel_pieces<T, CPL, RPG, 1>(a, grid, kl_shift, s);
  return launch_panel_pieces<T, CPL, RPG, 2>(a, grid, kl_shift, s);
}

// The panels of a launch, then its heavy rows (one workgroup each, by the per-row
// kernel with no light blocks); both write disjoint rows in stream order.
template <typename T, int CPL>
int launch_panel_typed(const GrgHipGatherArgs& a, int kl_shift, hipStream_t s) {
  const int chunk = a.kl * a.cpl;
  const unsigned gy = (unsigned)((a.k + chunk - 1) / chunk);
  if (a.num_panels > 0) {
    const dim3 grid((unsigned)a.num_panels, gy);
    const int rpg = a.rows_per_group ? a.rows_per_group : 1;
    int rc = 0;
    if (rpg == 1) rc = launch_panel_rows<T, CPL, 1>(a, grid, kl_shift, s);
    else if (rpg == 2) rc = launch_panel_rows<T, CPL, 2>(a, grid, kl_shift, s);
    else rc = launch_panel_rows<T, CPL, 4>(a, grid, kl_shift, s);
    if (rc) return rc;
  }
  if (a.num_heavy > 0) {
    GrgHipGatherArgs h = a;
    h.flags = 0;
    h.es = 1;
    h.unroll = 8;
    h.rows_per_group = 1;
    launch_typed<T, CPL, 8, 1>(h, dim3((unsigned)a.num_heavy, gy), 0, ilog2(a.kl), kl_shift, s);
  }
  return 0;
}

template <typename T, int CPL>
int launch_stream_typed(const GrgHipGatherArgs& a, int64_t want_blocks, int gy, int g_shift, int kl_shift, hipStream_t s) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t err = hipStreamGetDevice(s, &dev);
  if (err == hipSuccess) err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int u = a.unroll ? a.unroll : 8;
  const void* fn = u == 4 ? reinterpret_cast<const void*>(&row_stream_kernel<T, CPL, 4>)
                          : reinterpret_cast<const void*>(&row_stream_kernel<T, CPL, 8>);
  if (err == hipSuccess) err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0);
  if (err != hipSuccess) return fail(std::string("occupancy query: ") + hipGetErrorString(err), 2);
  // resident workgroups (over the column chunks) bound the grid; each wave then loops
  int64_t gx = (int64_t)cus * per_cu / (gy > 0 ? gy : 1);
  if (a.stream_blocks > 0 && a.stream_blocks < gx) gx = a.stream_blocks;
  if (gx > want_blocks) gx = want_blocks;
  if (gx < 1) gx = 1;
  dim3 grid((unsigned)gx, (unsigned)gy);
  if (u == 4) hipLaunchKernelGGL((row_stream_kernel<T, CPL, 4>), grid, dim3(kBlock), 0, s, a, g_shift, kl_shift);
  else hipLaunchKernelGGL((row_stream_kernel<T, CPL, 8>), grid, dim3(kBlock), 0, s, a, g_shift, kl_shift);
  return 0;
}

template <int CPL>
void launch_groups_typed(const GrgHipGatherArgs& a, dim3 grid, int kl_shift, hipStream_t s) {
  // unroll 4 / 8 / 16 select batches of 16 / 32 / 48 rounds
  const int u = a.unroll ? a.unroll : 8;
  if (u == 4) hipLaunchKernelGGL((group_gather_kernel<CPL, 16>), grid, dim3(kBlock), 0, s, a, kl_shift);
  else if (u == 16) hipLaunchKernelGGL((group_gather_kernel<CPL, 48>), grid, dim3(kBlock), 0, s, a, kl_shift);
  else hipLaunchKernelGGL((group_gather_kernel<CPL, 32>), grid, dim3(kBlock), 0, s, a, kl_shift);
}

int launch_one(const GrgHipGatherArgs& a, hipStream_t stream) {
  if (int rc = validate(a)) return rc;
  const int G = a.kl * a.es;
  const int g_shift = ilog2(G), kl_shift = ilog2(a.kl);
  int64_t n_light = 0, gx = 0;
  if (a.flags & GRG_HIP_FLAG_GROUPS) {
    gx = a.num_groups + a.num_heavy;
    if (gx == 0) return 0;
    if (gx > 0x7fffffffLL) return fail("grid too large");
    dim3 grid((unsigned)gx, (unsigned)((a.k + 15) / 16));
    switch (a.cpl) {
      case 1: launch_groups_typed<1>(a, grid, kl_shift, stream); break;
      case 2: launch_groups_typed<2>(a, grid, kl_shift, stream); break;
      default: launch_groups_typed<4>(a, grid, kl_shift, stream); break;
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return fail(std::string("kernel launch failed: ") + hipGetErrorString(err), 2);
    return 0;
  }
  if (a.flags & GRG_HIP_FLAG_STREAM) {
    const int64_t rows_per_block = (int64_t)kWavesPerBlock * (64 / G);
    const int64_t want = std::max<int64_t>((a.row_end - a.row_begin + rows_per_block - 1) / rows_per_block, a.num_heavy);
    if (want == 0) return 0;
    const int chunk = a.kl * a.cpl;
    const int gy = (a.k + chunk - 1) / chunk;
    int rc = 0;
    if (a.dtype == GRG_HIP_F64) {
      switch (a.cpl) {
        case 1: rc = launch_stream_typed<double, 1>(a, want, gy, g_shift, kl_shift, stream); break;
        case 2: rc = launch_stream_typed<double, 2>(a, want, gy, g_shift, kl_shift, stream); break;
        default: rc = launch_stream_typed<double, 4>(a, want, gy, g_shift, kl_shift, stream); break;
      }
    } else {
      switch (a.cpl) {
        case 1: rc = launch_stream_typed<float, 1>(a, want, gy, g_shift, kl_shift, stream); break;
        case 2: rc = launch_stream_typed<float, 2>(a, want, gy, g_shift, kl_shift, stream); break;
        case 8: rc = launch_stream_typed<float, 8>(a, want, gy, g_shift, kl_shift, stream); break;
        default: rc = launch_stream_typed<float, 4>(a, want, gy, g_shift, kl_shift, stream); break;
      }
    }
    if (rc) return rc;
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return fail(std::string("kernel launch failed: ") + hipGetErrorString(err), 2);
    return 0;
  }
  if (a.flags & GRG_HIP_FLAG_PANEL) {
    if (a.num_panels + a.num_heavy == 0) return 0;
    if (a.num_panels > 0x7fffffffLL || a.num_heavy > 0x7fffffffLL) return fail("grid too large");
    int rc = 0;
    if (a.dtype == GRG_HIP_F64) {
      switch (a.cpl) {
        case 1: rc = launch_panel_typed<double, 1>(a, kl_shift, stream); break;
        case 2: rc = launch_panel_typed<double, 2>(a, kl_shift, stream); break;
        default: rc = launch_panel_typed<double, 4>(a, kl_shift, stream); break;
      }
    } else {
      switch (a.cpl) {
        case 1: rc = launch_panel_typed<float, 1>(a, kl_shift, stream); break;
        case 2: rc = launch_panel_typed<float, 2>(a, kl_shift, stream); break;
        case 8: rc = launch_panel_typed<float, 8>(a, kl_shift, stream); break;
        default: rc = launch_panel_typed<float, 4>(a, kl_shift, stream); break;
      }
    }
    if (rc) return rc;
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return fail(std::string("kernel launch failed: ") + hipGetErrorString(err), 2);
    return 0;
  }
  if (a.flags & GRG_HIP_FLAG_TILED) {
    gx = a.num_tiles + a.num_heavy;
    if (gx == 0) return 0;
    if (gx > 0x7fffffffLL) return fail("grid too large");
    const int chunk = a.kl * a.cpl;
    dim3 grid((unsigned)gx, (unsigned)((a.k + chunk - 1) / chunk));
    int rc = 0;
    if (a.dtype == GRG_HIP_F64) {
      switch (a.cpl) {
        case 1: rc = launch_tiled_typed<double, 1>(a, grid, g_shift, kl_shift, stream); break;
        case 2: rc = launch_tiled_typed<double, 2>(a, grid, g_shift, kl_shift, stream); break;
        default: rc = launch_tiled_typed<double, 4>(a, grid, g_shift, kl_shift, stream); break;
      }
    } else {
      switch (a.cpl) {
        case 1: rc = launch_tiled_typed<float, 1>(a, grid, g_shift, kl_shift, stream); break;
        case 2: rc = launch_tiled_typed<float, 2>(a, grid, g_shift, kl_shift, stream); break;
        case 8: rc = launch_tiled_typed<float, 8>(a, grid, g_shift, kl_shift, stream); break;
        default: rc = launch_tiled_typed<float, 4>(a, grid, g_shift, kl_shift, stream); break;
      }
    }
    if (rc) return rc;
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return fail(std::string("kernel launch failed: ") + hipGetErrorString(err), 2);
    return 0;
  }
  if (a.flags & GRG_HIP_FLAG_PARTIALS) {
    gx = a.split_items;
  } else {
    const int64_t rows_per_block = (int64_t)kWavesPerBlock * (64 / G) * (a.rows_per_group ? a.rows_per_group : 1);
    n_light = (a.row_end - a.row_begin + rows_per_block - 1) / rows_per_block;
    gx = n_light + a.num_heavy;
  }
  if (gx == 0) return 0;
  if (gx > 0x7fffffffLL) return fail("grid too large");
  const int chunk = a.kl * a.cpl;
  dim3 grid((unsigned)gx, (unsigned)((a.k + chunk - 1) / chunk));
  if (a.dtype == GRG_HIP_F64) {
    switch (a.cpl) {
      case 1: launch_unrolled<double, 1>(a, grid, n_light, g_shift, kl_shift, stream); break;
      case 2: launch_unrolled<double, 2>(a, grid, n_light, g_shift, kl_shift, stream); break;
      default: launch_unrolled<double, 4>(a, grid, n_light, g_shift, kl_shift, stream); break;
    }
  } else {
    switch (a.cpl) {
      case 1: launch_unrolled<float, 1>(a, grid, n_light, g_shift, kl_shift, stream); break;
      case 8: launch_unrolled<float, 8>(a, grid, n_light, g_shift, kl_shift, stream); break;
      case 2: launch_unrolled<float, 2>(a, grid, n_light, g_shift, kl_shift, stream); break;
      default: launch_unrolled<float, 4>(a, grid, n_light, g_shift, kl_shift, stream); break;
    }
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(std::string("kernel launch failed: ") + hipGetErrorString(err), 2);
  return 0;
}

// Grid of one row-gather launch (as launch_one computes it).
int launch_grid(const GrgHipGatherArgs& a, int64_t* gx, int64_t* gy, int64_t* n_light, int* g_shift, int* kl_shift) {
  const int G = a.kl * a.es;
  *g_shift = ilog2(G);
  *kl_shift = ilog2(a.kl);
  *n_light = 0;
  if (a.flags & GRG_HIP_FLAG_PARTIALS) {
    *gx = a.split_items;
  } else {
    const int64_t rows_per_block = (int64_t)kWavesPerBlock * (64 / G);  // one row per lane group in job kernels
    *n_light = (a.row_end - a.row_begin + rows_per_block - 1) / rows_per_block;
    *gx = *n_light + a.num_heavy;
  }
  const int chunk = a.kl * a.cpl;
  *gy = (a.k + chunk - 1) / chunk;
  return 0;
}

// Split-row partials of grg_hip_gather_batch_static: one array per GPU, held for
// the life of the process. Its host-side name is only a handle: host code reaches
// the device copy through the address hipGetSymbolAddress returns for the device
// the stream belongs to (static_scratch below), never through the name itself.
__device__ __align__(256) double scratch[GRG_HIP_STATIC_SCRATCH_DOUBLES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the stream's GPU's copy of `scratch`, looked up once per
// device with that device current.
int static_scratch(hipStream_t stream, double** out) {
  int dev = 0;
  hipError_t err = hipStreamGetDevice(stream, &dev);
  if (err != hipSuccess) return fail(std::string("hipStreamGetDevice: ") + hipGetErrorString(err), 2);
  if (dev < 0 || dev >= kMaxDevices) return fail("device ordinal " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    err = hipGetDevice(&cur);
    if (err != hipSuccess) return fail(std::string("hipGetDevice: ") + hipGetErrorString(err), 2);
    if (cur != dev && (err = hipSetDevice(dev)) != hipSuccess)
      return fail(std::string("hipSetDevice: ") + hipGetErrorString(err), 2);
    void* ptr = nullptr;
    err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) {
      hipError_t back = hipSetDevice(cur);
      if (err == hipSuccess && back != hipSuccess) err = back;
    }
    if (err != hipSuccess || !ptr) return fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err), 2);
    g_scratch_ptr[dev] = ptr;
  }
  *out = static_cast<double*>(g_scratch_ptr[dev]);
  return 0;
}

}  // namespace

extern "C" {

int grg_hip_jobs_build(const GrgHipGatherArgs* args, const int32_t* phase_of, int32_t njobs, int32_t nphase,
                       GrgHipJob* jobs, int32_t* phase_job_off, int64_t* phase_blocks) {
  g_last_error.clear();
  if (njobs < 0 || nphase < 0 || (njobs > 0 && (!args || !phase_of || !jobs)) || !phase_job_off || (nphase > 0 && !phase_blocks))
    return fail("bad job table arguments");
  for (int32_t p = 0; p <= nphase; ++p) phase_job_off[p] = 0;
  for (int32_t p = 0; p < nphase; ++p) phase_blocks[p] = 0;
  int32_t cur = -1;
  for (int32_t i = 0; i < njobs; ++i) {
    const GrgHipGatherArgs& a = args[i];
    if (int rc = validate(a)) {
      g_last_error = "job " + std::to_string(i) + ": " + g_last_error;
      return rc;
    }
    if (a.flags & (GRG_HIP_FLAG_TILED | GRG_HIP_FLAG_PANEL | GRG_HIP_FLAG_STREAM | GRG_HIP_FLAG_GROUPS))
      return fail("job " + std::to_string(i) + ": tiled, panel, streamed and row-group launches are not run from job tables");
    if (a.dtype != args[0].dtype) return fail("jobs of one table share one value type");
    const int32_t p = phase_of[i];
    if (p < 0 || p >= nphase || p < cur) return fail("phase_of must be non-decreasing and in [0, nphase)");
    while (cur < p) phase_job_off[++cur] = i;
    GrgHipJob& J = jobs[i];
    J.args = a;
    int g_shift = 0, kl_shift = 0;
    launch_grid(a, &J.gx, &J.gy, &J.n_light, &g_shift, &kl_shift);
    J.g_shift = g_shift;
    J.kl_shift = kl_shift;
    J.block_begin = phase_blocks[p];
    phase_blocks[p] += J.gx * J.gy;
    if (phase_blocks[p] > 0x7fffffffLL) return fail("phase " + std::to_string(p) + " has too many workgroups");
  }
  while (cur < nphase) phase_job_off[++cur] = njobs;
  return 0;
}

int grg_hip_jobs_max_grid(int32_t dtype, void* stream, int32_t* max_workgroups) {
  g_last_error.clear();
  if (!max_workgroups) return fail("max_workgroups is null");
  int dev = 0;
  hipError_t err = hipStreamGetDevice(static_cast<hipStream_t>(stream), &dev);
  if (err != hipSuccess) return fail(std::string("hipStreamGetDevice: ") + hipGetErrorString(err), 2);
  int cus = 0, per_cu = 0;
  err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (err != hipSuccess) return fail(std::string("hipDeviceGetAttribute: ") + hipGetErrorString(err), 2);
  if (dtype == GRG_HIP_F64)
    err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&jobs_kernel<double>), kBlock, 0);
  else
    err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&jobs_kernel<float>), kBlock, 0);
  if (err != hipSuccess) return fail(std::string("hipOccupancyMaxActiveBlocksPerMultiprocessor: ") + hipGetErrorString(err), 2);
  *max_workgroups = cus * per_cu;
  return 0;
}

int grg_hip_jobs_run(const GrgHipJob* jobs, const int32_t* phase_job_off, const int64_t* phase_blocks,
                     const int64_t* phase_blocks_host, int32_t dtype, int32_t nphase, int32_t persistent,
                     int32_t grid, uint32_t* barrier, int32_t* status, void* stream) {
  g_last_error.clear();
  if (dtype != GRG_HIP_F64 && dtype != GRG_HIP_F32C) return fail("unknown dtype code " + std::to_string(dtype));
  if (nphase < 0 || (nphase > 0 && (!jobs || !phase_job_off || !phase_blocks || !phase_blocks_host)))
    return fail("bad job table arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (persistent) {
    if (!barrier || !status) return fail("a persistent run needs the barrier counter and the status word");
    int32_t max_grid = 0;
    if (int rc = grg_hip_jobs_max_grid(dtype, stream, &max_grid)) return rc;
    if (grid < 1 || grid > max_grid)
      return fail("persistent grid " + std::to_string(grid) + " outside [1, " + std::to_string(max_grid) + "]");
    if (nphase == 0) return 0;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(256), 0, s, barrier);
    if (dtype == GRG_HIP_F64)
      hipLaunchKernelGGL(jobs_kernel<double>, dim3(grid), dim3(kBlock), 0, s, jobs, phase_job_off, phase_blocks, 0, nphase, barrier, status);
    else
      hipLaunchKernelGGL(jobs_kernel<float>, dim3(grid), dim3(kBlock), 0, s, jobs, phase_job_off, phase_blocks, 0, nphase, barrier, status);
  } else {
    for (int32_t p = 0; p < nphase; ++p) {
      const int64_t nb = phase_blocks_host[p];
      if (nb <= 0) continue;
      if (nb > 0x7fffffffLL) return fail("phase has too many workgroups");
      if (dtype == GRG_HIP_F64)
        hipLaunchKernelGGL(jobs_kernel<double>, dim3((unsigned)nb), dim3(kBlock), 0, s, jobs, phase_job_off, phase_blocks, p, p + 1,
                           (uint32_t*)nullptr, (int32_t*)nullptr);
      else
        hipLaunchKernelGGL(jobs_kernel<float>, dim3((unsigned)nb), dim3(kBlock), 0, s, jobs, phase_job_off, phase_blocks, p, p + 1,
                           (uint32_t*)nullptr, (int32_t*)nullptr);
    }
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(std::string("kernel launch failed: ") + hipGetErrorString(err), 2);
  return 0;
}


int grg_hip_gather(const GrgHipGatherArgs* args, void* stream) {
  if (!args) return fail("args is null");
  g_last_error.clear();
  return launch_one(*args, static_cast<hipStream_t>(stream));
}

int grg_hip_gather_batch(const GrgHipGatherArgs* args, int32_t n, void* stream) {
  if (n < 0 || (n > 0 && !args)) return fail("bad batch");
  g_last_error.clear();
  for (int32_t i = 0; i < n; ++i) {
    if (int rc = launch_one(args[i], static_cast<hipStream_t>(stream))) {
      g_last_error = "launch " + std::to_string(i) + ": " + g_last_error;
      return rc;
    }
  }
  return 0;
}

int grg_hip_gather_batch_static(const GrgHipGatherArgs* args, int32_t n, void* stream) {
  if (n < 0 || (n > 0 && !args)) return fail("bad batch");
  g_last_error.clear();
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* part = nullptr;
  if (int rc = static_scratch(s, &part)) return rc;
  for (int32_t i = 0; i < n; ++i) {
    GrgHipGatherArgs a = args[i];
    if (a.split_ptr || (a.flags & GRG_HIP_FLAG_PARTIALS)) {
      // the launch's split-row partials (split_items rows of k doubles) go to the static array
      if (a.split_items < 0 || a.k < 1 || a.split_items > GRG_HIP_STATIC_SCRATCH_DOUBLES / a.k)
        return fail("launch " + std::to_string(i) + ": split-row partials need " + std::to_string(a.split_items) + " x " +
                    std::to_string(a.k) + " doubles, more than the static scratch holds (" +
                    std::to_string((long long)GRG_HIP_STATIC_SCRATCH_DOUBLES) + " doubles)");
      a.partials = part;
    }
    if (int rc = launch_one(a, s)) {
      g_last_error = "launch " + std::to_string(i) + ": " + g_last_error;
      return rc;
    }
  }
  return 0;
}

int grg_hip_row_weighted_sum(int32_t dtype, const void* x, int64_t rows, int64_t cols, int64_t ld, const void* w, void* out,
                             double* partials, int32_t n_blocks, void* stream) {
  g_last_error.clear();
  if (dtype != GRG_HIP_F64 && dtype != GRG_HIP_F32C) return fail("unknown dtype code " + std::to_string(dtype));
  if (rows < 1 || rows > kBlock) return fail("rows must be in [1, 256]");
  if (cols < 0 || ld < rows) return fail("cols must be >= 0 and ld >= rows");
  if (n_blocks < 1) return fail("n_blocks must be >= 1");
  if (!out || !partials || (cols > 0 && (!x || !w))) return fail("null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 fin((unsigned)((rows + kBlock - 1) / kBlock));
  if (dtype == GRG_HIP_F64) {
    hipLaunchKernelGGL(row_weighted_partials_kernel<double>, dim3(n_blocks), dim3(kBlock), 0, s, static_cast<const double*>(x),
                       static_cast<const double*>(w), rows, cols, ld, partials);
    hipLaunchKernelGGL(row_weighted_finish_kernel<double>, fin, dim3(kBlock), 0, s, partials, rows, n_blocks,
                       static_cast<double*>(out));
  } else {
    hipLaunchKernelGGL(row_weighted_partials_kernel<float>, dim3(n_blocks), dim3(kBlock), 0, s, static_cast<const float*>(x),
                       static_cast<const float*>(w), rows, cols, ld, partials);
    hipLaunchKernelGGL(row_weighted_finish_kernel<float>, fin, dim3(kBlock), 0, s, partials, rows, n_blocks,
                       static_cast<float*>(out));
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(std::string("kernel launch failed: ") + hipGetErrorString(err), 2);
  return 0;
}

const char* grg_hip_last_error(void) { return g_last_error.c_str(); }

int32_t grg_hip_abi_version(void) { return GRG_HIP_ABI_VERSION; }

int grg_hip_device_count(int32_t* count) {
  if (!count) return fail("count is null");
  int n = 0;
  hipError_t err = hipGetDeviceCount(&n);
  if (err != hipSuccess) {
    *count = 0;
    return fail(std::string("hipGetDeviceCount: ") + hipGetErrorString(err), 2);
  }
  *count = n;
  return 0;
}

int grg_hip_stream_device(void* stream, int32_t* device) {
  if (!device) return fail("device is null");
  int d = -1;
  hipError_t err = hipStreamGetDevice(static_cast<hipStream_t>(stream), &d);
  if (err != hipSuccess) return fail(std::string("hipStreamGetDevice: ") + hipGetErrorString(err), 2);
  *device = d;
  return 0;
}

int grg_hip_stream_capturing(void* stream, int32_t* capturing) {
  if (!capturing) return fail("capturing is null");
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  hipError_t err = hipStreamIsCapturing(static_cast<hipStream_t>(stream), &st);
  if (err != hipSuccess) return fail(std::string("hipStreamIsCapturing: ") + hipGetErrorString(err), 2);
  *capturing = st == hipStreamCaptureStatusNone ? 0 : 1;
  return 0;
}

}  // extern "C"
