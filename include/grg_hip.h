/* MI355X (gfx950) GRG matmul kernels -- C ABI of libgrg_hip.so.
 *
 * Drop-in boundary for pygrgl_spmv.backends.hip. The reference's GPU backends
 * reach device code through these Python-side entry points, which this ABI
 * replaces:
 *   - per-block SpMM launches: pygrgl_spmv/backends/triton/kernel.py:141-206
 *     (launch_block -> csr_spmm_add_kernel / csc_spmm_add_kernel) and
 *     pygrgl_spmv/backends/cusparse/backend.py:1607-1624 (cusparseSpMM);
 *   - the seed pass: pygrgl_spmv/backends/triton/backend.py:1116-1150
 *     (_seed_state: zero, init, input / selector scatter, miss);
 *   - the write-out pass: pygrgl_spmv/backends/triton/backend.py:1152-1219
 *     (_apply_endpoint_bias, _write_output: selector gather, miss output,
 *     endpoint bias, by_individual reduce, emit_all_nodes permutation).
 * Here every one of those is a "row gather" launch: for each output row,
 *   out[row, c] = (accumulate ? out[row, c] : 0)
 *               + sum over CSR segments s of  sum_{j in ptr_s[row]..ptr_s[row+1]} src_s[idx_s[j], c]
 *               + sum over map segments m with row in [lo_m, hi_m) of src_m[map_m[row - lo_m], c]
 *               + (row_scale ? row_scale[row] : 1) * (col_vec ? col_vec[c] : 1)   (if either is set)
 * with no atomics: the down sweep reads child-major adjacency (parents of a
 * node) and the selector transposes, so every output row is written once.
 *
 * Conventions: all pointers are device pointers (or NULL where optional);
 * sizes are element counts; matrices are row-major with a leading dimension
 * (elements between rows), k columns contiguous. The caller owns all memory;
 * nothing here allocates, and no global mutable state is kept. Launches are
 * asynchronous on the given hipStream_t (0 = legacy default stream) and may
 * be captured into a hipGraph. Functions return 0 on success and a non-zero
 * status otherwise, with a message available from grg_hip_last_error()
 * (thread-local).
 */
#ifndef GRG_HIP_H
#define GRG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRG_HIP_ABI_VERSION 17

/* uint32 words of the persistent job kernel's barrier buffer (160 KiB: its
 * counters sit on separate 256-byte lines). */
#define GRG_HIP_BARRIER_WORDS 40960

enum GrgHipDtype {
  GRG_HIP_F64 = 0,      /* FP64 values, FP64 accumulation */
  GRG_HIP_F32C = 1      /* FP32 values, two-float compensated (TwoSum) accumulation per row */
};

#define GRG_HIP_MAX_SEGS 3
#define GRG_HIP_MAX_MAPS 2

/* Workgroups are dispatched round-robin over the 8 XCDs; remap them so each
 * XCD (and its L2) works on one contiguous eighth of the rows. */
#define GRG_HIP_FLAG_XCD_REMAP 2
/* Partials launch of split heavy rows: grid = split_items workgroups; item i
 * sums entries [j * split_chunk, (j + 1) * split_chunk) of the concatenated
 * segments of heavy row h = split_row[i] (j = i - split_ptr[h]) and writes
 * the FP64 row partials[i * k + c]. No row terms, nothing else written. */
#define GRG_HIP_FLAG_PARTIALS 4
/* Tiled launch: see the tile fields of GrgHipGatherArgs. */
#define GRG_HIP_FLAG_TILED 8
/* Merged-entry light path: the per-row lane groups read a row's entries over
 * all segments from eptr / ekey (one pointer pair and one round of gathers per
 * row) instead of segment by segment; heavy rows still use the segments. */
#define GRG_HIP_FLAG_MERGED 16
/* Streamed light path: a grid of at most the resident workgroups (or
 * stream_blocks, when > 0) whose waves loop over the row batches, with the row
 * pointers, entry keys and gathers of three consecutive batches in flight at
 * once (software pipeline); heavy rows follow over the same grid. Needs
 * GRG_HIP_FLAG_MERGED or a single segment; unroll 4 or 8; rows_per_group is 1. */
#define GRG_HIP_FLAG_STREAM 32
/* Panel launch: see the panel fields of GrgHipGatherArgs. */
#define GRG_HIP_FLAG_PANEL 64
/* Light path of the per-row launch, and the tiles of a tiled launch: XCD x
 * (blocks b with b % 8 == x) takes runs of 2^((flags >> GRG_HIP_XCD_RUN_SHIFT)
 * & 15) consecutive blocks, the eight XCDs' runs side by side, so every XCD's
 * L2 sees consecutive rows while all XCDs move through one window of rows
 * together (their shared sources stay in the memory-side cache). Blocks past
 * the last full window keep their order. Also orders the panels of a panel
 * launch. Not combined with GRG_HIP_FLAG_XCD_REMAP or STREAM. No reference
 * counterpart. */
#define GRG_HIP_FLAG_XCD_RUNS 128
#define GRG_HIP_XCD_RUN_SHIFT 8
/* Per-row light path with one row per lane group (rows_per_group 1; merged
 * entries or one segment): the group loads each round's entry keys or source
 * indices together (one dword per lane per kl * es entries) and hands them out
 * with lane shuffles, instead of one load per lane and gather. Same sums, bit
 * for bit. Not for tiled, panel, streamed or partials launches. No reference
 * counterpart. */
#define GRG_HIP_FLAG_GROUP_KEYS 4096
/* Row-group launch: see the group fields of GrgHipGatherArgs. FP64 values only.
 * Combines with GRG_HIP_FLAG_XCD_REMAP or GRG_HIP_FLAG_XCD_RUNS (over the
 * launch's group workgroups), not with the other forms. No reference
 * counterpart. */
#define GRG_HIP_FLAG_GROUPS 8192
/* Bytes of one staged row (kl * cpl columns) in a tiled launch, at most. */
#define GRG_HIP_TILE_CHUNK_BYTES 128
/* Staged rows of one tile, at most (uint16 LDS indices). */
#define GRG_HIP_TILE_MAX_STAGED 1024

/* sum_{j in [ptr[row - ptr_base], ptr[row - ptr_base + 1])} src[idx[j] * ld + c] */
typedef struct GrgHipSeg {
  const int32_t* ptr;
  const int32_t* idx;
  const void* src;
  int64_t ld;
  int64_t ptr_base;
} GrgHipSeg;

/* for row in [row_lo, row_hi): src[(map ? map[row - row_lo] : row - row_lo) * ld + c] */
typedef struct GrgHipMap {
  const int32_t* map;
  const void* src;
  int64_t ld;
  int64_t row_lo;
  int64_t row_hi;
} GrgHipMap;

typedef struct GrgHipGatherArgs {
  int32_t dtype;        /* GrgHipDtype */
  int32_t accumulate;   /* 1: add into the existing out rows */
  int64_t row_begin;    /* rows [row_begin, row_end) are produced */
  int64_t row_end;
  int32_t k;            /* columns */
  int32_t cpl;          /* columns per lane: 1, 2 or 4 (8 for GRG_HIP_F32C); k, every ld and every base must be multiples */
  int32_t kl;           /* lanes across one column chunk (power of two, kl * cpl columns per chunk) */
  int32_t es;           /* lanes interleaving over one row's entries (power of two, kl * es <= 64) */
  int32_t unroll;       /* gathers in flight per lane, row and round: 4, 8 or 16 (0 means 8) */
  int32_t rows_per_group; /* consecutive rows per lane group: 1, 2 or 4 (0 means 1); with unroll:
                             (4|8|16, 1), (4|8, 2) or (4, 4) */
  int32_t flags;        /* GRG_HIP_FLAG_* launch options */
  void* out;            /* out row r lives at out + (r - out_base) * ld_out */
  int64_t ld_out;
  int64_t out_base;
  int32_t nseg;
  int32_t nmap;
  GrgHipSeg seg[GRG_HIP_MAX_SEGS];
  GrgHipMap map[GRG_HIP_MAX_MAPS];
  const void* row_scale;  /* indexed by row - scale_base */
  int64_t scale_base;
  const void* col_vec;    /* indexed by column */
  /* Rows whose total segment degree exceeds heavy_threshold (when >= 0) are
   * skipped by the per-row path and produced instead by one 256-thread
   * workgroup each, taken from heavy_rows[0..num_heavy) (global row ids; the
   * caller lists exactly those rows). */
  int64_t heavy_threshold;
  const int32_t* heavy_rows;
  int64_t num_heavy;
  /* Split heavy rows (optional). When split_ptr is set, heavy row h with a
   * non-empty item range [split_ptr[h], split_ptr[h + 1]) takes the sum of the
   * partial rows of that range (written by a preceding GRG_HIP_FLAG_PARTIALS
   * launch with the same arguments) in place of its segment sums; its row
   * terms are added as usual. */
  const int32_t* split_ptr;   /* num_heavy + 1 item offsets */
  const int32_t* split_row;   /* split_items heavy indices (partials launch) */
  int64_t split_items;
  int64_t split_chunk;        /* segment entries per item */
  void* partials;             /* double, split_items x k */
  /* Optional output row scale (no reference counterpart; used by the
   * standardized-genotype operator): every produced row is multiplied by
   * out_scale[row - out_scale_base] after all its terms are summed. */
  const void* out_scale;
  int64_t out_scale_base;
  /* Tiled launch (GRG_HIP_FLAG_TILED; no reference counterpart: an LDS-staged
   * form of the same row sums). Rows [row_begin, row_end) are cut into
   * num_tiles tiles of consecutive rows, tile t = rows [tile_row[t],
   * tile_row[t + 1]). Its distinct sources src_list[tile_src[t] ..
   * tile_src[t + 1]) (each (segment << 30) | source row) are copied once into
   * LDS together with its row pointers and entry indices, and row r sums the
   * staged rows lidx[j] for j in [eptr[r], eptr[r + 1]) (the row's entries over
   * all segments, in segment order; eptr is indexed by global row; lidx is
   * 16-byte aligned and readable 8 entries past the last one). A tile with no
   * staged sources sums its rows straight from the segments, one workgroup per
   * row. Heavy rows (as above) are skipped by the tiles and produced by the
   * heavy workgroups after them: grid = num_tiles + num_heavy. Over the staged
   * tiles of the launch, tile_lds_rows bounds the staged rows, tile_max_rows the
   * rows and tile_max_entries the entries of a tile (they size the LDS); kl * cpl
   * columns of the value type must fit in 128 bytes. */
  const int32_t* tile_row;
  const int32_t* tile_src;
  const uint32_t* src_list;
  const int32_t* eptr;
  const uint16_t* lidx;
  int64_t num_tiles;
  int64_t tile_lds_rows;
  int64_t tile_max_rows;
  int64_t tile_max_entries;
  /* Merged entry keys (GRG_HIP_FLAG_MERGED; no reference counterpart): row r's
   * entries over all segments, in segment order, are ekey[eptr[r] ..
   * eptr[r + 1]), each (segment << 30) | source row. */
  const uint32_t* ekey;
  /* Workgroups of a streamed launch (GRG_HIP_FLAG_STREAM) along the rows, at
   * most; 0 means as many as are resident at once. */
  int64_t stream_blocks;
  /* Panel launch (GRG_HIP_FLAG_PANEL; no reference counterpart: an LDS-staged
   * form for rows that share many sources). Rows [row_begin, row_end) are cut
   * into num_panels panels of panel_rows consecutive rows (the last one cut at
   * row_end), one workgroup of 1024 threads each: lane groups of kl lanes (es
   * is 1) own rows_per_group consecutive rows each (1, 2 or 4, at most kl),
   * so panel_rows = 1024 / kl * rows_per_group, and keep those rows' sums in
   * registers. A panel's distinct sources are staged in passes: pass g (panel
   * p's passes are [pan_pass[p], pan_pass[p + 1])) copies the source rows
   * pan_src[pass_src[g] .. pass_src[g + 1]) (each (segment << 30) | source
   * row; at most panel_staged of them; one column chunk of kl * cpl columns)
   * into LDS with its entry lists, and row i of the panel adds the staged rows
   * pan_lidx[pass_ent[g] + j] for j in [pass_rowoff[g * (panel_rows + 1) + i],
   * pass_rowoff[g * (panel_rows + 1) + i + 1]) (the first offset of a pass is
   * 0; every offset and pass_ent entry is a multiple of 4; position 255 is
   * padding and adds a zero row). unroll 4 keeps two passes in flight, unroll 8 three. Heavy rows (as
   * above) hold no panel entries and are produced after the panels by the
   * per-row kernel's heavy workgroups (a second launch on the stream).
   * Combines with GRG_HIP_FLAG_XCD_REMAP or GRG_HIP_FLAG_XCD_RUNS (over the
   * panels). panel_staged is in [1, 255] with panel_staged * kl <= 2048;
   * panel_max_entries bounds the entries of one pass (at most 8160) and
   * panel_max_passes the passes of one panel; pan_lidx is 16-byte aligned and
   * readable 16 bytes past its last entry. */
  const int32_t* pan_pass;
  const int32_t* pass_src;
  const int32_t* pass_ent;
  const uint16_t* pass_rowoff;
  const uint32_t* pan_src;
  const uint8_t* pan_lidx;
  int64_t num_panels;
  int32_t panel_rows;
  int32_t panel_staged;
  int32_t panel_max_entries;
  int32_t panel_max_passes; /* most passes of one panel (sizes the LDS) */
  /* Row-group launch (GRG_HIP_FLAG_GROUPS; no reference counterpart: rows that
   * share sources gathered once per group). Rows [row_begin, row_end) are cut
   * into num_groups groups of 16 consecutive rows (group g: rows row_begin +
   * 16 g + [0, 16), cut at row_end), one workgroup each. Group g's entries are
   * [grp_off[g], grp_off[g + 1]) (multiples of 4): grp_key[e] is a source
   * ((segment << 30) | source row, 0xffffffff for padding) and bit i of
   * grp_mask[e] says that row 16 g + i sums it. A row's value is the sum of the
   * sources whose mask has its bit, plus its row terms; the group stores the
   * rows whose bit is set in grp_store[g]. The sums run four entries at a time
   * as one FP64 16 x 16 x 4 MFMA per 16-column chunk (A = the 0/1 mask bits,
   * B = the four source rows), so every source row is read once per group.
   * Heavy rows (as above) hold no group entries and are produced by the heavy
   * workgroups after the groups: grid = num_groups + num_heavy by ceil(k / 16)
   * column chunks; kl * cpl must be 16 (the heavy rows' column chunk). The
   * workgroup's four waves take interleaved rounds of four entries (unroll
   * rounds in flight per wave) and add their tiles in wave order. */
  const int64_t* grp_off;
  const uint16_t* grp_store;
  const uint32_t* grp_key;
  const uint16_t* grp_mask;
  int64_t num_groups;
} GrgHipGatherArgs;

/* One row-gather launch (see above). Replaces one level of the reference's
 * per-block launch loop (triton/backend.py:1030-1046) with a single fused
 * launch over all resident source levels, and the seed / write-out passes. */
int grg_hip_gather(const GrgHipGatherArgs* args, void* stream);

/* n row-gather launches in order on one stream (a whole prepared sweep:
 * seed-fused levels then write-out, with the partials launch of a level's
 * split rows right before the level), from one host call. */
int grg_hip_gather_batch(const GrgHipGatherArgs* args, int32_t n, void* stream);

/* Doubles of the static split-row scratch of grg_hip_gather_batch_static (32 MiB
 * per GPU). */
#define GRG_HIP_STATIC_SCRATCH_DOUBLES (1LL << 22)

/* grg_hip_gather_batch without a caller workspace: the split-row partials of
 * every launch (args[i].partials is ignored) go to a file-scope device array of
 * GRG_HIP_STATIC_SCRATCH_DOUBLES doubles, one per GPU, kept until the process
 * ends; nothing is allocated on this path (no hipMalloc, hipFree or allocating
 * library call). A launch whose partials need more than the array holds
 * (split_items * k doubles) is rejected with an error before it runs. The
 * partials launch of a level writes every partial row its level launch reads,
 * so the array needs no clearing; calls that overlap on one GPU (other streams,
 * other threads) share the array and race. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid allocation). */
int grg_hip_gather_batch_static(const GrgHipGatherArgs* args, int32_t n, void* stream);

/* ---- Job tables: many row-gather launches in one kernel -------------------
 *
 * No reference counterpart. The reference runs runtime.grgs one GRG after the
 * other, one launch per block (triton/backend.py:970-1049); here the launches
 * of several sweeps (one per GRG, each with its own k and buffers) are laid
 * out as a table of jobs in phases: phase p holds the p-th launch of every
 * sweep, so one kernel runs a level of all GRGs at once. Jobs run the light
 * path with 8 gathers in flight and one row per lane group (the unroll and
 * rows_per_group of their descriptors are ignored); tiled descriptors are not
 * accepted. */
typedef struct GrgHipJob {
  GrgHipGatherArgs args;
  int64_t block_begin;  /* first virtual workgroup of the job within its phase */
  int64_t gx;           /* workgroups along the rows: light blocks + heavy rows, or partial items */
  int64_t gy;           /* column chunks */
  int64_t n_light;      /* light-path workgroups */
  int32_t g_shift;
  int32_t kl_shift;
} GrgHipJob;

/* Validate args[0..njobs) (one value type; phase_of non-decreasing in
 * [0, nphase)) and fill the host arrays jobs[njobs], phase_job_off[nphase + 1]
 * (phase p = jobs [phase_job_off[p], phase_job_off[p + 1])) and
 * phase_blocks[nphase] (virtual workgroups of phase p). The caller copies the
 * three arrays to device memory for grg_hip_jobs_run. */
int grg_hip_jobs_build(const GrgHipGatherArgs* args, const int32_t* phase_of, int32_t njobs, int32_t nphase,
                       GrgHipJob* jobs, int32_t* phase_job_off, int64_t* phase_blocks);

/* Workgroups of the persistent job kernel that are resident at once on the
 * stream's device (occupancy x compute units): the largest persistent grid. */
int grg_hip_jobs_max_grid(int32_t dtype, void* stream, int32_t* max_workgroups);

/* Run all phases of a job table (device copies of the arrays filled by
 * grg_hip_jobs_build; phase_blocks_host is the host copy of phase_blocks).
 * persistent = 0: one launch per phase, in stream order. persistent = 1: ONE
 * launch of `grid` workgroups (1 <= grid <= grg_hip_jobs_max_grid) that loop
 * over each phase's workgroups and meet at a grid-wide barrier between
 * phases; `barrier` (GRG_HIP_BARRIER_WORDS uint32 of device memory: the grid
 * and per-XCD counters of the two-tier barrier) is reset by a kernel on the
 * stream before the launch, and `status` (one int32 of device memory, set
 * to 0 by the caller) becomes 1 if a barrier wait ran past its bound (the
 * results are then invalid). Both are capturable into a hipGraph. */
int grg_hip_jobs_run(const GrgHipJob* jobs, const int32_t* phase_job_off, const int64_t* phase_blocks,
                     const int64_t* phase_blocks_host, int32_t dtype, int32_t nphase, int32_t persistent,
                     int32_t grid, uint32_t* barrier, int32_t* status, void* stream);

/* Thread-local description of the last failure ("" if none). */
/* Row-weighted sums for the standardized X^T X operator: out[j] = sum_m
 * x[j + m * ld] * w[m] for j < rows (at most 256), m < cols, with x, w and out
 * in the value type of dtype (FP32 for GRG_HIP_F32C) and FP64 sums in a fixed
 * order (deterministic). partials holds n_blocks * rows doubles of scratch
 * owned by the caller. Replaces the reference's CuPy RawKernel
 * row_weighted_sum (scripts/pca_bench/__main__.py:32-61). */
int grg_hip_row_weighted_sum(int32_t dtype, const void* x, int64_t rows, int64_t cols, int64_t ld, const void* w, void* out,
                             double* partials, int32_t n_blocks, void* stream);

const char* grg_hip_last_error(void);

/* GRG_HIP_ABI_VERSION of the loaded library. */
int32_t grg_hip_abi_version(void);

/* hipGetDeviceCount, for device-ordinal validation without torch
 * (replaces parse_cuda_device's cudaGetDeviceCount, backends/_cuda_stream.py:121-124). */
int grg_hip_device_count(int32_t* count);

/* hipStreamGetDevice (replaces backends/_cuda_stream.py:127-137). */
int grg_hip_stream_device(void* stream, int32_t* device);

/* *capturing = 1 while the stream is being captured into a hipGraph, else 0
 * (hipStreamIsCapturing). No reference counterpart: the runtime uses it to
 * refuse capturing sweeps that stream blocks through the pinned ring. */
int grg_hip_stream_capturing(void* stream, int32_t* capturing);

#ifdef __cplusplus
}
#endif

#endif /* GRG_HIP_H */
