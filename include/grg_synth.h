/* Seeded synthetic GRG generator (host C ABI, libgrg_synth.so).
 *
 * The reference builds .grg_spmv artifacts from real GRGs through the
 * third-party pygrgl module (pygrgl_spmv/grg/compile.py:385-460 compile_grg);
 * this generator produces the same compiled layout (stable height order,
 * per-(parent level, child level) binary CSR blocks with level-local columns,
 * mutation / missingness selectors, diploid coalescence counts) from a seeded
 * coalescent tree-sequence model, for workloads of any size.
 */
#ifndef GRG_SYNTH_H
#define GRG_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct GrgSynthParams {
  int64_t num_samples;     /* haploid samples S (even when ploidy == 2) */
  int64_t num_mutations;   /* mutations M */
  uint64_t seed;
  int32_t ploidy;          /* 1, or 2 (individual i = samples 2i, 2i+1; enables coalescence counts) */
  int32_t num_threads;     /* <= 0: all hardware threads */
  int64_t num_trees;       /* local trees along the genome */
  int64_t num_segments;    /* independent genome segments (parallelism); >= 1 */
  double keep_fraction;    /* probability that an unmutated clade is kept as a node */
  double window_fraction;  /* fraction of gaps re-shuffled between adjacent local trees */
  double missing_fraction; /* > 0: build a missingness pool over this fraction of samples */
  int64_t missing_pool;    /* number of missingness nodes in the pool */
} GrgSynthParams;

/* All arrays are malloc'ed by the library and released by grg_synth_free.
 * Node ids are internal (stable height order). Blocks are stored row-major
 * over (d, s) as index b = d * num_levels + s; only s < d can be non-empty. */
typedef struct GrgSynthResult {
  int64_t num_samples, num_mutations, num_nodes, num_edges, num_levels, ploidy;
  int64_t* level_offsets;       /* num_levels + 1 */
  int64_t* node_perm;           /* num_nodes: internal -> original node id */
  int64_t* block_edge_offset;   /* num_levels^2 + 1: start of block b in block_indices */
  int64_t* block_indptr_offset; /* num_levels^2 + 1: start of block b's indptr (rows(d) + 1 entries when s < d) */
  int32_t* block_indices;       /* num_edges: level-local child columns, sorted within each row */
  int32_t* block_indptr;        /* concatenated per-block indptr, relative to the block start */
  int64_t* mutation_node;       /* num_mutations: internal node id */
  int64_t* miss_node;           /* num_mutations: internal node id or -1 */
  int64_t* coalescence_counts;  /* num_nodes (internal order) or NULL when ploidy == 1 */
  double* mutation_positions;   /* num_mutations */
  double* mutation_times;       /* num_mutations */
} GrgSynthResult;

/* 0 on success; non-zero with grg_synth_last_error() set otherwise. */
int grg_synth_generate(const GrgSynthParams* params, GrgSynthResult* out);
void grg_synth_free(GrgSynthResult* result);
const char* grg_synth_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GRG_SYNTH_H */
