/* Host-side runtime preprocessing for the HIP backend (libgrg_host.so).
 *
 * No reference counterpart: the reference keeps the artifact's node order on
 * the device. HipRuntime renumbers nodes inside each level, privately on the
 * device, so that nodes sharing descendants sit next to each other and the
 * level gathers reuse cache lines; results are unchanged (the op surface maps
 * inputs and outputs through the permutation).
 */
#ifndef GRG_HOST_H
#define GRG_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One level block A[dst][src] in the artifact's CSR (level-local rows/columns). */
typedef struct GrgHostBlock {
  int32_t dst_level;
  int32_t src_level;
  int64_t nnz;
  const int64_t* indptr;   /* rows(dst) + 1 */
  const int32_t* indices;  /* nnz, level-local columns in src */
} GrgHostBlock;

/* Locality order: new_pos[old] = new internal id, a permutation inside every
 * level (level offsets are unchanged). Samples 0..num_samples-1 stay in
 * [0, num_samples) ordered by first visit of a depth-first walk from the top
 * levels; every other node is keyed by the mean key of its children (bottom
 * up), childless non-sample nodes go last in their level. Returns 0 on
 * success. */
int grg_host_locality_order(int64_t num_nodes, int64_t num_samples, int32_t num_levels, const int64_t* level_offsets,
                            const GrgHostBlock* blocks, int32_t num_blocks, int64_t* new_pos);

const char* grg_host_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GRG_HOST_H */
