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
 * up), childless non-sample nodes go last in their level. With
 * degree_window > 1, every level above 0 whose mean child count is at least
 * degree_min_mean is then cut into runs of degree_window consecutive nodes,
 * each run ordered by child count (stable), so that the rows a wave sums in
 * lockstep have similar lengths. Returns 0 on success. */
int grg_host_locality_order(int64_t num_nodes, int64_t num_samples, int32_t num_levels, const int64_t* level_offsets,
                            const GrgHostBlock* blocks, int32_t num_blocks, int64_t degree_window, double degree_min_mean,
                            int64_t* new_pos);

/* Min-child order: refines an order new_pos (a permutation inside every level,
 * as grg_host_locality_order returns) in place, level by level from level 1 up.
 * Every level above 0 with at least min_rows nodes and a mean child count of at
 * least min_mean is cut into windows of `window` consecutive positions; inside a
 * window the nodes are ordered (stably) by the position of their first child in
 * the refined order, so nodes sharing that child sit side by side and the rows a
 * workgroup sums together gather more of the same source rows. With
 * degree_window > 1, aligned runs of degree_window positions inside each window
 * are then ordered (stably) by child count, which keeps the set of rows of each
 * run. No reference counterpart. Returns 0 on success. */
int grg_host_cluster_order(int64_t num_nodes, int32_t num_levels, const int64_t* level_offsets, const GrgHostBlock* blocks,
                           int32_t num_blocks, int64_t window, int64_t degree_window, double min_mean, int64_t min_rows,
                           int64_t* new_pos);

/* Tiles of consecutive rows for the tiled launch of libgrg_hip.so (see the
 * tile fields of GrgHipGatherArgs in grg_hip.h), for every level of one sweep
 * configuration at once. No reference counterpart.
 *
 * Rows are the num_rows device-order nodes; level l is rows
 * [level_offsets[l], level_offsets[l + 1]). Row r gathers the entries of
 * nseg segments, segment s holding seg_idx[s][seg_ptr[s][r] .. seg_ptr[s][r + 1])
 * (source rows in [0, seg_src_rows[s]), < 2^30); eptr[r] (num_rows + 1
 * offsets, computed by the caller) is where row r's entries start in the
 * concatenation over segments.
 *
 * Per level, rows are taken in order. A row of degree > heavy_threshold[l]
 * (when >= 0; it is produced by a heavy workgroup), > max_staged or
 * > max_entries gets a tile of its own with nothing staged ("direct" tile).
 * Every other row joins the current tile unless that would raise the tile's
 * distinct sources above max_staged, its rows above max_rows or its entries
 * above max_entries. A tile never spans two levels; levels are tiled even
 * when their rows have no entries.
 *
 * Outputs (capacities in brackets):
 *   tile_row  [num_rows + num_levels + 1]: tile t = rows [tile_row[t], tile_row[t + 1]);
 *   tile_src  [num_rows + num_levels + 1]: tile t stages src_list[tile_src[t] .. tile_src[t + 1]);
 *   src_list  [eptr[num_rows]]: (segment << 30) | source row, first-use order within a tile;
 *   lidx      [eptr[num_rows]]: staged position of each entry (written for rows of staged tiles);
 *   level_tile_off [num_levels + 1]: level l's tiles are [level_tile_off[l], level_tile_off[l + 1]);
 *   level_max [3 * num_levels]: over the staged tiles of level l, the most staged
 *     sources (level_max[l]), rows (level_max[L + l]) and entries (level_max[2L + l]).
 * Returns 0 on success. */
int grg_host_build_tiles(int64_t num_rows, int32_t num_levels, const int64_t* level_offsets, int32_t nseg,
                         const int32_t* const* seg_ptr, const int32_t* const* seg_idx, const int64_t* seg_src_rows,
                         const int32_t* eptr, const int64_t* heavy_threshold, int32_t max_staged, int32_t max_rows,
                         int32_t max_entries, int32_t* tile_row, int32_t* tile_src, uint32_t* src_list, uint16_t* lidx,
                         int64_t* level_tile_off, int32_t* level_max);

/* Merged entry keys for the merged-entry launches of libgrg_hip.so: row r's
 * entries over all nseg segments, in segment order, as (segment << 30) |
 * source row, written at ekey[eptr[r] ..] (eptr as for grg_host_build_tiles,
 * starting at 0, with ekey holding eptr[num_rows] keys; source rows below 2^30). No reference counterpart. Returns 0 on success. */
int grg_host_merged_keys(int64_t num_rows, int32_t nseg, const int32_t* const* seg_ptr, const int32_t* const* seg_idx,
                         const int32_t* eptr, uint32_t* ekey);

/* Panels of the panel launch of libgrg_hip.so (GRG_HIP_FLAG_PANEL; see the
 * panel fields of GrgHipGatherArgs in grg_hip.h) for rows [row_begin, row_end)
 * of one level. No reference counterpart.
 *
 * Row r's entries are keys[eptr[r] .. eptr[r + 1]) (merged entry keys
 * (segment << 30) | source row, or plain source rows of one segment), each key
 * at most once per row. Panel p is rows row_begin + p * panel_rows + [0,
 * panel_rows), cut at row_end. Rows of degree > heavy_threshold (when >= 0)
 * contribute no entries (their heavy workgroups produce them). The panel's
 * distinct keys, ascending, are staged in passes of consecutive keys. Pass q
 * lists for every row of the panel the entries whose key is in the pass, as
 * 8-bit positions in the pass, ascending, padded to a multiple of four with
 * position 255 (the kernel keeps staged row 255 at zero). A row's entries are
 * therefore summed in ascending key order. A pass closes before its (max_staged
 * + 1)-th key or before its padded entries would exceed max_pass_entries.
 * panel_rows is in [1, 1024], max_staged in [1, 255], max_pass_entries in
 * [4 * panel_rows, 65535]; threads > 1 builds panels on that many host threads.
 *
 * Returns an opaque handle (NULL on error, see grg_host_last_error), read with
 * grg_host_panels_sizes / grg_host_panels_copy and released with
 * grg_host_panels_free. sizes[6] = {panels, passes, sources, entries, most
 * entries of one pass, most passes of one panel}. Copied arrays (sizes in
 * brackets):
 *   pan_pass    [panels + 1]: panel p's passes are [pan_pass[p], pan_pass[p + 1]);
 *   pass_src    [passes + 1]: pass g stages src_list[pass_src[g] .. pass_src[g + 1]);
 *   pass_ent    [passes + 1]: pass g's entries are lidx[pass_ent[g] .. pass_ent[g + 1]);
 *   pass_rowoff [passes * (panel_rows + 1)]: row i of pass g holds its pass's entries
 *               [pass_rowoff[g * (panel_rows + 1) + i], ... + i + 1]);
 *   src_list    [sources]: staged keys;
 *   lidx        [entries]: position of each entry's key in its pass. */
void* grg_host_build_panels(int64_t row_begin, int64_t row_end, const int32_t* eptr, const uint32_t* keys, int32_t panel_rows,
                            int32_t max_staged, int32_t max_pass_entries, int64_t heavy_threshold, int32_t threads);
int grg_host_panels_sizes(const void* panels, int64_t* sizes);
int grg_host_panels_copy(const void* panels, int32_t* pan_pass, int32_t* pass_src, int32_t* pass_ent, uint16_t* pass_rowoff,
                         uint32_t* src_list, uint8_t* lidx);
void grg_host_panels_free(void* panels);

/* Row groups of the row-group launch of libgrg_hip.so (GRG_HIP_FLAG_GROUPS; see
 * the group fields of GrgHipGatherArgs in grg_hip.h) for rows [row_begin,
 * row_end) of one level. No reference counterpart.
 *
 * Row r's entries are keys[eptr[r] .. eptr[r + 1]) (merged entry keys
 * (segment << 30) | source row, or plain source rows of one segment), each key
 * at most once per row (a repeated key is an error) and none equal to
 * 0xffffffff. Group g is rows row_begin + 16 g + [0, 16), cut at row_end. Rows
 * of degree > heavy_threshold (when >= 0) are left to their heavy workgroups:
 * they contribute no entries and are not stored by the group. A group lists
 * the distinct keys of its other rows, ascending, each with a 16-bit mask of
 * the rows (bit i = row 16 g + i) that list it, padded with (0xffffffff, 0)
 * entries to a multiple of 4. threads > 1 builds on that many host threads.
 *
 * Returns an opaque handle (NULL on error, see grg_host_last_error), read with
 * grg_host_groups_sizes / grg_host_groups_copy and released with
 * grg_host_groups_free. sizes[2] = {groups, entries}. Copied arrays:
 *   off   [groups + 1]: group g's entries are [off[g], off[g + 1]) (multiples of 4);
 *   store [groups]: bit i set when row 16 g + i is produced by the group;
 *   key   [entries], mask [entries]: the group entries. */
void* grg_host_build_groups(int64_t row_begin, int64_t row_end, const int32_t* eptr, const uint32_t* keys,
                            int64_t heavy_threshold, int32_t threads);
int grg_host_groups_sizes(const void* groups, int64_t* sizes);
int grg_host_groups_copy(const void* groups, int64_t* off, uint16_t* store, uint32_t* key, uint16_t* mask);
void grg_host_groups_free(void* groups);

const char* grg_host_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GRG_HOST_H */
