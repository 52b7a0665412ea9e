/*
 * mettle_hip.h — C ABI of the MI355X (gfx950) execution backend for mettle.
 *
 * This header is the drop-in boundary.  In the reference, the planner
 * (src/plan/...) reaches the GPU only through the free functions of
 * src/metal/mod.rs:96-609; every declaration below names the reference
 * function(s) it replaces.  All signatures use plain C types: device
 * pointers are `const void*` / `void*`, sizes are uint64_t, and streams are
 * opaque `mtl_stream*` handles (NULL = the library's default stream).
 *
 * Conventions
 *   - Every entry point returns an int32_t status (MTL_OK or an error code);
 *     the message of the last error on the calling thread is read with
 *     mtl_last_error().  Status codes mirror MettleError (src/error.rs:5-11)
 *     plus MTL_UNSUPPORTED, which tells the host to use DuckDB's CPU path.
 *   - Row ids and selection vectors are uint32_t (limit 2^32-1 rows per
 *     table, the same ceiling the reference has: helpers.rs:215-218).
 *   - A selection vector lives on the device together with a device-resident
 *     uint32_t count (the reference's GpuVector, gpu_buffer.rs:43-52).  A NULL
 *     selection means "all rows 0..n-1".  Kernels read the count on the
 *     device, so no host synchronisation is needed between operators.
 *   - Selection vectors produced here are stable: ascending row ids.  (The
 *     reference's atomic-counter compaction is unordered:
 *     kernels.metal:684-696.)
 *   - Integer arithmetic is exact two's complement with overflow detection
 *     (DuckDB raises on overflow); FLOAT/DOUBLE follow IEEE round-to-nearest.
 */
#ifndef METTLE_HIP_H
#define METTLE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Status codes — MettleError (src/error.rs:5-11) + fallback/overflow.    */
/* ---------------------------------------------------------------------- */
#define MTL_OK 0
#define MTL_ERR_INVALID_BUFFER_CONFIG 1 /* MettleError::InvalidBufferConfig */
#define MTL_ERR_DEVICE_UNAVAILABLE 2    /* MettleError::MetalUnavailable    */
#define MTL_ERR_NOT_INITIALIZED 3       /* MettleError::NotInitialized      */
#define MTL_ERR_PROCESSING 4            /* MettleError::Processing          */
#define MTL_UNSUPPORTED 5     /* shape not handled: run DuckDB CPU fallback  */
#define MTL_ERR_INVALID_ARGUMENT 6
#define MTL_ERR_OVERFLOW 7    /* integer overflow (DuckDB "Overflow in ...") */
#define MTL_ERR_BUDGET 8      /* request exceeds gpu_buffer_init budgets    */

/* ---------------------------------------------------------------------- */
/* Logical column types (DuckDB names).  Physical widths in brackets.     */
/* ---------------------------------------------------------------------- */
#define MTL_INTEGER 1   /* int32   [4]                                     */
#define MTL_BIGINT 2    /* int64   [8]                                     */
#define MTL_FLOAT 3     /* float   [4]                                     */
#define MTL_DOUBLE 4    /* double  [8]                                     */
#define MTL_DATE 5      /* int32 days since 1970-01-01 [4]                 */
#define MTL_TIMESTAMP 6 /* int64 microseconds since epoch [8]              */
#define MTL_DECIMAL 7   /* int64 scaled integer (scale kept by caller) [8] */
#define MTL_DICT8 8     /* uint8 dictionary code [1]                       */
#define MTL_DICT32 9    /* int32 dictionary code [4]                       */
#define MTL_HUGEINT 10  /* int128 {uint64 lower; int64 upper} [16]: SUM()  */
#define MTL_VARCHAR 11  /* result-only: dictionary-decoded strings         */

/* Compare ops — ScalarCompareOp codes (src/metal/shaders/mod.rs:27-36). */
#define MTL_CMP_EQ 0
#define MTL_CMP_LT 1
#define MTL_CMP_LE 2
#define MTL_CMP_GT 3
#define MTL_CMP_GE 4

/* Projection ops — ScalarProjectOp codes (src/metal/shaders/mod.rs:16-25). */
#define MTL_PROJ_ADD 0
#define MTL_PROJ_MUL 1
#define MTL_PROJ_SUB 2     /* col - s */
#define MTL_PROJ_SUBFROM 3 /* s - col */

/* Aggregate functions — AggFunc (src/plan/ir.rs:100-107). */
#define MTL_AGG_SUM 0
#define MTL_AGG_COUNT 1 /* COUNT(col); COUNT(*) passes values == NULL */
#define MTL_AGG_MIN 2
#define MTL_AGG_MAX 3
#define MTL_AGG_AVG 4

/* Join strategies — GpuJoinStrategy (src/plan/ir.rs:176-180). */
#define MTL_JOIN_DENSE 0
#define MTL_JOIN_HASH 1
#define MTL_JOIN_RADIX 2 /* LDS-staged partitions; pairs in partition order */

/* A scalar literal as raw bits: int64 for integer-like types, the IEEE bits
 * of a double for FLOAT/DOUBLE columns (rounded to float for FLOAT). */
typedef struct mtl_pred {
  const void *column; /* device pointer to the column (row-indexed)      */
  int32_t type;       /* MTL_INTEGER ... MTL_DICT32                       */
  int32_t op;         /* MTL_CMP_*                                        */
  int64_t scalar;     /* literal bits                                     */
} mtl_pred;

#define MTL_MAX_PREDS 8

typedef struct mtl_stream mtl_stream; /* HIP stream + scratch arena       */
typedef struct mtl_result mtl_result; /* host-side materialised result    */
typedef struct mtl_request mtl_request; /* in-flight serving request      */

/* ====================================================================== */
/* Lifecycle                                                              */
/* ====================================================================== */

/* Parse a human-readable size ("1.5 MB", "2gb", "64") with 1024-based units.
 * Replaces parse_human_bytes (src/config.rs:61-123). */
int32_t mtl_parse_bytes(const char *text, uint64_t *out_bytes);

/* Initialise the backend with budgets; pinned_bytes == 0 means cache +
 * processing.  Replaces BufferConfig::parse + state::initialize +
 * metal::validate_runtime (config.rs:11-52, state.rs:23-40,
 * metal/mod.rs:96-98).  device_name receives the HIP device name. */
int32_t mtl_init(uint64_t cache_bytes, uint64_t processing_bytes,
                 uint64_t pinned_bytes, char *device_name,
                 size_t device_name_len);

/* Same, parsing the three SQL string arguments of
 * CALL gpu_buffer_init(cache, processing, pinned_memory_size := ...)
 * (table_functions.rs:30-52).  pinned may be NULL. */
int32_t mtl_buffer_init(const char *cache, const char *processing,
                        const char *pinned, char *device_name,
                        size_t device_name_len);

/* Budgets after init; returns MTL_ERR_NOT_INITIALIZED before. */
int32_t mtl_budgets(uint64_t *cache_bytes, uint64_t *processing_bytes,
                    uint64_t *pinned_bytes);

/* Drop every cached column, result and scratch arena (state::reset_for_tests,
 * metal::reset_runtime_cache: state.rs:113-129, metal/mod.rs:106-117). */
int32_t mtl_shutdown(void);

/* Copy the calling thread's last error message (NUL-terminated). */
int32_t mtl_last_error(char *buf, size_t len);

/* Device count visible to HIP without initialising the device context. */
int32_t mtl_device_count(int32_t *out);

/* ====================================================================== */
/* Streams (replace begin_batch/end_batch command-buffer batching,        */
/* runtime/mod.rs:182-236).                                               */
/* ====================================================================== */
int32_t mtl_stream_create(mtl_stream **out);
int32_t mtl_stream_destroy(mtl_stream *s);
int32_t mtl_stream_sync(mtl_stream *s); /* = end_batch's wait              */
/* The hipStream_t behind a handle (for callers that share streams). */
void *mtl_stream_native(mtl_stream *s);

/* Device memory from the processing budget (GpuBuffer::new / upload).
 * After gpu_buffer_init every device allocation of the library (these, the
 * streams' scratch arenas, group tables, SQL buffers, the serving slots) is
 * charged to the processing budget; one that would exceed it fails with
 * MTL_ERR_BUDGET (the reference's budgets bounded only SQL text length:
 * runtime/mod.rs:339-351). */
int32_t mtl_alloc(uint64_t bytes, void **out_dev);
int32_t mtl_free(void *dev);
/* Device bytes charged to the processing budget now. */
int32_t mtl_processing_in_use(uint64_t *bytes);

/* Refcounted handles (GpuBuffer / GpuVector, gpu_buffer.rs:12-52): memory
 * from the processing budget, freed at the last release.  A mtl_vec is a
 * selection vector: capacity uint32 row ids and a device count word, which
 * mtl_vec_count reads (the one synchronising read, as GpuVector's count is
 * read only after end_batch, gpu_buffer.rs:94-103).  The operators take
 * their device pointers (mtl_buf_device, mtl_vec_rows, mtl_vec_count_ptr). */
typedef struct mtl_buf mtl_buf;
typedef struct mtl_vec mtl_vec;
int32_t mtl_buf_alloc(uint64_t bytes, mtl_buf **out);
int32_t mtl_buf_retain(mtl_buf *b);
int32_t mtl_buf_release(mtl_buf *b);
void *mtl_buf_device(const mtl_buf *b);
uint64_t mtl_buf_bytes(const mtl_buf *b);
int32_t mtl_vec_alloc(uint64_t capacity, mtl_vec **out);
int32_t mtl_vec_retain(mtl_vec *v);
int32_t mtl_vec_release(mtl_vec *v);
uint32_t *mtl_vec_rows(const mtl_vec *v);
uint32_t *mtl_vec_count_ptr(const mtl_vec *v);
uint64_t mtl_vec_capacity(const mtl_vec *v);
int32_t mtl_vec_count(const mtl_vec *v, mtl_stream *s, uint64_t *n);
/* mtl_filter over rows 0..n_rows-1 (in == NULL) or over the rows of `in`,
 * into `out` (capacity at least the input's): filter_to_indices_gv
 * (metal/mod.rs filter path, K24) with handles. */
int32_t mtl_filter_vec(const mtl_pred *preds, uint32_t n_preds,
                       uint64_t n_rows, const mtl_vec *in, mtl_vec *out,
                       mtl_stream *s);
int32_t mtl_memcpy_h2d(void *dst_dev, const void *src_host, uint64_t bytes,
                       mtl_stream *s);
int32_t mtl_memcpy_d2h(void *dst_host, const void *src_dev, uint64_t bytes,
                       mtl_stream *s);

/* ====================================================================== */
/* HBM column cache (replaces column_to_gpu_buffer / cached_input_buffer, */
/* runtime/mod.rs:398-565, and source_cache.rs:220-317).  Columns are     */
/* keyed by (table, column) plus a caller version, charged to the cache    */
/* budget, reference-counted while in use, evicted least-recently-used.   */
/* Uploads stream through two pinned chunks on a side stream (the staging  */
/* of pinned.rs:187-244).  Requires gpu_buffer_init.                       */
/* ====================================================================== */
/* Cached device copy of `bytes` host bytes; uploads on a miss or a version
 * change.  Takes a reference (release it with mtl_cache_release). */
int32_t mtl_cache_put(uint64_t table_id, uint32_t column_id, const void *host,
                      uint64_t bytes, uint64_t version, const void **out_dev);
/* Reference an already cached column; MTL_ERR_INVALID_ARGUMENT if absent. */
int32_t mtl_cache_get(uint64_t table_id, uint32_t column_id,
                      const void **out_dev, uint64_t *bytes);
int32_t mtl_cache_release(uint64_t table_id, uint32_t column_id);
/* Evict every column of a table (fails if one is referenced). */
int32_t mtl_cache_drop(uint64_t table_id);
int32_t mtl_cache_stats(uint64_t *entries, uint64_t *bytes, uint64_t *hits,
                        uint64_t *misses, uint64_t *evictions,
                        uint64_t *uploaded_bytes);
/* Evict everything (fails if a column is referenced); called by
 * mtl_shutdown. */
int32_t mtl_cache_reset(void);

/* ====================================================================== */
/* Operators (device pointers in, device pointers out)                    */
/* ====================================================================== */

/* Fused predicate conjunction + stable selection-vector compaction.
 * out_sel[0..*out_count) = ascending row ids r (from in_sel when given,
 * else 0..n_rows-1) for which every preds[i] holds.
 * Replaces compare_{i32,i64,f32,date32}_to_gpu + and_masks_gpu +
 * filter_to_indices_{gpu,gv} (metal/mod.rs:200-265, scalar_ops.rs:282-439,
 * gather_ops.rs:23-159; kernels K4/K5/K6/K23/K24/K38).
 * in_count: device count of in_sel (NULL -> in_capacity rows).  out_sel must
 * hold in_capacity (or n_rows) entries; out_count is a device uint32. */
int32_t mtl_filter(const mtl_pred *preds, uint32_t n_preds, uint64_t n_rows,
                   const uint32_t *in_sel, const uint32_t *in_count,
                   uint64_t in_capacity, uint32_t *out_sel,
                   uint32_t *out_count, mtl_stream *s);

/* mtl_filter with no allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * filter's scratch (pass bits, tile counts, group sums) lives in a 16 MiB
 * static device array, one per GPU, held until the process ends; the call
 * makes no hipMalloc, hipFree or other allocation.  Same arguments and
 * results as mtl_filter, at most 100,000,000 rows per call (more returns
 * MTL_ERR_INVALID_ARGUMENT).  s == NULL uses the default stream, which
 * must exist (mtl_init or an earlier call created it).  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (concurrent
 * streams or threads) race; serialise them. */
int32_t mtl_filter_noalloc(const mtl_pred *preds, uint32_t n_preds,
                           uint64_t n_rows, const uint32_t *in_sel,
                           const uint32_t *in_count, uint64_t in_capacity,
                           uint32_t *out_sel, uint32_t *out_count,
                           mtl_stream *s);

/* mtl_filter fused with one mtl_project of the passing rows: out_sel and
 * *out_count as mtl_filter; out_proj[i] = proj_column[out_sel[i]] op scalar
 * for i < *out_count, with mtl_project's types, ops and overflow flag.  The
 * projection is computed while the selection is written, so the selection
 * is not read back.  Replaces the scan path's filter_to_indices_gv followed
 * by project_f32_scalar_gv (gather_ops.rs:23-159, scalar_ops.rs:512-561;
 * kernels K23/K24 then K35). */
int32_t mtl_filter_project(const mtl_pred *preds, uint32_t n_preds, uint64_t n_rows,
                           const uint32_t *in_sel, const uint32_t *in_count,
                           uint64_t in_capacity, const void *proj_column,
                           int32_t proj_type, int32_t op, int64_t scalar,
                           uint32_t *out_sel, uint32_t *out_count, void *out_proj,
                           uint32_t *overflow_flag, mtl_stream *s);

/* Per-row mask (uint8 0/1) of a predicate conjunction over rows 0..n-1.
 * Replaces compare_*_values (metal/mod.rs:146-198; scalar_ops.rs:85-279). */
int32_t mtl_compare(const mtl_pred *preds, uint32_t n_preds, uint64_t n_rows,
                    uint8_t *out_mask, mtl_stream *s);

/* out[i] = column[sel[i]] for i < count; width 1, 4 or 8 bytes.
 * Replaces gather_{i32,f32,i32_as_f32}_gv (gather_ops.rs:162-236, K26-K28)
 * and reindex_gv (gather_ops.rs:250-295, K29) when column is a selection. */
int32_t mtl_gather(const void *column, uint32_t width, const uint32_t *sel,
                   const uint32_t *count, uint64_t capacity, void *out,
                   mtl_stream *s);

/* mtl_filter over rows 0..n_rows-1 for a conjunction of disjunctions:
 * predicates p and p+1 are OR-ed when bit p of or_next is set, and the runs
 * of OR-ed predicates are AND-ed, so `a AND (b OR c) AND d IN (1, 2)` is
 * preds {a, b, c, d=1, d=2} with or_next = 0b0010 | 0b1000.  No
 * counterpart in the reference, whose planner rejects OR and IN
 * (parser.rs:277-367); DuckDB ran those on the CPU. */
int32_t mtl_filter_or(const mtl_pred *preds, uint32_t n_preds, uint32_t or_next,
                      uint64_t n_rows, uint32_t *out_sel, uint32_t *out_count,
                      mtl_stream *s);

/* Scalar projection out[i] = column[sel[i]] op scalar (sel NULL -> i).
 * INTEGER/BIGINT/DATE(+/-)/TIMESTAMP(+/-)/DECIMAL: exact, sets
 * *overflow_flag (device uint32) to 1 on overflow.  FLOAT: IEEE f32.
 * DOUBLE: IEEE f64.  Replaces project_f32_values / project_f32_scalar_gv
 * (metal/mod.rs:146-156; scalar_ops.rs:32-82, 512-561; K3/K35), which
 * computed everything in f32. */
int32_t mtl_project(const void *column, int32_t type, const uint32_t *sel,
                    const uint32_t *count, uint64_t capacity, int32_t op,
                    int64_t scalar, void *out, uint32_t *overflow_flag,
                    mtl_stream *s);

/* ORDER BY one key column: out_rows[0..*out_n) = the selection (sel, or
 * rows 0..capacity-1 when sel is NULL) sorted by column[row] ascending, or
 * descending when desc != 0, ties in selection order (stable).  FLOAT and
 * DOUBLE follow DuckDB's total order (NaN largest, -0 == 0).  A device
 * count is read once on the host.  out_rows holds capacity entries.
 * Replaces the CPU sort_record_batch (helpers.rs:627-667). */
int32_t mtl_sort(const void *column, int32_t type, const uint32_t *sel,
                 const uint32_t *count, uint64_t capacity, int32_t desc,
                 uint32_t *out_rows, uint32_t *out_n, mtl_stream *s);

/* Bytes of the device table a join build needs (dense: slots x 4 B;
 * hash: next_pow2(max(2 x rows, 16)) x 16 B).  Mirrors the capacity rules of
 * choose_gpu_join_strategy / hash_join_capacity (join_plan.rs:29-56,
 * 300-303). */
uint64_t mtl_join_table_bytes(int32_t strategy, uint64_t build_rows,
                              uint64_t dense_slots);

/* Build side of an inner PK/FK equi-join on an INTEGER or BIGINT key.
 * Dense: table[key - min_key] = build row (slots = span).  Hash: open
 * addressing on 64-bit keys with a mixing hash.  *dup_flag (device uint32)
 * becomes 1 if a key repeats (the reference overwrote duplicates:
 * kernels.metal:202-206); the host then falls back.
 * Replaces hash_join_build_gv / dense_lookup_build_i32_gv /
 * dense_lookup_build_from_indices_i32_gv (join_ops.rs:22-70, 128-180,
 * 233-290; K31, K32, K33). */
int32_t mtl_join_build(int32_t strategy, const void *keys, int32_t key_type,
                       const uint32_t *sel, const uint32_t *count,
                       uint64_t capacity, int64_t min_key, uint64_t slots,
                       void *table, uint64_t table_bytes, uint32_t *dup_flag,
                       mtl_stream *s);

/* Probe side: for probe rows (sel or 0..n-1) whose key matches, emit the
 * pair (probe row, build row) in ascending probe order (MTL_JOIN_RADIX: in
 * an unspecified order).  `slots` is the dense key span, or for
 * MTL_JOIN_HASH and MTL_JOIN_RADIX the build capacity (rows) that was
 * passed to mtl_join_build.
 * MTL_JOIN_RADIX (build capacity <= 2^24 rows): build and probe rows are
 * partitioned by the top bits of the key's hash, about 2048 build rows per
 * partition; each partition's table is built in LDS (4096 slots) and probed
 * there, so a build far larger than the caches costs streaming passes, not
 * random HBM lookups (the LDS-staged partitions K31's global table,
 * kernels.metal:786-837, lacks).  Build *dup_flag bits: 2 a key repeats,
 * 16 a partition over 3072 rows (many repeats of one key): use HASH.
 * Replaces hash_join_probe_gv / dense_lookup_probe_i32_gv /
 * compact_matches_gv (join_ops.rs:72-125, 182-231; gather_ops.rs:298-393;
 * K30, K31, K32). */
int32_t mtl_join_probe(int32_t strategy, const void *table,
                       uint64_t table_bytes, int64_t min_key, uint64_t slots,
                       const void *probe_keys, int32_t key_type,
                       const uint32_t *sel, const uint32_t *count,
                       uint64_t capacity, uint32_t *out_probe,
                       uint32_t *out_build, uint32_t *out_count,
                       mtl_stream *s);

/* ---------------------------------------------------------------------- */
/* Grouped aggregate engine.                                              */
/* ---------------------------------------------------------------------- */
#define MTL_GROUP_MAX_KEYS 4
#define MTL_GROUP_MAX_AGGS 8

typedef struct mtl_agg_spec {
  int32_t func;       /* MTL_AGG_*                                        */
  int32_t type;       /* value column type (ignored when values == NULL)  */
  const void *values; /* device column; NULL = COUNT(*)                   */
  int32_t scale;      /* MTL_DECIMAL: digits after the point, 0..18; AVG  */
                      /* divides by count * 10^scale (DuckDB's DECIMAL    */
                      /* AVG -> DOUBLE); ignored for other types          */
} mtl_agg_spec;

typedef struct mtl_group_desc {
  uint64_t n_rows;            /* rows of the key / value columns          */
  const uint32_t *sel;        /* optional selection (device)              */
  const uint32_t *sel_count;  /* its device count (NULL = sel_capacity)   */
  uint64_t sel_capacity;
  uint32_t n_keys;            /* 1..MTL_GROUP_MAX_KEYS integer columns    */
  const void *key_col[MTL_GROUP_MAX_KEYS];
  int32_t key_type[MTL_GROUP_MAX_KEYS];
  int64_t key_min[MTL_GROUP_MAX_KEYS]; /* column stats: every key value   */
  int64_t key_max[MTL_GROUP_MAX_KEYS]; /* must lie in [min, max]          */
  uint32_t n_aggs;
  mtl_agg_spec aggs[MTL_GROUP_MAX_AGGS];
} mtl_group_desc;

typedef struct mtl_groups mtl_groups; /* device-resident group table     */

/* GROUP BY keys -> COUNT/SUM/MIN/MAX/AVG per group.  Keys are packed by
 * their [min, max] ranges (mixed radix, first key most significant); when
 * the packed domain is small the table is indexed directly, otherwise it
 * is an open-addressing hash table.  Hot keys are pre-aggregated in an
 * LDS cache per block.  SUM over integer types is exact (128-bit, DuckDB
 * HUGEINT); SUM/AVG over FLOAT/DOUBLE accumulate in double.
 * Replaces grouped_aggregate_exact_gv (metal/mod.rs:301-309,
 * aggregate_ops.rs:101-457; K13-K18), which used <= 4 i32 key lanes and
 * float32 atomics, and the CPU HashMap path of execute_aggregate_impl
 * (aggregate.rs:88-236) for BIGINT keys. */
int32_t mtl_group_agg(const mtl_group_desc *desc, mtl_groups **out,
                      mtl_stream *s);
/* Number of groups (synchronises the stream). */
int32_t mtl_groups_count(mtl_groups *g, uint64_t *n);
/* Copy every group to host buffers, ascending by packed key for a dense
 * table (arbitrary but complete for a hash table).  keys: n x n_keys int64
 * (row-major).  agg_out[a] receives per group: COUNT -> int64; SUM of an
 * integer type -> int128 {uint64 lower; int64 upper}; SUM/AVG of a float
 * type and AVG of an integer type -> double; MIN/MAX -> int64 (integer
 * types) or double (float types). */
int32_t mtl_groups_fetch(mtl_groups *g, int64_t *keys, void *const *agg_out,
                         uint64_t capacity, uint64_t *n_out);
/* Top-K groups by aggregate `agg` (desc != 0: largest first), ties broken
 * by the group keys ascending; same output layout as mtl_groups_fetch.
 * Replaces topk_dense_grouped_rows{,_parallel_v} (kernels.metal:1162-1390,
 * aggregate_ops.rs:572-686; K <= 64 there, K <= 512 here). */
int32_t mtl_groups_topk(mtl_groups *g, uint32_t agg, int32_t desc, uint32_t k,
                        int64_t *keys, void *const *agg_out, uint32_t *n_out);
/* Frees the group table.  Its device memory is kept on the table's stream
 * for the next mtl_group_agg there (reused when large enough and not over
 * four times too large; uncharged while kept) and returned by
 * mtl_stream_destroy or mtl_shutdown.  Release a table before destroying
 * its stream. */
int32_t mtl_groups_release(mtl_groups *g);

/* ---------------------------------------------------------------------- */
/* Star join -> grouped SUM -> top-K (the TPC-H Q3 shape).                */
/* ---------------------------------------------------------------------- */

/* Expression program over fact-table columns, evaluated per qualifying
 * fact row in int64 (exact, overflow-checked) or double.  Postfix ops: */
#define MTL_EXPR_COL 0   /* push column arg (index into cols)            */
#define MTL_EXPR_CONST 1 /* push constant arg (int64 or double bits)     */
#define MTL_EXPR_ADD 2
#define MTL_EXPR_SUB 3
#define MTL_EXPR_MUL 4
#define MTL_EXPR_MAX_OPS 16
#define MTL_EXPR_MAX_COLS 4

typedef struct mtl_expr {
  uint32_t n_ops;
  int32_t is_double; /* 0: int64 arithmetic, 1: double arithmetic       */
  int32_t op[MTL_EXPR_MAX_OPS];
  int64_t arg[MTL_EXPR_MAX_OPS];
  uint32_t n_cols;
  const void *col[MTL_EXPR_MAX_COLS];
  int32_t col_type[MTL_EXPR_MAX_COLS];
} mtl_expr;

#define MTL_STAR_MAX_TIES 3

typedef struct mtl_star_desc {
  /* dimension table D (optional: d_rows == 0 skips it) */
  uint64_t d_rows;
  const void *d_key; /* unique integer key of D            */
  int32_t d_key_type;
  int64_t d_key_min, d_key_max; /* column stats            */
  uint32_t d_n_preds;
  mtl_pred d_preds[MTL_MAX_PREDS];
  /* parent table P: grouped by its unique key (the group representative) */
  uint64_t p_rows;
  const void *p_key; /* unique integer key of P            */
  int32_t p_key_type;
  int64_t p_key_min, p_key_max;
  const void *p_fk; /* P -> D foreign key (ignored when d_rows == 0) */
  int32_t p_fk_type;
  uint32_t p_n_preds;
  mtl_pred p_preds[MTL_MAX_PREDS];
  /* fact table F */
  uint64_t f_rows;
  const void *f_fk; /* F -> P foreign key                  */
  int32_t f_fk_type;
  uint32_t f_n_preds;
  mtl_pred f_preds[MTL_MAX_PREDS];
  mtl_expr agg; /* SUM(agg) per P row                     */
  /* ORDER BY SUM(agg) DESC, then ties on P columns, then P key ASC */
  uint32_t k;
  uint32_t n_ties;
  const void *tie_col[MTL_STAR_MAX_TIES];
  int32_t tie_type[MTL_STAR_MAX_TIES];
  int32_t tie_desc[MTL_STAR_MAX_TIES];
} mtl_star_desc;

/* One result row of the star operator (host side). */
typedef struct mtl_star_row {
  uint32_t p_row;  /* row id in P (gather the group keys from it)   */
  uint32_t count;  /* number of fact rows summed                    */
  int64_t sum_i64; /* SUM when agg.is_double == 0                   */
  double sum_f64;  /* SUM when agg.is_double == 1                   */
} mtl_star_row;

/* Runs filter(D) -> semi-join bitmap -> filter(P) + probe -> filter(F) +
 * probe -> SUM(expr) per P row -> top-K, all device-resident with one
 * host synchronisation at the end (TARGET_ARCHITECTURE.md:11-17).
 * out_rows receives min(k, groups) rows in order; *out_n their number.
 * Replaces the GPU-native Q3 pipeline execute_gpu_native
 * (executor/fast_path/mod.rs:58-141) and its kernels K23/K24/K26/K29-K33/
 * K36/K37/K39-K44 (aggregate_ops.rs:30-98, 572-932). */
int32_t mtl_star_topk(const mtl_star_desc *desc, mtl_star_row *out_rows,
                      uint32_t *out_n, mtl_stream *s);

/* Same, but only enqueues the work; results land in the device buffer
 * out_dev (k x sizeof(mtl_star_row)) and out_n_dev (device, 2 x uint32:
 * [0] = rows, [1] = error bits, 0 when the run is valid; decode with
 * mtl_star_status).  Used by the serving loop and by graph capture. */
int32_t mtl_star_topk_async(const mtl_star_desc *desc, void *out_dev,
                            uint32_t *out_n_dev, mtl_stream *s);

/* Map the error bits of an asynchronous star run to a status (+ message). */
int32_t mtl_star_status(uint32_t error_bits);

/* ====================================================================== */
/* SQL lowering (replaces try_build_gpu_plan, parser.rs:20-120, and the    */
/* shape routing of try_execute / native_gpu_shape, executor/mod.rs:44-138 */
/* with its Q3 fast path, executor/fast_path/mod.rs:58-141).  Tables are   */
/* registered in a catalog of device-resident columns; the admitted SELECT */
/* subset lowers onto the operators above; anything else returns           */
/* MTL_UNSUPPORTED so the host runs DuckDB's CPU path.                     */
/* ====================================================================== */
int32_t mtl_catalog_add_table(const char *table, uint64_t n_rows);
/* Column stats: min/max bound every value (join and group keys need them);
 * unique marks a key column (the "one" side of a join).  scale: DECIMAL. */
int32_t mtl_catalog_add_column(const char *table, const char *column,
                               int32_t type, int32_t scale, const void *dev,
                               int64_t min_value, int64_t max_value,
                               int32_t unique);
/* Dictionary of a MTL_DICT8/MTL_DICT32 column: code i = values[i]. */
int32_t mtl_catalog_set_dictionary(const char *table, const char *column,
                                   const char *const *values, uint32_t n);
/* NULLs of a column: valid = device bytes, one per row, 1 = valid, 0 = NULL
 * (NULL: the column has no NULLs).  The value bytes of a NULL row must be
 * zero and min/max must bound the valid values (mtl_catalog_ingest_arrow
 * writes columns so).  Filters, joins, projections, GROUP BY and the
 * aggregates follow DuckDB's NULL semantics: a comparison with NULL is not
 * true, join keys that are NULL match nothing, aggregates skip NULLs (SUM,
 * MIN, MAX, AVG of no valid value are NULL, COUNT(col) counts valid values),
 * NULL keys form one group, ORDER BY puts NULLs last.  The reference pins no
 * column with NULLs (pinned.rs:295-413 rejects them), so its queries fall
 * back to DuckDB's CPU path. */
int32_t mtl_catalog_set_validity(const char *table, const char *column,
                                 const uint8_t *valid);
/* Drop one table, or every table when table == NULL (an ingested table's
 * cached columns are released). */
int32_t mtl_catalog_drop(const char *table);
/* Plan without running: buf receives the shape ("scan", "group", "star2",
 * "star3") or, on MTL_UNSUPPORTED, the reason. */
int32_t mtl_sql_explain(const char *sql, char *buf, size_t len);
/* Plan and run; the result holds host columns (mtl_result_*). */
int32_t mtl_sql_execute(const char *sql, mtl_result **out, mtl_stream *s);
int32_t mtl_result_shape(const mtl_result *r, uint32_t *n_cols,
                         uint64_t *n_rows);
/* Column i: its name, type (SUM of integers: MTL_HUGEINT; star SUM:
 * MTL_DECIMAL with `scale`), and n_rows values of that type's width. */
int32_t mtl_result_column(const mtl_result *r, uint32_t i, const char **name,
                          int32_t *type, int32_t *scale, const void **data);
/* NULLs of result column i: one byte per row (1 = valid), or *valid = NULL
 * when the column has none.  The value bytes of a NULL row are unspecified. */
int32_t mtl_result_validity(const mtl_result *r, uint32_t i,
                            const uint8_t **valid);
int32_t mtl_result_free(mtl_result *r);

/* Arrow C data interface structs, as the specification defines them
 * (arrow.apache.org/docs/format/CDataInterface.html). */
#ifndef ARROW_C_DATA_INTERFACE
#define ARROW_C_DATA_INTERFACE
struct ArrowSchema {
  const char *format;
  const char *name;
  const char *metadata;
  int64_t flags;
  int64_t n_children;
  struct ArrowSchema **children;
  struct ArrowSchema *dictionary;
  void (*release)(struct ArrowSchema *);
  void *private_data;
};
struct ArrowArray {
  int64_t length;
  int64_t null_count;
  int64_t offset;
  int64_t n_buffers;
  int64_t n_children;
  const void **buffers;
  struct ArrowArray **children;
  struct ArrowArray *dictionary;
  void (*release)(struct ArrowArray *);
  void *private_data;
};
#endif
/* The result as one Arrow record batch: a struct array ("+s") with one
 * child per column, in select order.  Types: INTEGER int32, BIGINT int64,
 * FLOAT float32, DOUBLE float64, DATE date32, TIMESTAMP timestamp[us],
 * DECIMAL decimal128(18, scale), HUGEINT sums decimal128(38, scale),
 * dictionary columns dictionary<uint8 | int32, utf8>.  NULLs are validity
 * bitmaps (the column's schema then carries ARROW_FLAG_NULLABLE).  The
 * caller owns both structs (release callbacks); they stay valid after
 * mtl_result_free.  Replaces the RecordBatch hand-back of gpu_processing
 * (processing.rs:77-96), whose computed columns were all Float64. */
int32_t mtl_result_export_arrow(const mtl_result *r, struct ArrowSchema *schema,
                                struct ArrowArray *array);

/* gpu_register_table (pinned.rs:187-453): register `table` from Arrow
 * record batches (the C data interface: schema is a struct "+s", one child
 * per column; every batch matches it).  Each column is converted on the
 * host and uploaded through the HBM column cache (mtl_cache_put, so
 * gpu_buffer_init comes first); the catalog keeps the cache references
 * until mtl_catalog_drop or the next ingestion of the table.  Types:
 * int8/16/32 and uint8/16 -> INTEGER; int64, uint32, uint64 -> BIGINT;
 * float32/64 -> FLOAT/DOUBLE; date32/date64 -> DATE; timestamp s/ms/us ->
 * TIMESTAMP (us); decimal128(p, s <= 18) -> DECIMAL when every value fits
 * int64; utf8, large_utf8 and dictionary<int, utf8> -> a dictionary column
 * whose sorted dictionary orders codes like the strings (DICT8 up to 256
 * values, else DICT32).  NULLs become validity bytes (see
 * mtl_catalog_set_validity).  Columns of other types are skipped, as the
 * reference skips them, and listed comma-separated in `skipped`.
 * unique_cols: key columns the caller declares unique (primary keys).
 * The caller keeps ownership of schema and batches (nothing is released). */
int32_t mtl_catalog_ingest_arrow(const char *table,
                                 const struct ArrowSchema *schema,
                                 struct ArrowArray *const *batches,
                                 uint32_t n_batches,
                                 const char *const *unique_cols,
                                 uint32_t n_unique, uint64_t *out_rows,
                                 uint64_t *out_bytes, char *skipped,
                                 size_t skipped_len);

/* gpu_rewrite_query(sql): the query printed back with every predicate
 * `col op literal` written as gpu_{eq,lt,le,gt,ge}_scalar(col, literal) and
 * every projection `col + literal` / `col * literal` as
 * gpu_{add,mul}_scalar(col, literal).  *needed = bytes including the NUL.
 * The rewritten text parses and runs through mtl_sql_execute unchanged.
 * (Named by the reference's README.md:57; absent from its code.) */
int32_t mtl_rewrite_query(const char *sql, char *buf, size_t len,
                          size_t *needed);

/* The scalar hooks over one vector of host values (a DuckDB DataChunk
 * column): gpu_add_scalar / gpu_mul_scalar write values op scalar in the
 * column's type (integer overflow -> MTL_ERR_OVERFLOW, as DuckDB raises);
 * gpu_{eq,lt,le,gt,ge}_scalar write uint8 0/1 (floats in DuckDB's total
 * order).  scalar_bits as in mtl_pred. */
#define MTL_HOOK_ADD 0
#define MTL_HOOK_MUL 1
#define MTL_HOOK_EQ 2
#define MTL_HOOK_LT 3
#define MTL_HOOK_LE 4
#define MTL_HOOK_GT 5
#define MTL_HOOK_GE 6
int32_t mtl_scalar_hook(int32_t hook, int32_t type, const void *values,
                        uint64_t n, int64_t scalar_bits, void *out);

/* ====================================================================== */
/* Serving loop: continuous admission of queries onto concurrent streams. */
/* Replaces the one-query-per-CALL materialize_query (processing.rs:      */
/* 103-143) and stage_query (state.rs:42-54), which ran every query       */
/* synchronously inside bind with no overlap and no admission control.    */
/* ====================================================================== */
#define MTL_Q_SCAN 0  /* filter -> optional projection (plain scans)   */
#define MTL_Q_STAR 1  /* star join -> SUM -> top-K (TPC-H Q3 shape)    */
#define MTL_Q_GROUP 2 /* grouped aggregate (+ top-K of the groups)     */
#define MTL_Q_SQL 3   /* SQL text, lowered and run like mtl_sql_execute */

typedef struct mtl_query {
  int32_t kind; /* MTL_Q_*                                                */
  /* MTL_Q_SCAN: rows of `n_rows` passing every predicate, then (when
   * proj_col != NULL) proj_col[row] op scalar over those rows.  The result
   * stays on the device in the server's slot; the host receives the row
   * count and an order-sensitive fingerprint of (row id, projected value)
   * pairs, so results can be compared alone and in a mix. */
  uint64_t n_rows;
  uint32_t n_preds;
  mtl_pred preds[MTL_MAX_PREDS];
  const void *proj_col;
  int32_t proj_type;
  int32_t proj_op;
  int64_t proj_scalar;
  /* MTL_Q_STAR */
  mtl_star_desc star;
  mtl_star_row *star_out; /* host buffer of star.k rows                  */
  /* MTL_Q_GROUP: mtl_group_agg over `group`; with group_k > 0 the top
   * group_k groups by aggregate group_agg (group_desc != 0: largest first),
   * else every group (at most group_capacity), written to the host buffers
   * in mtl_groups_fetch / mtl_groups_topk layout. */
  mtl_group_desc group;
  uint32_t group_agg;
  int32_t group_desc;
  uint32_t group_k;
  uint64_t group_capacity;
  int64_t *group_keys;                  /* capacity x n_keys            */
  void *group_out[MTL_GROUP_MAX_AGGS];  /* one buffer per aggregate     */
  /* MTL_Q_SQL: the statement (copied at submit) */
  const char *sql;
} mtl_query;

typedef struct mtl_query_result {
  int32_t status;       /* MTL_OK or the query's error status             */
  uint32_t n_rows;      /* MTL_Q_STAR: rows written to star_out;          */
                        /* MTL_Q_GROUP: groups written to the buffers     */
  uint64_t count;       /* MTL_Q_SCAN: selected rows                      */
  uint64_t fingerprint; /* MTL_Q_SCAN: hash of (row id, value) in order   */
  int64_t t_submit_ns;  /* steady-clock stamps: arrival (or submit),      */
  int64_t t_start_ns;   /* dispatch to a slot,                            */
  int64_t t_done_ns;    /* completion observed by the host                */
  char error[160];      /* message when status != MTL_OK                 */
  uint64_t groups;      /* MTL_Q_GROUP: number of groups                  */
  mtl_result *result;   /* MTL_Q_SQL: the rows (free: mtl_result_free)    */
  uint32_t batch;       /* queries of the batch it ran in (1: alone)      */
} mtl_query_result;

typedef struct mtl_server mtl_server;

/* n_slots concurrent streams, each with a worker thread.  A worker takes
 * every query that has arrived, in arrival order, up to max_batch, while
 * their device scratch fits scratch_budget (0 = the processing budget of
 * gpu_buffer_init, or 16 GiB before init), and runs them as one batch:
 * scans with one predicate on the same column and star queries over the
 * same tables share their table passes (each tile is read once for the
 * whole batch), with one host synchronisation for the batch; group-by and
 * SQL queries run one after another on the same stream.  n_slots = 1 and
 * max_batch = 1 is serial one-query-at-a-time dispatch, the reference's
 * one query per CALL (processing.rs:103-143). */
int32_t mtl_server_create(uint32_t n_slots, uint32_t max_batch,
                          uint64_t scratch_budget, mtl_server **out);
/* Queue a query (the descriptor is copied; columns and star_out must stay
 * valid until it completes).  not_before_ns > 0 holds the query until that
 * steady-clock time, which is then its arrival time. */
int32_t mtl_server_submit(mtl_server *srv, const mtl_query *q,
                          int64_t not_before_ns, uint64_t *ticket);
/* Queue a SQL statement (the text is copied); it is lowered at submit
 * (MTL_UNSUPPORTED statements complete at once with that status, the
 * signal for DuckDB's CPU path) and runs like mtl_sql_execute.  The result
 * handle arrives in mtl_query_result.result.  Replaces gpu_processing's
 * materialize_query / stage_query (processing.rs:103-143, state.rs:42-54)
 * as a request loop. */
int32_t mtl_server_submit_sql(mtl_server *srv, const char *sql,
                              int64_t not_before_ns, uint64_t *ticket);
/* Block until the query completes; releases the ticket. */
int32_t mtl_server_wait(mtl_server *srv, uint64_t ticket,
                        mtl_query_result *out);
/* Without blocking: *done = 1 and the result (the ticket released) when the
 * query has completed, else *done = 0. */
int32_t mtl_server_poll(mtl_server *srv, uint64_t ticket,
                        mtl_query_result *out, int32_t *done);
/* Steady-clock now, in the clock of the result stamps. */
int64_t mtl_now_ns(void);
/* Device scratch bytes a query needs at admission (its budget charge). */
uint64_t mtl_query_scratch_bytes(const mtl_query *q);
/* Waits for queued work, stops the workers, frees the slots. */
int32_t mtl_server_destroy(mtl_server *srv);

/* ====================================================================== */
/* Kernel timing: hipEvent pairs around every kernel launch, resolved at  */
/* mtl_stream_sync.  Counterpart of METTLE_PROFILE                        */
/* (src/plan/executor/cache.rs:10-21), which timed phases on the host.    */
/* ====================================================================== */
int32_t mtl_profile_enable(int32_t on);
int32_t mtl_profile_reset(void);
/* Entry `index` (0-based, sorted by name): kernel name, summed ms, count.
 * Returns MTL_ERR_INVALID_ARGUMENT past the last entry. */
int32_t mtl_profile_get(uint32_t index, char *name, size_t name_len,
                        double *total_ms, uint64_t *count);

/* ====================================================================== */
/* Autotuner.  No counterpart in the reference, which fixes these launch  */
/* parameters in code (join_plan.rs:29-56 load factors and cut-offs,      */
/* aggregate_ops.rs:780-786 threadgroup sizes).  Parameters:              */
/*   small_block, small_steps  the persistent small-scan kernel's block   */
/*                             size and 256-row steps per wave (rows per  */
/*                             lane / 4), read when a server is created;  */
/*   join_load_pct             hash join table fill (mtl_join_table_bytes)*/
/*   group_load_pct            group-by hash table fill;                  */
/*   part_bits                 group-by spill partitioning fan-out bits   */
/*                             per scatter level.                         */
/* Every candidate value gives identical results; only speed changes.     */
/* Set parameters at initialisation, before queries and servers run.      */
/* ====================================================================== */
#define MTL_TUNE_SMALL_SCAN 1u
#define MTL_TUNE_JOIN 2u
#define MTL_TUNE_GROUP 4u
#define MTL_TUNE_ALL 7u
int32_t mtl_tune_get(const char *name, int32_t *value);
/* value must be one of the parameter's candidates, else                   */
/* MTL_ERR_INVALID_ARGUMENT and the table is unchanged.                    */
int32_t mtl_tune_set(const char *name, int32_t value);
int32_t mtl_tune_reset(void); /* the defaults */
/* "name=value" lines ('#' starts a comment).  A file with an unknown name */
/* or a value outside the candidates is refused as a whole.                */
int32_t mtl_tune_save(const char *path);
int32_t mtl_tune_load(const char *path);
/* The table as one line: "name=value name=value ...". */
int32_t mtl_tune_describe(char *buf, size_t len);
/* Times every candidate of the selected groups (MTL_TUNE_*) on synthetic  */
/* device data, checks that each gives the default's result exactly       */
/* (MTL_ERR_PROCESSING otherwise), keeps the fastest in the table and     */
/* writes one "group: name=value time" line per candidate to report       */
/* (may be NULL). */
int32_t mtl_autotune(uint32_t groups, char *report, size_t len);

#ifdef __cplusplus
}
#endif
#endif /* METTLE_HIP_H */
