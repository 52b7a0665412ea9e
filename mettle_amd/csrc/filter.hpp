// Host-visible argument blocks of the filter/compaction kernels, shared by
// the operator entry points and the fused star pipeline.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mettle_hip.h"
#include "device.hpp"
#include "rowtile.hpp"

namespace mtl {

struct ZeroList {
  uint32_t* ptr[8];
  uint64_t words[8];
  uint32_t value[8];
  int n = 0;
  void add(uint32_t* p, uint64_t w, uint32_t v = 0) {
    ptr[n] = p;
    words[n] = w;
    value[n] = v;
    ++n;
  }
};
int32_t launch_zero(const ZeroList& zl, hipStream_t st);

// Host: mtl_pred -> compiled predicate (encoded literal + op mask).  Integer
// literals compare exactly against every integer width (DuckDB compares an
// INTEGER column with a BIGINT literal in BIGINT); FLOAT literals are
// rounded to float first.
inline dev::PredC compile_pred(const mtl_pred& p) {
  dev::PredC c;
  c.col = p.column;
  switch (p.type) {
    case MTL_BIGINT:
    case MTL_TIMESTAMP:
    case MTL_DECIMAL: c.kind = dev::PK_I64; break;
    case MTL_DICT8: c.kind = dev::PK_U8; break;
    case MTL_FLOAT: c.kind = dev::PK_F32; break;
    case MTL_DOUBLE: c.kind = dev::PK_F64; break;
    default: c.kind = dev::PK_I32; break;
  }
  static const uint32_t masks[5] = {2u, 1u, 3u, 4u, 6u};  // EQ LT LE GT GE
  c.opmask = masks[p.op];
  double d = __builtin_bit_cast(double, p.scalar);
  if (p.type == MTL_FLOAT) c.lit = dev::enc_dbl((double)(float)d);
  else if (p.type == MTL_DOUBLE) c.lit = dev::enc_dbl(d);
  else c.lit = dev::enc_int(p.scalar);
  // int32 interval of the passing values (EQ [L, L], LT [min, L-1], LE
  // [min, L], GT [L+1, max], GE [L, max]), clamped to int32
  const int64_t L = p.scalar, I32MIN = INT32_MIN, I32MAX = INT32_MAX;
  int64_t lo = I32MIN, hi = I32MAX;
  switch (p.op) {
    case 0: lo = L; hi = L; break;
    case 1: hi = L == INT64_MIN ? INT64_MIN : L - 1; break;
    case 2: hi = L; break;
    case 3: lo = L == INT64_MAX ? INT64_MAX : L + 1; break;
    default: lo = L; break;
  }
  lo = lo < I32MIN ? I32MIN : lo;
  hi = hi > I32MAX ? I32MAX : hi;
  if (lo > hi) {  // nothing passes
    c.lo32 = INT32_MIN;
    c.hi32 = INT32_MAX;
    c.neg32 = 1;
  } else {
    c.lo32 = (int32_t)lo;
    c.hi32 = (int32_t)hi;
    c.neg32 = 0;
  }
  return c;
}

struct FilterArgs {
  dev::PredC preds[MTL_MAX_PREDS];
  uint32_t n_preds;
  uint32_t vec_ok;  // bit p: column p is aligned for vector loads
  uint64_t n_rows;  // rows (no in_sel) or capacity of in_sel
  const uint32_t* in_sel;
  const uint32_t* in_count;
  uint64_t* sel_bm;       // optional: pass bits, 1 bit per row (tiles*4096)
  uint32_t* tile_counts;  // optional: pass count per 4096-row tile
  uint32_t* group_sums;   // optional: per 256-tile group, sum of its counts
                          // (group g at group_sums[g * kGroupStride])
  uint32_t* err;          // bit 3: mark key outside [min, max]
  // optional semi-join probe: keep the row only if bit (key - min) is set
  const uint64_t* probe_bm;
  const void* probe_key;
  int32_t probe_key_type;
  int64_t probe_min;
  uint64_t probe_span;
  // optional side effect: set bit (key - min) for every kept row
  uint64_t* mark_bm;
  const void* mark_key;
  int32_t mark_key_type;
  int64_t mark_min;
  uint64_t mark_span;
  uint32_t probe_vec, mark_vec;  // key columns aligned for vector loads
};
// One block per 4096-row tile, no inter-block dependency.  With sel_bm and
// tile_counts set, launch_compact turns the pass bits into a stable
// selection vector; with only mark_bm set it is a mark-only pass.

constexpr int kTileRows = 4096;
constexpr int kGroupShift = 8;  // 256 tiles to a group (see launch_compact)
// Group sums sit 256 bytes apart: atomics to one cache line serialise, so
// each group's 256 additions get a line of their own.
constexpr int kGroupStride = 64;
uint32_t filter_tiles(uint64_t n);
int32_t launch_filter(FilterArgs a, uint64_t capacity, hipStream_t st);
// Exclusive scan of tile counts in place (one block); *total = sum.
void scan_counts(uint32_t* counts, uint32_t tiles, uint32_t* total, hipStream_t st);
// Projection fused into the compaction: out[i] = col[row_i] op scalar.
struct CompactProj {
  const void* col;
  int32_t type;
  int32_t op;
  int64_t scalar;
  void* out;
  uint32_t* ovf;  // device flag, set to 1 on integer overflow
};
// Per-group fingerprint sums of a compaction's output (see fp_term), one
// uint64 per 256-tile group, kFpStride uint64 apart (a cache line each).
constexpr int kFpStride = 32;

// Expansion of sel_bm into ascending row ids (in_sel[k] when in_sel is
// given); *out_count = total.  Each tile's output offset comes from an
// exclusive scan of tile_counts (in place) or, when group_sums holds the
// sums of each 256-tile group of the counts (kGroupStride words apart), is
// computed by the expansion itself with no scan pass.
int32_t launch_compact(const uint64_t* sel_bm, uint32_t* tile_counts, uint32_t tiles,
                       const uint32_t* in_sel, uint32_t* out_sel, uint32_t* out_count,
                       hipStream_t st, const CompactProj* proj = nullptr,
                       const uint32_t* group_sums = nullptr, uint64_t* fp_groups = nullptr);

// Internal hooks of the serving loop on a filter (+ projection): extra
// regions cleared by the filter's own clearing launch, and the output
// fingerprint accumulated by the expansion (fp_groups / fp_n are outputs:
// the per-group sums, cleared and filled on the stream; the caller adds
// them up).
struct ScanExtras {
  ZeroList zero;
  bool fingerprint = false;
  uint64_t* fp_groups = nullptr;
  uint32_t fp_n = 0;
};
int32_t filter_project_ex(const mtl_pred* preds, uint32_t n_preds, uint64_t n_rows,
                          const CompactProj* proj, uint32_t* out_sel, uint32_t* out_count,
                          ScanExtras* ex, mtl_stream* s);

// ---------------------------------------------------------------------------
// Shared scans of a batch of queries (the serving loop's continuous
// batching).  Queries with one predicate on the same column, projecting the
// same column (or nothing), form a share group: each tile of the column is
// loaded once for all of them.  Each query keeps its own row count (a prefix
// of the resident table), operator, literal and outputs.
// ---------------------------------------------------------------------------
constexpr uint32_t kMQMax = dev::kMQQueries;  // queries per share group
struct ScanBatchItem {
  const mtl_query* q;
  uint32_t* out_sel;  // q->n_rows row ids
  void* out_proj;     // q->n_rows projected values (nullptr without projection)
};
// A query the shared scan can take: one predicate, rows < 2^32, a valid
// projection.
bool scan_batchable(const mtl_query& q);
// Same share group: same predicate column and type, same projection column
// and type.
bool scan_shareable(const mtl_query& a, const mtl_query& b);
// Pinned host bytes the enqueue stages its per-query arguments in.
size_t scan_batch_stage_bytes(uint32_t n);
// Enqueue the shared scan of n <= kMQMax shareable queries on s.  words: n x
// 4 uint32 per query (count, overflow flag, 64-bit fingerprint), cleared and
// filled on the stream; the fingerprint is the one the single-query serving
// path computes.  Scratch comes from the stream's arena (not reset here).
int32_t scan_batch_enqueue(const ScanBatchItem* items, uint32_t n, uint32_t* words,
                           void* host_stage, mtl_stream* s);

// ---------------------------------------------------------------------------
// Small scans of a batch (at most kSmallRows rows each; any columns, up to
// MTL_MAX_PREDS predicates, an optional projection) in one persistent
// launch, replayed from a HIP graph captured once per serving slot.
// ---------------------------------------------------------------------------
constexpr uint64_t kSmallRows = 1ull << 16;  // 16 tiles: one block per query
constexpr uint32_t kSmallMax = 256;          // small scans per batch
struct SmallGraph;
struct SmallScanItem {
  const mtl_query* q;
  uint32_t* out_sel;  // q->n_rows row ids
  void* out_proj;     // q->n_rows projected values (nullptr without projection)
};
bool small_scan_ok(const mtl_query& q);
// Pinned staging, device buffers and the instantiated graph (captured on s).
int32_t small_graph_create(mtl_stream* s, SmallGraph** out);
void small_graph_destroy(SmallGraph* g);
// Stage n <= kSmallMax small scans and launch the graph on s.  The staging
// is reused by the next call: the caller synchronises s in between.
int32_t small_scan_enqueue(SmallGraph* g, const SmallScanItem* items, uint32_t n,
                           mtl_stream* s);
// After s is synchronised: 4 uint32 per query (count, overflow flag,
// 64-bit fingerprint), in enqueue order.
const uint32_t* small_scan_results(const SmallGraph* g);
// The pieces the autotuner times directly: the staged block's size, its
// filling from items, and one launch of the kernel instantiation for
// (block, steps) on a device copy of the block (its ticket uploaded as 0).
size_t small_block_bytes();
int32_t small_scan_stage(void* h_blk, const SmallScanItem* items, uint32_t n);
int32_t small_scan_launch(int32_t block, int32_t steps, uint32_t grid, void* d_blk,
                          uint32_t* d_res, hipStream_t st);

}  // namespace mtl
