// Shared star passes of a batch of queries (the serving loop's continuous
// batching): star queries over the same three tables, each with its own
// predicate literals, row counts (prefixes of the resident tables) and
// K, are evaluated together.  The dimension, parent and fact tables are
// each streamed once for the whole batch; the per-query steps (ranks, slot
// map, top-K) run per query on the same stream.
#pragma once

#include "mtl_internal.hpp"

namespace mtl {

constexpr uint32_t kStarMQMax = 64;  // queries per share group
constexpr uint32_t kDimSlots = 8;    // distinct dimension predicates per share group

struct StarBatchItem {
  const mtl_star_desc* d;
  mtl_star_row* out;  // device, d->k rows
  uint32_t* out_n;    // device, [0] rows, [1] error bits (star_check_error)
};

// A query the shared passes can take: a dimension table, at most one
// predicate per table on a column of at most 4 bytes, join keys of at most
// 4 bytes, and a descriptor mtl_star_topk accepts.
bool star_batchable(const mtl_star_desc& d);
// Same share group: the same columns (keys, predicate columns, expression)
// and key ranges apart from the parent key's upper bound.  A group also
// holds at most kDimSlots distinct dimension predicates (star_dim_same).
bool star_shareable(const mtl_star_desc& a, const mtl_star_desc& b);
// Same dimension slot: the same dimension predicate and row count.
bool star_dim_same(const mtl_star_desc& a, const mtl_star_desc& b);
// Pinned host bytes the enqueue stages its per-query arguments in.
size_t star_batch_stage_bytes(uint32_t n);
// Enqueue n <= kStarMQMax shareable queries on s (scratch from its arena,
// not reset here).
int32_t star_batch_enqueue(const StarBatchItem* items, uint32_t n, void* host_stage,
                           mtl_stream* s);
// Status of the error bits a star run leaves in out_n[1].
int32_t star_check_error(uint32_t err);
// out[i] = the integer expression e over row i of its columns, for i < n;
// an overflow sets err[0] bit 0 (err is not cleared here).
int32_t expr_eval_i64(const mtl_expr& e, uint64_t n, int64_t* out, uint32_t* err,
                      hipStream_t st);

}  // namespace mtl
