// Many-to-many join helpers (keys repeating on both sides), used by the SQL
// join lowering: the build side sorted by key becomes runs of equal keys,
// the probe joins the runs' unique keys, and each (probe, run) pair expands
// to one pair per build row of the run.
#pragma once

#include "mtl_internal.hpp"

namespace mtl {

// Runs of equal keys in n >= 1 sorted keys (any integer column type):
// run_start[r] = first position of run r, run_key[r] = its key (int64),
// run_start[*n_runs] = n.  tile_buf: ceil(n / 4096) + 1 uint32; run_start:
// n + 1 entries; run_key: n entries.  *n_runs is device memory.
int32_t join_runs(const void* sorted_keys, int32_t type, uint64_t n, uint32_t* tile_buf,
                  uint32_t* run_start, int64_t* run_key, uint32_t* n_runs, hipStream_t st);
// Adds to *total (device, cleared by the caller) the rows the np pairs
// expand to; tile_buf: ceil(np / 4096) + 1 uint32, kept for the write.
int32_t join_expand_count(const uint32_t* pair_run, uint64_t np, const uint32_t* run_start,
                          uint32_t* tile_buf, unsigned long long* total, hipStream_t st);
// Writes the total (< 2^32) expanded pairs, in pair order and, within a
// pair, in sorted build order: out_probe[t] = the pair's probe value,
// out_build[t] = sorted_rows[run start + k].  off: np uint32 scratch.
int32_t join_expand_write(const uint32_t* pair_probe, const uint32_t* pair_run, uint64_t np,
                          const uint32_t* run_start, const uint32_t* sorted_rows,
                          uint32_t* tile_buf, uint32_t* scan_total, uint32_t* off,
                          uint64_t total, uint32_t* out_probe, uint32_t* out_build,
                          hipStream_t st);

}  // namespace mtl
