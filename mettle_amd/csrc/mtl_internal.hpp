// Internal runtime declarations shared by every translation unit of
// libmettle_hip.so.  Host-side: error reporting, the global runtime (budgets,
// device, default stream), stream handles with their scratch arenas.
// Device-side helpers live in device.hpp.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mettle_hip.h"

namespace mtl {

// ---------------------------------------------------------------------------
// Errors: thread-local message + status code (MettleError's Display strings,
// src/error.rs:14-22, are reproduced by the prefixes).
// ---------------------------------------------------------------------------
int32_t set_error(int32_t code, const char* fmt, ...)
    __attribute__((format(printf, 2, 3)));
const char* last_error();

#define MTL_HIP(call)                                                        \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess)                                                    \
      return ::mtl::set_error(MTL_ERR_PROCESSING, "processing error: %s: %s", \
                              #call, hipGetErrorString(_e));                 \
  } while (0)

#define MTL_TRY(expr)              \
  do {                             \
    int32_t _s = (expr);           \
    if (_s != MTL_OK) return _s;   \
  } while (0)

// ---------------------------------------------------------------------------
// Scratch arena: a bump allocator owned by one stream.  Work enqueued on the
// stream is ordered, so an operator may reset the arena at its start: the
// previous operator's kernels are complete before the next one's start.
// Chunks are only ever freed after a stream synchronisation.
// ---------------------------------------------------------------------------
struct Arena {
  struct Chunk {
    char* base;
    size_t cap;
  };
  std::vector<Chunk> chunks;  // chunks.back() is the active chunk
  size_t off = 0;
  size_t high_water = 0;

  void reset() { off = 0; }
  // Allocate `bytes` (256-byte aligned).  May hipMalloc a new, larger chunk;
  // pointers handed out earlier stay valid until release_retired().
  int32_t alloc(size_t bytes, void** out);
  // Free every chunk but the active one (call only after a stream sync).
  void release_retired();
  void release_all();
};

}  // namespace mtl

struct mtl_stream {
  hipStream_t stream = nullptr;
  bool owns_stream = false;
  mtl::Arena arena;
  // small pinned host staging for counts/results of synchronous entry points
  void* pinned_small = nullptr;
  size_t pinned_small_bytes = 0;
  std::mutex mu;  // one operator at a time per handle
  // the last released group table's memory (state words, hashed keys),
  // kept for the next group-by on this stream: reuse is stream-ordered, and
  // parked memory is not charged to the processing budget, like memory
  // returned to the device's pool
  void* parked_table[2] = {nullptr, nullptr};
  uint64_t parked_bytes[2] = {0, 0};
  // kernel timing (prof.hpp): pending (name, start, end) event pairs and a
  // pool of reusable events
  struct ProfPair {
    std::string name;
    hipEvent_t start, end;
  };
  std::vector<ProfPair> prof_pending;
  std::vector<hipEvent_t> prof_pool;
};

namespace mtl {

struct Runtime {
  std::mutex mu;
  bool initialized = false;
  int device = 0;
  std::string device_name;
  std::string arch_name;
  int cu_count = 0;
  uint64_t cache_bytes = 0;
  uint64_t processing_bytes = 0;
  uint64_t pinned_bytes = 0;
  mtl_stream* default_stream = nullptr;
  std::atomic<uint64_t> device_bytes_in_use{0};  // charged to the processing budget
  std::atomic<uint64_t> pinned_bytes_in_use{0};
};

// Processing-budget accounting: every device allocation of the library
// outside the column cache is charged here.  Before gpu_buffer_init there
// is no limit (the charge is still counted).
int32_t charge_processing(uint64_t bytes, const char* what);
void release_processing(uint64_t bytes);
// hipMalloc charged to the processing budget; dev_free releases the charge.
int32_t dev_alloc(uint64_t bytes, void** out, const char* what);
void dev_free(void* p, uint64_t bytes);

Runtime& runtime();
// Ensures a device context exists (HIP init + default stream) without
// requiring gpu_buffer_init; low-level operators only need the device.
int32_t ensure_device();
// Resolve NULL to the default stream.
int32_t resolve_stream(mtl_stream* s, mtl_stream** out);
// Catalog ingestion (ingest.hip): release the cache references an ingested
// table holds (every table when table == NULL); the catalog entry itself is
// dropped by catalog_drop_table (sql.hip), which releases nothing.
void ingest_release(const char* table);
void catalog_drop_table(const char* table);
// Pinned staging for small synchronous copies.
int32_t stream_pinned(mtl_stream* s, size_t bytes, void** out);
// Wait for a stream's work by polling it: a blocking synchronise can return
// tens of microseconds after the work ends, which the synchronous entry
// points (a query's result read back) would pay on every call.
inline hipError_t stream_wait(hipStream_t st) {
  for (;;) {
    const hipError_t e = hipStreamQuery(st);
    if (e != hipErrorNotReady) return e;
  }
}

// ---------------------------------------------------------------------------
// Tuned launch parameters: the autotuner's choices (tune.hip), or the
// defaults below until it runs or a saved table is loaded.  Read on the host
// when an operator plans its launches; every value gives the same results.
// ---------------------------------------------------------------------------
struct TuneParam {
  const char* name;
  int32_t value;
  int32_t candidates[8];  // allowed values (0 ends the list)
};
enum TuneId : int {
  T_SMALL_BLOCK,   // persistent small-scan kernel: threads per block
  T_SMALL_STEPS,   // its 256-row steps per wave and tile (rows per lane / 4)
  T_JOIN_LOAD,     // hash join table fill, percent of slots
  T_GROUP_LOAD,    // group-by hash table fill, percent of slots
  T_PART_BITS,     // group-by spill partitioning: fan-out bits per scatter level
  T_COUNT
};
TuneParam* tune_table();
inline int32_t tuned(TuneId id) { return tune_table()[id].value; }

// Device count of CUs (for grid sizing of persistent/top-k kernels).
inline int cu_count() { return runtime().cu_count > 0 ? runtime().cu_count : 256; }

// ---------------------------------------------------------------------------
// Type helpers (host side)
// ---------------------------------------------------------------------------
inline int type_width(int32_t t) {
  switch (t) {
    case MTL_INTEGER:
    case MTL_DATE:
    case MTL_DICT32:
    case MTL_FLOAT:
      return 4;
    case MTL_BIGINT:
    case MTL_TIMESTAMP:
    case MTL_DECIMAL:
    case MTL_DOUBLE:
      return 8;
    case MTL_DICT8:
      return 1;
    case MTL_HUGEINT:
      return 16;
    default:
      return 0;
  }
}
inline bool type_is_int(int32_t t) {
  return t == MTL_INTEGER || t == MTL_BIGINT || t == MTL_DATE ||
         t == MTL_TIMESTAMP || t == MTL_DECIMAL || t == MTL_DICT8 ||
         t == MTL_DICT32;
}

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
inline uint64_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

}  // namespace mtl
