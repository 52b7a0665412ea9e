// HBM column cache with pinned staging.
//
// Replaces the reference's column residency: column_to_gpu_buffer, a cache
// keyed by the host Vec's address and never evicted (runtime/mod.rs:398-469),
// cached_input_buffer's LRU bounded by cache_region_bytes (:496-565), and the
// host-side source caches (source_cache.rs:220-317).  Here a column is keyed
// by (table id, column id) plus a caller version, charged against the cache
// budget of gpu_buffer_init, reference-counted while queries use it, and
// evicted least-recently-used when room is needed.  Uploads stream through
// two pinned (hipHostMalloc) chunks on a side stream: the host copies chunk
// i + 1 into pinned memory while the DMA engine moves chunk i.  The
// reference copied DuckDB columns into host vectors with pinned.rs:187-244;
// the pinned chunks here are charged to the pinned budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <list>
#include <map>
#include <thread>
#include <vector>

#include "mtl_internal.hpp"

namespace mtl {

constexpr uint64_t kStageChunk = 32ull << 20;

struct CacheEntry {
  void* dev = nullptr;
  uint64_t bytes = 0;
  uint64_t version = 0;
  uint32_t refs = 0;
  std::list<std::pair<uint64_t, uint32_t>>::iterator lru;  // position in Cache::lru
};

struct Cache {
  std::mutex mu;
  std::map<std::pair<uint64_t, uint32_t>, CacheEntry> entries;
  std::list<std::pair<uint64_t, uint32_t>> lru;  // front = most recently used
  uint64_t bytes = 0;
  uint64_t hits = 0, misses = 0, evictions = 0, uploaded = 0;
  // staging
  hipStream_t side = nullptr;
  void* chunk[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  bool chunk_busy[2] = {false, false};
};

static Cache& cache() {
  static Cache c;
  return c;
}

static int32_t staging_init(Cache& c) {
  if (c.side) return MTL_OK;
  Runtime& rt = runtime();
  if (rt.pinned_bytes && rt.pinned_bytes < 2 * kStageChunk)
    return set_error(MTL_ERR_BUDGET, "pinned budget below the %llu-byte staging chunks",
                     (unsigned long long)(2 * kStageChunk));
  MTL_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
  for (int i = 0; i < 2; ++i) {
    MTL_HIP(hipHostMalloc(&c.chunk[i], kStageChunk, hipHostMallocDefault));
    MTL_HIP(hipEventCreateWithFlags(&c.done[i], hipEventDisableTiming));
  }
  rt.pinned_bytes_in_use += 2 * kStageChunk;
  return MTL_OK;
}

// memcpy into a pinned chunk on up to four host threads: one thread's copy
// from pageable memory (about 14 GB/s) would otherwise set the upload rate
// below what the DMA engine moves.
static void chunk_copy(void* dst, const void* src, uint64_t len) {
  constexpr uint64_t kSlice = 4ull << 20;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const uint64_t parts = std::min<uint64_t>({4, hw, (len + kSlice - 1) / kSlice});
  if (parts <= 1) {
    memcpy(dst, src, len);
    return;
  }
  const uint64_t step = (len / parts + 63) & ~63ull;
  std::vector<std::thread> th;
  for (uint64_t p = 1; p < parts; ++p) {
    const uint64_t off = p * step;
    if (off >= len) break;
    const uint64_t n = std::min(step, len - off);
    th.emplace_back([=] { memcpy((char*)dst + off, (const char*)src + off, n); });
  }
  memcpy(dst, src, std::min(step, len));
  for (auto& t : th) t.join();
}

// Host -> device through the pinned chunks (host memcpy overlapped with DMA).
static int32_t staged_upload(Cache& c, void* dev, const void* host, uint64_t bytes) {
  MTL_TRY(staging_init(c));
  const char* src = (const char*)host;
  char* dst = (char*)dev;
  int k = 0;
  for (uint64_t off = 0; off < bytes; off += kStageChunk, k ^= 1) {
    const uint64_t len = bytes - off < kStageChunk ? bytes - off : kStageChunk;
    if (c.chunk_busy[k]) MTL_HIP(hipEventSynchronize(c.done[k]));
    chunk_copy(c.chunk[k], src + off, len);
    MTL_HIP(hipMemcpyAsync(dst + off, c.chunk[k], len, hipMemcpyHostToDevice, c.side));
    MTL_HIP(hipEventRecord(c.done[k], c.side));
    c.chunk_busy[k] = true;
  }
  MTL_HIP(hipStreamSynchronize(c.side));
  c.chunk_busy[0] = c.chunk_busy[1] = false;
  c.uploaded += bytes;
  return MTL_OK;
}

static void evict(Cache& c, std::map<std::pair<uint64_t, uint32_t>, CacheEntry>::iterator it) {
  hipFree(it->second.dev);
  c.bytes -= it->second.bytes;
  runtime().device_bytes_in_use -= it->second.bytes;
  c.lru.erase(it->second.lru);
  c.entries.erase(it);
  ++c.evictions;
}

// Evict unreferenced entries, least recently used first, until `need` more
// bytes fit the budget; evicts nothing when that is impossible.
static bool make_room(Cache& c, uint64_t need, uint64_t budget) {
  if (c.bytes + need <= budget) return true;
  std::vector<std::pair<uint64_t, uint32_t>> victims;
  uint64_t freed = 0;
  for (auto pos = c.lru.rbegin(); pos != c.lru.rend() && c.bytes - freed + need > budget; ++pos) {
    const CacheEntry& e = c.entries.find(*pos)->second;
    if (e.refs) continue;
    victims.push_back(*pos);
    freed += e.bytes;
  }
  if (c.bytes - freed + need > budget) return false;
  for (const auto& k : victims) evict(c, c.entries.find(k));
  return true;
}

}  // namespace mtl

extern "C" {

int32_t mtl_cache_put(uint64_t table_id, uint32_t column_id, const void* host, uint64_t bytes,
                      uint64_t version, const void** out_dev) {
  using namespace mtl;
  if (!host || !out_dev) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  Runtime& rt = runtime();
  if (!rt.initialized)
    return set_error(MTL_ERR_NOT_INITIALIZED,
                     "not initialized: call gpu_buffer_init before gpu_processing");
  MTL_TRY(ensure_device());
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  const auto key = std::make_pair(table_id, column_id);
  auto it = c.entries.find(key);
  if (it != c.entries.end()) {
    CacheEntry& e = it->second;
    if (e.version == version && e.bytes == bytes) {
      ++c.hits;
      ++e.refs;
      c.lru.splice(c.lru.begin(), c.lru, e.lru);
      *out_dev = e.dev;
      return MTL_OK;
    }
    if (e.refs)
      return set_error(MTL_ERR_INVALID_ARGUMENT,
                       "column %llu.%u is in use and cannot be replaced",
                       (unsigned long long)table_id, column_id);
    evict(c, it);
  }
  ++c.misses;
  if (bytes > rt.cache_bytes || !make_room(c, bytes, rt.cache_bytes))
    return set_error(MTL_ERR_BUDGET,
                     "cache region full: %llu bytes needed, %llu of %llu in use by "
                     "referenced columns",
                     (unsigned long long)bytes, (unsigned long long)c.bytes,
                     (unsigned long long)rt.cache_bytes);
  void* dev = nullptr;
  MTL_HIP(hipMalloc(&dev, bytes ? bytes : 1));
  int32_t st = staged_upload(c, dev, host, bytes);
  if (st != MTL_OK) {
    hipFree(dev);
    return st;
  }
  c.lru.push_front(key);
  CacheEntry e;
  e.dev = dev;
  e.bytes = bytes;
  e.version = version;
  e.refs = 1;
  e.lru = c.lru.begin();
  c.entries.emplace(key, e);
  c.bytes += bytes;
  rt.device_bytes_in_use += bytes;
  *out_dev = dev;
  return MTL_OK;
}

int32_t mtl_cache_get(uint64_t table_id, uint32_t column_id, const void** out_dev,
                      uint64_t* bytes) {
  using namespace mtl;
  if (!out_dev) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.entries.find(std::make_pair(table_id, column_id));
  if (it == c.entries.end())
    return set_error(MTL_ERR_INVALID_ARGUMENT, "column %llu.%u is not cached",
                     (unsigned long long)table_id, column_id);
  ++c.hits;
  ++it->second.refs;
  c.lru.splice(c.lru.begin(), c.lru, it->second.lru);
  *out_dev = it->second.dev;
  if (bytes) *bytes = it->second.bytes;
  return MTL_OK;
}

int32_t mtl_cache_release(uint64_t table_id, uint32_t column_id) {
  using namespace mtl;
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.entries.find(std::make_pair(table_id, column_id));
  if (it == c.entries.end() || it->second.refs == 0)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "column %llu.%u is not referenced",
                     (unsigned long long)table_id, column_id);
  --it->second.refs;
  return MTL_OK;
}

int32_t mtl_cache_drop(uint64_t table_id) {
  using namespace mtl;
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.entries.lower_bound(std::make_pair(table_id, 0u));
  while (it != c.entries.end() && it->first.first == table_id) {
    if (it->second.refs)
      return set_error(MTL_ERR_INVALID_ARGUMENT, "table %llu has referenced columns",
                       (unsigned long long)table_id);
    auto next = std::next(it);
    evict(c, it);
    it = next;
  }
  return MTL_OK;
}

int32_t mtl_cache_stats(uint64_t* entries, uint64_t* bytes, uint64_t* hits, uint64_t* misses,
                        uint64_t* evictions, uint64_t* uploaded_bytes) {
  using namespace mtl;
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  if (entries) *entries = c.entries.size();
  if (bytes) *bytes = c.bytes;
  if (hits) *hits = c.hits;
  if (misses) *misses = c.misses;
  if (evictions) *evictions = c.evictions;
  if (uploaded_bytes) *uploaded_bytes = c.uploaded;
  return MTL_OK;
}

int32_t mtl_cache_reset(void) {
  using namespace mtl;
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  for (auto& kv : c.entries)
    if (kv.second.refs)
      return set_error(MTL_ERR_INVALID_ARGUMENT, "referenced columns remain");
  while (!c.entries.empty()) evict(c, c.entries.begin());
  c.hits = c.misses = c.evictions = c.uploaded = 0;
  return MTL_OK;
}

}  // extern "C"
