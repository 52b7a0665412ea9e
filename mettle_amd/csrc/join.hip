// Inner PK/FK equi-join: build + probe with stable (probe-ordered) output.
//
// Reference: hash_join_build/probe_i32_v (kernels.metal:786-837), dense
// lookup build/probe (:839-907), compact_matches_v (:769-784); host wrappers
// join_ops.rs:22-343 and the strategy rule choose_gpu_join_strategy
// (join_plan.rs:29-56).  Differences by design:
//   - keys are 64-bit (INTEGER keys are sign-extended), so BIGINT joins are
//     exact; the reference truncated BIGINT to i32 (runtime/mod.rs:424-431);
//   - the hash is a 64-bit finaliser (murmur3 fmix64) instead of the identity
//     `key & mask`, and EMPTY is INT64_MIN with a side slot for that key, so
//     no valid key collides with the sentinel (kernels.metal:187-193 used
//     0xFFFFFFFF, which is also the key -1);
//   - duplicate build keys are reported, not overwritten (:202-206);
//   - probe + match compaction are one pass with ascending probe order
//     (ballot ranks + decoupled look-back) instead of an atomic counter.
#include <algorithm>

#include "device.hpp"
#include "filter.hpp"
#include "mtl_internal.hpp"
#include "prof.hpp"

namespace mtl {
using namespace dev;

constexpr int64_t kEmpty = (int64_t)0x8000000000000000ull;

__device__ __forceinline__ uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

struct HashTable {
  int64_t* keys;  // cap + 1 entries; [cap] is the side slot for kEmpty
  int32_t* rows;  // cap + 1 entries
  uint64_t cap;   // power of two
};

__host__ __device__ inline uint64_t hash_cap_for(uint64_t rows) {
  uint64_t want = rows * 2 > 16 ? rows * 2 : 16, c = 1;
  while (c < want) c <<= 1;
  return c;
}
__host__ __device__ inline uint64_t hash_bytes(uint64_t cap) {
  uint64_t kb = ((cap + 1) * 8 + 255) & ~255ull;
  return kb + (cap + 1) * 4;
}
__host__ __device__ inline HashTable hash_view(void* base, uint64_t cap) {
  HashTable h;
  h.cap = cap;
  h.keys = (int64_t*)base;
  h.rows = (int32_t*)((char*)base + (((cap + 1) * 8 + 255) & ~255ull));
  return h;
}

__global__ void k_fill_table(int32_t strategy, void* table, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (strategy == MTL_JOIN_DENSE) {
      ((int32_t*)table)[i] = -1;
    } else {
      HashTable h = hash_view(table, n - 1);
      h.keys[i] = kEmpty;
      h.rows[i] = -1;
    }
  }
}

// err bits: 2 duplicate key, 8 key outside [min, min + slots)
__global__ void k_build(int32_t strategy, const void* keys, int32_t key_type,
                        const uint32_t* sel, const uint32_t* count, uint64_t cap_rows,
                        int64_t min_key, uint64_t slots, void* table, uint32_t* err) {
  const uint64_t n = sel ? (count ? (uint64_t)*count : cap_rows) : cap_rows;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t row = sel ? sel[i] : (uint32_t)i;
    int64_t key = load_key(keys, key_type, row);
    if (strategy == MTL_JOIN_DENSE) {
      uint64_t s = key_off(key, min_key);
      if (s >= slots) {
        atomicOr(err, 8u);
        continue;
      }
      int32_t old = atomicCAS((int32_t*)table + s, -1, (int32_t)row);
      if (old != -1) atomicOr(err, 2u);
    } else {
      HashTable h = hash_view(table, slots);
      if (key == kEmpty) {
        int32_t old = atomicCAS(&h.rows[h.cap], -1, (int32_t)row);
        if (old != -1) atomicOr(err, 2u);
        continue;
      }
      uint64_t mask = h.cap - 1, slot = fmix64((uint64_t)key) & mask;
      for (uint64_t probe = 0; probe < h.cap; ++probe) {
        unsigned long long old = atomicCAS((unsigned long long*)&h.keys[slot],
                                           (unsigned long long)kEmpty,
                                           (unsigned long long)key);
        if ((int64_t)old == kEmpty) {
          h.rows[slot] = (int32_t)row;
          break;
        }
        if ((int64_t)old == key) {
          atomicOr(err, 2u);
          break;
        }
        slot = (slot + 1) & mask;
      }
    }
  }
}

__device__ __forceinline__ int32_t lookup(int32_t strategy, const void* table,
                                          int64_t min_key, uint64_t slots, int64_t key) {
  if (strategy == MTL_JOIN_DENSE) {
    uint64_t s = key_off(key, min_key);
    return s < slots ? ((const int32_t*)table)[s] : -1;
  }
  HashTable h = hash_view((void*)table, slots);
  if (key == kEmpty) return h.rows[h.cap];
  uint64_t mask = h.cap - 1, slot = fmix64((uint64_t)key) & mask;
  for (uint64_t probe = 0; probe < h.cap; ++probe) {
    int64_t k = h.keys[slot];
    if (k == key) return h.rows[slot];
    if (k == kEmpty) return -1;
    slot = (slot + 1) & mask;
  }
  return -1;
}

// Probe + stable compaction: tile = 4 waves x PJ steps x 64 lanes x 4 rows.
constexpr int PBLOCK = 256;
constexpr int PJ = 2;
constexpr int PTILE = PBLOCK * PJ * 4;

struct ProbeArgs {
  int32_t strategy;
  const void* table;
  int64_t min_key;
  uint64_t slots;
  const void* keys;
  int32_t key_type;
  const uint32_t* sel;
  const uint32_t* count;
  uint64_t cap_rows;
  uint32_t* out_probe;
  uint32_t* out_build;
  uint32_t* out_count;
  uint64_t* status;
  uint32_t* ctr;
  uint32_t* err;
};

__global__ __launch_bounds__(PBLOCK) void k_probe(ProbeArgs a) {
  __shared__ uint32_t s_tile, s_prefix;
  __shared__ uint32_t s_wave[PBLOCK / WAVE];
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  if (tid == 0) s_tile = atomicAdd(a.ctr, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t n = a.sel ? (a.count ? (uint64_t)*a.count : a.cap_rows) : a.cap_rows;
  const uint64_t k0 = (uint64_t)tile * PTILE + (uint64_t)wave * (PJ * 256) + (uint64_t)lane * 4;
  uint32_t rid[PJ][4];
  int32_t bid[PJ][4];
  uint32_t bits = 0;
#pragma unroll
  for (int j = 0; j < PJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      uint64_t k = k0 + (uint64_t)j * 256 + v;
      rid[j][v] = 0;
      bid[j][v] = -1;
      if (k < n) {
        uint32_t row = a.sel ? a.sel[k] : (uint32_t)k;
        rid[j][v] = row;
        int32_t b = lookup(a.strategy, a.table, a.min_key, a.slots,
                           load_key(a.keys, a.key_type, row));
        bid[j][v] = b;
        if (b >= 0) bits |= 1u << (4 * j + v);
      }
    }
  const uint64_t lt = lanemask_lt();
  uint32_t lane_base[PJ], wave_total = 0;
#pragma unroll
  for (int j = 0; j < PJ; ++j) {
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      uint64_t b = __ballot((bits >> (4 * j + v)) & 1u);
      before += __popcll(b & lt);
      tot += __popcll(b);
    }
    lane_base[j] = wave_total + before;
    wave_total += tot;
  }
  if (lane == 0) s_wave[wave] = wave_total;
  __syncthreads();
  uint32_t woff = 0, total = 0;
#pragma unroll
  for (int w = 0; w < PBLOCK / WAVE; ++w) {
    if (w < wave) woff += s_wave[w];
    total += s_wave[w];
  }
  if (wave == 0) {
    uint32_t ex = lookback_wave(a.status, tile, total, a.err);
    if (lane == 0) s_prefix = ex;
  }
  __syncthreads();
  const uint32_t base = s_prefix + woff;
#pragma unroll
  for (int j = 0; j < PJ; ++j) {
    uint32_t pos = base + lane_base[j];
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if ((bits >> (4 * j + v)) & 1u) {
        a.out_probe[pos] = rid[j][v];
        a.out_build[pos] = (uint32_t)bid[j][v];
        ++pos;
      }
  }
  if (tile == gridDim.x - 1 && tid == 0) *a.out_count = s_prefix + total;
}

static uint32_t grid_rows(uint64_t n) {
  uint64_t b = ceil_div(n ? n : 1, 256 * 4);
  uint64_t cap = (uint64_t)cu_count() * 16;
  return (uint32_t)std::max<uint64_t>(1, std::min(b, cap));
}

}  // namespace mtl

using namespace mtl;

extern "C" {

uint64_t mtl_join_table_bytes(int32_t strategy, uint64_t build_rows, uint64_t dense_slots) {
  if (strategy == MTL_JOIN_DENSE) return dense_slots * 4;
  return hash_bytes(hash_cap_for(build_rows));
}

int32_t mtl_join_build(int32_t strategy, const void* keys, int32_t key_type,
                       const uint32_t* sel, const uint32_t* count, uint64_t capacity,
                       int64_t min_key, uint64_t slots, void* table, uint64_t table_bytes,
                       uint32_t* dup_flag, mtl_stream* s) {
  if (strategy != MTL_JOIN_DENSE && strategy != MTL_JOIN_HASH)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "join strategy %d", strategy);
  if (!type_is_int(key_type)) return set_error(MTL_UNSUPPORTED, "join key type %d", key_type);
  if (!table || !dup_flag) return set_error(MTL_ERR_INVALID_ARGUMENT, "table/dup_flag NULL");
  if (strategy == MTL_JOIN_HASH) {
    slots = hash_cap_for(capacity);
    if (table_bytes < hash_bytes(slots))
      return set_error(MTL_ERR_INVALID_ARGUMENT, "hash table needs %llu bytes",
                       (unsigned long long)hash_bytes(slots));
  } else if (table_bytes < slots * 4) {
    return set_error(MTL_ERR_INVALID_ARGUMENT, "dense table needs %llu bytes",
                     (unsigned long long)(slots * 4));
  }
  MTL_TRY(resolve_stream(s, &s));
  uint64_t fill_n = strategy == MTL_JOIN_DENSE ? slots : slots + 1;
  {
    ProfScope ps(s, "k_fill_table");
    hipLaunchKernelGGL(k_fill_table, dim3(grid_rows(fill_n)), dim3(256), 0, s->stream, strategy,
                       table, fill_n);
  }
  MTL_HIP(hipGetLastError());
  if (capacity == 0) return MTL_OK;
  {
    ProfScope ps(s, "k_build");
    hipLaunchKernelGGL(k_build, dim3(grid_rows(capacity)), dim3(256), 0, s->stream, strategy,
                       keys, key_type, sel, count, capacity, min_key, slots, table, dup_flag);
  }
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

int32_t mtl_join_probe(int32_t strategy, const void* table, uint64_t table_bytes,
                       int64_t min_key, uint64_t slots, const void* probe_keys,
                       int32_t key_type, const uint32_t* sel, const uint32_t* count,
                       uint64_t capacity, uint32_t* out_probe, uint32_t* out_build,
                       uint32_t* out_count, mtl_stream* s) {
  if (strategy != MTL_JOIN_DENSE && strategy != MTL_JOIN_HASH)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "join strategy %d", strategy);
  if (!type_is_int(key_type)) return set_error(MTL_UNSUPPORTED, "join key type %d", key_type);
  if (!out_probe || !out_build || !out_count)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "outputs are NULL");
  if (capacity > 0xFFFFFFFFull) return set_error(MTL_UNSUPPORTED, "more than 2^32-1 rows");
  if (strategy == MTL_JOIN_HASH) {
    // `slots` is the build capacity the table was built for
    uint64_t cap = hash_cap_for(slots);
    if (hash_bytes(cap) > table_bytes)
      return set_error(MTL_ERR_INVALID_ARGUMENT, "hash table too small");
    slots = cap;
  } else if (table_bytes < slots * 4) {
    return set_error(MTL_ERR_INVALID_ARGUMENT, "dense table too small");
  }
  MTL_TRY(resolve_stream(s, &s));
  std::lock_guard<std::mutex> g(s->mu);
  s->arena.reset();
  uint32_t tiles = (uint32_t)ceil_div(capacity ? capacity : 1, PTILE);
  void *status, *misc;
  MTL_TRY(s->arena.alloc((uint64_t)tiles * 8, &status));
  MTL_TRY(s->arena.alloc(256, &misc));
  ZeroList zl;
  zl.add((uint32_t*)status, (uint64_t)tiles * 2);
  zl.add((uint32_t*)misc, 64);
  MTL_TRY(launch_zero(zl, s->stream));
  ProbeArgs a{};
  a.strategy = strategy;
  a.table = table;
  a.min_key = min_key;
  a.slots = slots;
  a.keys = probe_keys;
  a.key_type = key_type;
  a.sel = sel;
  a.count = sel ? count : nullptr;
  a.cap_rows = capacity;
  a.out_probe = out_probe;
  a.out_build = out_build;
  a.out_count = out_count;
  a.status = (uint64_t*)status;
  a.ctr = (uint32_t*)misc;
  a.err = (uint32_t*)misc + 1;
  {
    ProfScope ps(s, "k_probe");
    hipLaunchKernelGGL(k_probe, dim3(tiles), dim3(PBLOCK), 0, s->stream, a);
  }
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

}  // extern "C"
