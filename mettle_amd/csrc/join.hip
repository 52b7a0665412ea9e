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
//   - matches come out in ascending probe order (match bitmask + tile counts,
//     scan, expand) instead of the atomic-counter order of compact_matches.
#include <algorithm>

#include "device.hpp"
#include "filter.hpp"
#include "join.hpp"
#include "mtl_internal.hpp"
#include "prof.hpp"
#include "rowtile.hpp"

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

__host__ __device__ inline uint64_t hash_bytes(uint64_t cap) {
  uint64_t kb = ((cap + 1) * 8 + 255) & ~255ull;
  return kb + (cap + 1) * 4;
}
// Slots for `rows` build rows at the tuned fill (T_JOIN_LOAD percent): the
// smallest power of two, at least 16, with rows <= load% of it.
inline uint64_t hash_cap_for(uint64_t rows) {
  const uint64_t load = (uint64_t)tuned(T_JOIN_LOAD);
  uint64_t c = 16;
  while (c * load < rows * 100) c <<= 1;
  return c;
}
// Slots of a table of `bytes` (what hash_cap_for gave when it was sized):
// build and probe derive it from the caller's allocation, so they agree
// whatever the tuned fill is at each call.  0: smaller than 16 slots.
inline uint64_t hash_cap_of_bytes(uint64_t bytes) {
  if (bytes < hash_bytes(16)) return 0;
  uint64_t c = 16;
  while (hash_bytes(c * 2) <= bytes) c <<= 1;
  return c;
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
  const uint64_t n = sel ? dev_count(count, cap_rows) : cap_rows;
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

// Probe in two parallel passes (no inter-block dependency):
//   k_probe_bits   : one 4096-row tile per block; bit k of the match mask is
//                    set when probe position k finds its key, plus the
//                    tile's match count;
//   (scan of the tile counts, filter.hip)
//   k_probe_expand : positions with a set bit look their key up again and
//                    write (probe row, build row) at ascending offsets.
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
  uint64_t* match_bm;
  uint32_t* tile_counts;
  uint32_t* out_probe;
  uint32_t* out_build;
};

__global__ __launch_bounds__(FBLOCK) void k_probe_bits(ProbeArgs a) {
  __shared__ uint32_t s_wave[FWAVES];
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  const uint64_t n = a.sel ? dev_count(a.count, a.cap_rows) : a.cap_rows;
  const uint64_t seg = (uint64_t)blockIdx.x * FTILE + (uint64_t)wave * (FJ * 256);
  const uint64_t k0 = seg + (uint64_t)lane * 4;
  int64_t key[FJ][4];
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      uint64_t k = k0 + (uint64_t)j * 256 + v;
      uint32_t row = k < n ? (a.sel ? a.sel[k] : (uint32_t)k) : 0u;
      key[j][v] = load_key(a.keys, a.key_type, row);
    }
  uint32_t bits = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      uint64_t k = k0 + (uint64_t)j * 256 + v;
      if (k < n && lookup(a.strategy, a.table, a.min_key, a.slots, key[j][v]) >= 0)
        bits |= 1u << (4 * j + v);
    }
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    uint64_t w = (uint64_t)((bits >> (4 * j)) & 0xFu) << (4 * (lane & 15));
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) w |= __shfl_xor(w, off, WAVE);
    if ((lane & 15) == 0) a.match_bm[(seg + (uint64_t)j * 256) / 64 + (lane >> 4)] = w;
  }
  uint32_t c = wave_sum_u32((uint32_t)__popc(bits));
  if (lane == 0) s_wave[wave] = c;
  __syncthreads();
  if (tid == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < FWAVES; ++w) t += s_wave[w];
    a.tile_counts[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(FBLOCK) void k_probe_expand(ProbeArgs a) {
  __shared__ uint32_t s_wave[FWAVES];
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  const uint64_t r0 = (uint64_t)blockIdx.x * FTILE + (uint64_t)tid * 16;
  const uint32_t m = (uint32_t)(a.match_bm[r0 / 64] >> (r0 % 64)) & 0xFFFFu;
  const uint32_t c = (uint32_t)__popc(m);
  uint32_t incl = wave_inclusive_u32(c);
  if (lane == WAVE - 1) s_wave[wave] = incl;
  __syncthreads();
  uint32_t pos = a.tile_counts[blockIdx.x] + incl - c;
#pragma unroll
  for (int w = 0; w < FWAVES; ++w)
    if (w < wave) pos += s_wave[w];
  uint32_t mm = m;
  while (mm) {
    int i = __ffs(mm) - 1;
    mm &= mm - 1;
    uint64_t k = r0 + (uint64_t)i;
    uint32_t row = a.sel ? a.sel[k] : (uint32_t)k;
    int32_t b = lookup(a.strategy, a.table, a.min_key, a.slots,
                       load_key(a.keys, a.key_type, row));
    a.out_probe[pos] = row;
    a.out_build[pos] = (uint32_t)b;
    ++pos;
  }
}

// ---------------------------------------------------------------------------
// Radix-partitioned join (MTL_JOIN_RADIX): LDS-staged partitions.  The
// build rows are partitioned by the top bits of the key's hash so that each
// partition (about kRxRows rows, at most kRxMax) fits a kRxSlots-slot hash
// table in LDS; the probe rows are partitioned the same way, and one block
// per (partition, slice of its probe rows) builds the partition's table in
// LDS and probes it there.  No probe touches a table in HBM: a large build
// costs a few streaming passes instead of two random HBM lookups per probe
// row.  The pairs come out grouped by partition (order unspecified), for
// callers that aggregate them.
//
// Table buffer: [0, 256) header (bits) | offsets (P + 1) | cursors (P) |
// keys (n int64) | rows (n int32), each part 256-byte aligned.
// ---------------------------------------------------------------------------
constexpr int RX_BLOCK = 256;
constexpr uint32_t kRxSlots = 4096;      // LDS table: 32 KB keys + 16 KB rows
constexpr uint32_t kRxMax = 3072;        // rows per partition (75 % fill)
constexpr uint64_t kRxRows = 2048;       // target rows per partition
constexpr uint32_t kRxMaxBits = 13;      // P <= 8192 (LDS histograms)

inline uint64_t al256(uint64_t b) { return (b + 255) & ~255ull; }
inline uint32_t rx_bits(uint64_t n) {
  uint32_t b = 0;
  while (b < kRxMaxBits && (kRxRows << b) < n) ++b;
  return b;
}
struct RxView {
  uint32_t bits;
  uint32_t* off;     // P + 1
  uint32_t* cursor;  // P
  int64_t* keys;
  int32_t* rows;
  int64_t* tkeys;    // first-level partitions (two-level scatter)
  int32_t* trows;
};
inline uint64_t rx_bytes(uint64_t n) {
  const uint64_t P = 1ull << rx_bits(n);
  return 256 + al256((P + 1) * 4) + al256(P * 4) + 2 * (al256(n * 8) + al256(n * 4));
}
inline RxView rx_view(void* base, uint64_t n) {
  RxView v;
  v.bits = rx_bits(n);
  const uint64_t P = 1ull << v.bits;
  char* p = (char*)base + 256;
  v.off = (uint32_t*)p;
  p += al256((P + 1) * 4);
  v.cursor = (uint32_t*)p;
  p += al256(P * 4);
  v.keys = (int64_t*)p;
  p += al256(n * 8);
  v.rows = (int32_t*)p;
  p += al256(n * 4);
  v.tkeys = (int64_t*)p;
  p += al256(n * 8);
  v.trows = (int32_t*)p;
  return v;
}

__device__ __forceinline__ uint32_t rx_part(int64_t key, uint32_t bits) {
  return bits ? (uint32_t)(fmix64((uint64_t)key) >> (64 - bits)) : 0u;
}

// Rows of a side: sel[i] (count-bounded) or i.
struct RxSide {
  const void* keys;
  int32_t type;
  const uint32_t* sel;
  const uint32_t* count;
  uint64_t cap;
};
__device__ __forceinline__ uint64_t rx_n(const RxSide& s) {
  return s.sel ? dev_count(s.count, s.cap) : s.cap;
}

// hist[p] += rows of partition p (hist cleared by the caller).
__global__ __launch_bounds__(RX_BLOCK) void k_rx_hist(RxSide s, uint32_t bits, uint32_t* hist) {
  __shared__ uint32_t h[1u << kRxMaxBits];
  const uint32_t P = 1u << bits;
  for (uint32_t i = threadIdx.x; i < P; i += RX_BLOCK) h[i] = 0;
  __syncthreads();
  const uint64_t n = rx_n(s);
  for (uint64_t i = (uint64_t)blockIdx.x * RX_BLOCK + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * RX_BLOCK) {
    const uint32_t row = s.sel ? s.sel[i] : (uint32_t)i;
    atomicAdd(&h[rx_part(load_key(s.keys, s.type, row), bits)], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < P; i += RX_BLOCK)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// One level of the partition scatter, staged through LDS so that a wave
// writes whole runs: bucket(key) = (fmix64(key) >> shift) & (nb - 1), at
// most kRxSplitMax buckets per level (8192 partitions are written in two
// levels, 64 then 128: one level of 8192 runs leaves every store on its own
// cache line).  Each block takes a range of rows: it counts its buckets,
// reserves its run in each with one atomic per bucket (cursor[base + b]),
// then, 2048 rows at a time, sorts the rows by bucket in LDS and writes
// every bucket's piece with consecutive lanes on consecutive addresses.
// Level 1 reads the side (RxSide); level 2 the first level's arrays, block
// b taking slice b % slices of first-level partition b / slices.
constexpr uint32_t kRxSplitMax = 128;
constexpr uint32_t kRxSub = 2048;

struct RxSplit {
  RxSide raw;              // level 1 input (in_keys == nullptr)
  const int64_t* in_keys;  // level 2 input
  const int32_t* in_rows;
  const uint32_t* in_off;  // level 2: first-level partition q starts at in_off[q * nb]
  uint32_t slices;         // level 2: blocks per first-level partition
  uint64_t chunk;          // level 1: rows per block
  uint32_t shift, nb;
  uint32_t* cursor;        // next free position of each bucket (base + bucket)
  int64_t* out_keys;
  int32_t* out_rows;
};

__global__ __launch_bounds__(RX_BLOCK) void k_rx_split(RxSplit a) {
  __shared__ uint32_t s_cnt[kRxSplitMax], s_cur[kRxSplitMax], s_start[kRxSplitMax];
  __shared__ int64_t s_key[kRxSub];
  __shared__ int32_t s_row[kRxSub];
  __shared__ uint8_t s_b[kRxSub];
  uint64_t lo, hi;
  uint32_t base = 0;
  if (a.in_keys) {
    const uint32_t q = blockIdx.x / a.slices, sl = blockIdx.x % a.slices;
    const uint64_t q0 = a.in_off[(uint64_t)q * a.nb], q1 = a.in_off[(uint64_t)(q + 1) * a.nb];
    lo = q0 + (q1 - q0) * sl / a.slices;
    hi = q0 + (q1 - q0) * (sl + 1) / a.slices;
    base = q * a.nb;
  } else {
    const uint64_t n = rx_n(a.raw);
    lo = (uint64_t)blockIdx.x * a.chunk;
    hi = lo + a.chunk < n ? lo + a.chunk : n;
  }
  if (lo >= hi) return;  // block-uniform
  auto fetch = [&](uint64_t i, int64_t& key, int32_t& row) {
    if (a.in_keys) {
      key = a.in_keys[i];
      row = a.in_rows[i];
    } else {
      row = (int32_t)(a.raw.sel ? a.raw.sel[i] : (uint32_t)i);
      key = load_key(a.raw.keys, a.raw.type, (uint32_t)row);
    }
  };
  auto bucket = [&](int64_t key) {
    return (uint32_t)(fmix64((uint64_t)key) >> a.shift) & (a.nb - 1);
  };
  for (uint32_t i = threadIdx.x; i < a.nb; i += RX_BLOCK) s_cnt[i] = 0;
  __syncthreads();
  for (uint64_t i = lo + threadIdx.x; i < hi; i += RX_BLOCK) {
    int64_t key;
    int32_t row;
    fetch(i, key, row);
    atomicAdd(&s_cnt[bucket(key)], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < a.nb; i += RX_BLOCK)
    s_cur[i] = s_cnt[i] ? atomicAdd(&a.cursor[base + i], s_cnt[i]) : 0u;
  __syncthreads();
  for (uint64_t t0 = lo; t0 < hi; t0 += kRxSub) {
    const uint32_t m = (uint32_t)(hi - t0 < kRxSub ? hi - t0 : kRxSub);
    for (uint32_t i = threadIdx.x; i < a.nb; i += RX_BLOCK) s_cnt[i] = 0;
    __syncthreads();
    constexpr int PER = kRxSub / RX_BLOCK;
    int64_t key[PER];
    int32_t row[PER];
    uint32_t b[PER], rank[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const uint32_t j = (uint32_t)u * RX_BLOCK + threadIdx.x;
      if (j < m) {
        fetch(t0 + j, key[u], row[u]);
        b[u] = bucket(key[u]);
        rank[u] = atomicAdd(&s_cnt[b[u]], 1u);
      }
    }
    __syncthreads();
    if (threadIdx.x < WAVE) {  // bucket starts within the staged rows (nb <= 128)
      const uint32_t c0 = threadIdx.x < a.nb ? s_cnt[threadIdx.x] : 0u;
      const uint32_t c1 = threadIdx.x + WAVE < a.nb ? s_cnt[threadIdx.x + WAVE] : 0u;
      const uint32_t i0 = wave_inclusive_u32(c0);
      const uint32_t tot0 = __shfl(i0, WAVE - 1, WAVE);
      const uint32_t i1 = wave_inclusive_u32(c1);
      if (threadIdx.x < a.nb) s_start[threadIdx.x] = i0 - c0;
      if (threadIdx.x + WAVE < a.nb) s_start[threadIdx.x + WAVE] = tot0 + i1 - c1;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const uint32_t j = (uint32_t)u * RX_BLOCK + threadIdx.x;
      if (j < m) {
        const uint32_t at = s_start[b[u]] + rank[u];
        s_key[at] = key[u];
        s_row[at] = row[u];
        s_b[at] = (uint8_t)b[u];
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < m; i += RX_BLOCK) {
      const uint32_t bb = s_b[i];
      const uint32_t at = s_cur[bb] + (i - s_start[bb]);
      a.out_keys[at] = s_key[i];
      a.out_rows[at] = s_row[i];
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < a.nb; i += RX_BLOCK) s_cur[i] += s_cnt[i];
    __syncthreads();
  }
}

// cursor[q] = off[q * stride] for q < n: the first level's cursors.
__global__ void k_rx_cursor1(const uint32_t* off, uint32_t stride, uint32_t n, uint32_t* cursor) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) cursor[q] = off[(uint64_t)q * stride];
}

// Partition side `side` (its P = 2^bits partitions' offsets `off` already
// scanned) into (keys, rows): one level up to kRxSplitMax partitions, else
// two (the first into (tkeys, trows)).  cursor: P words of scratch.
int32_t rx_partition(const RxSide& side, uint64_t cap, uint32_t bits, const uint32_t* off,
                     uint32_t* cursor, int64_t* keys, int32_t* rows, int64_t* tkeys,
                     int32_t* trows, mtl_stream* s) {
  const uint32_t P = 1u << bits;
  const uint32_t b1 = bits <= 7 ? bits : 6;  // first level: 2^b1 buckets
  const uint32_t b2 = bits - b1;
  const uint32_t grid = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(ceil_div(cap, RX_BLOCK * 16), (uint64_t)cu_count() * 4));
  const uint64_t chunk = std::max<uint64_t>(kRxSub, ceil_div(cap, grid));
  RxSplit a{};
  a.raw = side;
  a.chunk = chunk;
  if (b2 == 0) {  // one level straight into the partitions
    MTL_HIP(hipMemcpyAsync(cursor, off, P * 4ull, hipMemcpyDeviceToDevice, s->stream));
    a.shift = bits ? 64 - bits : 0;  // one partition: bucket = h & 0
    a.nb = P;
    a.cursor = cursor;
    a.out_keys = keys;
    a.out_rows = rows;
    ProfScope ps(s, "k_rx_split");
    hipLaunchKernelGGL(k_rx_split, dim3((uint32_t)ceil_div(std::max<uint64_t>(cap, 1), chunk)),
                       dim3(RX_BLOCK), 0, s->stream, a);
    MTL_HIP(hipGetLastError());
    return MTL_OK;
  }
  // level 1: the first-level partition q holds final partitions
  // [q 2^b2, (q + 1) 2^b2); its cursor is the first of them, read through
  // a stride of 2^b2 (the final cursors double as the first level's)
  const uint32_t nb2 = 1u << b2;
  a.shift = 64 - b1;
  a.nb = 1u << b1;
  a.cursor = cursor;  // bucket q uses cursor[q]: filled from off[q * nb2] below
  a.out_keys = tkeys;
  a.out_rows = trows;
  {
    hipLaunchKernelGGL(k_rx_cursor1, dim3(1), dim3(1u << b1), 0, s->stream, off, nb2, 1u << b1,
                       cursor);
    MTL_HIP(hipGetLastError());
    ProfScope ps(s, "k_rx_split");
    hipLaunchKernelGGL(k_rx_split, dim3((uint32_t)ceil_div(std::max<uint64_t>(cap, 1), chunk)),
                       dim3(RX_BLOCK), 0, s->stream, a);
    MTL_HIP(hipGetLastError());
  }
  // level 2: each first-level partition into its 2^b2 final partitions
  MTL_HIP(hipMemcpyAsync(cursor, off, P * 4ull, hipMemcpyDeviceToDevice, s->stream));
  RxSplit c{};
  c.in_keys = tkeys;
  c.in_rows = trows;
  c.in_off = off;
  c.slices = (uint32_t)std::max<uint64_t>(1, (uint64_t)cu_count() * 4 >> b1);
  c.shift = 64 - bits;
  c.nb = nb2;
  c.cursor = cursor;
  c.out_keys = keys;
  c.out_rows = rows;
  {
    ProfScope ps(s, "k_rx_split");
    hipLaunchKernelGGL(k_rx_split, dim3((1u << b1) * c.slices), dim3(RX_BLOCK), 0, s->stream, c);
    MTL_HIP(hipGetLastError());
  }
  return MTL_OK;
}

// The LDS table of one partition: kRxSlots keys (kEmpty = free; the key
// kEmpty itself lives in the side slot) and their rows.
struct RxLds {
  int64_t keys[kRxSlots];
  int32_t rows[kRxSlots];
  int32_t empty_row;
};

// Builds partition p's table in LDS; err bits: 2 duplicate key.
__device__ __forceinline__ void rx_build_lds(RxLds& t, const int64_t* keys, const int32_t* rows,
                                             uint32_t lo, uint32_t hi, uint32_t* err) {
  for (uint32_t i = threadIdx.x; i < kRxSlots; i += RX_BLOCK) t.keys[i] = kEmpty;
  if (threadIdx.x == 0) t.empty_row = -1;
  __syncthreads();
  for (uint32_t i = lo + threadIdx.x; i < hi; i += RX_BLOCK) {
    const int64_t key = keys[i];
    const int32_t row = rows[i];
    if (key == kEmpty) {
      if (atomicCAS(&t.empty_row, -1, row) != -1) atomicOr(err, 2u);
      continue;
    }
    uint32_t slot = (uint32_t)fmix64((uint64_t)key) & (kRxSlots - 1);
    for (uint32_t k = 0; k < kRxSlots; ++k) {
      const unsigned long long old = atomicCAS((unsigned long long*)&t.keys[slot],
                                               (unsigned long long)kEmpty,
                                               (unsigned long long)key);
      if ((int64_t)old == kEmpty) {
        t.rows[slot] = row;
        break;
      }
      if ((int64_t)old == key) {
        atomicOr(err, 2u);
        break;
      }
      slot = (slot + 1) & (kRxSlots - 1);
    }
  }
  __syncthreads();
}

__device__ __forceinline__ int32_t rx_lookup(const RxLds& t, int64_t key) {
  if (key == kEmpty) return t.empty_row;
  uint32_t slot = (uint32_t)fmix64((uint64_t)key) & (kRxSlots - 1);
  for (uint32_t k = 0; k < kRxSlots; ++k) {
    const int64_t x = t.keys[slot];
    if (x == key) return t.rows[slot];
    if (x == kEmpty) return -1;
    slot = (slot + 1) & (kRxSlots - 1);
  }
  return -1;
}

// One block per build partition: size check (err bit 16: a partition over
// kRxMax rows; only many repeats of one key cause it) and duplicate check.
__global__ __launch_bounds__(RX_BLOCK) void k_rx_check(const uint32_t* off, const int64_t* keys,
                                                      const int32_t* rows, uint32_t* err) {
  __shared__ RxLds t;
  const uint32_t lo = off[blockIdx.x], hi = off[blockIdx.x + 1];
  if (hi - lo > kRxMax) {
    if (threadIdx.x == 0) atomicOr(err, 16u);
    return;
  }
  rx_build_lds(t, keys, rows, lo, hi, err);
}

struct RxProbe {
  const uint32_t* boff;  // build partition offsets
  const int64_t* bkeys;
  const int32_t* brows;
  uint32_t bits;
  uint32_t slices;        // blocks per partition
  const uint32_t* poff;   // probe partition offsets (NULL: one partition, raw side)
  const int64_t* pkeys;   // partitioned probe keys
  const int32_t* prows;   // and their probe rows
  RxSide raw;             // the probe side when not partitioned
  uint32_t* out_probe;
  uint32_t* out_build;
  uint32_t* out_count;
  uint32_t* err;
};

// Matches collect in an LDS buffer and leave in runs of at least
// kRxOut - kRxStep pairs, one global reservation per run (a reservation per
// wave would serialise millions of atomics on out_count).  Each step takes
// kRxStep probe rows, eight per thread with their loads in flight together.
constexpr uint32_t kRxOut = 3072;
constexpr int kRxPer = 8;
constexpr uint32_t kRxStep = RX_BLOCK * kRxPer;

__global__ __launch_bounds__(RX_BLOCK) void k_rx_join(RxProbe a) {
  __shared__ RxLds t;
  __shared__ uint32_t s_wave[RX_BLOCK / WAVE];
  __shared__ uint32_t s_buf_p[kRxOut], s_buf_b[kRxOut];
  __shared__ uint32_t s_fill, s_base;
  const uint32_t p = blockIdx.x / a.slices, sl = blockIdx.x % a.slices;
  if (a.boff[p + 1] - a.boff[p] > kRxMax) {  // the build reported it (bit 16)
    if (threadIdx.x == 0) atomicOr(a.err, 16u);
    return;
  }
  rx_build_lds(t, a.bkeys, a.brows, a.boff[p], a.boff[p + 1], a.err);
  uint64_t lo, hi;
  if (a.poff) {
    lo = a.poff[p];
    hi = a.poff[p + 1];
  } else {
    lo = 0;
    hi = rx_n(a.raw);
  }
  const uint64_t len = hi - lo;
  const uint64_t s0 = lo + len * sl / a.slices, s1 = lo + len * (sl + 1) / a.slices;
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  if (threadIdx.x == 0) s_fill = 0;
  __syncthreads();
  auto flush = [&](uint32_t n) {  // every thread; s_fill == n
    if (threadIdx.x == 0) s_base = n ? atomicAdd(a.out_count, n) : 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += RX_BLOCK) {
      a.out_probe[s_base + i] = s_buf_p[i];
      a.out_build[s_base + i] = s_buf_b[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) s_fill = 0;
    __syncthreads();
  };
  for (uint64_t b = s0; b < s1; b += kRxStep) {  // the same iterations in every thread
    int64_t key[kRxPer];
    uint32_t prow[kRxPer];
#pragma unroll
    for (int u = 0; u < kRxPer; ++u) {
      const uint64_t i = b + (uint64_t)u * RX_BLOCK + threadIdx.x;
      key[u] = 0;
      prow[u] = 0;
      if (i < s1) {
        if (a.poff) {
          key[u] = a.pkeys[i];
          prow[u] = (uint32_t)a.prows[i];
        } else {
          prow[u] = a.raw.sel ? a.raw.sel[i] : (uint32_t)i;
          key[u] = load_key(a.raw.keys, a.raw.type, prow[u]);
        }
      }
    }
    int32_t m[kRxPer];
    uint32_t c = 0;
#pragma unroll
    for (int u = 0; u < kRxPer; ++u) {
      const uint64_t i = b + (uint64_t)u * RX_BLOCK + threadIdx.x;
      m[u] = i < s1 ? rx_lookup(t, key[u]) : -1;
      c += m[u] >= 0;
    }
    const uint32_t incl = wave_inclusive_u32(c);
    if (lane == WAVE - 1) s_wave[wave] = incl;
    __syncthreads();
    uint32_t at = s_fill + incl - c, step = 0;
#pragma unroll
    for (int w = 0; w < RX_BLOCK / WAVE; ++w) {
      at += w < wave ? s_wave[w] : 0;
      step += s_wave[w];
    }
#pragma unroll
    for (int u = 0; u < kRxPer; ++u)
      if (m[u] >= 0) {  // s_fill + step <= kRxOut: flushed below once over kRxOut - kRxStep
        s_buf_p[at] = prow[u];
        s_buf_b[at] = (uint32_t)m[u];
        ++at;
      }
    __syncthreads();
    if (threadIdx.x == 0) s_fill += step;
    __syncthreads();
    if (s_fill > kRxOut - kRxStep) flush(s_fill);
  }
  flush(s_fill);
}

int32_t rx_build(const void* keys, int32_t key_type, const uint32_t* sel, const uint32_t* count,
                 uint64_t capacity, void* table, uint32_t* err, mtl_stream* s) {
  const RxView v = rx_view(table, capacity);
  const uint32_t P = 1u << v.bits;
  MTL_HIP(hipMemsetAsync(v.off, 0, (P + 1) * 4ull, s->stream));
  RxSide side{keys, key_type, sel, sel ? count : nullptr, capacity};
  const uint32_t grid = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(ceil_div(capacity, RX_BLOCK * 16), (uint64_t)cu_count() * 4));
  {
    ProfScope ps(s, "k_rx_hist");
    hipLaunchKernelGGL(k_rx_hist, dim3(grid), dim3(RX_BLOCK), 0, s->stream, side, v.bits, v.off);
    MTL_HIP(hipGetLastError());
  }
  scan_counts(v.off, P + 1, v.off + P, s->stream);  // exclusive; off[P] = n after the scan
  MTL_HIP(hipGetLastError());
  MTL_TRY(rx_partition(side, capacity, v.bits, v.off, v.cursor, v.keys, v.rows, v.tkeys, v.trows,
                       s));
  {
    ProfScope ps(s, "k_rx_check");
    hipLaunchKernelGGL(k_rx_check, dim3(P), dim3(RX_BLOCK), 0, s->stream, v.off, v.keys, v.rows,
                       err);
    MTL_HIP(hipGetLastError());
  }
  return MTL_OK;
}

int32_t rx_probe(const void* table, uint64_t build_cap, const void* keys, int32_t key_type,
                 const uint32_t* sel, const uint32_t* count, uint64_t capacity,
                 uint32_t* out_probe, uint32_t* out_build, uint32_t* out_count, mtl_stream* s) {
  const RxView v = rx_view((void*)table, build_cap);
  const uint32_t P = 1u << v.bits;
  RxProbe a{};
  a.boff = v.off;
  a.bkeys = v.keys;
  a.brows = v.rows;
  a.bits = v.bits;
  a.raw = RxSide{keys, key_type, sel, sel ? count : nullptr, capacity};
  a.out_probe = out_probe;
  a.out_build = out_build;
  a.out_count = out_count;
  void* errp;
  MTL_TRY(s->arena.alloc(16, &errp));
  a.err = (uint32_t*)errp;
  MTL_HIP(hipMemsetAsync(out_count, 0, 4, s->stream));
  MTL_HIP(hipMemsetAsync(errp, 0, 4, s->stream));
  if (P > 1) {
    void *poff, *pcur, *pk, *pr, *tk = nullptr, *tr = nullptr;
    MTL_TRY(s->arena.alloc((P + 1) * 4ull, &poff));
    MTL_TRY(s->arena.alloc(P * 4ull, &pcur));
    MTL_TRY(s->arena.alloc(std::max<uint64_t>(capacity, 1) * 8, &pk));
    MTL_TRY(s->arena.alloc(std::max<uint64_t>(capacity, 1) * 4, &pr));
    if (v.bits > 7) {  // two levels: the first level's arrays
      MTL_TRY(s->arena.alloc(std::max<uint64_t>(capacity, 1) * 8, &tk));
      MTL_TRY(s->arena.alloc(std::max<uint64_t>(capacity, 1) * 4, &tr));
    }
    MTL_HIP(hipMemsetAsync(poff, 0, (P + 1) * 4ull, s->stream));
    const uint32_t grid = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(ceil_div(capacity, RX_BLOCK * 16), (uint64_t)cu_count() * 4));
    {
      ProfScope ps(s, "k_rx_hist");
      hipLaunchKernelGGL(k_rx_hist, dim3(grid), dim3(RX_BLOCK), 0, s->stream, a.raw, v.bits,
                         (uint32_t*)poff);
      MTL_HIP(hipGetLastError());
    }
    scan_counts((uint32_t*)poff, P + 1, (uint32_t*)poff + P, s->stream);
    MTL_HIP(hipGetLastError());
    MTL_TRY(rx_partition(a.raw, capacity, v.bits, (const uint32_t*)poff, (uint32_t*)pcur,
                         (int64_t*)pk, (int32_t*)pr, (int64_t*)tk, (int32_t*)tr, s));
    a.poff = (const uint32_t*)poff;
    a.pkeys = (const int64_t*)pk;
    a.prows = (const int32_t*)pr;
  }
  // enough blocks to fill the chip: slices of each partition's probe rows
  a.slices = (uint32_t)std::max<uint64_t>(1, (uint64_t)cu_count() * 4 / P);
  {
    ProfScope ps(s, "k_rx_join");
    hipLaunchKernelGGL(k_rx_join, dim3(P * a.slices), dim3(RX_BLOCK), 0, s->stream, a);
    MTL_HIP(hipGetLastError());
  }
  return MTL_OK;
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
  if (strategy == MTL_JOIN_RADIX) return rx_bytes(build_rows);
  return hash_bytes(hash_cap_for(build_rows));
}

int32_t mtl_join_build(int32_t strategy, const void* keys, int32_t key_type,
                       const uint32_t* sel, const uint32_t* count, uint64_t capacity,
                       int64_t min_key, uint64_t slots, void* table, uint64_t table_bytes,
                       uint32_t* dup_flag, mtl_stream* s) {
  if (strategy != MTL_JOIN_DENSE && strategy != MTL_JOIN_HASH && strategy != MTL_JOIN_RADIX)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "join strategy %d", strategy);
  if (!type_is_int(key_type)) return set_error(MTL_UNSUPPORTED, "join key type %d", key_type);
  if (!table || !dup_flag) return set_error(MTL_ERR_INVALID_ARGUMENT, "table/dup_flag NULL");
  if (strategy == MTL_JOIN_RADIX) {
    if (capacity > (kRxRows << kRxMaxBits))
      return set_error(MTL_UNSUPPORTED, "radix join build over %llu rows",
                       (unsigned long long)(kRxRows << kRxMaxBits));
    if (table_bytes < rx_bytes(capacity))
      return set_error(MTL_ERR_INVALID_ARGUMENT, "radix table needs %llu bytes",
                       (unsigned long long)rx_bytes(capacity));
    MTL_TRY(resolve_stream(s, &s));
    return rx_build(keys, key_type, sel, count, capacity, table, dup_flag, s);
  }
  if (strategy == MTL_JOIN_HASH) {
    // at least one empty slot besides the build rows
    slots = hash_cap_of_bytes(table_bytes);
    if (slots <= capacity)
      return set_error(MTL_ERR_INVALID_ARGUMENT, "hash table needs %llu bytes",
                       (unsigned long long)hash_bytes(hash_cap_for(capacity)));
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
  if (strategy != MTL_JOIN_DENSE && strategy != MTL_JOIN_HASH && strategy != MTL_JOIN_RADIX)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "join strategy %d", strategy);
  if (!type_is_int(key_type)) return set_error(MTL_UNSUPPORTED, "join key type %d", key_type);
  if (!out_probe || !out_build || !out_count)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "outputs are NULL");
  if (capacity > 0xFFFFFFFFull) return set_error(MTL_UNSUPPORTED, "more than 2^32-1 rows");
  if (strategy == MTL_JOIN_RADIX) {
    // `slots` is the build capacity passed to mtl_join_build
    if (slots > (kRxRows << kRxMaxBits) || table_bytes < rx_bytes(slots))
      return set_error(MTL_ERR_INVALID_ARGUMENT, "radix table too small");
    MTL_TRY(resolve_stream(s, &s));
    std::lock_guard<std::mutex> g(s->mu);
    s->arena.reset();
    return rx_probe(table, slots, probe_keys, key_type, sel, count, capacity, out_probe,
                    out_build, out_count, s);
  }
  if (strategy == MTL_JOIN_HASH) {
    // the slots the build used (from the same allocation); `slots` is the
    // build capacity, which must have left an empty slot
    const uint64_t cap = hash_cap_of_bytes(table_bytes);
    if (cap <= slots) return set_error(MTL_ERR_INVALID_ARGUMENT, "hash table too small");
    slots = cap;
  } else if (table_bytes < slots * 4) {
    return set_error(MTL_ERR_INVALID_ARGUMENT, "dense table too small");
  }
  MTL_TRY(resolve_stream(s, &s));
  std::lock_guard<std::mutex> g(s->mu);
  s->arena.reset();
  uint32_t tiles = filter_tiles(capacity);
  void *bm, *counts;
  MTL_TRY(s->arena.alloc((uint64_t)tiles * kTileRows / 8, &bm));
  MTL_TRY(s->arena.alloc((uint64_t)tiles * 4, &counts));
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
  a.match_bm = (uint64_t*)bm;
  a.tile_counts = (uint32_t*)counts;
  a.out_probe = out_probe;
  a.out_build = out_build;
  {
    ProfScope ps(s, "k_probe");
    hipLaunchKernelGGL(k_probe_bits, dim3(tiles), dim3(FBLOCK), 0, s->stream, a);
    MTL_HIP(hipGetLastError());
    scan_counts((uint32_t*)counts, tiles, out_count, s->stream);
    MTL_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_probe_expand, dim3(tiles), dim3(FBLOCK), 0, s->stream, a);
  }
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Many-to-many joins (keys repeating on both sides).  The reference's build
// overwrites a repeated key (kernels.metal:202-206), losing matches; here the
// build side is sorted by key (stable), its runs of equal keys become a
// unique-key table (run start, key), the probe side joins that table, and
// each (probe, run) pair expands to one pair per build row of the run.
// ---------------------------------------------------------------------------
namespace mtl {
using namespace dev;

constexpr int MM_BLOCK = 256;
constexpr int MM_STEPS = 16;
constexpr uint64_t MM_TILE = (uint64_t)MM_BLOCK * MM_STEPS;

__device__ __forceinline__ bool run_head(const void* keys, int32_t type, uint64_t i, uint64_t n) {
  return i < n && (i == 0 || load_key(keys, type, i) != load_key(keys, type, i - 1));
}

// Exclusive rank of this thread's flag among the block's flags of this step;
// *step_total receives the step's count.  Every thread of the block calls it.
__device__ __forceinline__ uint32_t block_rank(bool flag, uint32_t* s_wave, uint32_t* step_total) {
  const int tid = threadIdx.x, wave = tid / WAVE;
  const uint64_t bal = __ballot(flag);
  if (lane_id() == 0) s_wave[wave] = (uint32_t)__popcll(bal);
  __syncthreads();
  uint32_t before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < MM_BLOCK / WAVE; ++w) {
    before += w < wave ? s_wave[w] : 0;
    total += s_wave[w];
  }
  __syncthreads();
  *step_total = total;
  return before + (uint32_t)__popcll(bal & lanemask_lt());
}

__global__ __launch_bounds__(MM_BLOCK) void k_mm_head_count(const void* keys, int32_t type,
                                                           uint64_t n, uint32_t* tile_cnt) {
  __shared__ uint32_t s_wave[MM_BLOCK / WAVE];
  const uint64_t base = blockIdx.x * MM_TILE;
  uint32_t c = 0;
  for (int it = 0; it < MM_STEPS; ++it) c += run_head(keys, type, base + it * MM_BLOCK + threadIdx.x, n);
  c = wave_sum_u32(c);
  if (lane_id() == 0) s_wave[threadIdx.x / WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < MM_BLOCK / WAVE; ++w) t += s_wave[w];
    tile_cnt[blockIdx.x] = t;
  }
}

// run_start[r] = first sorted position of run r, run_key[r] = its key;
// run_start[runs] = n (runs = *n_runs, the scanned total).
__global__ __launch_bounds__(MM_BLOCK) void k_mm_head_write(const void* keys, int32_t type,
                                                           uint64_t n, const uint32_t* tile_off,
                                                           const uint32_t* n_runs,
                                                           uint32_t* run_start, int64_t* run_key) {
  __shared__ uint32_t s_wave[MM_BLOCK / WAVE];
  const uint64_t base = blockIdx.x * MM_TILE;
  uint32_t off = tile_off[blockIdx.x];
  for (int it = 0; it < MM_STEPS; ++it) {
    const uint64_t i = base + it * MM_BLOCK + threadIdx.x;
    const bool h = run_head(keys, type, i, n);
    uint32_t step = 0;
    const uint32_t r = off + block_rank(h, s_wave, &step);
    if (h) {
      run_start[r] = (uint32_t)i;
      run_key[r] = load_key(keys, type, i);
    }
    off += step;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) run_start[*n_runs] = (uint32_t)n;
}

__device__ __forceinline__ uint32_t run_len(const uint32_t* run_start, uint32_t r) {
  return run_start[r + 1] - run_start[r];
}

// Per tile of pairs: the number of build rows they expand to.
__global__ __launch_bounds__(MM_BLOCK) void k_mm_pair_count(const uint32_t* pair_run, uint64_t np,
                                                           const uint32_t* run_start,
                                                           uint32_t* tile_sum,
                                                           unsigned long long* total) {
  __shared__ unsigned long long s_wave[MM_BLOCK / WAVE];
  const uint64_t base = blockIdx.x * MM_TILE;
  unsigned long long c = 0;
  for (int it = 0; it < MM_STEPS; ++it) {
    const uint64_t j = base + it * MM_BLOCK + threadIdx.x;
    if (j < np) c += run_len(run_start, pair_run[j]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, WAVE);
  if (lane_id() == 0) s_wave[threadIdx.x / WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < MM_BLOCK / WAVE; ++w) t += s_wave[w];
    tile_sum[blockIdx.x] = t > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)t;
    atomicAdd(total, t);
  }
}

// off[j] = output position of pair j's first row (exclusive prefix).
__global__ __launch_bounds__(MM_BLOCK) void k_mm_offsets(const uint32_t* pair_run, uint64_t np,
                                                        const uint32_t* run_start,
                                                        const uint32_t* tile_off, uint32_t* off) {
  __shared__ uint32_t s_wave[MM_BLOCK / WAVE];
  const int wave = threadIdx.x / WAVE;
  const uint64_t base = blockIdx.x * MM_TILE;
  uint32_t acc = tile_off[blockIdx.x];
  for (int it = 0; it < MM_STEPS; ++it) {
    const uint64_t j = base + it * MM_BLOCK + threadIdx.x;
    const uint32_t c = j < np ? run_len(run_start, pair_run[j]) : 0;
    const uint32_t inc = wave_inclusive_u32(c);
    if (lane_id() == WAVE - 1) s_wave[wave] = inc;
    __syncthreads();
    uint32_t before = 0, step = 0;
#pragma unroll
    for (int w = 0; w < MM_BLOCK / WAVE; ++w) {
      before += w < wave ? s_wave[w] : 0;
      step += s_wave[w];
    }
    __syncthreads();
    if (j < np) off[j] = acc + before + inc - c;
    acc += step;
  }
}

// One thread per output row: its pair is the last j with off[j] <= t (every
// pair expands to at least one row, so a block's MM_BLOCK rows span at most
// MM_BLOCK pairs, found from the block's first row).
__global__ __launch_bounds__(MM_BLOCK) void k_mm_expand(const uint32_t* pair_probe,
                                                       const uint32_t* pair_run, uint64_t np,
                                                       const uint32_t* off,
                                                       const uint32_t* run_start,
                                                       const uint32_t* sorted_rows, uint64_t total,
                                                       uint32_t* out_probe, uint32_t* out_build) {
  __shared__ uint64_t s_lo;
  for (uint64_t b0 = (uint64_t)blockIdx.x * MM_BLOCK; b0 < total;
       b0 += (uint64_t)gridDim.x * MM_BLOCK) {
    if (threadIdx.x == 0) {
      uint64_t lo = 0, hi = np;  // last j with off[j] <= b0
      while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (off[mid] <= b0) lo = mid;
        else hi = mid;
      }
      s_lo = lo;
    }
    __syncthreads();
    const uint64_t t = b0 + threadIdx.x;
    if (t < total) {
      uint64_t lo = s_lo, hi = std::min<uint64_t>(np, s_lo + MM_BLOCK + 1);
      while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (off[mid] <= t) lo = mid;
        else hi = mid;
      }
      const uint32_t r = pair_run[lo];
      out_probe[t] = pair_probe[lo];
      out_build[t] = sorted_rows[run_start[r] + (uint32_t)(t - off[lo])];
    }
    __syncthreads();
  }
}

int32_t join_runs(const void* sorted_keys, int32_t type, uint64_t n, uint32_t* tile_buf,
                  uint32_t* run_start, int64_t* run_key, uint32_t* n_runs, hipStream_t st) {
  if (n == 0 || n > 0xFFFFFFFEull) return set_error(MTL_ERR_INVALID_ARGUMENT, "join runs: %llu rows", (unsigned long long)n);
  const uint32_t tiles = (uint32_t)((n + MM_TILE - 1) / MM_TILE);
  hipLaunchKernelGGL(k_mm_head_count, dim3(tiles), dim3(MM_BLOCK), 0, st, sorted_keys, type, n,
                     tile_buf);
  MTL_HIP(hipGetLastError());
  scan_counts(tile_buf, tiles, n_runs, st);
  MTL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mm_head_write, dim3(tiles), dim3(MM_BLOCK), 0, st, sorted_keys, type, n,
                     tile_buf, n_runs, run_start, run_key);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

int32_t join_expand_count(const uint32_t* pair_run, uint64_t np, const uint32_t* run_start,
                          uint32_t* tile_buf, unsigned long long* total, hipStream_t st) {
  if (np == 0) return MTL_OK;
  const uint32_t tiles = (uint32_t)((np + MM_TILE - 1) / MM_TILE);
  hipLaunchKernelGGL(k_mm_pair_count, dim3(tiles), dim3(MM_BLOCK), 0, st, pair_run, np,
                     run_start, tile_buf, total);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

int32_t join_expand_write(const uint32_t* pair_probe, const uint32_t* pair_run, uint64_t np,
                          const uint32_t* run_start, const uint32_t* sorted_rows,
                          uint32_t* tile_buf, uint32_t* scan_total, uint32_t* off,
                          uint64_t total, uint32_t* out_probe, uint32_t* out_build,
                          hipStream_t st) {
  if (np == 0 || total == 0) return MTL_OK;
  const uint32_t tiles = (uint32_t)((np + MM_TILE - 1) / MM_TILE);
  scan_counts(tile_buf, tiles, scan_total, st);
  MTL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mm_offsets, dim3(tiles), dim3(MM_BLOCK), 0, st, pair_run, np, run_start,
                     tile_buf, off);
  MTL_HIP(hipGetLastError());
  const uint64_t blocks = std::min<uint64_t>((total + MM_BLOCK - 1) / MM_BLOCK, 1u << 16);
  hipLaunchKernelGGL(k_mm_expand, dim3((uint32_t)blocks), dim3(MM_BLOCK), 0, st, pair_probe,
                     pair_run, np, off, run_start, sorted_rows, total, out_probe, out_build);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

}  // namespace mtl
