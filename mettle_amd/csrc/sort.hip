// Device ORDER BY: stable sort of a selection by one key column.
//
// Replaces the CPU comparator sort of the output batch, sort_record_batch
// (helpers.rs:627-667), for the ORDER BY of the scan path.  The keys are
// gathered through the selection and mapped to order-preserving unsigned
// integers (signed integers: sign bit flipped; floats: DuckDB's total order
// with NaN largest and -0 == 0; DESC: complemented), then sorted as (key,
// row id) pairs by the LSD radix sort below, which is stable: equal keys
// keep the selection's (ascending row) order, so the result is exactly
// ORDER BY key [DESC], row.
//
// The radix sort (8-bit digits, 4 passes for 4-byte keys, 8 for 8-byte):
//   k_rs_digits  : one read of the keys counts every digit position's 256
//                  values; a pass whose digit is the same for every key
//                  changes nothing and is skipped (dates, narrow integers
//                  and small tables skip most of their passes);
//   per pass:
//   k_rs_upsweep : each block counts the digit over its fixed range of
//                  tiles (a digit-major table of per-block counts);
//   k_rs_offsets : the output offset of each (digit, block): the keys of
//                  the smaller digits plus the digit's counts in earlier
//                  blocks (one block per digit scans its row);
//   k_rs_scatter : each block walks its tiles in order; a wave ranks its 64
//                  keys per step stably (the lanes with the same digit from
//                  eight ballots, plus a per-wave LDS count per digit), the
//                  tile is staged in LDS ordered by digit, and each digit's
//                  run is written out with consecutive threads on
//                  consecutive addresses.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "device.hpp"
#include "mtl_internal.hpp"
#include "prof.hpp"

namespace mtl {
using namespace dev;

template <typename K>
__global__ void k_sort_keys(const void* col, int32_t type, const uint32_t* sel,
                            const uint32_t* count, uint64_t cap, int32_t desc, K* keys,
                            uint32_t* rows) {
  const uint64_t n = dev_count(count, cap);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = sel ? sel[i] : (uint32_t)i;
    K e;
    if constexpr (sizeof(K) == 4)  // 4-byte integer types: sign bit flipped
      e = (uint32_t)(int32_t)load_key(col, type, r) ^ 0x80000000u;
    else
      e = enc_value(col, type, r);
    keys[i] = desc ? (K)~e : e;
    rows[i] = r;
  }
}

constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / WAVE;
constexpr int RS_STEPS = 16;                          // 64-key steps per wave and tile
constexpr int RS_TILE = RS_THREADS * RS_STEPS;       // 4096 keys
constexpr int RS_BINS = 256;
constexpr uint32_t RS_MAX_BLOCKS = 2048;

// Counts of every digit position: hist[p * 256 + digit p of key].
template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_digits(const K* __restrict__ keys, uint64_t n,
                                                          uint32_t* hist) {
  constexpr int P = (int)sizeof(K);
  __shared__ uint32_t h[P * RS_BINS];
  for (int i = threadIdx.x; i < P * RS_BINS; i += RS_THREADS) h[i] = 0;
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * RS_THREADS + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * RS_THREADS) {
    const K k = keys[i];
#pragma unroll
    for (int p = 0; p < P; ++p) atomicAdd(&h[p * RS_BINS + (uint32_t)((k >> (8 * p)) & 0xFF)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P * RS_BINS; i += RS_THREADS)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// Block b's digit counts over keys [b * per, min(n, (b + 1) * per)), stored
// digit-major: cnt[digit * nb + b].
template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_upsweep(const K* __restrict__ keys, uint64_t n,
                                                           uint32_t shift, uint64_t per,
                                                           uint32_t* cnt) {
  __shared__ uint32_t h[RS_WAVES][RS_BINS];
  const int wave = threadIdx.x / WAVE;
  for (int i = threadIdx.x; i < RS_WAVES * RS_BINS; i += RS_THREADS) (&h[0][0])[i] = 0;
  __syncthreads();
  const uint64_t b0 = (uint64_t)blockIdx.x * per, b1 = b0 + per < n ? b0 + per : n;
  for (uint64_t i = b0 + threadIdx.x; i < b1; i += RS_THREADS)
    atomicAdd(&h[wave][(uint32_t)((keys[i] >> shift) & 0xFF)], 1u);
  __syncthreads();
  const uint32_t d = threadIdx.x;  // RS_THREADS == RS_BINS: one digit per thread
  uint32_t t = 0;
#pragma unroll
  for (int w = 0; w < RS_WAVES; ++w) t += h[w][d];
  cnt[(uint64_t)d * gridDim.x + blockIdx.x] = t;
}

// Output offsets in place: cnt[d * nb + b] becomes the number of keys of
// every smaller digit (dh: this pass's digit totals) plus digit d's count
// in the blocks before b.  One block per digit, each scanning its row of nb
// counts with coalesced loads (a single block over the whole digit-major
// table took 0.6 ms at 2048 blocks).
__global__ __launch_bounds__(1024) void k_rs_offsets(uint32_t* cnt, uint32_t nb,
                                                     const uint32_t* __restrict__ dh) {
  __shared__ uint32_t s_wave[1024 / WAVE];
  const uint32_t d = blockIdx.x;
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  uint32_t v = (uint32_t)tid < d ? dh[tid] : 0u;  // d < 256 <= 1024 threads
  v = wave_sum_u32(v);
  if (lane == 0) s_wave[wave] = v;
  __syncthreads();
  uint32_t carry = 0;
#pragma unroll
  for (int w = 0; w < 1024 / WAVE; ++w) carry += s_wave[w];
  __syncthreads();
  uint32_t* row = cnt + (uint64_t)d * nb;
  for (uint32_t c0 = 0; c0 < nb; c0 += 1024) {
    const uint32_t i = c0 + (uint32_t)tid;
    const uint32_t x = i < nb ? row[i] : 0u;
    const uint32_t incl = wave_inclusive_u32(x);
    if (lane == WAVE - 1) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = 0, chunk = 0;
#pragma unroll
    for (int w = 0; w < 1024 / WAVE; ++w) {
      before += w < wave ? s_wave[w] : 0u;
      chunk += s_wave[w];
    }
    if (i < nb) row[i] = carry + before + incl - x;
    carry += chunk;
    __syncthreads();
  }
}

template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const K* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin,
                                                           K* __restrict__ kout,
                                                           uint32_t* __restrict__ vout, uint64_t n,
                                                           uint32_t shift, uint64_t per,
                                                           const uint32_t* __restrict__ off) {
  __shared__ K s_key[RS_TILE];
  __shared__ uint32_t s_val[RS_TILE];
  __shared__ uint32_t s_wcnt[RS_WAVES][RS_BINS];  // per wave: keys of each digit so far
  __shared__ uint32_t s_run[RS_BINS];             // the block's next output offset per digit
  __shared__ uint32_t s_start[RS_BINS];           // digit's first staged position in the tile
  __shared__ uint32_t s_scan[RS_WAVES];
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  const uint64_t lt = lanemask_lt();
  s_run[tid] = off[(uint64_t)tid * gridDim.x + blockIdx.x];
  const uint64_t b0 = (uint64_t)blockIdx.x * per, b1 = b0 + per < n ? b0 + per : n;
  for (uint64_t t0 = b0; t0 < b1; t0 += RS_TILE) {
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) s_wcnt[w][tid] = 0;
    __syncthreads();
    // this wave's 16 steps of 64 consecutive keys
    K key[RS_STEPS];
    uint32_t val[RS_STEPS], rank[RS_STEPS];
    const uint64_t w0 = t0 + (uint64_t)wave * (RS_STEPS * WAVE);
#pragma unroll
    for (int s = 0; s < RS_STEPS; ++s) {
      const uint64_t i = w0 + (uint64_t)s * WAVE + lane;
      key[s] = i < b1 ? kin[i] : (K)0;
      val[s] = i < b1 ? vin[i] : 0u;
    }
#pragma unroll
    for (int s = 0; s < RS_STEPS; ++s) {
      const uint64_t i = w0 + (uint64_t)s * WAVE + lane;
      const bool live = i < b1;
      const uint32_t d = (uint32_t)((key[s] >> shift) & 0xFF);
      // lanes (live) whose digit equals this lane's, from one ballot per bit
      uint64_t same = __ballot(live);
#pragma unroll
      for (int bit = 0; bit < 8; ++bit) {
        const uint64_t bb = __ballot((d >> bit) & 1u);
        same &= ((d >> bit) & 1u) ? bb : ~bb;
      }
      const uint32_t before = s_wcnt[wave][d];  // every lane reads before the update
      rank[s] = before + (uint32_t)__popcll(same & lt);
      // the highest lane of each digit group adds the group's size
      const bool last = live && (same >> lane) == 1ull;
      if (last) s_wcnt[wave][d] = before + (uint32_t)__popcll(same);
    }
    __syncthreads();
    // per digit (one per thread): offsets of the waves, the tile's total and
    // the digit's first position in the staged tile
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
      const uint32_t c = s_wcnt[w][tid];
      s_wcnt[w][tid] = tot;
      tot += c;
    }
    const uint32_t incl = wave_inclusive_u32(tot);
    if (lane == WAVE - 1) s_scan[wave] = incl;
    __syncthreads();
    uint32_t start = incl - tot;
    for (int w = 0; w < wave; ++w) start += s_scan[w];
    s_start[tid] = start;
    __syncthreads();
    // stage the tile ordered by digit
#pragma unroll
    for (int s = 0; s < RS_STEPS; ++s) {
      const uint64_t i = w0 + (uint64_t)s * WAVE + lane;
      if (i < b1) {
        const uint32_t d = (uint32_t)((key[s] >> shift) & 0xFF);
        const uint32_t p = s_start[d] + s_wcnt[wave][d] + rank[s];
        s_key[p] = key[s];
        s_val[p] = val[s];
      }
    }
    __syncthreads();
    // write each digit's run at the block's offset for that digit
    const uint32_t live_n = (uint32_t)(b1 - t0 < (uint64_t)RS_TILE ? b1 - t0 : RS_TILE);
    for (uint32_t p = tid; p < live_n; p += RS_THREADS) {
      const K k = s_key[p];
      const uint32_t d = (uint32_t)((k >> shift) & 0xFF);
      const uint64_t o = (uint64_t)s_run[d] + (p - s_start[d]);
      kout[o] = k;
      vout[o] = s_val[p];
    }
    __syncthreads();
    s_run[tid] += tot;
  }
}

static bool narrow_sort_type(int32_t t) {
  return t == MTL_INTEGER || t == MTL_DATE || t == MTL_DICT32 || t == MTL_DICT8;
}

template <typename K>
static int32_t sort_pairs(const void* col, int32_t type, const uint32_t* sel,
                          const uint32_t* count, uint64_t n, int32_t desc, uint32_t* out_rows,
                          mtl_stream* s) {
  constexpr int P = (int)sizeof(K);
  hipStream_t st = s->stream;
  Arena& ar = s->arena;
  void *ka, *kb, *r0, *r1, *hist, *dh;
  const uint64_t tiles = ceil_div(n, RS_TILE);
  const uint32_t nb = (uint32_t)std::min<uint64_t>(tiles, RS_MAX_BLOCKS);
  const uint64_t per = ceil_div(tiles, nb) * RS_TILE;
  MTL_TRY(ar.alloc(n * sizeof(K), &ka));
  MTL_TRY(ar.alloc(n * sizeof(K), &kb));
  MTL_TRY(ar.alloc(n * 4, &r0));
  MTL_TRY(ar.alloc(n * 4, &r1));
  MTL_TRY(ar.alloc((uint64_t)RS_BINS * nb * 4, &hist));
  MTL_TRY(ar.alloc((uint64_t)RS_BINS * P * 4, &dh));
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256 * 8), 4096));
  {
    ProfScope ps(s, "k_sort_keys");
    hipLaunchKernelGGL(k_sort_keys<K>, dim3(grid), dim3(256), 0, st, col, type, sel, count, n,
                       desc, (K*)ka, (uint32_t*)r0);
  }
  MTL_HIP(hipGetLastError());
  // which digit positions vary
  MTL_HIP(hipMemsetAsync(dh, 0, (size_t)P * RS_BINS * 4, st));
  {
    ProfScope ps(s, "k_rs_digits");
    const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256 * 16), 1024));
    hipLaunchKernelGGL(k_rs_digits<K>, dim3(g), dim3(RS_THREADS), 0, st, (const K*)ka, n,
                       (uint32_t*)dh);
  }
  MTL_HIP(hipGetLastError());
  std::vector<uint32_t> hh((size_t)P * RS_BINS);
  {
    void* host;
    MTL_TRY(stream_pinned(s, hh.size() * 4, &host));
    MTL_HIP(hipMemcpyAsync(host, dh, hh.size() * 4, hipMemcpyDeviceToHost, st));
    MTL_HIP(stream_wait(st));
    std::memcpy(hh.data(), host, hh.size() * 4);
  }
  std::vector<int> passes;
  for (int p = 0; p < P; ++p) {
    int nz = 0;
    for (int d = 0; d < RS_BINS; ++d) nz += hh[(size_t)p * RS_BINS + d] != 0;
    if (nz > 1) passes.push_back(p);
  }
  const int np = (int)passes.size();
  if (np == 0) {  // every key equal: the selection order is the result
    MTL_HIP(hipMemcpyAsync(out_rows, r0, n * 4, hipMemcpyDeviceToDevice, st));
    return MTL_OK;
  }
  // the row ids of pass j go to out_rows when (np - 1 - j) is even, so the
  // last pass writes out_rows
  const K* kin = (const K*)ka;
  const uint32_t* vin = (const uint32_t*)r0;
  for (int j = 0; j < np; ++j) {
    K* kout = (K*)(kin == (const K*)ka ? kb : ka);
    uint32_t* vout = ((np - 1 - j) % 2 == 0) ? out_rows : (uint32_t*)r1;
    const uint32_t shift = (uint32_t)(8 * passes[j]);
    {
      ProfScope ps(s, "k_rs_upsweep");
      hipLaunchKernelGGL(k_rs_upsweep<K>, dim3(nb), dim3(RS_THREADS), 0, st, kin, n, shift, per,
                         (uint32_t*)hist);
    }
    MTL_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rs_offsets, dim3(RS_BINS), dim3(1024), 0, st, (uint32_t*)hist, nb,
                       (const uint32_t*)dh + (size_t)passes[j] * RS_BINS);
    MTL_HIP(hipGetLastError());
    {
      ProfScope ps(s, "k_rs_scatter");
      hipLaunchKernelGGL(k_rs_scatter<K>, dim3(nb), dim3(RS_THREADS), 0, st, kin, vin, kout, vout,
                         n, shift, per, (const uint32_t*)hist);
    }
    MTL_HIP(hipGetLastError());
    kin = kout;
    vin = vout;
  }
  return MTL_OK;
}

}  // namespace mtl

extern "C" int32_t mtl_sort(const void* column, int32_t type, const uint32_t* sel,
                            const uint32_t* count, uint64_t capacity, int32_t desc,
                            uint32_t* out_rows, uint32_t* out_n, mtl_stream* s) {
  using namespace mtl;
  MTL_TRY(resolve_stream(s, &s));
  if (!column || !out_rows) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!type_width(type) || type == MTL_HUGEINT)
    return set_error(MTL_UNSUPPORTED, "ORDER BY on type %d", type);
  if (capacity > 0xFFFFFFFFull) return set_error(MTL_UNSUPPORTED, "more than 2^32-1 rows");
  std::lock_guard<std::mutex> lk(s->mu);
  s->arena.reset();
  uint64_t n = capacity;
  if (sel && count) {  // the row count lives on the device: read it once
    void* host;
    MTL_TRY(stream_pinned(s, 8, &host));
    MTL_HIP(hipMemcpyAsync(host, count, 4, hipMemcpyDeviceToHost, s->stream));
    MTL_HIP(stream_wait(s->stream));
    n = *(const uint32_t*)host;
    if (n > capacity) return set_error(MTL_ERR_INVALID_ARGUMENT, "count exceeds capacity");
  }
  if (out_n) *out_n = (uint32_t)n;
  if (n == 0) return MTL_OK;
  if (narrow_sort_type(type))
    return sort_pairs<uint32_t>(column, type, sel, nullptr, n, desc, out_rows, s);
  return sort_pairs<uint64_t>(column, type, sel, nullptr, n, desc, out_rows, s);
}
