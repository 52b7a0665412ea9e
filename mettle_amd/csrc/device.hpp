// Device-side helpers for gfx950 (CDNA4, wave64): atomics with explicit
// agent scope, wave reductions, a decoupled look-back tile scan (used only
// where the tile count is small), compiled predicates and key encodings.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mettle_hip.h"

namespace mtl {
namespace dev {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}

// Lane i of each 16-lane row receives lane i - OFF of the same row (DPP
// row_shr); the first OFF lanes of a row receive 0.  One VALU move, no LDS
// crossbar round trip.
template <int OFF>
__device__ __forceinline__ uint32_t row_shr(uint32_t v) {
  static_assert(OFF >= 1 && OFF <= 15, "row_shr takes 1..15");
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x110 + OFF, 0xF, 0xF, false);
}
// Lane i of each 16-lane row receives lane i + 1 (DPP row_shl:1); the last
// lane of a row receives 0.
__device__ __forceinline__ uint32_t row_shl1(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x101, 0xF, 0xF, false);
}

// Load at base + a 32-bit byte offset: with a wave-uniform base this is
// one SGPR base plus a 32-bit VGPR offset per lane, instead of a 64-bit
// VGPR address per lane (gathers keep many loads in flight at once).
template <typename T>
__device__ __forceinline__ T ld_off32(const T* base, uint32_t byte_off) {
  return *(const T*)((const char*)base + byte_off);
}

// Row count of a selection whose count lives on the device.  A count above
// the capacity the caller sized the buffers for (a failed scan reports
// 0xFFFFFFFF) is read as 0, so it never turns into reads of unwritten
// selection entries; the host reports the failure when it reads the count.
__device__ __forceinline__ uint64_t dev_count(const uint32_t* count, uint64_t cap) {
  if (!count) return cap;
  const uint64_t c = *count;
  return c <= cap ? c : 0;
}

// 64-bit finaliser of MurmurHash3 (fmix64): the mixing step of result
// fingerprints.
__host__ __device__ __forceinline__ uint64_t fp_mix(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}
// Fingerprint term of output position i holding (row, value): the sum of
// the terms over all positions (mod 2^64) depends on the order of the rows.
// 32-bit multiply-xorshift steps (a handful of VALU operations per output
// row, against about 45 for two 64-bit finalisers).
__host__ __device__ __forceinline__ uint64_t fp_term(uint64_t i, uint32_t row, uint64_t v) {
  uint32_t x = (uint32_t)i * 0x9E3779B1u + row;
  x ^= (uint32_t)v * 0x85EBCA77u;
  x ^= (uint32_t)(v >> 32) * 0xC2B2AE3Du;
  x ^= x >> 15;
  x *= 0x2C1B3C6Du;
  x ^= x >> 13;
  return x;
}

// Block-wide maxima (blockDim.x <= 1024, a multiple of 64); every thread
// of the block must call them and every thread gets the result.
__device__ __forceinline__ uint64_t block_max_u64(uint64_t v) {
  __shared__ uint64_t s_part[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t o = __shfl_xor(v, off, WAVE);
    v = o > v ? o : v;
  }
  __syncthreads();
  if (threadIdx.x % WAVE == 0) s_part[threadIdx.x / WAVE] = v;
  __syncthreads();
  uint64_t m = 0;
  for (uint32_t w = 0; w < blockDim.x / WAVE; ++w) m = s_part[w] > m ? s_part[w] : m;
  return m;
}
__device__ __forceinline__ uint32_t block_max_u32(uint32_t v) {
  return (uint32_t)block_max_u64(v);
}

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive_u32(uint32_t v) {
  int l = lane_id();
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    uint32_t o = __shfl_up(v, off, WAVE);
    if (l >= off) v += o;
  }
  return v;
}

// ---------------------------------------------------------------------------
// Decoupled look-back (single-pass prefix over tiles).  Status word per tile:
// low 32 bits = value, bits 32..33 = flag (0 not ready, 1 aggregate,
// 2 inclusive prefix).  Callers number tiles so that a tile only ever waits
// on tiles that are already running (an atomic ticket, or blockIdx where
// workgroups are dispatched in id order), so every wait ends.  A spin cap
// sets *err and stops waiting (defensive; it is never reached when the
// status array was cleared before the launch).
// ---------------------------------------------------------------------------
constexpr uint64_t kFlagAgg = 1ull << 32;
constexpr uint64_t kFlagInc = 2ull << 32;
constexpr uint32_t kSpinCap = 1u << 24;

// Executed by one full wave (all 64 lanes active).  Returns the exclusive
// prefix of `tile`, whose own total is `total`.
__device__ __forceinline__ uint32_t lookback_wave(uint64_t* status,
                                                  uint32_t tile,
                                                  uint32_t total,
                                                  uint32_t* err) {
  const int lane = lane_id();
  if (tile == 0) {
    if (lane == 0) st_agent(&status[0], kFlagInc | total);
    return 0;
  }
  if (lane == 0) st_agent(&status[tile], kFlagAgg | total);
  uint32_t excl = 0;
  int64_t top = (int64_t)tile - 1;
  uint32_t spins = 0;
  while (true) {
    int64_t idx = top - lane;
    uint64_t st = idx >= 0 ? ld_agent(&status[idx]) : kFlagInc;
    uint32_t flag = (uint32_t)(st >> 32);
    uint64_t inc = __ballot(flag == 2);
    uint64_t nr = __ballot(flag == 0);
    int first_inc = inc ? __ffsll((unsigned long long)inc) - 1 : 64;
    int first_nr = nr ? __ffsll((unsigned long long)nr) - 1 : 64;
    if (first_nr < first_inc) {
      // a closer predecessor has not published yet: wait for it
      if (++spins > kSpinCap) {
        if (lane == 0) atomicOr(err, 1u);
        __threadfence();  // the error bit is visible before this tile's status
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint32_t v = (uint32_t)st;
    uint32_t contrib = (lane <= first_inc) ? v : 0u;
    excl += wave_sum_u32(contrib);
    if (first_inc < 64) break;
    top -= 64;
  }
  if (lane == 0) st_agent(&status[tile], kFlagInc | (excl + total));
  return excl;
}

// ---------------------------------------------------------------------------
// Compiled predicates.  Column values and the literal are mapped to one
// order-preserving 64-bit encoding (signed integers: sign bit flipped;
// floats: DuckDB's total order with NaN largest and -0 == 0), so a single
// code path compares every type: pass = bit (value <, ==, > literal) of
// opmask.  Keeping one path instead of one per (type, op) keeps the scan
// kernels inside the instruction cache.
// ---------------------------------------------------------------------------
enum : uint32_t { PK_I32 = 0, PK_I64 = 1, PK_U8 = 2, PK_F32 = 3, PK_F64 = 4 };

struct PredC {
  const void* col;
  uint32_t kind;    // PK_*
  uint32_t opmask;  // bit 0: pass if value < literal, bit 1: ==, bit 2: >
  uint64_t lit;     // encoded literal
  // PK_I32: the same test as an int32 interval, pass iff
  // (x - lo32 <= hi32 - lo32, unsigned) != neg32 (an empty interval is the
  // full one negated)
  int32_t lo32, hi32;
  uint32_t neg32;
};

__device__ __forceinline__ bool pred_pass_i32(uint32_t x, const PredC& p) {
  return ((x - (uint32_t)p.lo32) <= ((uint32_t)p.hi32 - (uint32_t)p.lo32)) != (p.neg32 != 0);
}

__host__ __device__ __forceinline__ uint64_t enc_int(int64_t v) {
  return (uint64_t)v ^ 0x8000000000000000ull;
}
__host__ __device__ __forceinline__ uint64_t enc_dbl(double d) {
  if (d != d) return ~0ull;  // NaN: largest, all NaNs equal
  if (d == 0.0) d = 0.0;     // -0 == 0
  uint64_t b = __builtin_bit_cast(uint64_t, d);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ uint32_t pred_width(uint32_t kind) {
  return kind == PK_U8 ? 1u : ((kind == PK_I64 || kind == PK_F64) ? 8u : 4u);
}
// raw: the value's bytes zero-extended to 64 bits
__device__ __forceinline__ uint64_t enc_raw(uint64_t raw, uint32_t kind) {
  switch (kind) {
    case PK_I32: return enc_int((int64_t)(int32_t)(uint32_t)raw);
    case PK_I64: return enc_int((int64_t)raw);
    case PK_U8: return enc_int((int64_t)(raw & 0xFFu));
    case PK_F32: return enc_dbl((double)__builtin_bit_cast(float, (uint32_t)raw));
    default: return enc_dbl(__builtin_bit_cast(double, raw));
  }
}
__device__ __forceinline__ bool pred_pass(uint64_t x, const PredC& p) {
  uint32_t c = x < p.lit ? 0u : (x == p.lit ? 1u : 2u);
  return (p.opmask >> c) & 1u;
}
__device__ __forceinline__ uint64_t load_raw(const void* col, uint32_t width, uint64_t row) {
  switch (width) {
    case 1: return ((const uint8_t*)col)[row];
    case 4: return ((const uint32_t*)col)[row];
    default: return ((const uint64_t*)col)[row];
  }
}

// Read an integer-like key column as int64.
__device__ __forceinline__ int64_t load_key(const void* col, int32_t type,
                                            uint64_t row) {
  switch (type) {
    case MTL_BIGINT:
    case MTL_TIMESTAMP:
    case MTL_DECIMAL:
      return ((const int64_t*)col)[row];
    case MTL_DICT8:
      return (int64_t)((const uint8_t*)col)[row];
    default:
      return (int64_t)((const int32_t*)col)[row];
  }
}

// Offset of a key from a range minimum, in unsigned arithmetic (a signed
// subtraction overflows for full-range BIGINT keys).
__host__ __device__ __forceinline__ uint64_t key_off(int64_t key, int64_t min) {
  return (uint64_t)key - (uint64_t)min;
}

// Order-preserving unsigned encodings (smaller encoded = smaller value).
__device__ __forceinline__ uint64_t enc_i64(int64_t v) {
  return (uint64_t)v ^ 0x8000000000000000ull;
}
__device__ __forceinline__ uint64_t enc_f64(double d) {
  // DuckDB total order: NaN largest; -0.0 == 0.0
  if (d != d) return ~0ull;
  if (d == 0.0) d = 0.0;
  uint64_t b = (uint64_t)__double_as_longlong(d);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ uint64_t enc_value(const void* col, int32_t type,
                                              uint64_t row) {
  switch (type) {
    case MTL_FLOAT:
      return enc_f64((double)((const float*)col)[row]);
    case MTL_DOUBLE:
      return enc_f64(((const double*)col)[row]);
    default:
      return enc_i64(load_key(col, type, row));
  }
}

}  // namespace dev
}  // namespace mtl
